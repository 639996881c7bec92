/*
 * rift_host.h -- C ABI of the host-side pieces around the step: geometry of
 * the internal grid built from the DEM and corner rasters, and the
 * Simulator driver (keyword configuration, ESRI-ASCII rasters, CSV series,
 * gridded forcing series, hotstart).  These replace host C++ of the
 * reference that callers reach through Simulator (simulator.h:14-173).
 */
#ifndef RIFT_HOST_H
#define RIFT_HOST_H

#include "rift_kp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Desingularisation constant of InitSimulation (simulator.cu:281):
 * sqrtf(0.01f * fmaxf(1.f, fminf(dx, dy))), evaluated in float. */
double rift_kappa(double dx, double dy);

/* Internal-grid fields from the raw DEM and corner rasters, as InitSimulation
 * builds them (simulator.cu:376-431): face beds from two-point means of the
 * DEM with edge extension, h0 / n / K from four-corner means (K converted from
 * mm/hr to m/s).  dem, h0c, nc, Kc are b_ny x b_nx (row 0 = south); h0c, nc,
 * Kc may be NULL (n then takes n_const).  Outputs: BX (ny x (nx+1)),
 * BY ((ny+1) x nx), h0, n, K (ny x nx) with nx = b_nx+3, ny = b_ny+3; h0 and K
 * may be NULL. */
int rift_geometry(int b_nx, int b_ny, const double* dem, const double* h0c, const double* nc,
                  double n_const, const double* Kc, double* BX, double* BY, double* h0,
                  double* n, double* K);

#ifdef __cplusplus
}
#endif
#endif
