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

/* GridSeries::p_interp (GridSeries.cu:91-122): the internal-grid map
 * (cells, ny x nx doubles, nx = b_nx+3, ny = b_ny+3) of a forcing map given
 * at the DEM corners (b_ny x b_nx, row 0 = south): the four-corner mean
 * where more than two corners differ from nodata, times scale; every other
 * cell holds nodata when allow_nodata, else 0. */
/* Corner elevations behind the face beds (no reference counterpart: the
 * reference keeps only the face means, simulator.cu:384-414).  BX/BY as
 * rift_geometry writes them (nx = b_nx+3, ny = b_ny+3) are two-point means
 * of the corners D (b_ny x b_nx, row 0 = south); the means fix D up to a
 * checkerboard (-1)^(i+j) c, which changes no mean, and the D returned has
 * the checkerboard's mean taken out.  Returns RIFT_E_ARG (D untouched) when
 * the faces are not such means within 1e-9 relative or their ghost faces do
 * not repeat the faces two cells in.  The fused step forms every bed from
 * these corners. */
int rift_corners_from_faces(int nx, int ny, const double* BX, const double* BY, double* D);

int rift_map_cells(int b_nx, int b_ny, const double* corners, double scale, int allow_nodata,
                   double nodata, double* cells);

/* ---- File formats ------------------------------------------------------
 * CSV series, Source::ReadSource (source.cc:11-63): `hours,value` lines, '#'
 * comment lines, times converted to seconds in float, a zero-rate point
 * appended when the last rate is nonzero.  Writes min(n, cap) points and
 * returns the count in *n (call with cap 0 to size the buffers). */
int rift_read_series(const char* path, float* time_s, float* rate, int cap, int* n);
/* ESRI-ASCII raster, ReadGrid / SetOriginalGrid (io.cc:260-329): header
 * (ncols, nrows, xllcorner, yllcorner, cellsize, NODATA_value) then rows north
 * first.  data receives nrows x ncols doubles with row 0 = south (pass NULL to
 * read the header only); hdr4 = {xll, yll, cellsize, nodata}.  NODATA values
 * are an error unless allow_nodata, then they read as -9999. */
int rift_read_raster(const char* path, int allow_nodata, int* ncols, int* nrows, double* hdr4,
                     double* data, long cap);
/* writeGrid (io.cc:331-361): the interior of an internal-grid field (nx x ny,
 * row 0 = south; interior = 2 ghost cells off each side), header from the
 * internal lower-left corner and the raster cell size, 15 significant
 * digits, rows north first. */
int rift_write_grid(const char* path, const double* field, int nx, int ny, double h_xll,
                    double h_yll, double cellsize);

/* ---- Keyword configuration (ConfigFile, config.h:48-209) ---------------
 * `key = value` lines, `;` comments, unique keys.  Typed reads use stream
 * extraction as the reference's Convert::string_to_T does (config.h:34-44):
 * a string is the first word, a bool is 0/1.  Errors (missing file, bad
 * line, duplicate key, bad value, missing key) return RIFT_E_ARG with the
 * message in rift_last_error() instead of waiting on stdin and exiting. */
typedef struct rift_config rift_config;
int rift_config_open(const char* path, rift_config** out);
void rift_config_close(rift_config* c);
/* keyExists (config.h:192-195): 1 or 0. */
int rift_config_has(const rift_config* c, const char* key);
/* keyIsNumber (config.h:176-190): 1 when the first word holds only digits
 * and '.', else 0; RIFT_E_ARG for a missing key. */
int rift_config_is_number(const rift_config* c, const char* key, int* out);
/* getValueOfKey<std::string>: copies at most cap-1 bytes plus NUL; *len
 * (optional) receives the full length. */
int rift_config_string(const rift_config* c, const char* key, char* buf, int cap, int* len);
int rift_config_double(const rift_config* c, const char* key, double* out);
int rift_config_bool(const rift_config* c, const char* key, int* out);
/* getValueOfKey<std::complex<double>>: "(re,im)". */
int rift_config_complex(const rift_config* c, const char* key, double* re, double* im);

/* ---- Simulator driver (simulator.h:14-173, main.cc:40-67) -------------- */
typedef struct rift_sim rift_sim;

/* Simulator::OpenSimulation (simulator.cu:720-734): reads the keyword file
 * (same keywords and defaults as ReadUserParams, simulator.cu:25-265), the
 * rasters and series it names, builds the grid on the device, applies a
 * read_hotstart file when named.  flow_rate is kept as the reference keeps it
 * (stored, not used). */
int rift_sim_open(const char* config_path, double flow_rate, rift_sim** out);
/* Simulator::RunSimulation (simulator.cu:774-850): one time step plus the
 * gridded series updates and the periodic output; *t receives the new time. */
int rift_sim_step(rift_sim* s, double* t);
/* The loop of main.cc:57-63: steps while t < t_end (t_end < 0: tf) and the
 * run is not at steady state. */
int rift_sim_run(rift_sim* s, double t_end);
int rift_sim_time(rift_sim* s, double* t, double* dt, double* tf);
/* The device grid of the run, for field access through rift_kp.h. */
rift_grid* rift_sim_grid(rift_sim* s);
/* Simulator::setSourceValue (simulator.cu:891-901): total source discharge,
 * m^3/s, spread over the source cells. */
int rift_sim_set_source_value(rift_sim* s, double discharge);
/* Simulator::WriteHotstart / ReadHotstart (simulator.cu:912-1090), same
 * binary layout. */
int rift_sim_write_hotstart(rift_sim* s, const char* path);
int rift_sim_read_hotstart(rift_sim* s, const char* path);
/* Simulator::CloseSimulation (simulator.cu:735-772): summary maps, then the
 * handle is freed. */
int rift_sim_close(rift_sim* s);

#ifdef __cplusplus
}
#endif
#endif
