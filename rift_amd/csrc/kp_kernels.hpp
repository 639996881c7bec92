// kp_kernels.hpp -- device data layout and kernel launchers of the KP step.
//
// Layout in HBM (one handle = one domain):
//  - every cell field is FP32, row-major with pitch P floats and R rows,
//    P = roundup(16*nbx + 4, 64), R = 14*nby + 4, zero-filled, so that the
//    20 x 18 load window of any 16 x 14 tile (two-cell halo) is in bounds
//    without guards; cell (i, j) is at j*P + i;
//  - face beds BX (west face of cell i at column i) and BY (south face of row j
//    at row j) use the same pitch;
//  - per-tile arrays: wet / active flags (u8), the compacted active list
//    (int32), the tile speed maxima mx / my (FP32);
//  - the device scalars (t, dt in double, forcing rates, counters, FP64 mass
//    accumulators) live in one small struct, so a whole step runs without the
//    host and can be captured in a graph.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rift_kp.h"

namespace rift {

constexpr int kMaxSeries = 4096;   // points per scalar series

struct SeriesTable {
    int n;
    float time[kMaxSeries];
    float rate[kMaxSeries];
};

struct Scalars {
    double t, dt;
    // forcing of the current step (Source::interpolated_rate, source.h:10)
    float rain_rate, drain_rate, hydro_rate, src_cell_rate;
    int hydro_external;          // rate set by rift_set_source_value
    double hydro_value;          // external discharge, m^3/s
    int n_active;
    unsigned mx_bits, my_bits, ticket;
    int n_units;                     // active marching units of this step
    double dt_prev;                  // dt of the stage-1 sources (previous step's dt)
    unsigned long long wet_cells;    // interior cells of active tiles (this step)
    unsigned long long wcu_total;    // cumulative wet-cell updates since init
    // FP64 mass accounting (m^3 / (dx dy) until scaled on the host)
    double v_rain, v_src, v_drain, v_out;
    int pad;
};

struct Dev {
    float *h, *hu, *hv, *h0, *hu0, *hv0;
    float *dh, *dhu, *dhv, *G, *F, *F0, *dF, *K, *n, *BC, *BX, *BY;
    float *map[3][3];            // [rain|drain|surge][current|plane1|plane2]
    float *hmax, *qmax, *twet, *tpeak, *tdry;
    unsigned char *wet, *act;
    int* list;
    // fused step: per-tile buffer select (state of the tile's cells lives in
    // h/hu/hv when 0, in h0/hu0/hv0 when 1), pending flips, the stage-2
    // contributor word (count << 8 | wet contributors), the active units
    unsigned char *sel, *flip;
    unsigned* pending;
    int* ulist;
    int nstrips, nurows, nunits;
    long long soff;              // h0 - h (elements): offset of the second buffer
    int n_uniform;
    float n_const;
    float *mx, *my;              // per-tile maxima of the last flux pass
    float *mx1, *my1;            // fused step: stage-1 maxima of this step
    Scalars* sc;
    SeriesTable* ser;            // [3]: hyetograph, drain, hydrograph
    int P, nx, ny, nbx, nby, ntiles;
    float inv_dx, inv_dy, kappa, psi_dtheta, nodata;
    double dx, dy, courant, kappa_d;
    unsigned flags;
    int si0, si1, sj0, sj1;      // hydrograph source rectangle (inclusive)
    long nsrc;
};

__host__ __device__ inline size_t cell(const Dev& d, int i, int j)
{
    return (size_t)j * (size_t)d.P + (size_t)i;
}

// ComputeTimestep's arithmetic (simulator.cu:549-561) on one lane; also the
// FP64 accounting of the scalar inflows of the step.
__device__ inline void finish_timestep(const Dev& d, float vx, float vy)
{
    Scalars* sc = d.sc;
    const double max_x = (double)vx, max_y = (double)vy;
    const double c_x = d.dx / (double)fmaxf((float)(max_x / d.courant), (float)d.kappa_d);
    const double c_y = d.dy / (double)fmaxf((float)(max_y / d.courant), (float)d.kappa_d);
    double dt = (c_x < c_y) ? c_x : c_y;
    if (sc->t == 0.0) dt = (double)0.000001f;
    sc->dt_prev = sc->dt;
    sc->dt = dt;
    sc->t += dt;
    const double cells = (double)sc->wet_cells;
    if (d.flags & RIFT_RAIN_AVG) sc->v_rain += (double)sc->rain_rate * dt * cells;
    if (d.flags & RIFT_HYDROGRAPH) sc->v_src += (double)sc->src_cell_rate * dt * (double)d.nsrc;
}

// Launchers (kp_kernels.hip).  All run on `s` and read t/dt/counts from d.sc.
void launch_begin_step(const Dev& d, hipStream_t s);
void launch_track(const Dev& d, hipStream_t s);
void launch_grow(const Dev& d, hipStream_t s);
void launch_reset_counters(const Dev& d, hipStream_t s);
void launch_flux(const Dev& d, int stage, hipStream_t s, int grid);
void launch_timestep(const Dev& d, hipStream_t s);
void launch_integrate1(const Dev& d, hipStream_t s, int grid);
void launch_integrate2(const Dev& d, hipStream_t s, int grid);
void launch_boundaries(const Dev& d, hipStream_t s);
void launch_interp_map(const Dev& d, int which, float factor, hipStream_t s);
void launch_sum_interior(const Dev& d, const float* x, int mode, double* out, hipStream_t s);

// Fused step (kp_step2.hip).  One wavefront marches one unit: 60 output
// columns (lanes 2..61) by one 14-row tile row.
constexpr int kStrip = 60;
void launch_grow2(const Dev& d, hipStream_t s);
void launch_timestep2(const Dev& d, hipStream_t s);
void launch_march(const Dev& d, int stage, hipStream_t s);
void launch_canonicalize(const Dev& d, hipStream_t s);

}  // namespace rift
