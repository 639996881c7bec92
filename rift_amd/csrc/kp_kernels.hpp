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
    // fused step: units compacted by grow2 (read by stage 1; moved to
    // n_units by the dt kernel, which zeroes it for the next step), and the
    // forcing of the next step, evaluated by the dt kernel at the new time
    // and installed by the next grow2 when next_valid is set
    int n_units_acc;
    int next_valid;
    // volume budget of the fused step (m summed over cells): gridded rain,
    // drain, outflow through the domain boundary, dry-clamp and surge
    // exchange; b_flux1 = stage-1 boundary outflow rate (m/s summed over
    // cells), weighted with dt/2 by the dt kernel
    double b_rain_grid, b_drain, b_out, b_exch, b_flux1;
    int finished;                     // the domain reached its end time (ensembles)
    float nxt_rain, nxt_drain, nxt_hydro, nxt_cell;
    // device surge series (InterpolatedGridSeries, GridSeries.cu:291-384):
    // window state advanced by the dt kernel (p_in_time, p_next_time, p_done
    // and the map indices of its two planes), the window it evaluated for
    // the next step (valid flag, indices, factor), the window of the step
    // being run (installed by grow2), and the first map index a window
    // needed that was not loaded into its slot (-1: none)
    double sg_in, sg_next;
    int sg_done, sg_nk0, sg_nk1, sg_next_valid;
    float sg_nf;
    int sg_k0, sg_k1;
    float sg_f;
    int sg_missing;
    // the per-unit surge flags must be recomputed by the coming step
    // (window moved, or a map was loaded since the last dt kernel)
    int sg_udirty, sg_udirty_load;
    // single-domain fused step: full quads (all four strips of a quad
    // active in a tile row) and the active units outside them, compacted by
    // grow2 (read by stage 1), moved by the dt kernel for stage 2
    int n_quads, n_quads_acc, n_part, n_part_acc;
    // fused steps taken (the dt kernel counts); a single domain's grow tags
    // the units it activates with ustep + 1, so tags never need clearing
    int ustep;
};

constexpr int kDtMaxBlocks = 512;   // workgroups of the dt kernel
// Slots of the dt reduction formed in the stage-1 march (128 bytes apart),
// which grow and the march add into.
constexpr int kDtSlots = 256;

// Surge window of one step: map indices of the two planes and the factor.
struct SurgeWindow {
    int k0, k1;
    float f;
};

struct Dev {
    float *h, *hu, *hv, *h0, *hu0, *hv0;
    float *dh, *dhu, *dhv, *G, *F, *F0, *dF, *K, *n, *BC, *BX, *BY;
    // corner elevations behind the face beds (the DEM points; the fused
    // step forms BC, BX and BY from them, see k_beds_from_corners)
    float* DC;
    float *map[3][3];            // [rain|drain|surge][current|plane1|plane2]
    float *hmax, *qmax, *twet, *tpeak, *tdry;
    unsigned char *wet, *act;
    int* list;
    // fused step: per-tile buffer select (state of the tile's cells lives in
    // h/hu/hv when 0, in h0/hu0/hv0 when 1), pending flips, the active units
    unsigned char *sel, *flip;
    int* ulist;
    unsigned char* umark;        // per unit: covers an active tile (set by grow, cleared by compaction)
    int nstrips, nurows, nunits;
    // single domain: qlist = full quad units (q * nurows + r: strips 4q ..
    // 4q+3 all active in tile row r); ulist then holds the other active units
    int* qlist;
    int nquads;
    int* utag;                   // single domain: step tag of each unit's last activation
    long long soff;              // h0 - h (elements): offset of the second buffer
    int n_uniform;
    float n_const;
    double tf;                   // end time of an ensemble member (0: none)
    int band;                    // XCD-banded order of the marching shares
    int min_rows;                // least rows per marching share
    int quads;                   // single domain: full quads marched by one workgroup
    float *mx, *my;              // per-tile maxima of the last flux pass
    float *mx1, *my1;            // fused step: stage-1 maxima of this step
    Scalars* sc;
    unsigned* dtk;               // dt reductions: 8 per-XCD ticket counters, 128 bytes apart
    // single domain, dt formed by the stage-1 march: kDtSlots slots of 32
    // words (vx, vy bits, active tiles, -, cells as 64 bits)
    unsigned* gs;
    int dt_in_march;             // launch argument: the stage-1 march forms dt
    SeriesTable* ser;            // [3]: hyetograph, drain, hydrograph
    int P, nx, ny, nbx, nby, ntiles;
    float inv_dx, inv_dy, kappa, psi_dtheta, nodata;
    double dx, dy, courant, kappa_d;
    unsigned flags;
    int si0, si1, sj0, sj1;      // hydrograph source rectangle (inclusive)
    long nsrc;
    // device surge series: map k lives in slot k % sg_nslots, a plane of
    // the field layout (slot s at sgs + s * soff); sgflag[s * ntiles + t]
    // bit 0 = a cell of tile t's Track window (rows by*16+2 .. by*16+15,
    // grid.cu:1265) holds a value > 0 (not nodata), bit 1 = a cell of the
    // tile itself does; sgtag[s] = map index held by slot s (-1: none)
    float* sgs;
    unsigned char* sgflag;
    int* sgtag;
    unsigned char* sgu;          // per marching unit: surge > 0 possible this step
    int sg_series;               // 1: surge from the device series
    int sg_nslots;
    double sg_dt, sg_tmax;       // map interval (integer seconds) and surge_tf
};

// Slot of map k.
__host__ __device__ inline int sg_slot(const Dev& d, int k)
{
    return ((k % d.sg_nslots) + d.sg_nslots) % d.sg_nslots;
}

// The surge window in force now: the one the last dt kernel evaluated for
// the coming step, or (once grow2 installed it) the one of the running step.
__device__ inline SurgeWindow surge_window_now(const Scalars* sc)
{
    SurgeWindow w;
    if (sc->sg_next_valid) {
        w.k0 = sc->sg_nk0; w.k1 = sc->sg_nk1; w.f = sc->sg_nf;
    } else {
        w.k0 = sc->sg_k0; w.k1 = sc->sg_k1; w.f = sc->sg_f;
    }
    return w;
}

// InterpGrid_k's nodata rule and interpolant (grid.cu:292-311) in FP32, as
// k_interp_map evaluates it (no contraction).
__device__ __forceinline__ float surge_interp(float x0, float x1, float f, float nd)
{
    if (x1 == nd && x0 == nd) return nd;
    if (x1 == nd) return x0;
    if (x0 == nd) return x1;
    return f * (x1 - x0) + x0;
}

// InterpolatedGridSeries::p_update (GridSeries.cu:325-384) at the new time
// sc->t, after the first update: the window moves to the next pair of maps
// once t reaches p_next_time, freezes on the last map once t reaches
// surge_tf, and the factor (t - p_in_time)/(p_next_time - p_in_time) is
// formed in double as there.  The result is the next step's window.
__device__ inline void advance_surge_window(const Dev& d, Scalars* sc)
{
    const double t = sc->t;
    const int k0_old = sc->sg_nk0, k1_old = sc->sg_nk1;
    if (t >= d.sg_tmax) {
        if (!sc->sg_done) {
            sc->sg_in = sc->sg_next;
            sc->sg_next = sc->sg_in + d.sg_dt;
            sc->sg_nk0 = sc->sg_nk1;   // cudaMemcpy2D t1 -> t0
        }
        sc->sg_done = 1;
    } else if (t >= sc->sg_next) {
        sc->sg_nk0 = sc->sg_nk1;
        sc->sg_in = sc->sg_next;
        sc->sg_next = sc->sg_in + d.sg_dt;
        sc->sg_nk1 = (int)trunc(sc->sg_next / d.sg_dt);
    }
    sc->sg_udirty = (sc->sg_nk0 != k0_old || sc->sg_nk1 != k1_old || sc->sg_udirty_load) ? 1 : 0;
    sc->sg_udirty_load = 0;
    const double factor = (t - sc->sg_in) / (sc->sg_next - sc->sg_in);
    sc->sg_nf = (float)factor;
    sc->sg_next_valid = 1;
}

// Records the first map index of the installed window whose slot does not
// hold it (the caller loads maps between steps, so this is checked when a
// window comes into force, not when the dt kernel computes it).
__device__ inline void check_surge_slots(const Dev& d, Scalars* sc)
{
    if (sc->sg_missing >= 0) return;
    if (d.sgtag[sg_slot(d, sc->sg_k0)] != sc->sg_k0) sc->sg_missing = sc->sg_k0;
    else if (d.sgtag[sg_slot(d, sc->sg_k1)] != sc->sg_k1) sc->sg_missing = sc->sg_k1;
}

__host__ __device__ inline size_t cell(const Dev& d, int i, int j)
{
    return (size_t)j * (size_t)d.P + (size_t)i;
}

// Source::InterpolateRate (source.cc:65-73) in float: inside [t_i, t_i+1) the
// linear interpolant; outside the table the previous rate is kept (and the
// caller divides it by the unit again, as the reference does).
// The interval Source::InterpolateRate ends on: the last i with
// time[i] <= tf < time[i+1] (each match overwrites the rate, so the last one
// decides), or -1; and the rate of that interval (the previous rate if none).
__device__ inline int series_last_match(const SeriesTable& s, float tf)
{
    int m = -1;
    for (int i = 0; i < s.n - 1; ++i)
        if (tf >= s.time[i] && tf < s.time[i + 1]) m = i;
    return m;
}
__device__ inline float series_rate_at(const SeriesTable& s, int i, float tf, float rate)
{
    if (i < 0) return rate;
    float num = __fsub_rn(s.rate[i + 1], s.rate[i]);
    float den = __fsub_rn(s.time[i + 1], s.time[i]);
    float q = __fdiv_rn(num, den);
    float dd = __fsub_rn(tf, s.time[i]);
    return __fadd_rn(s.rate[i], __fmul_rn(q, dd));
}

struct Forcing {
    float rain, drain, hydro, cell;
};

// UpdateSource + setSourceValue (simulator.cu:500-538, 891-901) at the time
// held in sc->t, from the rates of the previous step.  match[k] (optional):
// series_last_match of series k at (float)t, found beforehand.
__device__ inline Forcing forcing_at(const Dev& d, const Scalars* sc, const int* match = nullptr)
{
    Forcing r{sc->rain_rate, sc->drain_rate, sc->hydro_rate, sc->src_cell_rate};
    const unsigned f = d.flags;
    const float tf = (float)sc->t;
    auto interp = [&](int k, float rate) {
        const SeriesTable& s = d.ser[k];
        return series_rate_at(s, match ? match[k] : series_last_match(s, tf), tf, rate);
    };
    if (f & RIFT_HYDROGRAPH) {
        double value;
        if (sc->hydro_external) {
            value = sc->hydro_value;
        } else {
            float q = interp(RIFT_SER_HYDROGRAPH, r.hydro);
            q = __fdiv_rn(q, 35.3147f);               // cfs -> m^3/s
            r.hydro = q;
            value = (double)q;
        }
        const double per_cell = value / d.dx / d.dy / (double)(d.nsrc > 0 ? d.nsrc : 1);
        r.cell = (float)per_cell;
    }
    if (f & RIFT_RAIN_AVG) {
        const float q = interp(RIFT_SER_HYETOGRAPH, r.rain);
        r.rain = __fdiv_rn(q, 3600.f * 1000.f);    // mm/hr -> m/s
    }
    if (f & RIFT_DRAIN_AVG) {
        const float q = interp(RIFT_SER_DRAIN, r.drain);
        r.drain = __fdiv_rn(q, 3600.f * 1000.f);
    }
    return r;
}

// ComputeTimestep's arithmetic (simulator.cu:549-561) on one lane; also the
// FP64 accounting of the scalar inflows of the step.
__device__ inline void finish_timestep(const Dev& d, float vx, float vy, Scalars* sc)
{
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
    // surge_series->update(t) at the end of RunSimulation (simulator.cu:827-829)
    if (d.sg_series) advance_surge_window(d, sc);
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
// device surge series: tile flags of one slot, Track of the series, and the
// window in force written into map plane 0 (per-phase path, downloads)
void launch_sg_flags(const Dev& d, int slot, hipStream_t s);
void launch_track_series(const Dev& d, hipStream_t s);
void launch_sg_materialize(const Dev& d, hipStream_t s);
void launch_sum_interior(const Dev& d, const float* x, int mode, double* out, hipStream_t s);
// BC, BX and BY planes from the corner plane DC (R rows)
void launch_beds_from_corners(const Dev& d, int R, hipStream_t s);

// Fused step (kp_step2.hip).  One wavefront marches one unit: 60 output
// columns (lanes 2..61) by one 14-row tile row.
constexpr int kStrip = 60;
// dt_in_march: grow adds its part of the dt reduction into Dev::gs for a
// stage-1 march launched with dt_in_march
void launch_grow2(const Dev& d, hipStream_t s, bool dt_in_march = false);
void launch_timestep2(const Dev& d, hipStream_t s);
// With dt_in_march (stage 1, single domain, after launch_grow2 with
// dt_in_march), the last workgroup of the stage-1 march forms dt and no dt
// kernel runs.
void launch_march(const Dev& d, int stage, hipStream_t s, int variant, int blocks, bool dt_in_march = false);
// workgroups of a marching kernel resident at once on a device of `cus` CUs
// (for the instantiation d's flags and surge series select)
int march_resident_blocks(const Dev& d, int stage, int variant, int cus);
constexpr int kMarchVariants = 4;
// Ensembles (devs: device array of the members' Dev)
void launch_grow2_ens(const Dev* devs, int nmem, int max_tiles, int max_units, hipStream_t s);
void launch_timestep2_ens(const Dev* devs, int nmem, int max_tiles, hipStream_t s);
// lean: every member runs the lean instantiation (march_lean of its flags)
void launch_march_ens(const Dev* devs, int nmem, unsigned flags, bool lean, int stage, int min_rows,
                      int blocks, hipStream_t s);
int march_ens_resident_blocks(int stage, bool textbook, bool budget, bool lean, int cus);
// the lean marching kernels serve these flags
bool march_lean(unsigned flags);
void launch_canonicalize(const Dev& d, hipStream_t s);

}  // namespace rift
