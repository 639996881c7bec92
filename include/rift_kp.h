/*
 * rift_kp.h -- C ABI of the MI355X (gfx950) Kurganov-Petrova shallow-water
 * step.  This is the drop-in boundary for RIFT's Grid host API (grid.h:6-69):
 * each entry point names the reference function it replaces.
 *
 * Conventions
 *  - One rift_grid is one simulation domain on one GPU.  All device memory is
 *    owned by the handle; host buffers are owned by the caller.
 *  - Every call returns RIFT_OK (0) or a negative RIFT_E* code; the message of
 *    the last failure on the calling thread is rift_last_error().
 *  - Work is ordered on the handle's HIP stream.  Only calls that hand data
 *    back to the host (rift_download, rift_get_*, rift_sync) block.
 *  - Host arrays use the reference's orientation and extents: row j = y,
 *    row 0 = south, internal extent nx x ny (= DEM points + 3 per side,
 *    grid_config.cpp:84-85), face beds BX (nx+1) x ny and BY nx x (ny+1)
 *    (grid.cu:402-403), all row-major doubles as in the reference.
 *  - State on the device is FP32 depth h = w - BC, hu, hv; time and dt are
 *    double scalars held on the device.
 */
#ifndef RIFT_KP_H
#define RIFT_KP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RIFT_OK 0
#define RIFT_E_ARG (-1)      /* invalid argument                         */
#define RIFT_E_HIP (-2)      /* HIP runtime failure                      */
#define RIFT_E_STATE (-3)    /* call out of order (e.g. before init)     */
#define RIFT_E_UNSUPPORTED (-4)

/* Feature flags: one bit per reference __constant__ flag (grid.cu:21-25)
 * plus two extensions that the reference does not have. */
#define RIFT_HYDROGRAPH (1u << 0)    /* dambreak                          */
#define RIFT_RAIN_AVG (1u << 1)      /* rainfall_averaged                 */
#define RIFT_RAIN_GRID (1u << 2)     /* rainfall_gridded                  */
#define RIFT_DRAIN_AVG (1u << 3)     /* drain_averaged                    */
#define RIFT_DRAIN_GRID (1u << 4)    /* drain_gridded                     */
#define RIFT_INFILTRATION (1u << 5)  /* infiltration (Green-Ampt)         */
#define RIFT_SURGE_GRID (1u << 6)    /* surge_gridded                     */
#define RIFT_EULER (1u << 7)         /* euler_integration                 */
#define RIFT_SAVE_MAX (1u << 8)      /* save_max                          */
#define RIFT_SAVE_ARRIVAL (1u << 9)  /* save_arrival_time                 */
#define RIFT_TILE_ACCEL (1u << 10)   /* tile_acceleration                 */
#define RIFT_H_INIT (1u << 11)       /* h_init                            */
#define RIFT_RAIN_INTERP (1u << 12)  /* hyetograph_interp (host series)   */
#define RIFT_MASS_BUDGET (1u << 13)  /* check_volume: FP64 volume budget   */
#define RIFT_TEXTBOOK (1u << 16)     /* extension: sign-aware minmod      */

/* Fields for rift_download / rift_upload (nx*ny doubles). */
enum rift_field {
    RIFT_W = 0,      /* water surface elevation w = h + BC               */
    RIFT_HU,
    RIFT_HV,
    RIFT_DEPTH,      /* h = w - BC (dev_h of the reference)              */
    RIFT_BC,
    RIFT_N,          /* Manning n                                        */
    RIFT_K,          /* Green-Ampt K, m/s                                */
    RIFT_F,          /* cumulative infiltration                          */
    RIFT_DW,         /* last flux divergence + sources, d(h)/dt          */
    RIFT_DHU,
    RIFT_DHV,
    RIFT_G,          /* Manning friction coefficient of the last flux    */
    RIFT_DF,
    RIFT_HMAX,
    RIFT_QMAX,
    RIFT_TWET,       /* arrival time (-9999 = never wet)                 */
    RIFT_TPEAK,
    RIFT_TDRY,
    RIFT_SURGE,      /* current surge depth map                          */
    RIFT_RAINMAP,    /* current gridded rain rate, m/s                   */
    RIFT_DRAINMAP,
    RIFT_Q,          /* unit discharge h*|u| (q_print)                   */
    RIFT_NFIELDS
};

/* Scalar series evaluated on the device each step (source.cc:65-73). */
enum rift_series { RIFT_SER_HYETOGRAPH = 0, RIFT_SER_DRAIN = 1, RIFT_SER_HYDROGRAPH = 2 };
/* Gridded forcing maps; plane 0 = current, 1 and 2 = interpolation planes. */
enum rift_map { RIFT_MAP_RAIN = 0, RIFT_MAP_DRAIN = 1, RIFT_MAP_SURGE = 2 };

typedef struct rift_desc {
    int nx, ny;          /* internal extent incl. ghosts (h_nx, h_ny)        */
    double dx, dy;       /* cell size, m                                     */
    double kappa;        /* desingularisation constant                       */
    double courant;      /* courant_limit                                    */
    double psi, dtheta;  /* Green-Ampt suction head (m), moisture deficit    */
    double nodata;       /* GridConfig::nodata                               */
    unsigned flags;      /* RIFT_* flags                                     */
    int device;          /* HIP device ordinal                               */
} rift_desc;

typedef struct rift_grid rift_grid;

/* Message of the last failed call on this thread ("" if none). */
const char* rift_last_error(void);

/* Replaces SetDeviceConstants (grid.cu:321-345) and AllocateGrid
 * (grid.cu:347-490): geometry, flags and constants go into the handle; every
 * device buffer is allocated and zeroed. */
int rift_create(const rift_desc* desc, rift_grid** out);

/* Replaces FreeGrid (grid.cu:1413-1451). */
int rift_destroy(rift_grid* g);

/* Static fields, replacing the cudaMemcpy2D uploads of simulator.cu:433-450.
 * BX (nx+1)*ny and BY nx*(ny+1) are the face beds; n, K (m/s) and h0 are
 * nx*ny interior fields (K and h0 may be NULL when unused). */
int rift_upload_static(rift_grid* g, const double* BX, const double* BY, const double* n,
                       const double* K, const double* h0);

/* Replaces InitGrid (grid.cu:556-561) followed by the ApplyBoundaries of
 * OpenSimulation (simulator.cu:731): BC from the face beds, w = BC + h0,
 * every tile wet and active, t = t0, dt = 1e-7f (simulator.cu:35). */
int rift_init(rift_grid* g, double t0);

/* Scalar forcing series (hyetograph mm/hr, drain mm/hr, hydrograph cfs) with
 * times in seconds as the reference's Source holds them (source.h:8-11).  The
 * device evaluates Source::InterpolateRate at (float)t and the unit division
 * of UpdateSource (simulator.cu:500-538) at the start of every step. */
int rift_set_series(rift_grid* g, int which, const float* time_s, const float* value, int n);

/* Hydrograph source cells, replacing InitializeSources' index list
 * (simulator.cu:858-889): idx = j*nx + i.  The cells must form a rectangle
 * inside the domain. */
int rift_set_source_cells(rift_grid* g, const int* idx, long n);

/* External 1-D/2-D coupling: replaces setSourceValue + updateSources
 * (simulator.cu:891-907).  `value` is the discharge in m^3/s; it is used
 * instead of the hydrograph series until rift_set_series(HYDROGRAPH) is called
 * again. */
int rift_set_source_value(rift_grid* g, double value);

/* Gridded forcing map (nx*ny doubles, internal grid) into plane 0..2 of `which`;
 * replaces the cudaMemcpy2D calls of GridSeries (GridSeries.cu:218-223, 331-377). */
int rift_set_map(rift_grid* g, int which, int plane, const double* host);
/* InterpGrid (grid.cu:313-318): plane0 = factor*(plane2-plane1)+plane1 with
 * the nodata rule. */
int rift_interp_map(rift_grid* g, int which, double factor);
/* Copy plane `src` into plane `dst` of a map (cudaMemcpy2D DtoD of
 * GridSeries.cu:345-378). */
int rift_copy_map(rift_grid* g, int which, int dst, int src);

/* Device-resident interpolated surge series: replaces InterpolatedGridSeries
 * (GridSeries.cu:291-384) -- its two device planes, the per-step
 * InterpGrid launch (grid.cu:313-318) and the host's p_update calls after
 * every step (simulator.cu:827-829).  Map k (the spatially interpolated map
 * of file index k, nx*ny doubles as p_interp leaves it, GridSeries.cu:91-
 * 122) is held in slot k % nslots of device memory.  The fused step's dt
 * kernel advances p_update's window at each new time on the device and the
 * marching kernels and TrackSurge read the two planes of the window,
 * interpolating per cell, so a step needs no host call and no full-grid
 * interpolation pass.  nslots >= number of maps keeps the whole series
 * resident; with fewer slots the caller loads map k before the window
 * needs it (two slots suffice when maps are loaded as p_update reads them).
 * Only which = RIFT_MAP_SURGE (handle created with RIFT_SURGE_GRID).
 * Call order: rift_init (or rift_set_time), then rift_map_series_start. */
int rift_map_series_create(rift_grid* g, int which, int nslots, int deltat, double tmax);
int rift_map_series_load(rift_grid* g, int which, int index, const double* host);
/* The first p_update at time t (GridSeries.cu:333-352): window maps
 * trunc(t/deltat) and the next one. */
int rift_map_series_start(rift_grid* g, int which, double t);
/* The window in force (map indices, factor) and the first map index a
 * window needed that was not loaded in its slot (-1: none); blocking. */
int rift_map_series_state(rift_grid* g, int which, int* k0, int* k1, double* factor, int* missing);

/* One or more complete time steps: the body of Simulator::RunSimulation
 * (simulator.cu:780-816) -- forcing update, Grow, two flux/integrate stages
 * (one in Euler mode), boundaries -- entirely on the device. */
int rift_step(rift_grid* g, int nsteps);

/* The phases of a step, one call each, for stage-level parity checks. */
int rift_update_source(rift_grid* g);    /* UpdateSource, simulator.cu:500  */
int rift_grow(rift_grid* g);             /* Grow, grid.cu:1279              */
int rift_compute_fluxes(rift_grid* g);   /* ComputeFluxes, grid.cu:841      */
int rift_compute_timestep(rift_grid* g); /* ComputeTimestep, simulator.cu:540 */
int rift_integrate_1(rift_grid* g);      /* Integrate_1, grid.cu:1026       */
int rift_integrate_2(rift_grid* g);      /* Integrate_2, grid.cu:1173       */
int rift_apply_boundaries(rift_grid* g); /* ApplyBoundaries, grid.cu:1408   */

/* Options (rift_set_option).  RIFT_OPT_STEP_PATH: 0 = fused marching step
 * (default), 1 = one kernel per reference phase (Grow, ComputeFluxes,
 * Integrate_1, ApplyBoundaries, ...); both give the same results. */
#define RIFT_OPT_STEP_PATH 1
/* RIFT_OPT_GRAPHS: 1 (default) = replay the fused step from captured HIP
 * graphs (16 steps per launch), 0 = launch its kernels one by one (an
 * ensemble with any member at 0 launches its steps one by one too). */
#define RIFT_OPT_GRAPHS 2
/* RIFT_OPT_VARIANT: launch variant of the fused step's marching kernels,
 * 0 .. RIFT_NVARIANTS-1 (default 0): bit 0 = stage 1 at 5 waves per SIMD
 * instead of 4 (a smaller register budget), bit 1 = stage 2 at 4 instead of
 * 3, bit 2 = work shares dealt to workgroups in launch order
 * instead of by XCD, bits 3-4 = least rows a wave marches in one share (the
 * cells per lane of a share: 14, 7, 28 or 42), bit 5 = every active unit
 * marched by a wave of its own (no workgroups of four neighbouring strips).
 * Every variant computes the same results bit for bit. */
#define RIFT_OPT_VARIANT 3
#define RIFT_NVARIANTS 64
/* RIFT_OPT_MASS_BUDGET: 1 = step with the FP64 volume budget kernels (as
 * RIFT_MASS_BUDGET at creation), 0 = without; switching it on resets the
 * budget (rift_budget_reset).  The state arithmetic is the same either way. */
#define RIFT_OPT_MASS_BUDGET 4
/* RIFT_OPT_DT_KERNEL: 0 (default) = the last workgroup of the fused step's
 * stage-1 march forms dt (ComputeTimestep, simulator.cu:540-566) and the
 * next step's forcing; 1 = a dt kernel of its own between the two marches.
 * Both give the same dt, state and forcing bit for bit. */
#define RIFT_OPT_DT_KERNEL 5
int rift_set_option(rift_grid* g, int option, int value);

/* Autotuner (no counterpart in the reference, whose 16 x 14 launch shape is
 * fixed, grid.cu:18-19): times every launch variant of the fused step on a
 * private copy of the handle's current state (nsteps graph-replayed steps
 * each, after two warm-up steps), times the six fastest twice more, selects
 * the one with the least total on the handle and returns it in *best;
 * ms_per_step (optional, RIFT_NVARIANTS entries) receives each variant's
 * time in the first round.  The handle's own state is not advanced.
 * Needs device memory for a second copy of the domain. */
int rift_autotune(rift_grid* g, int nsteps, int* best, double* ms_per_step);

/* Wait for all queued work of the handle. */
int rift_sync(rift_grid* g);

/* Simulation time and last dt (blocking). */
int rift_get_time(rift_grid* g, double* t, double* dt);
int rift_set_time(rift_grid* g, double t, double dt);

/* Field copies to/from the host (nx*ny doubles; blocking).  Replaces the
 * cudaMemcpy2D transfers of PrintData, PrintSummaryData and the hotstart
 * reader/writer (simulator.cu:568-697, 912-1090).  Ghost cells of RIFT_W
 * hold the reference's values: a ghost's surface is the bed of the adjacent
 * interior cell plus the outflow depth (ApplyBoundaries, grid.cu:1325-1411);
 * RIFT_DEPTH reports that outflow depth itself there. */
int rift_download(rift_grid* g, int field, double* host);
int rift_upload(rift_grid* g, int field, const double* host);

/* Device-resident field windows (no reference counterpart: they let a
 * training or screening pipeline collect solver maps without a host round
 * trip).  Copies the FP32 window [i0, i0+w) x [j0, j0+h) of a stored field
 * (RIFT_DEPTH, RIFT_HU, ..., RIFT_HMAX, RIFT_TWET, ...; not the derived
 * RIFT_W, RIFT_Q, RIFT_BC) to device memory dst, row-major h x w floats,
 * ordered on the handle's stream; rift_sync waits for it. */
int rift_copy_field(rift_grid* g, int field, int i0, int j0, int w, int h, float* dst);

/* Device memory owned by the caller (no reference counterpart), for the
 * dst of rift_copy_field and the arrays of rift_unet.h.  rift_memcpy is
 * synchronous; kind 0 = host to device, 1 = device to host, 2 = device to
 * device. */
int rift_device_alloc(int device, size_t bytes, void** ptr);
int rift_device_free(void* ptr);
int rift_memcpy(void* dst, const void* src, size_t bytes, int kind);

/* Tile state: wet flags and active flags (ntiles bytes each) and the active
 * count, replacing the wet_blocks/active_blocks copies of the hotstart. */
int rift_get_tiles(rift_grid* g, unsigned char* wet, unsigned char* active, int* n_active);
int rift_set_tiles(rift_grid* g, const unsigned char* wet, const unsigned char* active);
/* Per-tile maximal wave speeds mx, my (ntiles floats each). */
int rift_get_tile_speeds(rift_grid* g, float* mx, float* my);
/* Tile-grid geometry: nbx, nby (16 x 14 tiles, grid.cu:18-19). */
int rift_get_tile_dims(rift_grid* g, int* nbx, int* nby);
/* The descriptor the handle was created with (the values SetDeviceConstants
 * stores in constant memory, grid.cu:321-345). */
int rift_get_desc(rift_grid* g, rift_desc* out);
/* Interpolated scalar rates used by the last step: hyetograph (m/s), drain
 * (m/s), hydrograph (per-cell m/s). */
int rift_get_rates(rift_grid* g, float* out3);

/* FP64 sum of a field over interior cells (SumReduce, grid.cu:264-290). */
int rift_sum_interior(rift_grid* g, int field, double* out);
/* rift_sum_interior without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid it).  Per-block
 * FP64 partials and the completion ticket live in a static device array of
 * 32.9 KB per GPU (held until the process ends), cleared by a kernel, not a
 * memset; the partials are summed in a fixed order, so the result does not
 * depend on completion order.  At most 8192 x 8192 interior cells (larger
 * domains: RIFT_E_UNSUPPORTED).  Calls that overlap on one GPU share the
 * array and race; the call blocks until the sum is on the host. */
int rift_sum_interior_noalloc(rift_grid* g, int field, double* out);
/* FP64 mass balance accumulated on the device since rift_init:
 * out[0] = stored water volume (m^3), out[1] = infiltrated volume,
 * out[2] = rain volume added, out[3] = source volume added,
 * out[4] = drain volume removed, out[5] = boundary outflow volume. */
int rift_mass_balance(rift_grid* g, double* out6);
/* FP64 volume budget (m^3) since the last rift_budget_reset (rift_init
 * resets), accumulated on the device by the fused step when the handle was
 * created with RIFT_MASS_BUDGET (no reference counterpart: the reference
 * prints a host-side sum, simulator.cu:609-637, 820-842):
 * out[0] stored water, out[1] infiltrated, out[2] stored + infiltrated at the
 * reset, out[3] rain in, out[4] source in, out[5] drain out, out[6] outflow
 * through the domain boundary, out[7] exchange of the dry clamps and the
 * surge override, out[8] residual = change of stored + infiltrated minus
 * (rain + source - drain - outflow + exchange). */
int rift_budget(rift_grid* g, double* out9);
int rift_budget_reset(rift_grid* g);

/* The handle's HIP stream (hipStream_t), for callers that interleave work. */
int rift_get_stream(rift_grid* g, void** stream);

/* Kernel timing: when enabled, the flux/stage kernels are bracketed by HIP
 * events on the handle's stream; rift_kernel_stats returns launches and total
 * milliseconds per kernel class (0 = stage-1 flux, 1 = stage-2 flux/update,
 * 2 = integrate, 3 = boundaries, 4 = grow, 5 = dt). */
int rift_profile_enable(rift_grid* g, int on);
int rift_kernel_stats(rift_grid* g, int kernel, long* launches, double* total_ms);

/* Wet-cell updates (interior cells of active tiles) of the last step and
 * cumulated since rift_init, read back from the device. */
int rift_wet_cell_updates(rift_grid* g, long* last_step, long* total);

/* ---- Ensembles (no reference counterpart: the reference holds one domain
 * in process-global state, grid.cu:21-40) ----------------------------------
 * Many domains (members) of different sizes, inputs and forcing advanced
 * together: each phase of a step is one launch over all members, and the
 * marching kernels deal equal row shares of every member's active units to
 * the device's resident waves.  Each member keeps its own t, dt (CFL of its
 * own speeds), forcing and tile sets, and its cells see exactly the
 * arithmetic of a single-domain step (bit-identical results).  Members are
 * ordinary handles for set-up and field access (rift_ensemble_member); they
 * share the ensemble's stream.  They must agree on RIFT_TEXTBOOK,
 * RIFT_MASS_BUDGET and RIFT_EULER (Euler is not supported), and use one
 * device. */
typedef struct rift_ensemble rift_ensemble;
int rift_ensemble_create(int n, const rift_desc* descs, rift_ensemble** out);
int rift_ensemble_destroy(rift_ensemble* e);
int rift_ensemble_size(rift_ensemble* e);
rift_grid* rift_ensemble_member(rift_ensemble* e, int i);
/* End time of member i: from the step whose start time is >= tf on, the
 * member is left as it is (0 = never).  Also returns a member that
 * rift_ensemble_finished saw ended to the launches. */
int rift_ensemble_set_tf(rift_ensemble* e, int i, double tf);
/* nsteps batched steps: one step of every member that has not ended (the
 * members rift_ensemble_finished saw at their end time are not launched). */
int rift_ensemble_step(rift_ensemble* e, int nsteps);
/* Members whose time has reached their end time (blocking: one batch of
 * reads); they are left out of the following steps' launches. */
int rift_ensemble_finished(rift_ensemble* e, int* n_finished);
/* Steps until every member has ended (each needs an end time), checking
 * every check_every steps; max_steps < 0: no limit. */
int rift_ensemble_run(rift_ensemble* e, int check_every, long max_steps, long* steps_done);

#ifdef __cplusplus
}
#endif
#endif
