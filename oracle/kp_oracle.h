/*
 * kp_oracle.h -- CPU restatement of RIFT's Kurganov-Petrova step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; nothing in the product (rift_amd/, include/) links, loads or calls
 * it.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * use it.
 *
 * Two builds of one source (oracle/Makefile):
 *   liboracle.so    real = double.  `mode` selects
 *                     ORC_FAITHFUL  the reference as written: sign-blind minmod
 *                                   from the low-word type-pun (grid.cu:62-76),
 *                                   float-narrowed fminf/fmaxf/fabsf/sqrtf/powf
 *                                   on doubles (grid.cu:75,83,134-135,...),
 *                                   float constants widened (grid.cu:28-30);
 *                     ORC_TEXTBOOK  sign-aware minmod and double math.
 *   liboracle32.so  real = float, OpenMP over tiles: the CPU throughput baseline
 *                   (same per-cell update; minmod = min(|a|,|b|,|c|) when all
 *                   three are nonzero, which is the sign-blind rule without the
 *                   low-word cases that only FP64 values can hit).
 *
 * Layout: row-major, row j = y (row 0 = south), column i = x, internal extent
 * nx = b_nx+3 by ny = b_ny+3 with two ghost cells per side
 * (grid_config.cpp:84-85).  Face beds: BX is (nx+1) x ny, west face of cell i
 * at column i; BY is nx x (ny+1), south face of cell j at row j
 * (grid.cu:402-403).  Tiles are 16 x 14 cells; tile id = by*nbx+bx and its
 * cell (tx,ty) is (bx*16+tx+2, by*14+ty+2) (grid.cu:18-19, 576-580).
 */
#ifndef KP_ORACLE_H
#define KP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum orc_mode { ORC_FAITHFUL = 0, ORC_TEXTBOOK = 1 };

/* Feature flags: one bit per reference __constant__ flag (grid.cu:21-25). */
enum orc_flags {
    ORC_HYDROGRAPH    = 1u << 0,  /* dambreak: point/rectangle source       */
    ORC_RAIN_AVG      = 1u << 1,  /* rainfall_averaged (hyetograph CSV)     */
    ORC_RAIN_GRID     = 1u << 2,  /* rainfall_gridded (map series)          */
    ORC_DRAIN_AVG     = 1u << 3,
    ORC_DRAIN_GRID    = 1u << 4,
    ORC_INFILTRATION  = 1u << 5,  /* Green-Ampt                              */
    ORC_SURGE_GRID    = 1u << 6,
    ORC_EULER         = 1u << 7,
    ORC_SAVE_MAX      = 1u << 8,
    ORC_SAVE_ARRIVAL  = 1u << 9,
    ORC_TILE_ACCEL    = 1u << 10,
    ORC_H_INIT        = 1u << 11,
    ORC_RAIN_INTERP   = 1u << 12  /* hyetograph_interp                       */
};

enum orc_field {
    ORC_W = 0, ORC_HU, ORC_HV, ORC_DEPTH, ORC_BC, ORC_N, ORC_K, ORC_F,
    ORC_DW, ORC_DHU, ORC_DHV, ORC_G, ORC_DF, ORC_HMAX, ORC_QMAX, ORC_TWET,
    ORC_TPEAK, ORC_TDRY, ORC_SURGE, ORC_RAINMAP, ORC_DRAINMAP, ORC_QDIAG,
    ORC_NFIELDS
};

enum orc_series_id { ORC_SER_HYETO = 0, ORC_SER_DRAIN = 1, ORC_SER_HYDRO = 2 };
enum orc_map_id    { ORC_MAP_RAIN = 0, ORC_MAP_DRAIN = 1, ORC_MAP_SURGE = 2 };

typedef struct orc_params {
    int    b_nx, b_ny;      /* DEM points (ncols, nrows)                    */
    double dx, dy;          /* internal cell size, m                        */
    double kappa;           /* desingularisation constant                   */
    double courant;         /* courant_limit                                */
    double psi, dtheta;     /* Green-Ampt                                   */
    double t0;              /* start time                                   */
    double nodata;          /* GridConfig::nodata                           */
    double n_const;         /* uniform Manning n when n_corner is NULL      */
    unsigned flags;         /* orc_flags                                     */
    int    mode;            /* orc_mode                                      */
} orc_params;

typedef struct orc_sim orc_sim;

/* --- scalar building blocks (grid.cu:62-214), exported for unit pins ----- */
double orc_signf(double f);
double orc_minmod(int mode, double a, double b, double c);
double orc_limiter(int mode, double U, double Up1, double Um1);
double orc_velocity(int mode, double h, double q, double kappa);
double orc_kp_flux(double Uij, double Fij, double Ukh, double Fkh, double cm, double cp);
/* out[6] = wp, hup, hvp, wm, hum, hvm */
void   orc_construct(int mode, const double c3[9], double Bp, double Bm, double out[6]);
/* left = face state of the cell west/south of the face (its E/N side), right =
 * state of the cell east/north (its W/S side); out[4] = H0, H1, H2, speed.
 * dir 0 = x (getHX), 1 = y (getHY). */
void   orc_face(int mode, int dir, const double left[3], const double right[3],
                double Bface, double kappa, double out[4]);
double orc_kappa(double dx, double dy);
/* Tile dilation of a wet pattern (Grow_k, grid.cu:1188-1222, race-free rule).
 * active_out[t] = t if active else -t. */
void   orc_grow_tiles(int nbx, int nby, const unsigned char* wet, int* active_out);
/* TrackRainfall_k (kind 0: rain > epsilon) / TrackSurge_k (kind 1: surge
 * != nodata and > 0), grid.cu:1224-1276, on the tiles inactive in the last
 * Grow (aval < 0) that are dry; map = nx*ny current map, wet updated. */
void   orc_track_tiles(int nx, int ny, int nbx, int nby, const double* map, int kind, double nodata,
                       const int* aval, unsigned char* wet);

/* --- whole simulation ----------------------------------------------------- */
/* dem: b_nx*b_ny raw DEM (row 0 = south).  h0/n/K corner maps may be NULL.
 * K is in mm/hr as in the input raster. */
orc_sim* orc_create(const orc_params* p, const double* dem, const double* h0_corner,
                    const double* n_corner, const double* K_corner);
void     orc_destroy(orc_sim* s);
int      orc_nx(const orc_sim* s);
int      orc_ny(const orc_sim* s);
int      orc_ntiles(const orc_sim* s);
int      orc_nbx(const orc_sim* s);

/* CSV series given in the reader's units (hours, value); the reader's float
 * conversion and zero-append rule are applied (source.cc:11-63). */
int      orc_set_series_hours(orc_sim* s, int which, const float* hours, const float* value, int n);
/* Hydrograph cell indices idx = j*nx+i (simulator.cu:858-889). */
int      orc_set_sources(orc_sim* s, const int* idx, long n);
/* Map series from corner rasters (b_nx*b_ny each, nodata = params.nodata),
 * nmaps consecutive maps with file index 0..nmaps-1, deltat truncated to int
 * seconds (GridSeries.cuh:29).  which = orc_map_id. */
int      orc_set_map_series(orc_sim* s, int which, const double* maps, int nmaps,
                            int deltat, double tmax);
/* The same with the maps of file indices first .. first+nmaps-1 only (a run
 * continued from a hotstart needs only the maps around its time); a map
 * outside the range reads as a missing file. */
int      orc_set_map_series_range(orc_sim* s, int which, int first, const double* maps, int nmaps,
                                  int deltat, double tmax);

/* One RunSimulation() call (simulator.cu:773-853). */
void     orc_step(orc_sim* s);
void     orc_run(orc_sim* s, int nsteps);
/* The separate phases, for stage-level parity checks. */
void     orc_update_source(orc_sim* s);
void     orc_grow(orc_sim* s);
void     orc_compute_fluxes(orc_sim* s);
void     orc_compute_timestep(orc_sim* s);
void     orc_integrate_1(orc_sim* s);
void     orc_integrate_2(orc_sim* s);
void     orc_apply_boundaries(orc_sim* s);
void     orc_update_maps(orc_sim* s);

void     orc_get_time(const orc_sim* s, double* t, double* dt);
void     orc_set_time(orc_sim* s, double t, double dt);
/* Full nx*ny copies (BX/BY via orc_get_faces). */
int      orc_get_field(const orc_sim* s, int field, double* out);
int      orc_set_field(orc_sim* s, int field, const double* in);
void     orc_get_faces(const orc_sim* s, double* bx, double* by);
/* wet[ntiles], active[ntiles] in the reference's signed encoding (unordered). */
void     orc_get_tiles(const orc_sim* s, unsigned char* wet, int* active, int* n_active);
void     orc_set_tiles(orc_sim* s, const unsigned char* wet, const int* active);
void     orc_get_tile_speeds(const orc_sim* s, double* mx, double* my);
/* Interpolated scalar rates after the last UpdateSource (float members of
 * Source, source.h:8-11): out[3] = hyetograph, drain, hydrograph. */
void     orc_get_rates(const orc_sim* s, float* out);
/* Mass balance: out[0] = sum interior depth * dx*dy, out[1] = sum F*dx*dy,
 * out[2] = cumulative rain volume added, out[3] = cumulative source volume. */
void     orc_mass(const orc_sim* s, double* out);
/* FP64 volume budget (m^3) since the last orc_budget_reset (orc_create
 * resets): out[0] stored water, out[1] infiltrated, out[2] stored +
 * infiltrated at the reset, out[3] rain in, out[4] source in, out[5] drain
 * out, out[6] outflow through the faces of the active region, out[7]
 * exchange of the dry clamps and the surge override, out[8] residual
 * (change of stored + infiltrated minus the budget terms). */
void     orc_budget(const orc_sim* s, double* out);
void     orc_budget_reset(orc_sim* s);

#ifdef __cplusplus
}
#endif
#endif
