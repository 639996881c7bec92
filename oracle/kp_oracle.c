/*
 * kp_oracle.c -- CPU restatement of RIFT's Kurganov-Petrova central-upwind
 * shallow-water step, including the reference's quirks.
 *
 * TEST INFRASTRUCTURE ONLY (see kp_oracle.h).  Every function cites the
 * reference file:line it restates.  Build: oracle/Makefile.
 *
 * Parity pins for each part live in tests/test_oracle_pins.py (closed forms,
 * exact bit patterns, brute force over tile patterns, Stoker's dam-break
 * solution, lake at rest, uniform rain, friction decay, Green-Ampt root).
 *
 * Parity unpinned (documented, cannot be pinned without NVIDIA hardware):
 * the device result of the strict-aliasing type-pun in signf (grid.cu:63-68;
 * defined here as the little-endian low-word rule) and CUDA --use_fast_math
 * sqrtf/powf values (here: correctly rounded host libm float functions).
 */
#include "kp_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef ORC_FLOAT
typedef float real;
#else
typedef double real;
#endif

/* grid.cu:28-30: float literals widened into double __constant__s. */
#define G_GRAV   ((real)9.80665f)
#define EPS      ((real)1.19209e-07f)
#define THETA    ((real)1.3f)
#define TILE_W   16   /* BLOCK_COLS, grid.cu:19 */
#define TILE_H   14   /* BLOCK_ROWS, grid.cu:18 */

/* ---------------------------------------------------------------------------
 * float-narrowed libm calls.  The reference applies fminf/fmaxf/fabsf/sqrtf/
 * powf to doubles (grid.cu:75,83,134-135,164,183-184,213,756-758,800,...): the
 * arguments are rounded to float and the float result widened back.
 * ------------------------------------------------------------------------- */
#ifdef ORC_FLOAT
static inline real n_fmax(int m, real a, real b) { (void)m; return fmaxf(a, b); }
static inline real n_fmin(int m, real a, real b) { (void)m; return fminf(a, b); }
static inline real n_fabs(int m, real a) { (void)m; return fabsf(a); }
static inline real n_sqrt(int m, real a) { (void)m; return sqrtf(a); }
static inline real n_pow(int m, real a, real b) { (void)m; return powf(a, b); }
#else
static inline real n_fmax(int m, real a, real b)
{ return m == ORC_FAITHFUL ? (double)fmaxf((float)a, (float)b) : fmax(a, b); }
static inline real n_fmin(int m, real a, real b)
{ return m == ORC_FAITHFUL ? (double)fminf((float)a, (float)b) : fmin(a, b); }
static inline real n_fabs(int m, real a)
{ return m == ORC_FAITHFUL ? (double)fabsf((float)a) : fabs(a); }
static inline real n_sqrt(int m, real a)
{ return m == ORC_FAITHFUL ? (double)sqrtf((float)a) : sqrt(a); }
static inline real n_pow(int m, real a, real b)
{ return m == ORC_FAITHFUL ? (double)powf((float)a, (float)b) : pow(a, b); }
#endif

/* ---------------------------------------------------------------------------
 * signf, grid.cu:62-71.  `(int&)f` reads the low 32-bit word of the double on
 * a little-endian machine.  The result is 0 when the low 31 bits are zero and
 * otherwise 1.0 with the low word's bit 31 OR-ed into 1.0's (zero) low word,
 * i.e. 1 or 1 + 2^-21.  It is never negative.
 * ------------------------------------------------------------------------- */
double orc_signf(double f)
{
    uint64_t bits;
    memcpy(&bits, &f, sizeof bits);
    uint32_t lo = (uint32_t)(bits & 0xffffffffu);
    if ((lo & 0x7fffffffu) == 0) return 0.0;
    double r = 1.0;
    uint64_t rb;
    memcpy(&rb, &r, sizeof rb);
    rb |= (uint64_t)(lo & 0x80000000u);
    memcpy(&r, &rb, sizeof r);
    return r;
}

/* minmod, grid.cu:73-76 (faithful); textbook = sign-aware three-way minmod. */
static inline real minmod_r(int m, real a, real b, real c)
{
    if (m == ORC_TEXTBOOK) {
        if (a > 0 && b > 0 && c > 0) return fmin(fmin(a, b), c);
        if (a < 0 && b < 0 && c < 0) return fmax(fmax(a, b), c);
        return 0;
    }
#ifdef ORC_FLOAT
    /* FP32: the low-word cases cannot occur for float operands widened to
     * double in a way the FP32 arithmetic tracks; a zero operand makes the
     * min zero, so the rule reduces to min(|a|,|b|,|c|). */
    if (a == 0 || b == 0 || c == 0) return 0;
    return fminf(fminf(fabsf(a), fabsf(b)), fabsf(c));
#else
    double sa = orc_signf(a), sb = orc_signf(b), sc = orc_signf(c);
    return 0.25f * sa * (sa + sb) * (sb + sc) *
           n_fmin(m, n_fmin(m, n_fabs(m, a), n_fabs(m, b)), n_fabs(m, c));
#endif
}

double orc_minmod(int mode, double a, double b, double c)
{ return (double)minmod_r(mode, (real)a, (real)b, (real)c); }

/* limiter, grid.cu:78-80 */
static inline real limiter_r(int m, real U, real Up1, real Um1)
{
    return (real)0.5f * minmod_r(m, THETA * (Up1 - U), (real)0.5f * (Up1 - Um1), THETA * (U - Um1));
}

double orc_limiter(int mode, double U, double Up1, double Um1)
{ return (double)limiter_r(mode, (real)U, (real)Up1, (real)Um1); }

/* getVelocity, grid.cu:82-84 */
static inline real velocity_r(int m, real h, real q, real kappa)
{
    return (real)2.f * h * q / (h * h + n_fmax(m, h * h, kappa));
}

double orc_velocity(int mode, double h, double q, double kappa)
{ return (double)velocity_r(mode, (real)h, (real)q, (real)kappa); }

/* computeH, grid.cu:86-89 */
static inline real kp_flux_r(real Uij, real Fij, real Ukh, real Fkh, real cm, real cp)
{
    return (cp * Fij - cm * Fkh) / (cp - cm) + (cp * cm) / (cp - cm) * (Ukh - Uij);
}

double orc_kp_flux(double Uij, double Fij, double Ukh, double Fkh, double cm, double cp)
{ return (double)kp_flux_r((real)Uij, (real)Fij, (real)Ukh, (real)Fkh, (real)cm, (real)cp); }

/* Construct, grid.cu:91-116: MUSCL face states plus the positivity fix. */
static inline void construct_r(int m, real w, real hu, real hv,
                               real wp1, real hup1, real hvp1,
                               real wm1, real hum1, real hvm1,
                               real Bp, real Bm, real out[6])
{
    real wz = limiter_r(m, w, wp1, wm1);
    real huz = limiter_r(m, hu, hup1, hum1);
    real hvz = limiter_r(m, hv, hvp1, hvm1);
    real wp = w + wz, wm = w - wz;
    real hup = hu + huz, hum = hu - huz;
    real hvp = hv + hvz, hvm = hv - hvz;
    real wbar = wp + wm;
    if (wp < Bp) { wp = Bp; wm = wbar - Bp; }
    if (wm < Bm) { wp = wbar - Bm; wm = Bm; }
    out[0] = wp; out[1] = hup; out[2] = hvp;
    out[3] = wm; out[4] = hum; out[5] = hvm;
}

void orc_construct(int mode, const double c3[9], double Bp, double Bm, double out[6])
{
    real o[6];
    construct_r(mode, (real)c3[0], (real)c3[1], (real)c3[2], (real)c3[3], (real)c3[4],
                (real)c3[5], (real)c3[6], (real)c3[7], (real)c3[8], (real)Bp, (real)Bm, o);
    for (int k = 0; k < 6; ++k) out[k] = (double)o[k];
}

/* getHX / getHY, grid.cu:118-165 / 167-214.  L = face state on the west/south
 * side of the face (that cell's E/N value), R = the east/north cell's W/S value.
 * dir 0: x (speeds from u, fluxes F), dir 1: y (speeds from v, fluxes G). */
static inline real face_r(int m, int dir, real wp, real hup, real hvp,
                          real wp1m, real hup1m, real hvp1m, real Bp, real kappa,
                          real H[3])
{
    real hp = wp - Bp;
    real hp1m = wp1m - Bp;
    real up = velocity_r(m, hp, hup, kappa);
    real up1m = velocity_r(m, hp1m, hup1m, kappa);
    real vp = velocity_r(m, hp, hvp, kappa);
    real vp1m = velocity_r(m, hp1m, hvp1m, kappa);

    hup = hp * up;
    hup1m = hp1m * up1m;
    hvp = hp * vp;
    hvp1m = hp1m * vp1m;

    real cp = (dir == 0) ? up : vp;
    real cq = (dir == 0) ? up1m : vp1m;
    /* sqrtf of a negative depth is NaN; fmaxf/fminf then return the other
     * operand (IEEE maxNum), so that side's speed drops out. */
    real ap = n_fmax(m, n_fmax(m, cp + n_sqrt(m, G_GRAV * hp), cq + n_sqrt(m, G_GRAV * hp1m)), (real)0.f);
    real am = n_fmin(m, n_fmin(m, cp - n_sqrt(m, G_GRAV * hp), cq - n_sqrt(m, G_GRAV * hp1m)), (real)0.f);

    real Fp0, Fp1, Fp2, Fq0, Fq1, Fq2;
    if (hp > (real)0.f) {
        if (dir == 0) {
            Fp0 = hup;
            Fp1 = hup * hup / hp + (real)0.5f * G_GRAV * hp * hp;
            Fp2 = hup * hvp / hp;
        } else {
            Fp0 = hvp;
            Fp1 = hup * hvp / hp;
            Fp2 = hvp * hvp / hp + (real)0.5f * G_GRAV * hp * hp;
        }
    } else {
        Fp0 = Fp1 = Fp2 = 0;
    }
    if (hp1m > (real)0.f) {
        if (dir == 0) {
            Fq0 = hup1m;
            Fq1 = hup1m * hup1m / hp1m + (real)0.5f * G_GRAV * hp1m * hp1m;
            Fq2 = hup1m * hvp1m / hp1m;
        } else {
            Fq0 = hvp1m;
            Fq1 = hup1m * hvp1m / hp1m;
            Fq2 = hvp1m * hvp1m / hp1m + (real)0.5f * G_GRAV * hp1m * hp1m;
        }
    } else {
        Fq0 = Fq1 = Fq2 = 0;
    }

    if (ap - am > (real)0.f) {
        H[0] = kp_flux_r(wp, Fp0, wp1m, Fq0, am, ap);
        H[1] = kp_flux_r(hup, Fp1, hup1m, Fq1, am, ap);
        H[2] = kp_flux_r(hvp, Fp2, hvp1m, Fq2, am, ap);
    } else {
        H[0] = H[1] = H[2] = 0;
    }
    return n_fmax(m, ap, -am);
}

void orc_face(int mode, int dir, const double left[3], const double right[3],
              double Bface, double kappa, double out[4])
{
    real H[3];
    real sp = face_r(mode, dir, (real)left[0], (real)left[1], (real)left[2],
                     (real)right[0], (real)right[1], (real)right[2],
                     (real)Bface, (real)kappa, H);
    out[0] = H[0]; out[1] = H[1]; out[2] = H[2]; out[3] = sp;
}

/* simulator.cu:281: kappa = sqrtf(0.01f*fmaxf(1.f, fminf(dx, dy))) in float. */
double orc_kappa(double dx, double dy)
{
    return (double)sqrtf(0.01f * fmaxf(1.f, fminf((float)dx, (float)dy)));
}

/* Grow_k, grid.cu:1188-1222, with the intended (race-free) neighbour rule:
 * out-of-grid neighbours count as dry; -0 == 0 keeps tile 0 active. */
void orc_grow_tiles(int nbx, int nby, const unsigned char* wet, int* active_out)
{
    for (int by = 0; by < nby; ++by) {
        for (int bx = 0; bx < nbx; ++bx) {
            int t = by * nbx + bx;
            int a = wet[t] ||
                    (bx > 0 && wet[t - 1]) || (bx < nbx - 1 && wet[t + 1]) ||
                    (by > 0 && wet[t - nbx]) || (by < nby - 1 && wet[t + nbx]);
            active_out[t] = a ? t : -t;
        }
    }
}

/* ---------------------------------------------------------------------------
 * Simulation state
 * ------------------------------------------------------------------------- */
typedef struct series {
    int n;
    float* time;   /* seconds, float (source.h:8)  */
    float* rate;
    float interpolated;   /* Source::interpolated_rate (source.h:10) */
} series;

typedef struct map_series {
    int configured, interpolated, allow_nodata, nmaps, deltat;
    int first;             /* file index of maps[0] */
    double tmax, scale, in_time, next_time, sum_cache;
    int done, cur_init;
    double* maps;          /* nmaps * b_nx*b_ny */
    real *cur, *p0, *p1, *ibuf;
} map_series;

struct orc_sim {
    int b_nx, b_ny, nx, ny, nbx, nby, ntiles;
    double dx, dy, kappa, courant, psi, dtheta, nodata;
    unsigned flags;
    int mode;
    double t, dt;
    real *w, *hu, *hv, *w_old, *hu_old, *hv_old, *dw, *dhu, *dhv, *G, *BC, *n;
    real *F, *F_old, *dF, *K, *BX, *BY;
    real *hdiag, *qdiag, *h_max, *q_max, *t_wet, *t_peak, *t_dry;
    unsigned char* clamp;   /* deferred dry-clamp marks of ComputeFluxes */
    unsigned char* wet;
    int* aval;              /* active_blocks values indexed by tile */
    int* list;              /* current active list */
    int nlist;
    double *mx, *my;
    series ser[3];
    long nsrc;
    int* src_idx;
    double* src_rate;
    map_series ms[3];
    double v_rain, v_src;
    /* FP64 volume budget since orc_budget_reset (m^3): stored + infiltrated
     * volume at the reset, then rain and source in, drain and outflow out
     * (through the faces of the active region: the domain boundary and faces
     * to inactive tiles), and the exchange of the dry clamps and the surge
     * override.  Each flux pass leaves its per-cell rate sums (m/s summed
     * over cells) and its clamp change (m summed over cells) in st_*; the
     * integration stage that follows weights them with its share of dt. */
    double b_v0, b_rain, b_src, b_drain, b_out, b_exch;
    double st_rain, st_src, st_drain, st_out, st_clamp;
};

#define IX(s, i, j) ((size_t)(j) * (size_t)(s)->nx + (size_t)(i))
#define IBX(s, i, j) ((size_t)(j) * (size_t)((s)->nx + 1) + (size_t)(i))
#define IBY(s, i, j) ((size_t)(j) * (size_t)(s)->nx + (size_t)(i))

static real* zalloc(size_t n) { return (real*)calloc(n ? n : 1, sizeof(real)); }

int orc_nx(const orc_sim* s) { return s->nx; }
int orc_ny(const orc_sim* s) { return s->ny; }
int orc_ntiles(const orc_sim* s) { return s->ntiles; }
int orc_nbx(const orc_sim* s) { return s->nbx; }

/* ApplyBoundaries{X,Y}_k, grid.cu:1325-1406: critical-depth outflow ghosts. */
void orc_apply_boundaries(orc_sim* s)
{
    const int m = s->mode, nx = s->nx, ny = s->ny;
    const real third = (real)(1.f / 3.f);
    for (int j = 2; j <= ny - 3; ++j) {
        size_t c = IX(s, 2, j);
        real HL = n_fmin(m, n_pow(m, s->hu[c] * s->hu[c] / G_GRAV, third), s->w[c] - s->BC[c]);
        HL = (HL > EPS) ? HL : (real)0.f;
        s->w[IX(s, 0, j)] = s->w[IX(s, 1, j)] = s->BC[c] + HL;
        s->hu[IX(s, 0, j)] = s->hu[IX(s, 1, j)] = s->hu[c];
        s->hv[IX(s, 0, j)] = s->hv[IX(s, 1, j)] = s->hv[c];
        c = IX(s, nx - 3, j);
        real HR = n_fmin(m, n_pow(m, s->hu[c] * s->hu[c] / G_GRAV, third), s->w[c] - s->BC[c]);
        HR = (HR > EPS) ? HR : (real)0.f;
        s->w[IX(s, nx - 1, j)] = s->w[IX(s, nx - 2, j)] = s->BC[c] + HR;
        s->hu[IX(s, nx - 1, j)] = s->hu[IX(s, nx - 2, j)] = s->hu[c];
        s->hv[IX(s, nx - 1, j)] = s->hv[IX(s, nx - 2, j)] = s->hv[c];
    }
    for (int i = 2; i <= nx - 3; ++i) {
        size_t c = IX(s, i, 2);
        real HB = n_fmin(m, n_pow(m, s->hv[c] * s->hv[c] / G_GRAV, third), s->w[c] - s->BC[c]);
        s->w[IX(s, i, 0)] = s->w[IX(s, i, 1)] = s->BC[c] + HB;
        s->hu[IX(s, i, 0)] = s->hu[IX(s, i, 1)] = s->hu[c];
        s->hv[IX(s, i, 0)] = s->hv[IX(s, i, 1)] = s->hv[c];
        c = IX(s, i, ny - 3);
        real HT = n_fmin(m, n_pow(m, s->hv[c] * s->hv[c] / G_GRAV, third), s->w[c] - s->BC[c]);
        s->w[IX(s, i, ny - 1)] = s->w[IX(s, i, ny - 2)] = s->BC[c] + HT;
        s->hu[IX(s, i, ny - 1)] = s->hu[IX(s, i, ny - 2)] = s->hu[c];
        s->hv[IX(s, i, ny - 1)] = s->hv[IX(s, i, ny - 2)] = s->hv[c];
    }
}

/* --- map series (GridSeries.cu) ------------------------------------------ */

/* GridSeries::p_interp, GridSeries.cu:91-122: four-corner mean where more
 * than two corners hold data, times the series scale. */
static int ms_read(orc_sim* s, map_series* ms, int index)
{
    if (index < ms->first || index >= ms->first + ms->nmaps) return -1;   /* missing file */
    const double* b = ms->maps + (size_t)(index - ms->first) * (size_t)s->b_nx * (size_t)s->b_ny;
    size_t N = (size_t)s->nx * (size_t)s->ny;
    real fill = ms->allow_nodata ? (real)s->nodata : (real)0;
    for (size_t k = 0; k < N; ++k) ms->ibuf[k] = fill;
    for (int j = 2; j < s->ny - 2; ++j) {
        for (int i = 2; i < s->nx - 2; ++i) {
            int jt = j - 2, it = i - 2, nnd = 0;
            double vsum = 0.0;
            for (int ni = 0; ni < 2; ++ni) {
                for (int nj = 0; nj < 2; ++nj) {
                    double v = b[(size_t)(jt + nj) * s->b_nx + (it + ni)];
                    if (v != s->nodata) { nnd++; vsum += v; }
                }
            }
            if (nnd > 2) {
                double v = vsum / (double)nnd;
                v *= ms->scale;
                ms->ibuf[IX(s, i, j)] = (real)v;
            }
        }
    }
    return 0;
}

/* FillGrid_k, grid.cu:492-503: fill from (2,2) over the tile-grid extent. */
static void ms_fill(orc_sim* s, real* x, real v)
{
    for (int j = 2; j < s->ny; ++j)
        for (int i = 2; i < s->nx; ++i) x[IX(s, i, j)] = v;
}

/* SumReduce, grid.cu:219-290: per-tile row sums then column sums. */
static double sum_interior_tiles(const orc_sim* s, const real* x)
{
    double total = 0.0;
    for (int by = 0; by < s->nby; ++by) {
        for (int bx = 0; bx < s->nbx; ++bx) {
            double rows[TILE_H];
            for (int ty = 0; ty < TILE_H; ++ty) {
                double acc = 0.0;
                for (int tx = 0; tx < TILE_W; ++tx) {
                    int i = bx * TILE_W + tx + 2, j = by * TILE_H + ty + 2;
                    double v = (i > s->nx - 3 || j > s->ny - 3) ? 0.0 : (double)x[IX(s, i, j)];
                    acc = (tx == 0) ? v : acc + v;
                }
                rows[ty] = acc;
            }
            double acc = rows[0];
            for (int ty = 1; ty < TILE_H; ++ty) acc += rows[ty];
            total += acc;
        }
    }
    return total;
}

static void ms_copy_cur(orc_sim* s, map_series* ms)
{
    memcpy(ms->cur, ms->ibuf, sizeof(real) * (size_t)s->nx * (size_t)s->ny);
}

/* GridSeries::p_update, GridSeries.cu:144-178 */
static void ms_base_update(orc_sim* s, map_series* ms, double t)
{
    int sendit = 0;
    if (ms->in_time < 0.0) {
        ms->in_time = t;
        if (ms->in_time < ms->tmax) {
            ms_read(s, ms, (int)trunc(ms->in_time / (double)ms->deltat));
            sendit = 1;
        }
    }
    if (t >= ms->tmax) {
        if (!ms->done) {
            size_t N = (size_t)s->nx * (size_t)s->ny;
            for (size_t k = 0; k < N; ++k) ms->ibuf[k] = 0;
            sendit = 1;
        }
        ms->done = 1;
    } else if (t >= ms->in_time + (double)ms->deltat) {
        ms->in_time += (double)ms->deltat;
        ms_read(s, ms, (int)trunc(ms->in_time / (double)ms->deltat));
        sendit = 1;
    }
    if (sendit) ms_copy_cur(s, ms);
}

/* HyetographGridSeries::p_update, GridSeries.cu:261-283: the rate in
 * [t, t+dt) comes from the map at t+dt; its interior sum is cached. */
static void ms_held_update(orc_sim* s, map_series* ms, double t)
{
    if (ms->in_time < 0.0) ms_base_update(s, ms, t);
    if (t >= ms->tmax) {
        ms_base_update(s, ms, t);
    } else if (t >= ms->in_time) {
        ms->in_time += (double)ms->deltat;
        ms_read(s, ms, (int)trunc(ms->in_time / (double)ms->deltat));
        ms_copy_cur(s, ms);
        ms->sum_cache = sum_interior_tiles(s, ms->cur);
    }
}

/* InterpolatedGridSeries::p_update, GridSeries.cu:325-384, and InterpGrid_k,
 * grid.cu:292-311 (over i in [2,nx-1], j in [2,ny-1]). */
static void ms_interp_update(orc_sim* s, map_series* ms, double t)
{
    size_t N = (size_t)s->nx * (size_t)s->ny;
    if (!ms->cur_init) {
        real v = ms->allow_nodata ? (real)s->nodata : (real)0;
        ms_fill(s, ms->cur, v);
        ms_fill(s, ms->p0, v);
        ms_fill(s, ms->p1, v);
        ms->cur_init = 1;
    }
    if (ms->in_time < 0.0) {
        ms->in_time = t;
        int index = (int)trunc(ms->in_time / (double)ms->deltat);
        ms_read(s, ms, index);
        memcpy(ms->p0, ms->ibuf, sizeof(real) * N);
        ms->in_time = index * (double)ms->deltat;
        ms->next_time = ms->in_time + (double)ms->deltat;
        ms_read(s, ms, index + 1);
        memcpy(ms->p1, ms->ibuf, sizeof(real) * N);
    } else if (t >= ms->tmax) {
        if (!ms->done) {
            ms->in_time = ms->next_time;
            ms->next_time = ms->in_time + (double)ms->deltat;
            memcpy(ms->p0, ms->p1, sizeof(real) * N);
        }
        ms->done = 1;
    } else if (t >= ms->next_time) {
        memcpy(ms->p0, ms->p1, sizeof(real) * N);
        ms->in_time = ms->next_time;
        ms->next_time = ms->in_time + (double)ms->deltat;
        ms_read(s, ms, (int)trunc(ms->next_time / (double)ms->deltat));
        memcpy(ms->p1, ms->ibuf, sizeof(real) * N);
    }
    real factor = (real)((t - ms->in_time) / (ms->next_time - ms->in_time));
    const real nd = (real)s->nodata;
#pragma omp parallel for schedule(static)
    for (int j = 2; j < s->ny; ++j) {
        for (int i = 2; i < s->nx; ++i) {
            size_t c = IX(s, i, j);
            real x0 = ms->p0[c], x1 = ms->p1[c];
            if (x1 == nd && x0 == nd) ms->cur[c] = nd;
            else if (x1 == nd) ms->cur[c] = x0;
            else if (x0 == nd) ms->cur[c] = x1;
            else ms->cur[c] = factor * (x1 - x0) + x0;
        }
    }
}

static void ms_update(orc_sim* s, int which, double t)
{
    map_series* ms = &s->ms[which];
    if (!ms->configured) return;
    if (ms->interpolated) ms_interp_update(s, ms, t);
    else ms_held_update(s, ms, t);
}

int orc_set_map_series(orc_sim* s, int which, const double* maps, int nmaps, int deltat, double tmax)
{
    return orc_set_map_series_range(s, which, 0, maps, nmaps, deltat, tmax);
}

int orc_set_map_series_range(orc_sim* s, int which, int first, const double* maps, int nmaps, int deltat,
                             double tmax)
{
    if (which < 0 || which > 2 || nmaps <= 0 || deltat <= 0 || first < 0) return -1;
    map_series* ms = &s->ms[which];
    ms->first = first;
    size_t per = (size_t)s->b_nx * (size_t)s->b_ny;
    size_t N = (size_t)s->nx * (size_t)s->ny;
    free(ms->maps);
    ms->maps = (double*)malloc(sizeof(double) * per * (size_t)nmaps);
    memcpy(ms->maps, maps, sizeof(double) * per * (size_t)nmaps);
    ms->nmaps = nmaps;
    ms->deltat = deltat;
    ms->tmax = tmax;
    ms->in_time = -9999.0;
    ms->next_time = 0.0;
    ms->done = 0;
    ms->sum_cache = 0.0;
    if (!ms->cur) { ms->cur = zalloc(N); ms->p0 = zalloc(N); ms->p1 = zalloc(N); ms->ibuf = zalloc(N); }
    /* simulator.cu:338-374: rain maps are held (HyetographGridSeries) unless
     * hyetograph_interp; drain maps are held; surge maps are interpolated with
     * scale 1 and nodata allowed. */
    if (which == ORC_MAP_RAIN) {
        ms->interpolated = (s->flags & ORC_RAIN_INTERP) ? 1 : 0;
        ms->allow_nodata = 0;
        ms->scale = 1.0 / 3600.0 / 1000.0;
    } else if (which == ORC_MAP_DRAIN) {
        ms->interpolated = 0;
        ms->allow_nodata = 0;
        ms->scale = 1.0 / 3600.0 / 1000.0;
    } else {
        ms->interpolated = 1;
        ms->allow_nodata = 1;
        ms->scale = 1.0;
    }
    /* allow_no_data(true) clears p_current_dev_init (GridSeries.cuh:52-56). */
    ms->cur_init = ms->allow_nodata ? 0 : 1;
    ms->configured = 1;
    ms_update(s, which, s->t);
    return 0;
}

/* --- scalar series (source.cc) -------------------------------------------- */

/* Source::ReadSource, source.cc:11-63: hours -> seconds in float; a zero rate is
 * appended one mean interval after the last point when the last rate != 0. */
int orc_set_series_hours(orc_sim* s, int which, const float* hours, const float* value, int n)
{
    if (which < 0 || which > 2 || n < 1) return -1;
    series* sr = &s->ser[which];
    free(sr->time); free(sr->rate);
    sr->time = (float*)malloc(sizeof(float) * (size_t)(n + 1));
    sr->rate = (float*)malloc(sizeof(float) * (size_t)(n + 1));
    for (int k = 0; k < n; ++k) {
        float v = hours[k];
        sr->time[k] = v * 60 * 60;
        sr->rate[k] = value[k];
    }
    sr->n = n;
    if (sr->rate[n - 1] != 0.f) {
        float dt_estimate = 0.f;
        for (int k = 1; k < n; ++k) dt_estimate += sr->time[k] - sr->time[k - 1];
        dt_estimate /= (float)n - 1;
        sr->time[n] = sr->time[n - 1] + dt_estimate;
        sr->rate[n] = 0.f;
        sr->n = n + 1;
    }
    sr->interpolated = 0.f;
    return 0;
}

/* Source::InterpolateRate, source.cc:65-73 (float arithmetic). */
static void series_interp(series* sr, float sim_time)
{
    for (int i = 0; i < sr->n - 1; ++i) {
        if (sim_time >= sr->time[i] && sim_time < sr->time[i + 1]) {
            float num = sr->rate[i + 1] - sr->rate[i];
            float den = sr->time[i + 1] - sr->time[i];
            float q = num / den;
            float d = sim_time - sr->time[i];
            float p = q * d;
            sr->interpolated = sr->rate[i] + p;
        }
    }
}

int orc_set_sources(orc_sim* s, const int* idx, long n)
{
    free(s->src_idx); free(s->src_rate);
    s->nsrc = n;
    s->src_idx = (int*)malloc(sizeof(int) * (size_t)(n ? n : 1));
    s->src_rate = (double*)calloc((size_t)(n ? n : 1), sizeof(double));
    memcpy(s->src_idx, idx, sizeof(int) * (size_t)n);
    return 0;
}

/* UpdateSource + setSourceValue, simulator.cu:500-538, 891-901. */
void orc_update_source(orc_sim* s)
{
    if (s->flags & ORC_HYDROGRAPH) {
        series* sr = &s->ser[ORC_SER_HYDRO];
        if (sr->n > 0) series_interp(sr, (float)s->t);
        sr->interpolated /= 35.3147f;
        double value = (double)sr->interpolated;
        for (long k = 0; k < s->nsrc; ++k) {
            s->src_rate[k] = value / s->dx / s->dy / (double)s->nsrc;
            s->v_src += s->src_rate[k] * s->dt * s->dx * s->dy;
        }
    }
    if (s->flags & ORC_RAIN_AVG) {
        series* sr = &s->ser[ORC_SER_HYETO];
        if (sr->n > 0) series_interp(sr, (float)s->t);
        sr->interpolated /= (3600.f * 1000.f);
        s->v_rain += (double)sr->interpolated * s->dt *
                     (double)((s->nx - 4) * (s->ny - 4)) * s->dx * s->dy;
    }
    if (s->flags & ORC_DRAIN_AVG) {
        series* sr = &s->ser[ORC_SER_DRAIN];
        if (sr->n > 0) series_interp(sr, (float)s->t);
        sr->interpolated /= (3600.f * 1000.f);
    }
}

/* --- tiles ---------------------------------------------------------------- */

static void rebuild_list(orc_sim* s)
{
    s->nlist = 0;
    for (int t = 0; t < s->ntiles; ++t)
        if (s->aval[t] > -1) s->list[s->nlist++] = t;
}

/* TrackRainfall_k / TrackSurge_k, grid.cu:1224-1276: a tile inactive in the
 * last Grow (active value < 0; tile 0 never, -0 == 0) and dry becomes wet
 * when a cell of its window holds gridded rain > epsilon (kind 0) or surge
 * != nodata and > 0 (kind 1).  The window's rows use the BLOCK_COLS stride
 * (j = by*16 + ty + 2, ty < 14) as the reference's do; cells outside
 * [2, nx-3] x [2, ny-3] are skipped.  map: nx*ny values of the current map. */
static void track_tiles(int nx, int ny, int nbx, int nby, const real* map, int kind, real nodata,
                        const int* aval, unsigned char* wet)
{
    for (int tile = 0; tile < nbx * nby; ++tile) {
        if (aval[tile] >= 0 || wet[tile]) continue;
        int bx = tile % nbx, by = tile / nbx;
        int hit = 0;
        for (int ty = 0; ty < TILE_H && !hit; ++ty) {
            for (int tx = 0; tx < TILE_W; ++tx) {
                int i = bx * TILE_W + tx + 2;
                int j = by * TILE_W + ty + 2;
                if (i < 2 || j < 2 || i > nx - 3 || j > ny - 3) continue;
                real v = map[(size_t)j * (size_t)nx + (size_t)i];
                if (kind == 0 ? (v > EPS) : (v != nodata && v > 0)) { hit = 1; break; }
            }
        }
        if (hit) wet[tile] = 1;
    }
}

void orc_track_tiles(int nx, int ny, int nbx, int nby, const double* map, int kind, double nodata,
                     const int* aval, unsigned char* wet)
{
    const size_t N = (size_t)nx * (size_t)ny;
    real* m = (real*)malloc(sizeof(real) * (N ? N : 1));
    for (size_t k = 0; k < N; ++k) m[k] = (real)map[k];
    track_tiles(nx, ny, nbx, nby, m, kind, (real)nodata, aval, wet);
    free(m);
}

/* Grow, grid.cu:1279-1322, with TrackRainfall_k/TrackSurge_k (1224-1276):
 * previously inactive, dry tiles become wet when gridded rain > epsilon or
 * surge > 0 in their cells; rows use the BLOCK_COLS stride of the reference. */
void orc_grow(orc_sim* s)
{
    if (s->flags & (ORC_RAIN_GRID | ORC_SURGE_GRID)) {
        for (int pass = 0; pass < 2; ++pass) {
            if (pass == 0 && !(s->flags & ORC_RAIN_GRID)) continue;
            if (pass == 1 && !(s->flags & ORC_SURGE_GRID)) continue;
            const real* map = (pass == 0) ? s->ms[ORC_MAP_RAIN].cur : s->ms[ORC_MAP_SURGE].cur;
            if (!map) continue;
            track_tiles(s->nx, s->ny, s->nbx, s->nby, map, pass, (real)s->nodata, s->aval, s->wet);
        }
    }
    orc_grow_tiles(s->nbx, s->nby, s->wet, s->aval);
    rebuild_list(s);
}

/* --- ComputeFluxes_k, grid.cu:564-839 ------------------------------------- */

typedef struct stage_sums {
    double rain, src, drain, out;
} stage_sums;

/* active_blocks value >= 0 marks an active tile (tile 0 always, -0 == 0) */
static int tile_active(const orc_sim* s, int bx, int by)
{
    if (bx < 0 || by < 0 || bx >= s->nbx || by >= s->nby) return 0;
    return s->aval[by * s->nbx + bx] >= 0;
}

static void flux_tile(orc_sim* s, int tile, double dtd, stage_sums* sums)
{
    const int m = s->mode, nx = s->nx, ny = s->ny;
    const real kappa = (real)s->kappa, dx = (real)s->dx, dy = (real)s->dy;
    const real dt = (real)dtd;
    const int bx = tile % s->nbx, by = tile / s->nbx;
    const int i0 = bx * TILE_W + 2, j0 = by * TILE_H + 2;
    int i1 = i0 + TILE_W, j1 = j0 + TILE_H;
    if (i1 > nx - 2) i1 = nx - 2;   /* interior cells only: i <= nx-3 */
    if (j1 > ny - 2) j1 = ny - 2;
    real smx = 0, smy = 0;
    if (i0 >= i1 || j0 >= j1) { s->mx[tile] = 0; s->my[tile] = 0; return; }

    real st[TILE_W + 2][6];       /* face states of cells i0-1 .. i1 */
    real Hf[TILE_W + 1][3], sp[TILE_W + 1];

    /* x-pass, row by row */
    for (int j = j0; j < j1; ++j) {
        for (int c = i0 - 1; c <= i1; ++c) {
            size_t q = IX(s, c, j), ql = IX(s, c - 1, j), qr = IX(s, c + 1, j);
            construct_r(m, s->w[q], s->hu[q], s->hv[q], s->w[qr], s->hu[qr], s->hv[qr],
                        s->w[ql], s->hu[ql], s->hv[ql],
                        s->BX[IBX(s, c + 1, j)], s->BX[IBX(s, c, j)], st[c - (i0 - 1)]);
        }
        for (int f = i0; f <= i1; ++f) {          /* face f = west face of cell f */
            const real* L = st[f - 1 - (i0 - 1)];
            const real* R = st[f - (i0 - 1)];
            sp[f - i0] = face_r(m, 0, L[0], L[1], L[2], R[3], R[4], R[5],
                                s->BX[IBX(s, f, j)], kappa, Hf[f - i0]);
        }
        /* the tile's west / east faces on the edge of the active region */
        if (i0 == 2 || !tile_active(s, bx - 1, by)) sums->out -= (double)(Hf[0][0] / dx);
        if (i1 == nx - 2 || !tile_active(s, bx + 1, by)) sums->out += (double)(Hf[i1 - i0][0] / dx);
        for (int i = i0; i < i1; ++i) {
            size_t q = IX(s, i, j);
            s->dw[q] = -(Hf[i + 1 - i0][0] - Hf[i - i0][0]) / dx;
            s->dhu[q] = -(Hf[i + 1 - i0][1] - Hf[i - i0][1]) / dx;
            s->dhv[q] = -(Hf[i + 1 - i0][2] - Hf[i - i0][2]) / dx;
            smx = n_fmax(m, smx, sp[i + 1 - i0]);
        }
    }
    /* y-pass, column by column */
    real sty[TILE_H + 2][6];
    real Hg[TILE_H + 1][3], spy[TILE_H + 1];
    for (int i = i0; i < i1; ++i) {
        for (int c = j0 - 1; c <= j1; ++c) {
            size_t q = IX(s, i, c), ql = IX(s, i, c - 1), qr = IX(s, i, c + 1);
            construct_r(m, s->w[q], s->hu[q], s->hv[q], s->w[qr], s->hu[qr], s->hv[qr],
                        s->w[ql], s->hu[ql], s->hv[ql],
                        s->BY[IBY(s, i, c + 1)], s->BY[IBY(s, i, c)], sty[c - (j0 - 1)]);
        }
        for (int f = j0; f <= j1; ++f) {
            const real* L = sty[f - 1 - (j0 - 1)];
            const real* R = sty[f - (j0 - 1)];
            spy[f - j0] = face_r(m, 1, L[0], L[1], L[2], R[3], R[4], R[5],
                                 s->BY[IBY(s, i, f)], kappa, Hg[f - j0]);
        }
        if (j0 == 2 || !tile_active(s, bx, by - 1)) sums->out -= (double)(Hg[0][0] / dy);
        if (j1 == ny - 2 || !tile_active(s, bx, by + 1)) sums->out += (double)(Hg[j1 - j0][0] / dy);
        for (int j = j0; j < j1; ++j) {
            size_t q = IX(s, i, j);
            s->dw[q] += -(Hg[j + 1 - j0][0] - Hg[j - j0][0]) / dy;
            s->dhu[q] += -(Hg[j + 1 - j0][1] - Hg[j - j0][1]) / dy;
            s->dhv[q] += -(Hg[j + 1 - j0][2] - Hg[j - j0][2]) / dy;
            smy = n_fmax(m, smy, spy[j + 1 - j0]);
        }
    }
    /* sources, grid.cu:742-811 */
    const series* hy = &s->ser[ORC_SER_HYETO];
    const series* dr = &s->ser[ORC_SER_DRAIN];
    const real five_thirds = (real)(5.f / 3.f);
    for (int j = j0; j < j1; ++j) {
        for (int i = i0; i < i1; ++i) {
            size_t q = IX(s, i, j);
            real S1 = (s->BX[IBX(s, i + 1, j)] - s->BX[IBX(s, i, j)]) / dx;
            real S2 = (s->BY[IBY(s, i, j + 1)] - s->BY[IBY(s, i, j)]) / dy;
            real hC = s->w[q] - s->BC[q];
            s->clamp[q] = (hC > EPS) ? 0 : 1;
            S1 *= -G_GRAV * hC;
            S2 *= -G_GRAV * hC;
            {
                real nn = s->n[q];
                real C = (real)-4.f * G_GRAV * nn * nn *
                         n_sqrt(m, s->hu[q] * s->hu[q] + s->hv[q] * s->hv[q]) *
                         n_pow(m, hC, five_thirds) /
                         n_pow(m, (hC * hC + n_fmax(m, hC * hC, kappa)), (real)2.f);
                s->G[q] = C;
            }
            real S0 = 0;
            if (s->flags & ORC_RAIN_AVG) S0 += (real)hy->interpolated;
            if (s->flags & ORC_RAIN_GRID) S0 += s->ms[ORC_MAP_RAIN].cur[q];
            sums->rain += (double)S0;
            if (s->flags & ORC_DRAIN_AVG) {
                real d = (real)dr->interpolated;
                real taken = (d > hC / dt) ? hC / dt : d;
                S0 -= taken;
                sums->drain += (double)taken;
            }
            if (s->flags & ORC_DRAIN_GRID) {
                real d = s->ms[ORC_MAP_DRAIN].cur[q];
                real taken = (d > hC / dt) ? hC / dt : d;
                S0 -= taken;
                sums->drain += (double)taken;
            }
            if ((s->flags & ORC_HYDROGRAPH) || s->nsrc > 0) {
                long cell = (long)j * nx + i;
                for (long k = 0; k < s->nsrc; ++k)
                    if (cell == s->src_idx[k]) {
                        S0 += (real)s->src_rate[k];
                        sums->src += s->src_rate[k];
                    }
            }
            if (s->flags & ORC_INFILTRATION) {
                real Kq = s->K[q], Fq = s->F[q];
                real p1 = Kq * dt - (real)2.f * Fq;
                real p2 = Kq * (Fq + (real)s->psi * (real)s->dtheta);
                real rate = (p1 + n_sqrt(m, p1 * p1 + (real)8.f * p2 * dt)) / ((real)2.f * dt);
                rate = (rate > hC / dt) ? hC / dt : rate;
                rate = (rate > (real)0.f) ? rate : (real)0.f;
                s->dF[q] = rate;
                S0 -= rate;
            }
            s->dw[q] += S0;
            s->dhu[q] += S1;
            s->dhv[q] += S2;
        }
    }
    s->mx[tile] = (double)smx;
    s->my[tile] = (double)smy;
}

void orc_compute_fluxes(orc_sim* s)
{
    const int nl = s->nlist;
    const double dt = s->dt;
    double r_rain = 0, r_src = 0, r_drain = 0, r_out = 0, r_clamp = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : r_rain, r_src, r_drain, r_out)
    for (int k = 0; k < nl; ++k) {
        stage_sums sm = {0, 0, 0, 0};
        flux_tile(s, s->list[k], dt, &sm);
        r_rain += sm.rain;
        r_src += sm.src;
        r_drain += sm.drain;
        r_out += sm.out;
    }
    /* The dry clamp w <- BC (grid.cu:748-749) is written after every tile has
     * read its stencil: the reference's cross-tile halo race is resolved as
     * "reads see the values from before the kernel". */
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : r_clamp)
    for (int k = 0; k < nl; ++k) {
        int tile = s->list[k];
        int bx = tile % s->nbx, by = tile / s->nbx;
        for (int ty = 0; ty < TILE_H; ++ty) {
            for (int tx = 0; tx < TILE_W; ++tx) {
                int i = bx * TILE_W + tx + 2, j = by * TILE_H + ty + 2;
                if (i > s->nx - 3 || j > s->ny - 3) continue;
                size_t q = IX(s, i, j);
                if (s->clamp[q]) {
                    r_clamp += (double)(s->BC[q] - s->w[q]);
                    s->w[q] = s->BC[q];
                }
            }
        }
    }
    s->st_rain = r_rain;
    s->st_src = r_src;
    s->st_drain = r_drain;
    s->st_out = r_out;
    s->st_clamp = r_clamp;
}

/* Books the last flux pass into the budget with the weight its rates carry
 * in the stage's update (dt for Euler's single stage, dt/2 for each Heun
 * stage) and its clamp change with clamp_weight. */
static void book_stage(orc_sim* s, double rate_weight, double clamp_weight, double exch)
{
    const double area = s->dx * s->dy;
    s->b_rain += rate_weight * s->st_rain * area;
    s->b_src += rate_weight * s->st_src * area;
    s->b_drain += rate_weight * s->st_drain * area;
    s->b_out += rate_weight * s->st_out * area;
    s->b_exch += (clamp_weight * s->st_clamp + exch) * area;
    s->st_rain = s->st_src = s->st_drain = s->st_out = s->st_clamp = 0;
}

/* ComputeTimestep, simulator.cu:540-566: max over every tile's (possibly
 * stale) speed, float fmaxf on the host, dt = 1e-6f on the first step. */
void orc_compute_timestep(orc_sim* s)
{
    double max_x = s->mx[0], max_y = s->my[0];
    for (int t = 1; t < s->ntiles; ++t) {
        if (s->mx[t] > max_x) max_x = s->mx[t];
        if (s->my[t] > max_y) max_y = s->my[t];
    }
    double c_x = s->dx / (double)fmaxf((float)(max_x / s->courant), (float)s->kappa);
    double c_y = s->dy / (double)fmaxf((float)(max_y / s->courant), (float)s->kappa);
    s->dt = (c_x < c_y) ? c_x : c_y;
    if (s->t == 0.f) s->dt = (double)0.000001f;
    s->t += s->dt;
}

/* Per-cell wet rule, grid.cu:1125-1145 (and 991-998 for surge in Euler). */
static int cell_wet(const orc_sim* s, int i, int j, real hC, int with_surge)
{
    int wet = (hC > EPS) ? 1 : 0;
    if ((s->flags & ORC_HYDROGRAPH) || s->nsrc > 0) {
        long cell = (long)j * s->nx + i;
        for (long k = 0; k < s->nsrc; ++k)
            if (cell == s->src_idx[k]) wet = 1;
    }
    if ((s->flags & ORC_RAIN_AVG) && s->ser[ORC_SER_HYETO].interpolated > 0.f) wet = 1;
    if ((s->flags & ORC_RAIN_GRID) && s->ms[ORC_MAP_RAIN].cur[IX(s, i, j)] > 0) wet = 1;
    if (with_surge && (s->flags & ORC_SURGE_GRID)) {
        real v = s->ms[ORC_MAP_SURGE].cur[IX(s, i, j)];
        if (v != (real)s->nodata && v > 0) wet = 1;
    }
    return wet;
}

/* Tile wet flag: only thread (0,0) writes it, so tiles whose origin cell is
 * outside the interior keep their old flag (grid.cu:1013-1022, 1160-1170). */
static void tile_wet_flag(orc_sim* s, int tile, int any)
{
    int bx = tile % s->nbx, by = tile / s->nbx;
    if (bx * TILE_W + 2 > s->nx - 3 || by * TILE_H + 2 > s->ny - 3) return;
    s->wet[tile] = (unsigned char)(((s->flags & ORC_TILE_ACCEL) ? 0 : 1) | any);
}

static void diagnostics(orc_sim* s, size_t q, real hC, int euler)
{
    const int m = s->mode;
    const real kappa = (real)s->kappa;
    const real t = (real)s->t;
    if (s->flags & ORC_SAVE_MAX) {
        real uC = velocity_r(m, hC, s->hu[q], kappa);
        real vC = velocity_r(m, hC, s->hv[q], kappa);
        if (euler) {
            s->t_peak[q] = (hC > s->h_max[q]) ? t : s->t_peak[q];
            s->t_dry[q] = (hC > (real)0.0762) ? t : s->t_dry[q];
        }
        s->h_max[q] = n_fmax(m, s->h_max[q], hC);
        s->q_max[q] = n_fmax(m, s->q_max[q], hC * n_sqrt(m, vC * vC + uC * uC));
    }
    s->hdiag[q] = hC;
    {
        real uC = velocity_r(m, hC, s->hu[q], kappa);
        real vC = velocity_r(m, hC, s->hv[q], kappa);
        s->qdiag[q] = hC * n_sqrt(m, vC * vC + uC * uC);
    }
    if (s->flags & ORC_SAVE_ARRIVAL) {
        if (s->t_wet[q] < (real)0.f && hC > EPS) s->t_wet[q] = t;
    }
}

/* Integrate_1_k, grid.cu:856-1024 */
void orc_integrate_1(orc_sim* s)
{
    const real dt = (real)s->dt;
    const int euler = (s->flags & ORC_EULER) ? 1 : 0;
    const real nd = (real)s->nodata;
    double exch = 0;   /* surge override and dry clamp, m summed over cells */
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : exch)
    for (int k = 0; k < s->nlist; ++k) {
        int tile = s->list[k];
        int bx = tile % s->nbx, by = tile / s->nbx, any = 0;
        for (int ty = 0; ty < TILE_H; ++ty) {
            for (int tx = 0; tx < TILE_W; ++tx) {
                int i = bx * TILE_W + tx + 2, j = by * TILE_H + ty + 2;
                if (i < 2 || j < 2 || i > s->nx - 3 || j > s->ny - 3) continue;
                size_t q = IX(s, i, j);
                s->w_old[q] = s->w[q];
                s->hu_old[q] = s->hu[q];
                s->hv_old[q] = s->hv[q];
                if (s->flags & ORC_SURGE_GRID) {
                    real sg = s->ms[ORC_MAP_SURGE].cur[q];
                    if (sg != nd && sg > (real)0.0) {
                        real next_h = s->w[q] + s->dw[q] * dt - s->BC[q];
                        real deltaw = sg - next_h;
                        real before = s->dw[q];
                        s->dw[q] = deltaw / dt;
                        exch += (double)((s->dw[q] - before) * dt);
                    }
                }
                s->w[q] += s->dw[q] * dt;
                s->hu[q] += s->dhu[q] * dt;
                s->hu[q] /= ((real)1.0f - dt * s->G[q]);
                s->hv[q] += s->dhv[q] * dt;
                s->hv[q] /= ((real)1.0f - dt * s->G[q]);
                real hC = (s->w[q] - s->BC[q] > EPS) ? s->w[q] - s->BC[q] : (real)0.f;
                if (!(hC > EPS)) exch += (double)(s->BC[q] - s->w[q]);
                s->w[q] = (hC > EPS) ? s->w[q] : s->BC[q];
                if (s->flags & ORC_INFILTRATION) {
                    s->F_old[q] = s->F[q];
                    s->F[q] += s->dF[q] * dt;
                }
                if (euler) {
                    diagnostics(s, q, hC, 1);
                    any |= cell_wet(s, i, j, s->w[q] - s->BC[q], 1);
                }
            }
        }
        if (euler) tile_wet_flag(s, tile, any);
    }
    /* the stage-1 flux pass's clamp persists in w_old (weight 1); the rates
     * and this stage's exchange enter U^{n+1} with weight 1/2 (Heun) or 1 */
    const double wgt = euler ? 1.0 : 0.5;
    book_stage(s, wgt * s->dt, 1.0, wgt * exch);
}

/* Integrate_2_k, grid.cu:1043-1171 */
void orc_integrate_2(orc_sim* s)
{
    const real dt = (real)s->dt;
    double exch = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : exch)
    for (int k = 0; k < s->nlist; ++k) {
        int tile = s->list[k];
        int bx = tile % s->nbx, by = tile / s->nbx, any = 0;
        for (int ty = 0; ty < TILE_H; ++ty) {
            for (int tx = 0; tx < TILE_W; ++tx) {
                int i = bx * TILE_W + tx + 2, j = by * TILE_H + ty + 2;
                if (i > s->nx - 3 || j > s->ny - 3) continue;
                size_t q = IX(s, i, j);
                s->w[q] = (real)0.5f * (s->w_old[q] + s->w[q] + s->dw[q] * dt);
                s->hu[q] = (real)0.5f * (s->hu_old[q] + s->hu[q] + s->dhu[q] * dt);
                s->hv[q] = (real)0.5f * (s->hv_old[q] + s->hv[q] + s->dhv[q] * dt);
                real hC = (s->w[q] - s->BC[q] > EPS) ? s->w[q] - s->BC[q] : (real)0.f;
                if (!(hC > EPS)) exch += (double)(s->BC[q] - s->w[q]);
                s->w[q] = (hC > EPS) ? s->w[q] : s->BC[q];
                if (s->flags & ORC_INFILTRATION)
                    s->F[q] = (real)0.5f * (s->F_old[q] + s->F[q] + s->dF[q] * dt);
                diagnostics(s, q, hC, 0);
                any |= cell_wet(s, i, j, hC, 0);
            }
        }
        tile_wet_flag(s, tile, any);
    }
    /* the stage-2 flux pass's clamp acts on U* (weight 1/2) */
    book_stage(s, 0.5 * s->dt, 0.5, exch);
}

void orc_update_maps(orc_sim* s)
{
    if (s->flags & ORC_RAIN_GRID) {
        ms_update(s, ORC_MAP_RAIN, s->t);
        s->v_rain += s->ms[ORC_MAP_RAIN].sum_cache * s->dt * s->dx * s->dy;
    }
    if (s->flags & ORC_DRAIN_GRID) ms_update(s, ORC_MAP_DRAIN, s->t);
    if (s->flags & ORC_SURGE_GRID) ms_update(s, ORC_MAP_SURGE, s->t);
}

/* RunSimulation, simulator.cu:773-853 */
void orc_step(orc_sim* s)
{
    if (s->flags & (ORC_HYDROGRAPH | ORC_RAIN_AVG | ORC_DRAIN_AVG)) orc_update_source(s);
    orc_grow(s);
    orc_compute_fluxes(s);
    orc_compute_timestep(s);
    orc_integrate_1(s);
    orc_apply_boundaries(s);
    if (!(s->flags & ORC_EULER)) {
        orc_compute_fluxes(s);
        orc_integrate_2(s);
        orc_apply_boundaries(s);
    }
    orc_update_maps(s);
}

void orc_run(orc_sim* s, int nsteps)
{
    for (int k = 0; k < nsteps; ++k) orc_step(s);
}

/* --- construction: InitSimulation (simulator.cu:267-493) + InitGrid_k ------ */
orc_sim* orc_create(const orc_params* p, const double* dem, const double* h0c,
                    const double* nc, const double* Kc)
{
    orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
    s->b_nx = p->b_nx; s->b_ny = p->b_ny;
    s->nx = p->b_nx + 3; s->ny = p->b_ny + 3;          /* grid_config.cpp:84-85 */
    s->nbx = (int)ceil((double)s->nx / (double)TILE_W); /* grid.cu:336-337 */
    s->nby = (int)ceil((double)s->ny / (double)TILE_H);
    s->ntiles = s->nbx * s->nby;
    s->dx = p->dx; s->dy = p->dy; s->kappa = p->kappa; s->courant = p->courant;
    s->psi = p->psi; s->dtheta = p->dtheta; s->nodata = p->nodata;
    s->flags = p->flags; s->mode = p->mode;
    const int nx = s->nx, ny = s->ny, bnx = s->b_nx;
    size_t N = (size_t)nx * (size_t)ny;
    real** fields[] = {&s->w, &s->hu, &s->hv, &s->w_old, &s->hu_old, &s->hv_old, &s->dw,
                       &s->dhu, &s->dhv, &s->G, &s->BC, &s->n, &s->F, &s->F_old, &s->dF,
                       &s->K, &s->hdiag, &s->qdiag, &s->h_max, &s->q_max, &s->t_wet,
                       &s->t_peak, &s->t_dry};
    for (size_t k = 0; k < sizeof fields / sizeof fields[0]; ++k) *fields[k] = zalloc(N);
    s->BX = zalloc((size_t)(nx + 1) * (size_t)ny);
    s->BY = zalloc((size_t)nx * (size_t)(ny + 1));
    s->clamp = (unsigned char*)calloc(N, 1);
    s->wet = (unsigned char*)calloc((size_t)s->ntiles, 1);
    s->aval = (int*)malloc(sizeof(int) * (size_t)s->ntiles);
    s->list = (int*)malloc(sizeof(int) * (size_t)s->ntiles);
    s->mx = (double*)calloc((size_t)s->ntiles, sizeof(double));
    s->my = (double*)calloc((size_t)s->ntiles, sizeof(double));

    /* simulator.cu:384-431: face beds from two-point means of the raw DEM with
     * edge extension; h0, n, K from four-corner means; K in mm/hr -> m/s. */
    real* h0 = zalloc(N);
    for (int j = 2; j < ny - 2; ++j) {
        for (int i = 2; i < nx - 2; ++i) {
            int jt = j - 2, it = i - 2;
            size_t ll = (size_t)jt * bnx + it, ul = (size_t)(jt + 1) * bnx + it;
            size_t ur = (size_t)(jt + 1) * bnx + (it + 1), lr = (size_t)jt * bnx + (it + 1);
            s->BX[IBX(s, i + 1, j)] = (real)(0.5f * (dem[ur] + dem[lr]));
            s->BX[IBX(s, i, j)] = (real)(0.5f * (dem[ll] + dem[ul]));
            s->BY[IBY(s, i, j + 1)] = (real)(0.5f * (dem[ul] + dem[ur]));
            s->BY[IBY(s, i, j)] = (real)(0.5f * (dem[ll] + dem[lr]));
            if (i == 2) {
                s->BX[IBX(s, i - 1, j)] = s->BX[IBX(s, i + 1, j)];
                s->BX[IBX(s, i - 2, j)] = s->BX[IBX(s, i, j)];
            }
            if (i == nx - 3) s->BX[IBX(s, i + 2, j)] = s->BX[IBX(s, i, j)];
            if (j == 2) {
                s->BY[IBY(s, i, j - 1)] = s->BY[IBY(s, i, j + 1)];
                s->BY[IBY(s, i, j - 2)] = s->BY[IBY(s, i, j)];
            }
            if (j == ny - 3) s->BY[IBY(s, i, j + 2)] = s->BY[IBY(s, i, j)];
            if (p->flags & ORC_H_INIT)
                h0[IX(s, i, j)] = (real)(0.25f * (h0c[ur] + h0c[lr] + h0c[ll] + h0c[ul]));
            if (nc) s->n[IX(s, i, j)] = (real)(0.25f * (nc[ur] + nc[lr] + nc[ll] + nc[ul]));
            else s->n[IX(s, i, j)] = (real)p->n_const;
            if (p->flags & ORC_INFILTRATION) {
                double k = 0.25f * (Kc[ur] + Kc[lr] + Kc[ll] + Kc[ul]);
                k /= (3600.f * 1000.f);
                s->K[IX(s, i, j)] = (real)k;
            }
        }
    }
    /* InitGrid_k, grid.cu:506-553: cells i in [2,nx-1], j in [2,ny-1]. */
    for (int j = 2; j < ny; ++j) {
        for (int i = 2; i < nx; ++i) {
            size_t q = IX(s, i, j);
            real BE = s->BX[IBX(s, i + 1, j)], BW = s->BX[IBX(s, i, j)];
            real BN = s->BY[IBY(s, i, j + 1)], BS = s->BY[IBY(s, i, j)];
            s->BC[q] = (real)0.25f * (BE + BW + BN + BS);
            s->w[q] = s->w_old[q] = s->BC[q];
            if (p->flags & ORC_H_INIT) { s->w[q] += h0[q]; s->w_old[q] += h0[q]; }
            if (p->flags & ORC_SAVE_ARRIVAL) s->t_wet[q] = (real)-9999.f;
        }
    }
    free(h0);
    for (int by = 0; by < s->nby; ++by) {
        for (int bx = 0; bx < s->nbx; ++bx) {
            int t = by * s->nbx + bx;
            if (bx * TILE_W + 2 < nx && by * TILE_H + 2 < ny) { s->wet[t] = 1; s->aval[t] = t; }
            else { s->wet[t] = 0; s->aval[t] = -1; }
        }
    }
    rebuild_list(s);
    s->t = p->t0;
    s->dt = (double)0.0000001f;                          /* simulator.cu:35 */
    orc_apply_boundaries(s);                              /* simulator.cu:731 */
    orc_budget_reset(s);
    return s;
}

void orc_destroy(orc_sim* s)
{
    if (!s) return;
    real* fields[] = {s->w, s->hu, s->hv, s->w_old, s->hu_old, s->hv_old, s->dw, s->dhu,
                      s->dhv, s->G, s->BC, s->n, s->F, s->F_old, s->dF, s->K, s->hdiag,
                      s->qdiag, s->h_max, s->q_max, s->t_wet, s->t_peak, s->t_dry, s->BX, s->BY};
    for (size_t k = 0; k < sizeof fields / sizeof fields[0]; ++k) free(fields[k]);
    free(s->clamp); free(s->wet); free(s->aval); free(s->list); free(s->mx); free(s->my);
    for (int k = 0; k < 3; ++k) {
        free(s->ser[k].time); free(s->ser[k].rate);
        free(s->ms[k].maps); free(s->ms[k].cur); free(s->ms[k].p0);
        free(s->ms[k].p1); free(s->ms[k].ibuf);
    }
    free(s->src_idx); free(s->src_rate);
    free(s);
}

/* --- accessors --------------------------------------------------------------- */
static real* field_ptr(const orc_sim* s, int f)
{
    switch (f) {
    case ORC_W: return s->w;
    case ORC_HU: return s->hu;
    case ORC_HV: return s->hv;
    case ORC_BC: return s->BC;
    case ORC_N: return s->n;
    case ORC_K: return s->K;
    case ORC_F: return s->F;
    case ORC_DW: return s->dw;
    case ORC_DHU: return s->dhu;
    case ORC_DHV: return s->dhv;
    case ORC_G: return s->G;
    case ORC_DF: return s->dF;
    case ORC_HMAX: return s->h_max;
    case ORC_QMAX: return s->q_max;
    case ORC_TWET: return s->t_wet;
    case ORC_TPEAK: return s->t_peak;
    case ORC_TDRY: return s->t_dry;
    case ORC_SURGE: return s->ms[ORC_MAP_SURGE].cur;
    case ORC_RAINMAP: return s->ms[ORC_MAP_RAIN].cur;
    case ORC_DRAINMAP: return s->ms[ORC_MAP_DRAIN].cur;
    case ORC_QDIAG: return s->qdiag;
    default: return NULL;
    }
}

int orc_get_field(const orc_sim* s, int f, double* out)
{
    size_t N = (size_t)s->nx * (size_t)s->ny;
    if (f == ORC_DEPTH) {
        for (size_t k = 0; k < N; ++k) out[k] = (double)(s->w[k] - s->BC[k]);
        return 0;
    }
    real* p = field_ptr(s, f);
    if (!p) return -1;
    for (size_t k = 0; k < N; ++k) out[k] = (double)p[k];
    return 0;
}

int orc_set_field(orc_sim* s, int f, const double* in)
{
    size_t N = (size_t)s->nx * (size_t)s->ny;
    real* p = field_ptr(s, f);
    if (!p) return -1;
    for (size_t k = 0; k < N; ++k) p[k] = (real)in[k];
    return 0;
}

void orc_get_faces(const orc_sim* s, double* bx, double* by)
{
    size_t nbx = (size_t)(s->nx + 1) * (size_t)s->ny, nby = (size_t)s->nx * (size_t)(s->ny + 1);
    for (size_t k = 0; k < nbx; ++k) bx[k] = (double)s->BX[k];
    for (size_t k = 0; k < nby; ++k) by[k] = (double)s->BY[k];
}

void orc_get_time(const orc_sim* s, double* t, double* dt) { *t = s->t; *dt = s->dt; }
void orc_set_time(orc_sim* s, double t, double dt) { s->t = t; s->dt = dt; }

void orc_get_tiles(const orc_sim* s, unsigned char* wet, int* active, int* n_active)
{
    memcpy(wet, s->wet, (size_t)s->ntiles);
    memcpy(active, s->aval, sizeof(int) * (size_t)s->ntiles);
    *n_active = s->nlist;
}

void orc_set_tiles(orc_sim* s, const unsigned char* wet, const int* active)
{
    memcpy(s->wet, wet, (size_t)s->ntiles);
    memcpy(s->aval, active, sizeof(int) * (size_t)s->ntiles);
    rebuild_list(s);
}

void orc_get_tile_speeds(const orc_sim* s, double* mx, double* my)
{
    memcpy(mx, s->mx, sizeof(double) * (size_t)s->ntiles);
    memcpy(my, s->my, sizeof(double) * (size_t)s->ntiles);
}

void orc_get_rates(const orc_sim* s, float* out)
{
    out[0] = s->ser[ORC_SER_HYETO].interpolated;
    out[1] = s->ser[ORC_SER_DRAIN].interpolated;
    out[2] = s->ser[ORC_SER_HYDRO].interpolated;
}

static double total_volume(const orc_sim* s)
{
    double v = 0.0;
    for (int j = 2; j <= s->ny - 3; ++j)
        for (int i = 2; i <= s->nx - 3; ++i) {
            size_t q = IX(s, i, j);
            v += (double)(s->w[q] - s->BC[q]);
            if (s->flags & ORC_INFILTRATION) v += (double)s->F[q];
        }
    return v * s->dx * s->dy;
}

void orc_budget_reset(orc_sim* s)
{
    s->b_v0 = total_volume(s);
    s->b_rain = s->b_src = s->b_drain = s->b_out = s->b_exch = 0;
}

void orc_budget(const orc_sim* s, double* out)
{
    double m[4];
    orc_mass(s, m);
    const double total = m[0] + ((s->flags & ORC_INFILTRATION) ? m[1] : 0.0);
    out[0] = m[0];
    out[1] = m[1];
    out[2] = s->b_v0;
    out[3] = s->b_rain;
    out[4] = s->b_src;
    out[5] = s->b_drain;
    out[6] = s->b_out;
    out[7] = s->b_exch;
    out[8] = (total - s->b_v0) - (s->b_rain + s->b_src - s->b_drain - s->b_out + s->b_exch);
}

void orc_mass(const orc_sim* s, double* out)
{
    double vh = 0.0, vf = 0.0;
    for (int j = 2; j <= s->ny - 3; ++j) {
        for (int i = 2; i <= s->nx - 3; ++i) {
            size_t q = IX(s, i, j);
            vh += (double)(s->w[q] - s->BC[q]);
            vf += (double)s->F[q];
        }
    }
    out[0] = vh * s->dx * s->dy;
    out[1] = vf * s->dx * s->dy;
    out[2] = s->v_rain;
    out[3] = s->v_src;
}
