// kp_kernels.hip -- gfx950 kernels of the Kurganov-Petrova step.
//
// One step (Heun), all on one stream, no host round trip:
//   begin_step   forcing rates at (float)t                         (1 lane)
//   [track]      gridded rain/surge wakes dry inactive tiles        (maps only)
//   grow         counter reset, tile dilation + ballot/prefix-sum
//                compaction -> list, count
//   flux(1)      fused flux + sources per active tile -> dU, G, dF, mx/my
//   timestep     shuffle max over all tile speeds -> dt, t           (last block)
//   integrate1   U* = U + dt dU with semi-implicit friction, surge, dry clamp
//   boundaries   critical-depth outflow ghosts
//   flux(2), integrate2 (Heun average, diagnostics, tile wet flags), boundaries
// The flux/integrate kernels loop over the device-side active list
// (grid-stride), so launch sizes never depend on a host copy of the count.
#include <hip/hip_runtime.h>

#include "kp_device.hpp"
#include "kp_kernels.hpp"

#include "../../include/rift_kp.h"

namespace rift {

__global__ void k_reset_counters(Dev d)
{
    if (threadIdx.x != 0) return;
    d.sc->n_active = 0;
    d.sc->n_units = 0;
    d.sc->n_quads = 0;
    d.sc->n_part = 0;
    d.sc->wet_cells = 0ull;
}

__global__ void k_begin_step(Dev d)
{
    if (threadIdx.x != 0) return;
    Scalars* sc = d.sc;
    sc->n_active = 0;
    sc->n_units = 0;
    sc->n_units_acc = 0;
    sc->n_quads = sc->n_quads_acc = 0;
    sc->n_part = sc->n_part_acc = 0;
    sc->wet_cells = 0ull;
    sc->next_valid = 0;
    if (!(d.flags & (RIFT_HYDROGRAPH | RIFT_RAIN_AVG | RIFT_DRAIN_AVG))) return;
    const Forcing f = forcing_at(d, sc);
    sc->rain_rate = f.rain;
    sc->drain_rate = f.drain;
    sc->hydro_rate = f.hydro;
    sc->src_cell_rate = f.cell;
}

// ---------------------------------------------------------------------------
// TrackRainfall_k / TrackSurge_k (grid.cu:1224-1276): a dry tile that was not
// active in the last step becomes wet when gridded rain > epsilon or surge > 0
// in its cells.  Rows use the reference's BLOCK_COLS stride (by*16).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_track(Dev d)
{
    const int tid = threadIdx.x;
    // a surge series on the device is tracked by k_track_series
    const bool rain = d.flags & RIFT_RAIN_GRID, surge = (d.flags & RIFT_SURGE_GRID) && !d.sg_series;
    const float* rmap = d.map[RIFT_MAP_RAIN][0];
    const float* smap = d.map[RIFT_MAP_SURGE][0];
    for (int t = blockIdx.x; t < d.ntiles; t += gridDim.x) {
        if (d.act[t] || d.wet[t]) continue;   // uniform per block
        const int bx = t % d.nbx, by = t / d.nbx;
        bool hit = false;
        if (tid < kTileW * kTileH) {
            int i = bx * kTileW + (tid % kTileW) + 2;
            int j = by * kTileW + (tid / kTileW) + 2;
            if (i <= d.nx - 3 && j <= d.ny - 3) {
                size_t q = cell(d, i, j);
                if (rain && rmap[q] > kEps) hit = true;
                if (surge) {
                    float v = smap[q];
                    if (v != d.nodata && v > 0.f) hit = true;
                }
            }
        }
        if (__syncthreads_or(hit) && tid == 0) d.wet[t] = 1;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Device surge series.  A slot's tile flags say where its map can wake a
// tile or override a cell at all: InterpGrid_k's value is a convex
// combination of the two planes (or one of them), and f*(x1-x0)+x0 rounds to
// a value <= 0 whenever neither plane holds a value > 0 at the cell (both
// planes <= 0: |f (x1-x0)| <= |x0| after rounding, which is monotone).  So a
// tile whose flags are clear in both slots of a window has no cell with
// surge > 0, and skipping it is exact.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sg_flags(Dev d, int slot)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.ntiles) return;
    const float* m = d.sgs + (size_t)slot * (size_t)d.soff;
    const int bx = t % d.nbx, by = t / d.nbx;
    const float nd = d.nodata;
    unsigned char bits = 0;
    const int i0 = bx * kTileW + 2, i1 = min(i0 + kTileW - 1, d.nx - 3);
    // TrackSurge_k's window: rows by*BLOCK_COLS + ty + 2, ty < BLOCK_ROWS
    for (int j = by * kTileW + 2; j <= min(by * kTileW + kTileH + 1, d.ny - 3); ++j)
        for (int i = i0; i <= i1; ++i) {
            const float v = m[cell(d, i, j)];
            if (v != nd && v > 0.f) bits |= 1;
        }
    // the tile's own cells (the surge override of Integrate_1_k)
    for (int j = by * kTileH + 2; j <= min(by * kTileH + kTileH + 1, d.ny - 3); ++j)
        for (int i = i0; i <= i1; ++i) {
            const float v = m[cell(d, i, j)];
            if (v != nd && v > 0.f) bits |= 2;
        }
    d.sgflag[(size_t)slot * d.ntiles + t] = bits;
}

// TrackSurge_k (grid.cu:1250-1276) over the device series: one lane per
// tile; a dry tile that was inactive in the last Grow and whose slots can
// hold surge > 0 in its Track window evaluates the window cell by cell.
__global__ __launch_bounds__(256) void k_track_series(Dev d)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.ntiles || d.act[t] || d.wet[t]) return;
    const SurgeWindow w = surge_window_now(d.sc);
    const int s0 = sg_slot(d, w.k0), s1 = sg_slot(d, w.k1);
    if (!((d.sgflag[(size_t)s0 * d.ntiles + t] | d.sgflag[(size_t)s1 * d.ntiles + t]) & 1)) return;
    const float* m0 = d.sgs + (size_t)s0 * (size_t)d.soff;
    const float* m1 = d.sgs + (size_t)s1 * (size_t)d.soff;
    const int bx = t % d.nbx, by = t / d.nbx;
    const int i0 = bx * kTileW + 2, i1 = min(i0 + kTileW - 1, d.nx - 3);
    for (int j = by * kTileW + 2; j <= min(by * kTileW + kTileH + 1, d.ny - 3); ++j)
        for (int i = i0; i <= i1; ++i) {
            const size_t q = cell(d, i, j);
            const float v = surge_interp(m0[q], m1[q], w.f, d.nodata);
            if (v != d.nodata && v > 0.f) {
                d.wet[t] = 1;
                return;
            }
        }
}

// The window in force written into surge plane 0 over i in [2, nx-1],
// j in [2, ny-1] (InterpGrid_k's range): what the per-phase kernels and the
// field downloads read.
__global__ __launch_bounds__(256) void k_sg_materialize(Dev d)
{
    const int i = blockIdx.x * 16 + (threadIdx.x & 15) + 2;
    const int j = blockIdx.y * 16 + (threadIdx.x >> 4) + 2;
    if (i >= d.nx || j >= d.ny) return;
    const SurgeWindow w = surge_window_now(d.sc);
    const size_t q = cell(d, i, j);
    const float* m0 = d.sgs + (size_t)sg_slot(d, w.k0) * (size_t)d.soff;
    const float* m1 = d.sgs + (size_t)sg_slot(d, w.k1) * (size_t)d.soff;
    d.map[RIFT_MAP_SURGE][0][q] = surge_interp(m0[q], m1[q], w.f, d.nodata);
}

// ---------------------------------------------------------------------------
// Grow (grid.cu:1188-1222, 1279-1322): a tile is active when it or an
// edge neighbour is wet (tile 0 always: -0 == 0 passes the >= 0 test).  The
// active list is compacted with a wavefront ballot and one atomic per wave;
// its order is irrelevant because tiles are independent.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_grow(Dev d)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    bool a = false;
    unsigned long long cells = 0ull;
    if (t < d.ntiles) {
        const int bx = t % d.nbx, by = t / d.nbx, nbx = d.nbx;
        const unsigned char* w = d.wet;
        a = w[t] || (bx > 0 && w[t - 1]) || (bx < nbx - 1 && w[t + 1]) ||
            (by > 0 && w[t - nbx]) || (by < d.nby - 1 && w[t + nbx]) || t == 0;
        d.act[t] = a ? 1 : 0;
        if (a) {
            int i0 = bx * kTileW + 2, j0 = by * kTileH + 2;
            int ni = min(i0 + kTileW, d.nx - 2) - i0, nj = min(j0 + kTileH, d.ny - 2) - j0;
            if (ni > 0 && nj > 0) cells = (unsigned long long)(ni * nj);
        }
    }
    const int lane = threadIdx.x & 63;
    unsigned long long m = __ballot(a);
    int pre = __popcll(m & ((1ull << lane) - 1ull));
    int cnt = __popcll(m);
    for (int off = 32; off > 0; off >>= 1) cells += __shfl_xor(cells, off);
    int base = 0;
    if (lane == 0 && cnt) {
        base = atomicAdd(&d.sc->n_active, cnt);
        atomicAdd(&d.sc->wet_cells, cells);
        atomicAdd(&d.sc->wcu_total, cells);
    }
    base = __shfl(base, 0);
    if (a) d.list[base + pre] = t;
}

// ---------------------------------------------------------------------------
// Fused flux + sources for one 16 x 14 tile per workgroup iteration
// (ComputeFluxes_k, grid.cu:564-839).  The 20 x 18 window of depth, momenta
// and cell bed is staged in LDS; x face states, x fluxes, y face states and y
// fluxes are each computed once per face by one lane.
// ---------------------------------------------------------------------------
constexpr int WX = kTileW + 4, WY = kTileH + 4;   // 20 x 18 window
constexpr int NXC = kTileW + 2;                    // x constructs per row (18)
constexpr int NXF = kTileW + 1;                    // x faces per row (17)
constexpr int NYF = kTileH + 1;                    // y faces per column (15)

template <bool Textbook>
__global__ __launch_bounds__(256) void k_flux(Dev d)
{
    __shared__ float sH[WX * WY], sX[WX * WY], sY[WX * WY], sB[WX * WY];
    __shared__ float sSt[256 * 6];
    __shared__ float sFx[kTileH * NXF * 3], sSx[kTileH * NXF];
    __shared__ float sFy[NYF * kTileW * 3], sSy[NYF * kTileW];
    __shared__ float sRed[2][4];

    const int tid = threadIdx.x;
    const int nact = d.sc->n_active;
    const float dt = (float)d.sc->dt;
    const float inv_dt = 1.f / dt;
    const float kappa = d.kappa;
    const unsigned fl = d.flags;
    const float rain_rate = d.sc->rain_rate, drain_rate = d.sc->drain_rate;
    const float src_rate = d.sc->src_cell_rate;

    for (int k = blockIdx.x; k < nact; k += gridDim.x) {
        const int tile = d.list[k];
        const int bx = tile % d.nbx, by = tile / d.nbx;
        const int i0 = bx * kTileW + 2, j0 = by * kTileH + 2;
        const size_t base = cell(d, i0 - 2, j0 - 2);

        for (int e = tid; e < WX * WY; e += 256) {
            int y = e / WX, x = e - y * WX;
            size_t q = base + (size_t)y * d.P + x;
            sH[e] = d.h[q];
            sX[e] = d.hu[q];
            sY[e] = d.hv[q];
            sB[e] = d.BC[q];
        }
        __syncthreads();

        // x face states of cells i0-1 .. i0+16 in rows j0 .. j0+13
        if (tid < kTileH * NXC) {
            int ry = tid / NXC, cx = tid - ry * NXC;
            int e = (ry + 2) * WX + (cx + 1);
            size_t qb = cell(d, i0 - 1 + cx, j0 + ry);
            FaceStates s = construct<Textbook>(sH[e - 1], sH[e], sH[e + 1], sB[e - 1], sB[e],
                                               sB[e + 1], sX[e - 1], sX[e], sX[e + 1],
                                               sY[e - 1], sY[e], sY[e + 1], d.BX[qb + 1], d.BX[qb]);
            float* o = &sSt[tid * 6];
            o[0] = s.hp; o[1] = s.qxp; o[2] = s.qyp; o[3] = s.hm; o[4] = s.qxm; o[5] = s.qym;
        }
        __syncthreads();
        // x faces i0 .. i0+16 (face f between cells f-1 and f)
        if (tid < kTileH * NXF) {
            int ry = tid / NXF, fx = tid - ry * NXF;
            const float* L = &sSt[(ry * NXC + fx) * 6];
            const float* R = &sSt[(ry * NXC + fx + 1) * 6];
            Flux f = face_flux<0>(L[0], L[1], L[2], R[3], R[4], R[5], kappa);
            sFx[tid * 3 + 0] = f.f0; sFx[tid * 3 + 1] = f.f1; sFx[tid * 3 + 2] = f.f2;
            sSx[tid] = f.speed;
        }
        __syncthreads();
        // y face states of rows j0-1 .. j0+14 in columns i0 .. i0+15
        {
            int rr = tid / kTileW, cc = tid - rr * kTileW;
            int e = (rr + 1) * WX + (cc + 2);
            size_t qb = cell(d, i0 + cc, j0 - 1 + rr);
            FaceStates s = construct<Textbook>(sH[e - WX], sH[e], sH[e + WX], sB[e - WX], sB[e],
                                               sB[e + WX], sX[e - WX], sX[e], sX[e + WX],
                                               sY[e - WX], sY[e], sY[e + WX], d.BY[qb + d.P], d.BY[qb]);
            float* o = &sSt[tid * 6];
            o[0] = s.hp; o[1] = s.qxp; o[2] = s.qyp; o[3] = s.hm; o[4] = s.qxm; o[5] = s.qym;
        }
        __syncthreads();
        // y faces j0 .. j0+14 (face g between rows g-1 and g)
        if (tid < NYF * kTileW) {
            int gy = tid / kTileW, cc = tid - gy * kTileW;
            const float* L = &sSt[(gy * kTileW + cc) * 6];
            const float* R = &sSt[((gy + 1) * kTileW + cc) * 6];
            Flux f = face_flux<1>(L[0], L[1], L[2], R[3], R[4], R[5], kappa);
            sFy[tid * 3 + 0] = f.f0; sFy[tid * 3 + 1] = f.f1; sFy[tid * 3 + 2] = f.f2;
            sSy[tid] = f.speed;
        }
        __syncthreads();

        float smx = 0.f, smy = 0.f;
        if (tid < kTileW * kTileH) {
            const int ty = tid / kTileW, tx = tid - ty * kTileW;
            const int i = i0 + tx, j = j0 + ty;
            if (i <= d.nx - 3 && j <= d.ny - 3) {
                const int xe = ty * NXF + tx + 1, xw = ty * NXF + tx;
                const int yn = (ty + 1) * kTileW + tx, ys = ty * kTileW + tx;
                smx = sSx[xe];
                smy = sSy[yn];
                const float dh = flux_divergence(sFx[xe * 3 + 0], sFx[xw * 3 + 0], sFy[yn * 3 + 0],
                                                 sFy[ys * 3 + 0], d.inv_dx, d.inv_dy);
                const float dhu = flux_divergence(sFx[xe * 3 + 1], sFx[xw * 3 + 1], sFy[yn * 3 + 1],
                                                  sFy[ys * 3 + 1], d.inv_dx, d.inv_dy);
                const float dhv = flux_divergence(sFx[xe * 3 + 2], sFx[xw * 3 + 2], sFy[yn * 3 + 2],
                                                  sFy[ys * 3 + 2], d.inv_dx, d.inv_dy);

                const int e = (ty + 2) * WX + (tx + 2);
                const size_t q = cell(d, i, j);
                const float hC = sH[e];                      // w - BC, unclamped
                const float S1 = -kG * hC * ((d.BX[q + 1] - d.BX[q]) * d.inv_dx);
                const float S2 = -kG * hC * ((d.BY[q + d.P] - d.BY[q]) * d.inv_dy);
                d.G[q] = manning_g(d.n[q], sX[e], sY[e], hC, kappa);
                float S0 = 0.f;
                if (fl & RIFT_RAIN_AVG) S0 += rain_rate;
                if (fl & RIFT_RAIN_GRID) S0 += d.map[RIFT_MAP_RAIN][0][q];
                if (fl & RIFT_DRAIN_AVG) {
                    float cap = hC * inv_dt;
                    S0 -= (drain_rate > cap) ? cap : drain_rate;
                }
                if (fl & RIFT_DRAIN_GRID) {
                    float cap = hC * inv_dt, dr = d.map[RIFT_MAP_DRAIN][0][q];
                    S0 -= (dr > cap) ? cap : dr;
                }
                if ((fl & RIFT_HYDROGRAPH) && i >= d.si0 && i <= d.si1 && j >= d.sj0 && j <= d.sj1)
                    S0 += src_rate;
                if (fl & RIFT_INFILTRATION) {
                    float f = green_ampt(d.K[q], d.F[q], d.psi_dtheta, hC, dt, inv_dt);
                    d.dF[q] = f;
                    S0 -= f;
                }
                d.dh[q] = dh + S0;
                d.dhu[q] = dhu + S1;
                d.dhv[q] = dhv + S2;
            }
        }
        // tile maxima of the east / north face speeds of interior cells
        for (int off = 32; off > 0; off >>= 1) {
            smx = fmaxf(smx, __shfl_xor(smx, off));
            smy = fmaxf(smy, __shfl_xor(smy, off));
        }
        if ((tid & 63) == 0) { sRed[0][tid >> 6] = smx; sRed[1][tid >> 6] = smy; }
        __syncthreads();
        if (tid == 0) {
            d.mx[tile] = fmaxf(fmaxf(sRed[0][0], sRed[0][1]), fmaxf(sRed[0][2], sRed[0][3]));
            d.my[tile] = fmaxf(fmaxf(sRed[1][0], sRed[1][1]), fmaxf(sRed[1][2], sRed[1][3]));
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// ComputeTimestep (simulator.cu:540-566) on the device: max over every tile's
// (possibly stale) speed by wavefront shuffles + one atomic per block; the
// last block forms dt exactly as the host does (double, float fmaxf), sets
// t += dt and resets the accumulators for the next step.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_timestep(Dev d)
{
    __shared__ float red[2][4];
    __shared__ bool last;
    float vx = 0.f, vy = 0.f;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < d.ntiles; t += gridDim.x * blockDim.x) {
        vx = fmaxf(vx, d.mx[t]);
        vy = fmaxf(vy, d.my[t]);
        // active tiles accumulate the next flux pass's maxima from zero
        if (d.act[t]) {
            d.mx[t] = 0.f;
            d.my[t] = 0.f;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        vx = fmaxf(vx, __shfl_xor(vx, off));
        vy = fmaxf(vy, __shfl_xor(vy, off));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][w] = vx; red[1][w] = vy; }
    __syncthreads();
    if (threadIdx.x == 0) {
        vx = fmaxf(fmaxf(red[0][0], red[0][1]), fmaxf(red[0][2], red[0][3]));
        vy = fmaxf(fmaxf(red[1][0], red[1][1]), fmaxf(red[1][2], red[1][3]));
        Scalars* sc = d.sc;
        atomicMax(&sc->mx_bits, __float_as_uint(vx));   // speeds are >= 0
        atomicMax(&sc->my_bits, __float_as_uint(vy));
        __threadfence();
        unsigned ticket = atomicAdd(&sc->ticket, 1u);
        last = (ticket == gridDim.x - 1);
    }
    __syncthreads();
    if (last && threadIdx.x == 0) {
        Scalars* sc = d.sc;
        __threadfence();
        const float vx_all = __uint_as_float(atomicAdd(&sc->mx_bits, 0u));
        const float vy_all = __uint_as_float(atomicAdd(&sc->my_bits, 0u));
        finish_timestep(d, vx_all, vy_all, sc);
        sc->mx_bits = 0u;
        sc->my_bits = 0u;
        sc->ticket = 0u;
    }
}

// ---------------------------------------------------------------------------
// Integrate_1_k (grid.cu:856-1024).  The dry clamp that ComputeFluxes_k
// writes (w <- BC when w - BC <= epsilon, grid.cu:748-749) is applied here to
// the value read, so the flux kernel never writes the state it reads.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool in_source(const Dev& d, int i, int j)
{
    return (d.flags & RIFT_HYDROGRAPH) && i >= d.si0 && i <= d.si1 && j >= d.sj0 && j <= d.sj1;
}

__device__ __forceinline__ bool origin_interior(const Dev& d, int bx, int by)
{
    return bx * kTileW + 2 <= d.nx - 3 && by * kTileH + 2 <= d.ny - 3;
}

__device__ __forceinline__ void diagnostics(const Dev& d, size_t q, float hC, float hu, float hv,
                                            float t, bool euler)
{
    if (d.flags & RIFT_SAVE_MAX) {
        float fv = vel_factor(hC, d.kappa);
        float uC = fv * hu, vC = fv * hv;
        float hm = d.hmax[q];
        if (euler) {
            d.tpeak[q] = (hC > hm) ? t : d.tpeak[q];
            if (hC > 0.0762f) d.tdry[q] = t;
        }
        d.hmax[q] = fmaxf(hm, hC);
        d.qmax[q] = fmaxf(d.qmax[q], hC * fsqrt(vC * vC + uC * uC));
    }
    if (d.flags & RIFT_SAVE_ARRIVAL) {
        float tw = d.twet[q];
        if (tw < 0.f && hC > kEps) d.twet[q] = t;
    }
}

__device__ __forceinline__ bool cell_wet(const Dev& d, int i, int j, size_t q, float hC,
                                         bool with_surge)
{
    bool w = hC > kEps;
    w |= in_source(d, i, j);
    if ((d.flags & RIFT_RAIN_AVG) && d.sc->rain_rate > 0.f) w = true;
    if ((d.flags & RIFT_RAIN_GRID) && d.map[RIFT_MAP_RAIN][0][q] > 0.f) w = true;
    if (with_surge && (d.flags & RIFT_SURGE_GRID)) {
        float v = d.map[RIFT_MAP_SURGE][0][q];
        if (v != d.nodata && v > 0.f) w = true;
    }
    return w;
}

__global__ __launch_bounds__(256) void k_integrate1(Dev d)
{
    const int tid = threadIdx.x;
    const int nact = d.sc->n_active;
    const float dt = (float)d.sc->dt, inv_dt = 1.f / dt;
    const float tnow = (float)d.sc->t;
    const bool euler = d.flags & RIFT_EULER;
    for (int k = blockIdx.x; k < nact; k += gridDim.x) {
        const int tile = d.list[k];
        const int bx = tile % d.nbx, by = tile / d.nbx;
        bool any = false;
        if (tid < kTileW * kTileH) {
            const int i = bx * kTileW + (tid % kTileW) + 2, j = by * kTileH + (tid / kTileW) + 2;
            if (i <= d.nx - 3 && j <= d.ny - 3) {
                const size_t q = cell(d, i, j);
                float h = d.h[q];
                h = (h > kEps) ? h : 0.f;
                float hu = d.hu[q], hv = d.hv[q];
                d.h0[q] = h; d.hu0[q] = hu; d.hv0[q] = hv;
                float dh = d.dh[q];
                if (d.flags & RIFT_SURGE_GRID) {
                    float sg = d.map[RIFT_MAP_SURGE][0][q];
                    if (sg != d.nodata && sg > 0.f) {
                        float next_h = fmaf(dh, dt, h);
                        dh = (sg - next_h) * inv_dt;
                        d.dh[q] = dh;
                    }
                }
                float hn = fmaf(dh, dt, h);
                const float rden = frcp(fmaf(-dt, d.G[q], 1.f));
                hu = fmaf(d.dhu[q], dt, hu) * rden;
                hv = fmaf(d.dhv[q], dt, hv) * rden;
                float hc = (hn > kEps) ? hn : 0.f;
                d.h[q] = hc; d.hu[q] = hu; d.hv[q] = hv;
                if (d.flags & RIFT_INFILTRATION) {
                    float F = d.F[q];
                    d.F0[q] = F;
                    d.F[q] = fmaf(d.dF[q], dt, F);
                }
                if (euler) {
                    diagnostics(d, q, hc, hu, hv, tnow, true);
                    any = cell_wet(d, i, j, q, hc, true);
                }
            }
        }
        if (euler) {
            bool a = __syncthreads_or(any);
            if (tid == 0 && origin_interior(d, bx, by))
                d.wet[tile] = (!(d.flags & RIFT_TILE_ACCEL) || a) ? 1 : 0;
        }
    }
}

// Integrate_2_k (grid.cu:1043-1171): Heun average, dry clamp, diagnostics and
// the tile wet flag (written only when the tile's origin cell is interior).
__global__ __launch_bounds__(256) void k_integrate2(Dev d)
{
    const int tid = threadIdx.x;
    const int nact = d.sc->n_active;
    const float dt = (float)d.sc->dt;
    const float tnow = (float)d.sc->t;
    for (int k = blockIdx.x; k < nact; k += gridDim.x) {
        const int tile = d.list[k];
        const int bx = tile % d.nbx, by = tile / d.nbx;
        bool any = false;
        if (tid < kTileW * kTileH) {
            const int i = bx * kTileW + (tid % kTileW) + 2, j = by * kTileH + (tid / kTileW) + 2;
            if (i <= d.nx - 3 && j <= d.ny - 3) {
                const size_t q = cell(d, i, j);
                float hs = d.h[q];
                hs = (hs > kEps) ? hs : 0.f;
                float hn = 0.5f * fmaf(d.dh[q], dt, d.h0[q] + hs);
                float hu = 0.5f * fmaf(d.dhu[q], dt, d.hu0[q] + d.hu[q]);
                float hv = 0.5f * fmaf(d.dhv[q], dt, d.hv0[q] + d.hv[q]);
                float hc = (hn > kEps) ? hn : 0.f;
                d.h[q] = hc; d.hu[q] = hu; d.hv[q] = hv;
                if (d.flags & RIFT_INFILTRATION)
                    d.F[q] = 0.5f * fmaf(d.dF[q], dt, d.F0[q] + d.F[q]);
                diagnostics(d, q, hc, hu, hv, tnow, false);
                any = cell_wet(d, i, j, q, hc, false);
            }
        }
        bool a = __syncthreads_or(any);
        if (tid == 0 && origin_interior(d, bx, by))
            d.wet[tile] = (!(d.flags & RIFT_TILE_ACCEL) || a) ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// ApplyBoundaries{X,Y}_k (grid.cu:1325-1406): critical-depth outflow.  Ghost
// cells keep the bed of the interior cell they copy, so the ghost depth is
// the boundary depth H itself (w_ghost = BC_int + H as in the reference).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_boundaries(Dev d)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = d.nx, ny = d.ny;
    if (k >= 2 && k <= ny - 3) {
        const int j = k;
        size_t c = cell(d, 2, j);
        float hu = d.hu[c], hv = d.hv[c];
        float H = fminf(critical_depth(hu), d.h[c]);
        H = (H > kEps) ? H : 0.f;
        d.h[cell(d, 0, j)] = H; d.h[cell(d, 1, j)] = H;
        d.hu[cell(d, 0, j)] = hu; d.hu[cell(d, 1, j)] = hu;
        d.hv[cell(d, 0, j)] = hv; d.hv[cell(d, 1, j)] = hv;
        c = cell(d, nx - 3, j);
        hu = d.hu[c]; hv = d.hv[c];
        H = fminf(critical_depth(hu), d.h[c]);
        H = (H > kEps) ? H : 0.f;
        d.h[cell(d, nx - 1, j)] = H; d.h[cell(d, nx - 2, j)] = H;
        d.hu[cell(d, nx - 1, j)] = hu; d.hu[cell(d, nx - 2, j)] = hu;
        d.hv[cell(d, nx - 1, j)] = hv; d.hv[cell(d, nx - 2, j)] = hv;
    }
    if (k >= 2 && k <= nx - 3) {
        const int i = k;
        size_t c = cell(d, i, 2);
        float hu = d.hu[c], hv = d.hv[c];
        float H = fminf(critical_depth(hv), d.h[c]);   // no epsilon clamp in y
        d.h[cell(d, i, 0)] = H; d.h[cell(d, i, 1)] = H;
        d.hu[cell(d, i, 0)] = hu; d.hu[cell(d, i, 1)] = hu;
        d.hv[cell(d, i, 0)] = hv; d.hv[cell(d, i, 1)] = hv;
        c = cell(d, i, ny - 3);
        hu = d.hu[c]; hv = d.hv[c];
        H = fminf(critical_depth(hv), d.h[c]);
        d.h[cell(d, i, ny - 1)] = H; d.h[cell(d, i, ny - 2)] = H;
        d.hu[cell(d, i, ny - 1)] = hu; d.hu[cell(d, i, ny - 2)] = hu;
        d.hv[cell(d, i, ny - 1)] = hv; d.hv[cell(d, i, ny - 2)] = hv;
    }
}

// InterpGrid_k (grid.cu:292-311): plane0 = f*(p2-p1)+p1 with the nodata rule,
// over i in [2, nx-1], j in [2, ny-1].
__global__ __launch_bounds__(256) void k_interp_map(Dev d, int which, float factor)
{
    const int i = blockIdx.x * 16 + (threadIdx.x & 15) + 2;
    const int j = blockIdx.y * 16 + (threadIdx.x >> 4) + 2;
    if (i >= d.nx || j >= d.ny) return;
    const size_t q = cell(d, i, j);
    float x0 = d.map[which][1][q], x1 = d.map[which][2][q];
    const float nd = d.nodata;
    float r;
    if (x1 == nd && x0 == nd) r = nd;
    else if (x1 == nd) r = x0;
    else if (x0 == nd) r = x1;
    else r = factor * (x1 - x0) + x0;
    d.map[which][0][q] = r;
}

// FP64 sum of a field over interior cells (SumReduce, grid.cu:219-290).
__global__ __launch_bounds__(256) void k_sum_interior(Dev d, const float* x, double* out)
{
    __shared__ double red[4];
    const int ni = d.nx - 4, nj = d.ny - 4;
    const long total = (long)ni * nj;
    double acc = 0.0;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (long)gridDim.x * blockDim.x) {
        int j = (int)(k / ni) + 2, i = (int)(k % ni) + 2;
        acc += (double)x[cell(d, i, j)];
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, red[0] + red[1] + red[2] + red[3]);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_begin_step(const Dev& d, hipStream_t s)
{
    hipLaunchKernelGGL(k_begin_step, dim3(1), dim3(64), 0, s, d);
}

void launch_track(const Dev& d, hipStream_t s)
{
    const bool series = (d.flags & RIFT_SURGE_GRID) && d.sg_series;
    if ((d.flags & RIFT_RAIN_GRID) || ((d.flags & RIFT_SURGE_GRID) && !series)) {
        int grid = d.ntiles < 8192 ? d.ntiles : 8192;
        hipLaunchKernelGGL(k_track, dim3(grid), dim3(256), 0, s, d);
    }
    if (series) launch_track_series(d, s);
}

void launch_track_series(const Dev& d, hipStream_t s)
{
    hipLaunchKernelGGL(k_track_series, dim3((d.ntiles + 255) / 256), dim3(256), 0, s, d);
}

void launch_sg_flags(const Dev& d, int slot, hipStream_t s)
{
    hipLaunchKernelGGL(k_sg_flags, dim3((d.ntiles + 255) / 256), dim3(256), 0, s, d, slot);
}

void launch_sg_materialize(const Dev& d, hipStream_t s)
{
    dim3 grid((d.nx + 15) / 16, (d.ny + 15) / 16);
    hipLaunchKernelGGL(k_sg_materialize, grid, dim3(256), 0, s, d);
}

void launch_reset_counters(const Dev& d, hipStream_t s)
{
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, d);
}

void launch_grow(const Dev& d, hipStream_t s)
{
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, d);
    hipLaunchKernelGGL(k_grow, dim3((d.ntiles + 255) / 256), dim3(256), 0, s, d);
}

void launch_flux(const Dev& d, int stage, hipStream_t s, int grid)
{
    (void)stage;
    if (d.flags & RIFT_TEXTBOOK)
        hipLaunchKernelGGL(k_flux<true>, dim3(grid), dim3(256), 0, s, d);
    else
        hipLaunchKernelGGL(k_flux<false>, dim3(grid), dim3(256), 0, s, d);
}

void launch_timestep(const Dev& d, hipStream_t s)
{
    int blocks = (d.ntiles + 4095) / 4096;
    if (blocks < 1) blocks = 1;
    if (blocks > 64) blocks = 64;
    hipLaunchKernelGGL(k_timestep, dim3(blocks), dim3(256), 0, s, d);
}

void launch_integrate1(const Dev& d, hipStream_t s, int grid)
{
    hipLaunchKernelGGL(k_integrate1, dim3(grid), dim3(256), 0, s, d);
}

void launch_integrate2(const Dev& d, hipStream_t s, int grid)
{
    hipLaunchKernelGGL(k_integrate2, dim3(grid), dim3(256), 0, s, d);
}

void launch_boundaries(const Dev& d, hipStream_t s)
{
    int n = (d.nx > d.ny ? d.nx : d.ny);
    hipLaunchKernelGGL(k_boundaries, dim3((n + 255) / 256), dim3(256), 0, s, d);
}

void launch_interp_map(const Dev& d, int which, float factor, hipStream_t s)
{
    dim3 grid((d.nx + 15) / 16, (d.ny + 15) / 16);
    hipLaunchKernelGGL(k_interp_map, grid, dim3(256), 0, s, d, which, factor);
}

// BC, BX and BY of every plane cell from the corner plane (period-2 beyond
// the corners, so that ghost faces repeat the faces two cells in and ghost
// cells take the bed of the interior cell they copy from): the per-phase
// kernels read the same FP32 beds the fused step forms from the corners.
__global__ __launch_bounds__(256) void k_beds_from_corners(Dev d, int R)
{
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= d.P || j >= R) return;
    const int i1 = (i + 1 < d.P) ? i + 1 : i - 1;
    const int j1 = (j + 1 < R) ? j + 1 : j - 1;
    const float* r0 = d.DC + (size_t)j * d.P;
    const float* r1 = d.DC + (size_t)j1 * d.P;
    const float p0 = r0[i] + r0[i1], p1 = r1[i] + r1[i1];
    const size_t q = (size_t)j * d.P + i;
    d.BC[q] = bed_cell(p0, p1);
    d.BY[q] = bed_south(p0);
    d.BX[q] = bed_west(r0[i], r1[i]);
}

void launch_beds_from_corners(const Dev& d, int R, hipStream_t s)
{
    hipLaunchKernelGGL(k_beds_from_corners, dim3((d.P + 255) / 256, R), dim3(256), 0, s, d, R);
}

void launch_sum_interior(const Dev& d, const float* x, int mode, double* out, hipStream_t s)
{
    (void)mode;
    hipLaunchKernelGGL(k_sum_interior, dim3(512), dim3(256), 0, s, d, x, out);
}

}  // namespace rift
