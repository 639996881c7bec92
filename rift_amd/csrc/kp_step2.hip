// kp_step2.hip -- the fused Heun step: two marching kernels per step.
//
//   grow2      tile dilation (+ pending buffer flips, per-tile resets) and
//              compaction of the active marching units (strip-major)
//   march<1>   stage-1 flux + sources of U^n            -> dU, tile speeds
//   timestep   dt from all tile speeds                   (kp_kernels.hip)
//   march<2>   U* = Integrate_1(U^n, dU) on the fly, flux + sources of U*,
//              U^{n+1} = (U^n + U* + dt L(U*))/2, infiltration, diagnostics,
//              tile wet flags                            -> U^{n+1} in the
//              other buffer of the tile
//
// A wavefront marches down a strip: lane L holds column c = 60*s + L of
// strip s, keeping a window of three rows in registers (two rows of halo
// above and below the rows it outputs); x neighbours come from DPP
// wavefront shifts, so no LDS is used.  Lanes 2..61 produce output columns
// 60s+2 .. 60s+61.  A unit is one strip by one 14-row tile row; each wave
// marches an equal share of the active units' rows (see k_march).
//
// Ghost cells are never stored: a lane or row outside the interior loads the
// interior cell the reference's ApplyBoundaries copies from (grid.cu:1325-
// 1406) and applies the same critical-depth transform, which is what the
// ghost holds whenever the reference computes a flux.
//
// Stage 2 cannot update U in place (neighbouring units read it as halo), so
// every 16 x 14 tile owns a select bit: its current state lives in h/hu/hv or
// in h0/hu0/hv0, and stage 2 writes the other one; the bit flips at the next
// grow2.  Inactive tiles are never touched, so no copies are needed.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kp_device.hpp"
#include "kp_kernels.hpp"

#include "../../include/rift_kp.h"

namespace rift {

// lane L receives lane L-1's value (wave_shr:1) / lane L+1's (wave_shl:1);
// lane 0 (63) receives 0 (bound_ctrl: the missing source reads as zero, so
// no destination has to be zeroed first and the move can fold into its user)
__device__ __forceinline__ float from_left(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float from_right(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xF, 0xF, true));
}

__device__ __forceinline__ float wave_max(float v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    return v;
}

__device__ __forceinline__ bool dilated(const Dev& d, int t)
{
    const int bx = t % d.nbx, by = t / d.nbx, nbx = d.nbx;
    const unsigned char* w = d.wet;
    return w[t] || (bx > 0 && w[t - 1]) || (bx < nbx - 1 && w[t + 1]) ||
           (by > 0 && w[t - nbx]) || (by < d.nby - 1 && w[t + nbx]) || t == 0;
}

// tile columns covered by the output columns of strip s
__device__ __forceinline__ void strip_tiles(const Dev& d, int s, int& txa, int& txb)
{
    const int c0 = kStrip * s + 2;
    const int c1 = min(c0 + kStrip - 1, d.nx - 3);
    txa = (c0 - 2) / kTileW;
    txb = (c1 - 2) / kTileW;
}

// tile columns under all 64 lanes of strip s (halo lanes included), the
// columns clamped to the interior as the lanes clamp their loads
__device__ __forceinline__ void strip_lane_tiles(const Dev& d, int s, int& txa, int& txb)
{
    const int c0 = max(kStrip * s, 2);
    const int c1 = min(kStrip * s + 63, d.nx - 3);
    txa = (c0 - 2) / kTileW;
    txb = (c1 - 2) / kTileW;
}

// ---------------------------------------------------------------------------
// grow2, in two launches.  Tiles (one thread each): pending buffer flips,
// dilation (Grow_k), reset of an active tile's maxima, and a mark on the
// marching units whose output columns cover the tile.  Units (one thread
// each, after every tile is marked): the unit is active when marked;
// compacted by wavefront ballot + popcount prefix + one atomic per wave, and
// the mark cleared for the next step.  The marching kernels walk units, the
// dt kernel counts the active tiles and cells.  Thread 0 of the tile launch
// installs the forcing and the surge window the last dt kernel evaluated.
// ---------------------------------------------------------------------------
// A domain with an end time (ensemble members) stops stepping once its time
// has reached it: no units, no dt.
__device__ __forceinline__ bool finished(const Dev& d)
{
    return d.tf > 0.0 && d.sc->t >= d.tf;
}

// Returns whether tile t is active in this step (false past the tiles).
template <bool Mark>
__device__ __forceinline__ bool grow_tiles_body(const Dev& d, int t)
{
    if (t == 0) {
        Scalars* sc = d.sc;
        sc->finished = finished(d) ? 1 : 0;
        if (sc->next_valid) {
            sc->rain_rate = sc->nxt_rain;
            sc->drain_rate = sc->nxt_drain;
            sc->hydro_rate = sc->nxt_hydro;
            sc->src_cell_rate = sc->nxt_cell;
            sc->next_valid = 0;
        }
        // the surge window of this step (k_track_series read it already)
        if (sc->sg_next_valid) {
            sc->sg_k0 = sc->sg_nk0;
            sc->sg_k1 = sc->sg_nk1;
            sc->sg_f = sc->sg_nf;
            sc->sg_next_valid = 0;
            if (d.sg_series) check_surge_slots(d, sc);
        }
    }
    if (t >= d.ntiles) return false;
    if (d.flip[t]) {   // stage 2 of the last step wrote the other buffer
        d.sel[t] ^= 1;
        d.flip[t] = 0;
    }
    const bool a = dilated(d, t);
    d.act[t] = a ? 1 : 0;
    if (!a) return false;
    d.mx[t] = 0.f;
    d.my[t] = 0.f;
    d.mx1[t] = 0.f;
    d.my1[t] = 0.f;
    // units (s, by) whose output columns [60s+2, 60s+61] overlap the tile's
    // columns [16bx+2, 16bx+17]: s from 16bx/60 to (16bx+15)/60, each kept
    // only when the tile lies in that strip's tile range (strip_tiles)
    const int bx = t % d.nbx, by = t / d.nbx;
    if (by >= d.nurows) return true;
    const int s0 = (kTileW * bx) / kStrip, s1 = min((kTileW * bx + kTileW - 1) / kStrip, d.nstrips - 1);
    const int tag = Mark ? 0 : d.sc->ustep + 1;
    for (int s = s0; s <= s1; ++s) {
        int txa, txb;
        strip_tiles(d, s, txa, txb);
        if (bx >= txa && bx <= txb) {
            if constexpr (Mark) d.umark[s * d.nurows + by] = 1;
            else d.utag[s * d.nurows + by] = tag;
        }
    }
    return true;
}

// Surge relevance of every unit for the window in force: can it override a
// cell under any of the strip's 64 lanes (halo lanes included) in the
// unit's tile row?  Recomputed only when the window's maps or a slot's
// contents changed (sg_udirty, set by the dt kernel and the map loads).
__device__ __forceinline__ void unit_surge_flags(const Dev& d, int u, int s, int r)
{
    const Scalars* sc = d.sc;
    const unsigned char* f0 = d.sgflag + (size_t)sg_slot(d, sc->sg_nk0) * d.ntiles;
    const unsigned char* f1 = d.sgflag + (size_t)sg_slot(d, sc->sg_nk1) * d.ntiles;
    int la, lb;
    strip_lane_tiles(d, s, la, lb);
    bool rel = false;
    for (int tx = la; tx <= lb; ++tx) rel |= ((f0[r * d.nbx + tx] | f1[r * d.nbx + tx]) & 2) != 0;
    d.sgu[u] = rel ? 1 : 0;
}

// List appends of a workgroup of up to 1024 threads: for each of N lists,
// the flagged threads get consecutive slots (waves in order, lanes in order
// within a wave) from one atomic on the list's counter per workgroup, not
// one per wave (same-address atomics serialise: ~1300 of them, one per
// wave, cost several microseconds); slot[k] is this thread's slot in list k
// (valid where flag[k]).  Every thread of the workgroup must call it.
template <int N>
__device__ __forceinline__ void wg_append(const bool (&flag)[N], int* const (&counter)[N], int (&slot)[N])
{
    __shared__ int s_cnt[16][N];
    __shared__ int s_base[N];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long lt = (1ull << lane) - 1ull;
    unsigned long long m[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        m[k] = __ballot(flag[k]);
        if (lane == 0) s_cnt[wave][k] = __popcll(m[k]);
    }
    __syncthreads();
    if (threadIdx.x < N) {
        const int k = threadIdx.x, nw = (int)(blockDim.x >> 6);
        int tot = 0;
        for (int w = 0; w < nw; ++w) tot += s_cnt[w][k];
        s_base[k] = tot ? atomicAdd(counter[k], tot) : 0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; ++k) {
        int b = s_base[k];
        for (int w = 0; w < wave; ++w) b += s_cnt[w][k];
        slot[k] = b + __popcll(m[k] & lt);
    }
}

__device__ __forceinline__ void compact_units_body(const Dev& d, int u)
{
    const bool is_unit = u < d.nunits;
    bool ua = false;
    if (is_unit) {
        ua = d.umark[u] != 0;
        if (ua) d.umark[u] = 0;
        if (finished(d)) ua = false;
        if ((d.flags & RIFT_SURGE_GRID) && d.sg_series && d.sc->sg_udirty) {
            const int s = u / d.nurows;
            unit_surge_flags(d, u, s, u - s * d.nurows);
        }
    }
    int* const ctr[1] = {&d.sc->n_units_acc};
    const bool f[1] = {ua};
    int slot[1];
    wg_append<1>(f, ctr, slot);
    if (ua) d.ulist[slot[0]] = u;
}

// Single domain, one thread per unit u = (s, r), strip-major, after every
// tile is tagged: the unit is active when grow tagged it in this step; its
// quad (strips 4q .. 4q+3, q = s / 4) is full when all four strips are
// active in tile row r.  A full quad is listed once (by its first strip's
// thread) in qlist, the other active units in ulist; runs of one strip stay
// in order within each wave's 64 units.

__device__ __forceinline__ void compact_quads_body(const Dev& d, int u)
{
    bool own = false, full = false;
    int s = 0, r = 0;
    if (u < d.nunits) {
        s = u / d.nurows;
        r = u - s * d.nurows;
        if (!finished(d)) {
            const int tag = d.sc->ustep + 1;
            own = d.utag[u] == tag;
            const int s0 = s & ~3;
            if (own && d.quads && s0 + 3 < d.nstrips) {
                full = true;
                for (int k = s0; k < s0 + 4; ++k) full = full && d.utag[k * d.nurows + r] == tag;
            }
        }
        if ((d.flags & RIFT_SURGE_GRID) && d.sg_series && d.sc->sg_udirty) unit_surge_flags(d, u, s, r);
    }
    const bool part = own && !full;
    const bool head = full && (s & 3) == 0;
    int* const ctr[3] = {&d.sc->n_units_acc, &d.sc->n_part_acc, &d.sc->n_quads_acc};
    const bool f[3] = {own, part, head};
    int slot[3];
    wg_append<3>(f, ctr, slot);
    if (part) d.ulist[slot[1]] = u;
    if (head) d.qlist[slot[2]] = (s >> 2) * d.nurows + r;
}

// ---------------------------------------------------------------------------
// marching kernels
// ---------------------------------------------------------------------------
// Global loads of one row, issued a row ahead.  The beds come from the
// corner plane: the loads of row j carry corner row j + 1 when marching down
// and corner row j when marching up, the row a step further on, so that
// finishing a row needs only corner rows already loaded (its own loads and
// those of the row finished before it).  Each lane loads the corners of its
// column and the next, whose sum is the row's pair sum.
struct Raw {
    float h, hu, hv;
    float dc, dc1;               // corners of columns c and c + 1
    float dh, dhu, dhv, sg, n;   // stage 2
};

struct Row {
    float h, hu, hv;        // state used by the flux (U^n in stage 1, U* in 2)
    float b, by, bx;        // beds: cell, south face, west face
    float hraw;             // stage 2: U^n depth as stored (stage-1 sources)
    float hun, hvn;         // stage 2: U^n momenta of the cell (own-cell Heun)
    float d1;               // stage 2: the stage-1 dh of the cell (volume budget)
};

struct Src {                 // own-cell inputs of an output row
    float K, F, rain, drain;
};

// Per-lane constants of one unit.  The rows of a unit touch tile rows r-1,
// r, r+1 only; their buffer selects and activity flags are read once, and
// those of the tile row after r+1 in marching direction (r+2 down, r-2 up)
// are loaded one tile row ahead: the loads retire in order, so a flag loaded
// at the crossing itself would make the wave wait for every row prefetch in
// flight.
// Cells are addressed as a row pointer (wave-uniform, scalar registers) plus
// a 32-bit byte offset per lane (the column, and one plane more for a tile
// whose state lives in the second buffer), which the global loads and stores
// take directly (scalar base + vector offset).
struct Lane {
    unsigned c4, sc4;        // byte offsets of the output / clamped column
    unsigned pl4;            // bytes of one plane (second state buffer)
    int txo, r;
    bool gx;
    // byte offsets of the lane's state cell in tile rows r-1, r, r+1: the
    // clamped column, plus one plane where the tile's state lives in the
    // second buffer (formed at tile-row crossings, not per row)
    unsigned so0, so1, so2;
    bool act0, act1, act2;   // activity of tile rows r-1, r, r+1
    unsigned soN;            // the same for tile row r+2 (down) or r-2 (up)
    bool actN;
    // surge series: the strip's tiles of tile rows r-1, r, r+1 (and the next
    // one) can hold surge > 0 in this step's window (wave-uniform)
    bool sg0, sg1, sg2, sgN;
};

__device__ __forceinline__ float ld(const float* row, unsigned byte_off)
{
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(row) + byte_off);
}
__device__ __forceinline__ float& at(float* row, unsigned byte_off)
{
    return *reinterpret_cast<float*>(reinterpret_cast<char*>(row) + byte_off);
}

__device__ __forceinline__ int tile_row_slot(const Lane& L, int sr)
{
    const int tr = (sr - 2) / kTileH;
    return tr - L.r + 1;     // 0, 1 or 2 (uniform across the wave)
}

template <class T>
__device__ __forceinline__ T pick(int slot, T a, T b, T c)
{
    return slot == 0 ? a : (slot == 1 ? b : c);
}

__device__ __forceinline__ size_t row_of(const Dev& d, int j)
{
    return (size_t)j * (size_t)d.P;
}

// Feature flags an instantiation of the marching path may see.  The lean
// kernels are compiled without gridded forcing, point sources and summary
// maps, so that the pointers and branches of those features take no
// registers in the common case.
constexpr unsigned kLeanMask = ~(RIFT_RAIN_GRID | RIFT_DRAIN_GRID | RIFT_SURGE_GRID | RIFT_HYDROGRAPH |
                                 RIFT_SAVE_MAX | RIFT_SAVE_ARRIVAL);
constexpr unsigned kFullMask = ~0u;
// The coastal feature set: the lean one plus the surge series and the
// hydrograph rectangle (a surge map is always read from the device series).
constexpr unsigned kCoastMask = ~(RIFT_RAIN_GRID | RIFT_DRAIN_GRID | RIFT_SAVE_MAX | RIFT_SAVE_ARRIVAL);
// Instantiations of the single-domain marching kernels.
enum MarchKind { kKindFull = 0, kKindLean = 1, kKindCoast = 2 };
constexpr unsigned kind_mask(int kind)
{
    return kind == kKindLean ? kLeanMask : (kind == kKindCoast ? kCoastMask : kFullMask);
}
// The same features for the ensemble kernels.  Bit 31 is no flag, so the
// masking is identical; the distinct value keeps the one-body row loop and
// the scalar-base addressing there (their member lookup leaves no registers
// for three bodies).
constexpr unsigned kLeanEnsMask = kLeanMask & ~(1u << 31);

// A lane's column offset as the loads and stores of a row use it.  Left to
// itself the compiler hoists field + column out of the row loop as a 64-bit
// per-lane address (two VGPRs per field) and adds the row offset with a
// 64-bit vector add per access.  Made opaque once per row, the offset stays a
// 32-bit lane offset to a scalar row base: lean stage 1 then needs 87 VGPRs
// instead of 103 (2 % faster), and the full and ensemble kernels stop
// spilling.  Lean stage 2 keeps the hoisted form: with its 12 fields a row,
// forming the scalar bases cost it more (scalar adds, spilled scalar
// registers, 2.3x the dependency waits) than the vector adds saved.
template <int Stage, unsigned FM>
__device__ __forceinline__ unsigned lane_off(unsigned x)
{
    if constexpr (!(Stage == 2 && (FM == kLeanMask || FM == kCoastMask))) asm volatile("" : "+v"(x));
    return x;
}

// Scalars of a marching launch; sg0/sg1/sgf: the two surge planes of the
// step's window and its factor (device surge series).
struct StepScalars {
    float dt, inv_dt, dtp, inv_dtp, kappa, rain_rate, drain_rate, src_rate, tnow, inv_dx, inv_dy;
    unsigned fl;
    const float* sg0;
    const float* sg1;
    float sgf;
};

// The surge planes of the running step's window (installed by grow2).
__device__ __forceinline__ void surge_scalars(const Dev& d, StepScalars& S)
{
    S.sg0 = S.sg1 = nullptr;
    S.sgf = 0.f;
    if (d.sg_series) {
        const Scalars* sc = d.sc;
        S.sg0 = d.sgs + (size_t)sg_slot(d, sc->sg_k0) * (size_t)d.soff;
        S.sg1 = d.sgs + (size_t)sg_slot(d, sc->sg_k1) * (size_t)d.soff;
        S.sgf = sc->sg_f;
    }
}

template <int Stage, bool Up, unsigned FM>
__device__ __forceinline__ Raw fetch(const Dev& d, const StepScalars& S, const Lane& L, int j)
{
    const int sr = min(max(j, 2), d.ny - 3);
    const int slot = tile_row_slot(L, sr);
    const size_t rb = row_of(d, sr);
    const unsigned sc4 = lane_off<Stage, FM>(L.sc4), c4 = lane_off<Stage, FM>(L.c4);
    const unsigned lo = lane_off<Stage, FM>(pick(slot, L.so0, L.so1, L.so2));
    Raw r;
    r.h = ld(d.h + rb, lo);
    r.hu = ld(d.hu + rb, lo);
    r.hv = ld(d.hv + rb, lo);
    const float* dcr = d.DC + row_of(d, Up ? j : j + 1);
    r.dc = ld(dcr, c4);
    r.dc1 = ld(dcr, c4 + 4u);
    if constexpr (Stage == 2) {
        r.dh = ld(d.dh + rb, sc4);
        r.dhu = ld(d.dhu + rb, sc4);
        r.dhv = ld(d.dhv + rb, sc4);
        // surge only where this step's window can hold surge > 0 in the
        // strip's tiles of the row's tile row (a uniform branch)
        if ((d.flags & FM & RIFT_SURGE_GRID) && pick(slot, L.sg0, L.sg1, L.sg2)) {
            if (FM == kCoastMask || d.sg_series)
                r.sg = surge_interp(ld(S.sg0 + rb, sc4), ld(S.sg1 + rb, sc4), S.sgf, d.nodata);
            else r.sg = ld(d.map[RIFT_MAP_SURGE][0] + rb, sc4);
        } else {
            r.sg = 0.f;
        }
        r.n = d.n_uniform ? d.n_const : ld(d.n + rb, sc4);
    }
    return r;
}

template <int Stage, unsigned FM>
__device__ __forceinline__ Src fetch_src(const Dev& d, const Lane& L, int j)
{
    const size_t rb = row_of(d, j);
    const unsigned c4 = lane_off<Stage, FM>(L.c4);
    Src s;
    s.K = (d.flags & RIFT_INFILTRATION) ? ld(d.K + rb, c4) : 0.f;
    s.F = (d.flags & RIFT_INFILTRATION) ? ld(d.F + rb, c4) : 0.f;
    s.rain = (d.flags & FM & RIFT_RAIN_GRID) ? ld(d.map[RIFT_MAP_RAIN][0] + rb, c4) : 0.f;
    s.drain = (d.flags & FM & RIFT_DRAIN_GRID) ? ld(d.map[RIFT_MAP_DRAIN][0] + rb, c4) : 0.f;
    return s;
}

// Row j from its loads; dcp, pp: corner and pair sum of the corner row the
// row finished before it loaded (the other corner row of row j), replaced by
// this row's.
template <int Stage, bool Up, unsigned FM>
__device__ __forceinline__ Row finish(const Dev& d, const Lane& L, const Raw& raw, int j, float dt,
                                      float inv_dt, float& dcp, float& pp)
{
    const int sr = min(max(j, 2), d.ny - 3);
    const bool gy = (j < 2) || (j > d.ny - 3);
    float h = raw.h, hu = raw.hu, hv = raw.hv;
    Row r;
    {
        const float pc = raw.dc + raw.dc1;
        const float plo = Up ? pc : pp, phi = Up ? pp : pc;
        r.b = bed_cell(plo, phi);
        r.by = bed_south(plo);
        r.bx = Up ? bed_west(raw.dc, dcp) : bed_west(dcp, raw.dc);
        dcp = raw.dc;
        pp = pc;
    }
    if constexpr (Stage == 2) {
        r.hraw = h;
        r.hun = hu;
        r.hvn = hv;
        r.d1 = raw.dh;
        if (pick(tile_row_slot(L, sr), L.act0, L.act1, L.act2)) {
            // Integrate_1_k (grid.cu:856-922) of this cell from U^n and dU;
            // G is recomputed from U^n exactly as the stage-1 flux forms it
            const float G = manning_g(raw.n, hu, hv, h, d.kappa);
            const float hcl = (h > kEps) ? h : 0.f;
            float dh = raw.dh;
            if ((d.flags & FM & RIFT_SURGE_GRID) && raw.sg != d.nodata && raw.sg > 0.f)
                dh = (raw.sg - fmaf(dh, dt, hcl)) * inv_dt;
            const float hnew = fmaf(dh, dt, hcl);
            const float rden = frcp(fmaf(-dt, G, 1.f));
            hu = fmaf(raw.dhu, dt, hu) * rden;
            hv = fmaf(raw.dhv, dt, hv) * rden;
            h = (hnew > kEps) ? hnew : 0.f;
        }
    }
    (void)inv_dt;
    if (L.gx) {      // ApplyBoundariesX_k: epsilon-clamped critical depth
        const float H = fminf(critical_depth(hu), h);
        h = (H > kEps) ? H : 0.f;
    } else if (gy) { // ApplyBoundariesY_k: no clamp
        h = fminf(critical_depth(hv), h);
    }
    r.h = h;
    r.hu = hu;
    r.hv = hv;
    return r;
}

// Lane constants for strip s, tile row rr.
// Surge relevance of unit (strip s, tile row r), clamped to the tile rows.
template <unsigned FM>
__device__ __forceinline__ bool unit_surge(const Dev& d, int s, int r)
{
    if constexpr (!(FM & RIFT_SURGE_GRID)) {
        return false;
    } else {
        if (!(d.flags & RIFT_SURGE_GRID)) return false;
        if (!d.sg_series) return true;
        r = min(max(r, 0), d.nurows - 1);
        return d.sgu[s * d.nurows + r] != 0;
    }
}

template <bool Up, unsigned FM>
__device__ __forceinline__ Lane make_lane(const Dev& d, int s, int rr, int lane)
{
    Lane L;
    const int c = kStrip * s + lane;
    const int sc = min(max(c, 2), d.nx - 3);
    L.c4 = 4u * (unsigned)c;
    L.sc4 = 4u * (unsigned)sc;
    L.pl4 = 4u * (unsigned)d.soff;
    L.gx = (c < 2) || (c > d.nx - 3);
    L.txo = (sc - 2) / kTileW;
    L.r = rr;
    const int Tm = max(rr - 1, 0) * d.nbx + L.txo, T0 = rr * d.nbx + L.txo;
    const int Tp = min(rr + 1, d.nby - 1) * d.nbx + L.txo;
    L.so0 = L.sc4 + (d.sel[Tm] ? L.pl4 : 0u);
    L.so1 = L.sc4 + (d.sel[T0] ? L.pl4 : 0u);
    L.so2 = L.sc4 + (d.sel[Tp] ? L.pl4 : 0u);
    L.act0 = d.act[Tm] != 0; L.act1 = d.act[T0] != 0; L.act2 = d.act[Tp] != 0;
    const int TN = (Up ? max(rr - 2, 0) : min(rr + 2, d.nby - 1)) * d.nbx + L.txo;
    L.soN = L.sc4 + (d.sel[TN] ? L.pl4 : 0u);
    L.actN = d.act[TN] != 0;
    L.sg0 = unit_surge<FM>(d, s, rr - 1);
    L.sg1 = unit_surge<FM>(d, s, rr);
    L.sg2 = unit_surge<FM>(d, s, rr + 1);
    L.sgN = unit_surge<FM>(d, s, Up ? rr - 2 : rr + 2);
    return L;
}

// Move the lane's tile-row window to tile row rr = L.r + 1 (down) or - 1.
template <bool Up, unsigned FM>
__device__ __forceinline__ void shift_lane(const Dev& d, Lane& L, int s, int rr)
{
    L.r = rr;
    if constexpr (!Up) {
        L.so0 = L.so1; L.so1 = L.so2; L.so2 = L.soN;
        L.act0 = L.act1; L.act1 = L.act2; L.act2 = L.actN;
        L.sg0 = L.sg1; L.sg1 = L.sg2; L.sg2 = L.sgN;
    } else {
        L.so2 = L.so1; L.so1 = L.so0; L.so0 = L.soN;
        L.act2 = L.act1; L.act1 = L.act0; L.act0 = L.actN;
        L.sg2 = L.sg1; L.sg1 = L.sg0; L.sg0 = L.sgN;
    }
    const int TN = (Up ? max(rr - 2, 0) : min(rr + 2, d.nby - 1)) * d.nbx + L.txo;
    L.soN = L.sc4 + (d.sel[TN] ? L.pl4 : 0u);
    L.actN = d.act[TN] != 0;
    L.sgN = unit_surge<FM>(d, s, Up ? rr - 2 : rr + 2);
}


// Per-tile results of the rows of tile row rr a wave marched in strip s:
// speed maxima (stage 1 into mx1/my1 for this step's dt, stage 2 into mx/my,
// which later steps see while the tile is inactive) and the tile wet flags
// of Integrate_2_k, written by thread (0,0) of a tile, i.e. only for a tile
// whose origin cell is interior (grid.cu:1160-1170): stage 1 resets the flag
// of an active tile (to 1 without tile acceleration), stage 2 sets it where
// a cell is wet.  Several waves may share a tile; every write is idempotent.
// wv: the largest tile maxima the wave flushed so far (uniform).
template <int Stage>
__device__ __forceinline__ void flush_tiles(const Dev& d, int s, int rr, const Lane& L, bool out_lane,
                                            float mxl, float myl, bool wetl, int lane, float2& wv)
{
    float* const tmx = (Stage == 1) ? d.mx1 : d.mx;
    float* const tmy = (Stage == 1) ? d.my1 : d.my;
    int txa, txb;
    strip_tiles(d, s, txa, txb);
    for (int tx = txa; tx <= txb; ++tx) {
        const int T = rr * d.nbx + tx;
        // every tile of the strip holds output lanes, whose L.act1 is the
        // tile's activity: no load (and no wait on the row prefetches)
        const bool seg = out_lane && L.txo == tx;
        if (__ballot(seg && L.act1) == 0ull) continue;   // uniform
        const float m1 = wave_max(seg ? mxl : 0.f);
        const float m2 = wave_max(seg ? myl : 0.f);
        wv.x = fmaxf(wv.x, m1);
        wv.y = fmaxf(wv.y, m2);
        const bool anyw = __ballot(seg && wetl) != 0ull;
        if (lane == 0) {
            atomicMax(reinterpret_cast<unsigned*>(tmx) + T, __float_as_uint(m1));
            atomicMax(reinterpret_cast<unsigned*>(tmy) + T, __float_as_uint(m2));
            const bool origin = tx * kTileW + 2 <= d.nx - 3 && rr * kTileH + 2 <= d.ny - 3;
            if constexpr (Stage == 1) {
                if (origin) d.wet[T] = (d.flags & RIFT_TILE_ACCEL) ? 0 : 1;
            } else {
                d.flip[T] = 1;
                if (origin && anyw) d.wet[T] = 1;
            }
        }
    }
}

// Per-lane sums of the volume budget over a run (m summed over cells; the
// stage-1 outflow as a rate, m/s, weighted by the dt kernel).
struct BudgetAcc {
    float rain, drain, out, exch;
};

// Sources, the stage update and the stores of output row j (lanes of active
// tiles); returns the cell's wet flag (stage 2).  With Budget, also the
// cell's budget terms: outflow through the domain boundary faces of the cell
// (fE, fw, gn, gs: mass fluxes of its east, west, north, south faces), and in
// stage 2 the gridded rain, the drain of both stages and the exchange of the
// dry clamps and the surge override of this step.
template <int Stage, bool Budget, unsigned FM>
__device__ __forceinline__ bool output_row(const Dev& d, const StepScalars& S, const Lane& L, int c,
                                           int j, const Row& C, const Src& src, float dh, float dhu,
                                           float dhv, float bxw, float bxe, float bys, float byn,
                                           float fE, float fw, float gn, float gs, BudgetAcc& acc)
{
    const unsigned fl = S.fl & FM;
    const size_t rbo = row_of(d, j);   // output row (uniform)
    const unsigned c4 = lane_off<Stage, FM>(L.c4);
    const float hC = C.h;
    const float S1 = -kG * hC * ((bxe - bxw) * S.inv_dx);
    const float S2 = -kG * hC * ((byn - bys) * S.inv_dy);
    float S0 = 0.f;
    if (fl & RIFT_RAIN_AVG) S0 += S.rain_rate;
    if (fl & RIFT_RAIN_GRID) S0 += src.rain;
    if (fl & RIFT_DRAIN_AVG) {
        const float cap = hC * S.inv_dt;
        S0 -= (S.drain_rate > cap) ? cap : S.drain_rate;
    }
    if (fl & RIFT_DRAIN_GRID) {
        const float cap = hC * S.inv_dt;
        S0 -= (src.drain > cap) ? cap : src.drain;
    }
    const bool srcc = (fl & RIFT_HYDROGRAPH) && c >= d.si0 && c <= d.si1 && j >= d.sj0 && j <= d.sj1;
    if (srcc) S0 += S.src_rate;
    if constexpr (Budget) {
        float outflow = 0.f;
        if (c == 2) outflow -= fw * S.inv_dx;
        if (c == d.nx - 3) outflow += fE * S.inv_dx;
        if (j == 2) outflow -= gs * S.inv_dy;
        if (j == d.ny - 3) outflow += gn * S.inv_dy;
        if constexpr (Stage == 1) {
            acc.out += outflow;
        } else {
            const float half_dt = 0.5f * S.dt;
            acc.out = fmaf(half_dt, outflow, acc.out);
            if (fl & RIFT_RAIN_GRID) acc.rain = fmaf(S.dt, src.rain, acc.rain);
            // what the drain took in both stages: stage 1 capped with the
            // stored depth and the previous dt, stage 2 with U*
            float taken = 0.f;
            const float cap1 = C.hraw * S.inv_dtp;
            if (fl & RIFT_DRAIN_AVG) {
                taken += (S.drain_rate > cap1) ? cap1 : S.drain_rate;
                taken += (S.drain_rate > hC * S.inv_dt) ? hC * S.inv_dt : S.drain_rate;
            }
            if (fl & RIFT_DRAIN_GRID) {
                taken += (src.drain > cap1) ? cap1 : src.drain;
                taken += (src.drain > hC * S.inv_dt) ? hC * S.inv_dt : src.drain;
            }
            acc.drain = fmaf(half_dt, taken, acc.drain);
        }
    }
    float Fnew = 0.f;
    if (fl & RIFT_INFILTRATION) {
        if constexpr (Stage == 1) {
            S0 -= green_ampt(src.K, src.F, d.psi_dtheta, hC, S.dt, S.inv_dt);
        } else {
            // the stage-1 rate (previous dt, stored depth), then
            // F* = F + dF dt of Integrate_1_k (grid.cu:917-923)
            const float f1 = green_ampt(src.K, src.F, d.psi_dtheta, C.hraw, S.dtp, S.inv_dtp);
            const float Fs = fmaf(f1, S.dt, src.F);
            const float f2 = green_ampt(src.K, Fs, d.psi_dtheta, hC, S.dt, S.inv_dt);
            S0 -= f2;
            Fnew = 0.5f * fmaf(f2, S.dt, src.F + Fs);
        }
    }
    dh += S0;
    dhu += S1;
    dhv += S2;
    if constexpr (Stage == 1) {
        at(d.dh + rbo, c4) = dh;
        at(d.dhu + rbo, c4) = dhu;
        at(d.dhv + rbo, c4) = dhv;
        return false;
    } else {
        // Integrate_2_k (grid.cu:1077-1123); U^n depth with the dry clamp of
        // ComputeFluxes_k
        const float h_n = (C.hraw > kEps) ? C.hraw : 0.f;
        const float hn = 0.5f * fmaf(dh, S.dt, h_n + C.h);
        const float hun = 0.5f * fmaf(dhu, S.dt, C.hun + C.hu);
        const float hvn = 0.5f * fmaf(dhv, S.dt, C.hvn + C.hv);
        const float hc = (hn > kEps) ? hn : 0.f;
        if constexpr (Budget) {
            // clamp of U^n, half of (stage-1 clamp + surge override) carried
            // by U*, clamp of U^{n+1}
            acc.exch += (h_n - C.hraw) + 0.5f * (C.h - fmaf(C.d1, S.dt, h_n)) + (hc - hn);
        }
        // into the buffer the tile does not read (an output lane's column
        // is its clamped column: c4 + pl4 - (so1 - c4))
        const unsigned lo = 2u * c4 + L.pl4 - L.so1;
        at(d.h + rbo, lo) = hc;
        at(d.hu + rbo, lo) = hun;
        at(d.hv + rbo, lo) = hvn;
        if (fl & RIFT_INFILTRATION) at(d.F + rbo, c4) = Fnew;
        if (fl & RIFT_SAVE_MAX) {
            const float fv = vel_factor(hc, S.kappa);
            const float uC = fv * hun, vC = fv * hvn;
            float& hm = at(d.hmax + rbo, c4);
            float& qm = at(d.qmax + rbo, c4);
            hm = fmaxf(hm, hc);
            qm = fmaxf(qm, hc * fsqrt(vC * vC + uC * uC));
        }
        if (fl & RIFT_SAVE_ARRIVAL) {
            float& tw = at(d.twet + rbo, c4);
            if (tw < 0.f && hc > kEps) tw = S.tnow;
        }
        bool w = hc > kEps || srcc;
        if ((fl & RIFT_RAIN_AVG) && S.rain_rate > 0.f) w = true;
        if ((fl & RIFT_RAIN_GRID) && src.rain > 0.f) w = true;
        return w;
    }
}

// FP64 wave sum of one lane value, added to *dst by lane 0.
__device__ __forceinline__ void wave_add_f64(double* dst, float v, int lane)
{
    double x = (double)v;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if (lane == 0 && x != 0.0) atomicAdd(dst, x);
}

template <int K>
struct Phase {
    static constexpr int value = K;
};

// Loop form: the row body instantiated once per window phase (register
// renaming) or once with the window advanced by copies.  Three copies of the
// body lengthen live ranges: the lean kernels have the registers for it, the
// full and ensemble kernels (more pointers and flags live) would spill.
template <unsigned FM>
constexpr bool phased_rows() { return FM == kLeanMask || FM == kCoastMask; }

// Issue priority of a wave that has marched `done` of the `total` rows of
// its share: 3 at its start down to 0 near its end.  The SIMD's arbiter
// prefers the older of two ready waves of equal priority, so with equal
// shares the waves dispatched first run ahead and the last ones finish
// alone; a wave that has run ahead yields to those behind it at each step
// down.  Steps at the quarters of the share (sparse dealing: shares of
// several runs), or at 1/2, 3/4 and 7/8 (one band per wave of a fully
// active domain; measured 2 % faster there, 1 % slower on the sparse
// coastal workload).
__device__ __forceinline__ void pace_wave(int done, int total, int& prio, bool dense)
{
    const int p = dense ? ((2 * done < total) ? 3 : (4 * done < 3 * total) ? 2 : (8 * done < 7 * total) ? 1 : 0)
                        : ((4 * done < total) ? 3 : (2 * done < total) ? 2 : (4 * done < 3 * total) ? 1 : 0);
    if (p == prio) return;
    prio = p;
    switch (p) {
    case 3: __builtin_amdgcn_s_setprio(3); break;
    case 2: __builtin_amdgcn_s_setprio(2); break;
    case 1: __builtin_amdgcn_s_setprio(1); break;
    default: __builtin_amdgcn_s_setprio(0); break;
    }
}

// One run of a wave: output rows [ja, jb) of strip s, marched downwards
// (increasing j) or upwards.  The window holds the output row C, the next
// row D (in marching direction) and the raw row after it; the y flux of the
// face behind C is carried from the previous row.  Crossing into another
// tile row flushes the per-tile results and shifts the lane's tile-row
// window.
//
// The window lives in three slots per quantity: row j + m*step of output row
// j sits in slot (k + m) mod 3, where k is the phase of row j.  The loop body
// is instantiated once per phase, so that advancing the window renames
// registers at compile time instead of copying ~40 values per row.
//
// sdone / stotal: rows of the wave's share marched before this run and in
// all (stotal 0: no pacing); dense: the share is one band of a fully active
// domain (pace_wave's schedule).
template <int Stage, bool Textbook, bool Up, bool Budget, unsigned FM>
__device__ float2 march_run(const Dev& d, const StepScalars& S, int s, int ja, int jb, int lane, int sdone = 0,
                            int stotal = 0, bool dense = false)
{
    constexpr int step = Up ? -1 : 1;
    const int c = kStrip * s + lane;
    const bool out_lane = lane >= 2 && lane <= kStrip + 1 && c <= d.nx - 3;
    int j = Up ? jb - 1 : ja;
    const int jstop = Up ? ja - 1 : jb;
    int rr = (j - 2) / kTileH;
    Lane L = make_lane<Up, FM>(d, s, rr, lane);
    const float dt = S.dt, inv_dt = S.inv_dt, kappa = S.kappa;

    Row R[3];          // finished rows C, D (and E once formed)
    Raw W[3];          // raw rows in flight
    Src SR[3];         // own-cell inputs of the output row and the next
    FaceStates Y[3];   // y face states of C (and D once formed)
    Flux G[3];         // y flux of the face behind C (and between C and D)

    // prologue (phase 0): rows j-2 .. j+2 (both directions).  The first
    // finished row, A, also needs the corner row before its own loads' one
    // (row j-2 down, j+3 up).
    float dcp, pp;
    {
        const float* xr = d.DC + row_of(d, Up ? j + 3 : j - 2);
        const unsigned c4 = lane_off<Stage, FM>(L.c4);
        dcp = ld(xr, c4);
        pp = dcp + ld(xr, c4 + 4u);
    }
    const Row A = finish<Stage, Up, FM>(d, L, fetch<Stage, Up, FM>(d, S, L, j - 2 * step), j - 2 * step, dt,
                                        inv_dt, dcp, pp);
    const Raw rB = fetch<Stage, Up, FM>(d, S, L, j - step);
    const Raw rC = fetch<Stage, Up, FM>(d, S, L, j);
    const Raw rD = fetch<Stage, Up, FM>(d, S, L, j + step);
    W[2] = fetch<Stage, Up, FM>(d, S, L, j + 2 * step);
    SR[0] = fetch_src<Stage, FM>(d, L, j);
    const Row B = finish<Stage, Up, FM>(d, L, rB, j - step, dt, inv_dt, dcp, pp);
    R[0] = finish<Stage, Up, FM>(d, L, rC, j, dt, inv_dt, dcp, pp);
    R[1] = finish<Stage, Up, FM>(d, L, rD, j + step, dt, inv_dt, dcp, pp);
    // the row behind C (its south face bed is read as by_behind below)
    R[2].by = B.by;
    // differences across the face between rows D and E (north minus south),
    // carried to the next row, whose construction reads the same face
    FaceDiff DY;
    {
        const Row& C = R[0];
        const Row& D = R[1];
        if constexpr (!Up) {
            const FaceStates yB = construct<Textbook>(A.h, B.h, C.h, A.b, B.b, C.b, A.hu, B.hu, C.hu,
                                                      A.hv, B.hv, C.hv, C.by, B.by);
            const FaceDiff dBC = face_diff(B.h, C.h, B.b, C.b, B.hu, C.hu, B.hv, C.hv);
            DY = face_diff(C.h, D.h, C.b, D.b, C.hu, D.hu, C.hv, D.hv);
            Y[0] = construct_d<Textbook>(C.h, C.b, C.hu, C.hv, DY, dBC, D.by, C.by);
            G[0] = face_flux<1>(yB.hp, yB.qxp, yB.qyp, Y[0].hm, Y[0].qxm, Y[0].qym, kappa);
        } else {
            // rows: A = j+2, B = j+1 (north of C), D = j-1
            const FaceStates yB = construct<Textbook>(C.h, B.h, A.h, C.b, B.b, A.b, C.hu, B.hu, A.hu,
                                                      C.hv, B.hv, A.hv, A.by, B.by);
            const FaceDiff dCB = face_diff(C.h, B.h, C.b, B.b, C.hu, B.hu, C.hv, B.hv);
            DY = face_diff(D.h, C.h, D.b, C.b, D.hu, C.hu, D.hv, C.hv);
            Y[0] = construct_d<Textbook>(C.h, C.b, C.hu, C.hv, dCB, DY, B.by, C.by);
            G[0] = face_flux<1>(Y[0].hp, Y[0].qxp, Y[0].qyp, yB.hm, yB.qxm, yB.qym, kappa);
        }
    }

    float mxl = 0.f, myl = 0.f;
    float2 wv = make_float2(0.f, 0.f);   // maxima of the tiles flushed (uniform)
    bool wetl = false;
    BudgetAcc acc{0.f, 0.f, 0.f, 0.f};
    const int jfirst = j;
    int prio = -1;
    // output row j in phase k; returns whether the run is done
    auto row = [&](auto phase) __attribute__((always_inline)) -> bool {
        constexpr int kc = decltype(phase)::value;
        constexpr int kd = (kc + 1) % 3, ke = (kc + 2) % 3;
        if (stotal > 0) pace_wave(sdone + (Up ? jfirst - j : j - jfirst), stotal, prio, dense);
        const int rj = (j - 2) / kTileH;   // uniform
        if (rj != rr) {
            flush_tiles<Stage>(d, s, rr, L, out_lane, mxl, myl, wetl, lane, wv);
            mxl = 0.f;
            myl = 0.f;
            wetl = false;
            shift_lane<Up, FM>(d, L, s, rj);
            rr = rj;
        }
        // prefetch: the row after next and the next output row's inputs
        // (clamped to the domain; unused past the end of the run)
        const int jn = Up ? max(j - 3, 0) : j + 3;
        const Raw rN = fetch<Stage, Up, FM>(d, S, L, jn);
        SR[kd] = fetch_src<Stage, FM>(d, L, min(max(j + step, 2), d.ny - 3));

        const float by_behind = R[ke].by;    // BY of row j - step
        R[ke] = finish<Stage, Up, FM>(d, L, W[ke], j + 2 * step, dt, inv_dt, dcp, pp);
        W[kc] = rN;
        const Row& C = R[kc];
        const Row& D = R[kd];
        const Row& E = R[ke];
        if constexpr (!Up) {
            const FaceDiff dDE = face_diff(D.h, E.h, D.b, E.b, D.hu, E.hu, D.hv, E.hv);
            Y[kd] = construct_d<Textbook>(D.h, D.b, D.hu, D.hv, dDE, DY, E.by, D.by);
            DY = dDE;
            G[kd] = face_flux<1>(Y[kc].hp, Y[kc].qxp, Y[kc].qyp, Y[kd].hm, Y[kd].qxm, Y[kd].qym, kappa);
        } else {
            const FaceDiff dED = face_diff(E.h, D.h, E.b, D.b, E.hu, D.hu, E.hv, D.hv);
            Y[kd] = construct_d<Textbook>(D.h, D.b, D.hu, D.hv, DY, dED, C.by, D.by);
            DY = dED;
            G[kd] = face_flux<1>(Y[kd].hp, Y[kd].qxp, Y[kd].qyp, Y[kc].hm, Y[kc].qxm, Y[kc].qym, kappa);
        }
        // north / south faces of row j
        const Flux& gn = Up ? G[kc] : G[kd];
        const Flux& gs = Up ? G[kd] : G[kc];
        const Src& src = SR[kc];

        // x pass of row j
        const float bxw = C.bx;
        const float bxe = from_right(bxw);
        // the west face's w difference is the west neighbour's east one
        const FaceDiff dE = face_diff(C.h, from_right(C.h), C.b, from_right(C.b), C.hu, from_right(C.hu), C.hv,
                                      from_right(C.hv));
        const FaceDiff dW{from_left(dE.w), C.hu - from_left(C.hu), C.hv - from_left(C.hv)};
        const FaceStates xs = construct_d<Textbook>(C.h, C.b, C.hu, C.hv, dE, dW, bxe, bxw);
        const Flux fE = face_flux<0>(xs.hp, xs.qxp, xs.qyp, from_right(xs.hm), from_right(xs.qxm),
                                     from_right(xs.qym), kappa);
        const float fw0 = from_left(fE.f0), fw1 = from_left(fE.f1), fw2 = from_left(fE.f2);

        if (out_lane && L.act1) {
            const float dh = flux_divergence(fE.f0, fw0, gn.f0, gs.f0, S.inv_dx, S.inv_dy);
            const float dhu = flux_divergence(fE.f1, fw1, gn.f1, gs.f1, S.inv_dx, S.inv_dy);
            const float dhv = flux_divergence(fE.f2, fw2, gn.f2, gs.f2, S.inv_dx, S.inv_dy);
            mxl = fmaxf(mxl, fE.speed);
            myl = fmaxf(myl, gn.speed);
            // beds of the south / north faces of row j
            const float bys = C.by, byn = Up ? by_behind : D.by;
            wetl |= output_row<Stage, Budget, FM>(d, S, L, c, j, C, src, dh, dhu, dhv, bxw, bxe, bys, byn,
                                                  fE.f0, fw0, gn.f0, gs.f0, acc);
        }
        j += step;
        return j == jstop;
    };
    if (j != jstop) {
        if constexpr (phased_rows<FM>()) {
            for (;;) {
                if (row(Phase<0>{})) break;
                if (row(Phase<1>{})) break;
                if (row(Phase<2>{})) break;
            }
        } else {
            // one phase, the window advanced by copies
            for (;;) {
                if (row(Phase<0>{})) break;
                const float by_c = R[0].by;   // the row behind the next output row
                R[0] = R[1];
                R[1] = R[2];
                R[2].by = by_c;
                W[2] = W[0];
                SR[0] = SR[1];
                Y[0] = Y[1];
                G[0] = G[1];
            }
        }
    }
    flush_tiles<Stage>(d, s, rr, L, out_lane, mxl, myl, wetl, lane, wv);
    if constexpr (Budget) {
        Scalars* sc = d.sc;
        if constexpr (Stage == 1) {
            wave_add_f64(&sc->b_flux1, acc.out, lane);
        } else {
            wave_add_f64(&sc->b_out, acc.out, lane);
            wave_add_f64(&sc->b_exch, acc.exch, lane);
            if (S.fl & FM & RIFT_RAIN_GRID) wave_add_f64(&sc->b_rain_grid, acc.rain, lane);
            if (S.fl & FM & (RIFT_DRAIN_AVG | RIFT_DRAIN_GRID)) wave_add_f64(&sc->b_drain, acc.drain, lane);
        }
    }
    return wv;
}

// Work partition.  The launch holds exactly the waves that fit on the
// device at once, so that every wave runs in a single round: units handed
// out one per wave leave a second, nearly empty round whose lone waves are
// latency-bound.  Workgroup b runs on XCD b % 8 (each with its own L2); with
// Dev::band the waves are dealt so that XCD x holds the x-th eighth of the
// work order below.
//
// Every unit active (dense): each strip is cut into m equal row shares, one
// per wave, ordered share-major, so the waves marching the same rows of
// neighbouring strips are neighbours in the order, run in step and share
// their halo columns in L2.  Even shares march down, odd shares up: two
// shares meet at a boundary either both at their first rows or both at
// their last, so the halo rows around it are read once from HBM.
//
// Otherwise (sparse): full quads (four neighbouring strips all active in a
// tile row) go to workgroups, the other active units to single waves, the
// workgroups split between the two in proportion to their rows.  Each list
// forms a sequence of rows; workgroup (wave) i marches the i-th equal share
// of at least Dev::min_rows rows, downwards, continuing from one entry into
// the next while they are vertically adjacent.  The four waves of a quad's
// workgroup march the same rows of its four strips side by side on one CU,
// so the 128-byte lines two neighbouring strips' windows share (a strip's
// 64-column window starts 240 bytes after the last one's) come from HBM
// once: dealt one strip per wave, every wave read 2.8 lines a row for the
// 240 bytes of its 60 output columns.  Partial quads stay with single waves:
// a workgroup per quad idled 19 % of its waves on the coastal-surge
// workload at 3 h and ran 9 % slower despite 21 % less traffic.
// dt formed by the stage-1 march (defined with the dt kernel below).
__device__ __forceinline__ void march1_dt(const Dev& d, float2 wv);

// Returns the largest tile maxima this wave flushed (uniform).
template <int Stage, bool Textbook, bool Budget, int Kind>
__device__ __forceinline__ float2 march_body(const Dev& d)
{
    constexpr unsigned FM = kind_mask(Kind);
    float2 wv = make_float2(0.f, 0.f);
    auto keep = [&](float2 v) {
        wv.x = fmaxf(wv.x, v.x);
        wv.y = fmaxf(wv.y, v.y);
    };
    const int lane = threadIdx.x & 63;
    // the wavefront index is uniform across the wave: in a scalar register,
    // every unit and row index below is scalar too (SALU address math,
    // scalar branches)
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // stage 1 runs before the dt kernel moves the unit count
    const int nunits = min((Stage == 1) ? d.sc->n_units_acc : d.sc->n_units, d.nunits);
    if (nunits <= 0) return wv;

    const int G = (int)gridDim.x;
    int pos = (int)blockIdx.x;
    if (d.band) {
        const int nb = min(G, 8), x = (int)blockIdx.x % nb, q = G / nb, r = G % nb;
        pos = x * q + min(x, r) + (int)blockIdx.x / nb;
    }
    const int W = 4 * G;
    const int wp = 4 * pos + wid;

    StepScalars S;
    S.dt = (float)d.sc->dt;
    S.inv_dt = 1.f / S.dt;
    S.dtp = (float)d.sc->dt_prev;
    S.inv_dtp = 1.f / S.dtp;
    S.kappa = d.kappa;
    S.fl = d.flags;
    S.rain_rate = d.sc->rain_rate;
    S.drain_rate = d.sc->drain_rate;
    S.src_rate = d.sc->src_cell_rate;
    S.tnow = (float)d.sc->t;
    S.inv_dx = d.inv_dx;
    S.inv_dy = d.inv_dy;
    surge_scalars(d, S);

    const int rows = d.ny - 4;   // interior rows of a strip
    const int mlim = max(1, rows / max(d.min_rows, 1));
    const int m = min(W / d.nstrips, mlim);
    if (nunits == d.nunits && m >= 2) {
        const int mi = wp / d.nstrips, s = wp - mi * d.nstrips;
        if (mi >= m) return wv;
        const int ja = 2 + (int)((long long)rows * mi / m);
        const int jb = 2 + (int)((long long)rows * (mi + 1) / m);
        if (ja >= jb) return wv;
        if (mi & 1) keep(march_run<Stage, Textbook, true, Budget, FM>(d, S, s, ja, jb, lane, 0, jb - ja, true));
        else keep(march_run<Stage, Textbook, false, Budget, FM>(d, S, s, ja, jb, lane, 0, jb - ja, true));
        return wv;
    }

    // full quads to the first GA workgroups, the other active units to the
    // waves of the rest, in proportion to their rows
    const int nqf = (Stage == 1) ? d.sc->n_quads_acc : d.sc->n_quads;
    const int npart = (Stage == 1) ? d.sc->n_part_acc : d.sc->n_part;
    const long long ua = 4LL * nqf, ub = npart;
    int GA = (int)((G * ua + (ua + ub) / 2) / max(ua + ub, 1LL));
    if (ua > 0) GA = max(GA, 1);
    if (ub > 0) GA = min(GA, G - 1);
    const bool quad = pos < GA;
    const int* list = quad ? d.qlist : d.ulist;
    const int nentries = quad ? nqf : npart;
    const int nwork = quad ? GA : 4 * (G - GA);   // workgroups / waves sharing it
    const int me = quad ? pos : 4 * (pos - GA) + wid;
    const long long total = (long long)nentries * kTileH;
    const long long nshare = max(1LL, min((long long)nwork, total / max(d.min_rows, 1)));
    // shares [i0, i1) of this workgroup / wave (at most one when nshare <= nwork)
    const long long i0 = ((long long)me * nshare + nwork - 1) / nwork;
    const long long i1 = ((long long)(me + 1) * nshare + nwork - 1) / nwork;
    if (i0 >= i1) return wv;
    const long long p0 = total * i0 / nshare, pend = total * i1 / nshare;
    if (p0 >= pend) return wv;
    int k = (int)(p0 / kTileH);
    int a = (int)(p0 - (long long)k * kTileH);
    for (;;) {
        // a maximal run of vertically adjacent entries (units of one strip,
        // or quads of one column of quads)
        const int u = __builtin_amdgcn_readfirstlane(list[k]);
        const int c = u / d.nurows, rr = u - c * d.nurows;
        const int ja = rr * kTileH + 2 + a;
        int last = u, kk = k;
        while ((long long)(kk + 1) * kTileH < pend) {
            const int un = __builtin_amdgcn_readfirstlane(list[kk + 1]);
            if (un != last + 1 || un - c * d.nurows <= 0 || un - c * d.nurows >= d.nurows) break;
            last = un;
            ++kk;
        }
        const int rlast = last - c * d.nurows;
        const long long pe = min(pend, (long long)(kk + 1) * kTileH);
        const int jb = min(rlast * kTileH + 2 + (int)(pe - (long long)kk * kTileH), d.ny - 2);
        // a quad's wave w marches strip 4c + w
        const int s = quad ? 4 * c + wid : c;
        const int sdone = (int)((long long)k * kTileH + a - p0);
        if (ja < jb)
            keep(march_run<Stage, Textbook, false, Budget, FM>(d, S, s, ja, jb, lane, sdone, (int)(pend - p0)));
        k = kk + 1;
        a = 0;
        if ((long long)k * kTileH >= pend) break;
    }
    return wv;
}

template <int Stage, bool Textbook, int Waves, bool Budget, int Kind>
__global__ __launch_bounds__(256, Waves) void k_march(Dev d)
{
    const float2 wv = march_body<Stage, Textbook, Budget, Kind>(d);
    if constexpr (Stage == 1) {
        if (d.dt_in_march) march1_dt(d, wv);
    }
}

// ---------------------------------------------------------------------------
// canonical layout for the per-phase kernels and transfers: pending flips
// applied, every tile's state gathered into h/hu/hv, selects cleared.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_apply_flips(Dev d)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < d.ntiles && d.flip[t]) {
        d.sel[t] ^= 1;
        d.flip[t] = 0;
    }
}

__global__ __launch_bounds__(256) void k_gather(Dev d, int rows)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)rows * d.P;
    if (k >= total) return;
    const int j = (int)(k / d.P), i = (int)(k % d.P);
    const int ic = min(max(i, 2), d.nx - 3), jc = min(max(j, 2), d.ny - 3);
    const int T = ((jc - 2) / kTileH) * d.nbx + (ic - 2) / kTileW;
    if (d.sel[T]) {
        d.h[k] = d.h0[k];
        d.hu[k] = d.hu0[k];
        d.hv[k] = d.hv0[k];
    }
}

__global__ __launch_bounds__(256) void k_clear_sel(Dev d)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < d.ntiles) d.sel[t] = 0;
}

// dt of the fused step: max over every tile of the stage-1 maxima (active
// tiles) or the stage-2 maxima of the last step a tile was active (others),
// then ComputeTimestep's arithmetic.  The same reduction counts the active
// tiles and their interior cells.  The dt kernel (RIFT_OPT_DT_KERNEL 1, and
// ensembles) reduces kDtTilesPerBlock consecutive tiles per workgroup (four
// independent loads per lane in flight) into a dt slot (dt_slot_add); by
// default grow and the stage-1 march add into the slots instead
// (march1_dt).  The last workgroup to finish (ticket) reduces the slots,
// forms dt, moves the unit count to stage 2 and evaluates the next step's
// forcing at the new time (UpdateSource runs at the start of the next
// RunSimulation, simulator.cu:780-784, with the t of this step).  Few
// workgroups meet on one address: with one atomic per workgroup and
// quantity on a single line, the 300k tiles of an 8192^2 domain took 41 us
// over 32 workgroups.
constexpr int kDtTilesPerBlock = 1024;

__host__ __device__ inline int dt_blocks(int ntiles)
{
    const int b = (ntiles + kDtTilesPerBlock - 1) / kDtTilesPerBlock;
    return b < 1 ? 1 : (b > kDtMaxBlocks ? kDtMaxBlocks : b);
}

// Fixed-order reduction of (max, max, sum, sum) over a workgroup of 256:
// shuffles within each wave, then lane 0 of the workgroup combines the four
// waves in order.  Returns the result in thread 0.
struct DtRed {
    float vx, vy;
    unsigned long long cells;
    int nact;
};
__device__ __forceinline__ DtRed dt_block_reduce(DtRed r, DtRed* lds4)
{
    for (int off = 32; off > 0; off >>= 1) {
        r.vx = fmaxf(r.vx, __shfl_xor(r.vx, off));
        r.vy = fmaxf(r.vy, __shfl_xor(r.vy, off));
        r.cells += __shfl_xor(r.cells, off);
        r.nact += __shfl_xor(r.nact, off);
    }
    if ((threadIdx.x & 63) == 0) lds4[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        r = lds4[0];
        for (int w = 1; w < 4; ++w) {
            r.vx = fmaxf(r.vx, lds4[w].vx);
            r.vy = fmaxf(r.vy, lds4[w].vy);
            r.cells += lds4[w].cells;
            r.nact += lds4[w].nact;
        }
    }
    return r;
}

// Per-XCD ticket x of the dt reductions: a counter on its own 128-byte line.
__device__ __forceinline__ unsigned* dt_xticket(const Dev& d, int x)
{
    return d.dtk + 32 * x;
}

// The dt reductions of the fused step (the dt kernel's, or the stage-1
// march's) meet in kDtSlots slots of Dev::gs, one 128-byte line each: a
// workgroup adds its maxima and counts into slot blockIdx % kDtSlots by
// atomics and takes a two-level ticket (the workgroups of each XCD,
// blockIdx % 8, count on a line of their own, the last of each on the
// scalars' ticket: same-address atomics serialise, ~300 on one line took
// ~6 us); the last workgroup of the grid reads and clears the slots, one per
// thread.  Every value crossing workgroups travels in read-modify-write
// atomics, and a workgroup takes its ticket only after its additions have
// returned: no device-scope fence, which would write the XCD's L2 back in
// every workgroup.  Thread 0 of each workgroup; true in the last one.  The
// ticket counters are back at zero once the tail has run (the tail clears
// the scalars' ticket).
__device__ __forceinline__ bool dt_slot_add(const Dev& d, const DtRed& r)
{
    unsigned* g = d.gs + 32 * (blockIdx.x % kDtSlots);
    const unsigned o1 = r.vx > 0.f ? atomicMax(g, __float_as_uint(r.vx)) : 0u;
    const unsigned o2 = r.vy > 0.f ? atomicMax(g + 1, __float_as_uint(r.vy)) : 0u;
    const unsigned o3 = r.nact ? atomicAdd(g + 2, (unsigned)r.nact) : 0u;
    const unsigned long long o4 = r.nact ? atomicAdd(reinterpret_cast<unsigned long long*>(g + 4), r.cells) : 0ull;
    // the ticket is taken once the additions have returned
    asm volatile("" ::"v"(o1), "v"(o2), "v"(o3), "v"(o4) : "memory");
    const int G = (int)gridDim.x, ng = min(G, 8), x = (int)blockIdx.x % ng;
    const int nx = G / ng + (x < G % ng ? 1 : 0);
    unsigned* xt = dt_xticket(d, x);
    if (atomicAdd(xt, 1u) != (unsigned)(nx - 1)) return false;
    atomicExch(xt, 0u);
    return atomicAdd(&d.sc->ticket, 1u) == (unsigned)(ng - 1);
}

// Slot `threadIdx.x` of the dt slots, read and cleared (the last workgroup,
// every thread; kDtSlots == its threads).
__device__ __forceinline__ DtRed dt_slot_take(const Dev& d)
{
    static_assert(kDtSlots == 256, "one dt slot per thread of the last workgroup");
    unsigned* g = d.gs + 32 * threadIdx.x;
    const unsigned bx = atomicExch(g, 0u), by = atomicExch(g + 1, 0u), na = atomicExch(g + 2, 0u);
    const unsigned long long nc = atomicExch(reinterpret_cast<unsigned long long*>(g + 4), 0ull);
    return DtRed{__uint_as_float(bx), __uint_as_float(by), nc, (int)na};
}

// Copy of the device scalars into LDS (every thread of the workgroup takes
// part; visible after the next barrier).
__device__ __forceinline__ void stage_scalars(const Dev& d, Scalars& ssc)
{
    constexpr int kWords = (int)(sizeof(Scalars) / sizeof(unsigned));
    static_assert(sizeof(Scalars) % sizeof(unsigned) == 0, "Scalars is copied in 32-bit words");
    for (int k = threadIdx.x; k < kWords; k += blockDim.x)
        reinterpret_cast<unsigned*>(&ssc)[k] = reinterpret_cast<const unsigned*>(d.sc)[k];
}

// The tail of the dt reduction, on the one workgroup that completes it:
// `tot` (thread 0) holds the maxima over every tile and the counts of the
// active tiles and cells; ssc is the scalars' copy in LDS, written back at
// the end.  ComputeTimestep's arithmetic, the unit counts moved to stage 2,
// and the next step's forcing at the new time (UpdateSource runs at the
// start of the next RunSimulation, simulator.cu:780-784, with the t of this
// step).
__device__ __forceinline__ void dt_tail(const Dev& d, const DtRed& tot, Scalars& ssc)
{
    __shared__ bool ended;
    __shared__ float s_tf;
    __shared__ int s_match[3];
    const int tid = threadIdx.x;
    Scalars* sc = d.sc;
    constexpr int kWords = (int)(sizeof(Scalars) / sizeof(unsigned));
    if (tid == 0) {
        ssc.ticket = 0u;
        ended = ssc.finished != 0;
        if (ended) {
            // an ended domain keeps its time; the step's counters are cleared
            ssc.n_units = 0;
            ssc.n_units_acc = 0;
            ssc.n_quads = ssc.n_quads_acc = 0;
            ssc.n_part = ssc.n_part_acc = 0;
            ssc.next_valid = 0;
            ssc.b_flux1 = 0.0;
        } else {
            ssc.n_active = tot.nact;
            ssc.wet_cells = tot.cells;
            ssc.wcu_total += tot.cells;
            finish_timestep(d, tot.vx, tot.vy, &ssc);
            s_tf = (float)ssc.t;
        }
        s_match[0] = s_match[1] = s_match[2] = -1;
    }
    __syncthreads();
    if (ended) {
        for (int k = tid; k < kWords; k += blockDim.x)
            reinterpret_cast<unsigned*>(sc)[k] = reinterpret_cast<const unsigned*>(&ssc)[k];
        return;
    }
    // the next step's forcing: the workgroup finds the interval of each
    // series (the last match decides, as in InterpolateRate's loop)
    const unsigned fl = d.flags;
    const bool use[3] = {(fl & RIFT_RAIN_AVG) != 0, (fl & RIFT_DRAIN_AVG) != 0, (fl & RIFT_HYDROGRAPH) != 0};
    const float tf = s_tf;
    for (int k = 0; k < 3; ++k) {
        if (!use[k]) continue;
        const SeriesTable& st = d.ser[k];
        const int n = st.n;
        for (int i = tid; i < n - 1; i += blockDim.x)
            if (tf >= st.time[i] && tf < st.time[i + 1]) atomicMax(&s_match[k], i);
    }
    __syncthreads();
    if (tid == 0) {
        if (fl & RIFT_MASS_BUDGET) {
            ssc.b_out += 0.5 * ssc.dt * ssc.b_flux1;
            ssc.b_flux1 = 0.0;
        }
        ssc.n_units = ssc.n_units_acc;
        ssc.n_units_acc = 0;
        ssc.n_quads = ssc.n_quads_acc;
        ssc.n_quads_acc = 0;
        ssc.n_part = ssc.n_part_acc;
        ssc.n_part_acc = 0;
        ssc.ustep += 1;
        const int match[3] = {s_match[0], s_match[1], s_match[2]};
        const Forcing f = forcing_at(d, &ssc, match);
        ssc.nxt_rain = f.rain;
        ssc.nxt_drain = f.drain;
        ssc.nxt_hydro = f.hydro;
        ssc.nxt_cell = f.cell;
        ssc.next_valid = 1;
    }
    __syncthreads();
    for (int k = tid; k < kWords; k += blockDim.x)
        reinterpret_cast<unsigned*>(sc)[k] = reinterpret_cast<const unsigned*>(&ssc)[k];
}

// Interior cells of tile t (the cells a step updates when it is active).
__device__ __forceinline__ unsigned long long tile_cells(const Dev& d, int t)
{
    const int bx = t % d.nbx, by = t / d.nbx;
    const int i0 = bx * kTileW + 2, j0 = by * kTileH + 2;
    const int i1 = min(i0 + kTileW, d.nx - 2), j1 = min(j0 + kTileH, d.ny - 2);
    return (i1 > i0 && j1 > j0) ? (unsigned long long)((i1 - i0) * (j1 - j0)) : 0ull;
}

__device__ __forceinline__ void timestep2_body(const Dev& d)
{
    __shared__ DtRed red4[4];
    __shared__ bool last;
    DtRed r{0.f, 0.f, 0ull, 0};
    const int tid = threadIdx.x;
    // the tail runs on a copy of the device scalars in LDS, staged by every
    // workgroup up front (in flight with its tile loads; only the last one
    // uses it) and written back by the last one: no round trip in the tail.
    // Nothing changes the scalars while this kernel runs but the ticket,
    // which the tail clears.
    __shared__ Scalars ssc;
    stage_scalars(d, ssc);
    static_assert(kDtTilesPerBlock == 4 * 256, "four consecutive tiles per lane");
    // four consecutive tiles per lane: one 32-bit load of the activity bytes
    // and one 128-bit load of each maxima array, all in flight together (the
    // per-tile arrays carry 64 bytes of padding, so the last lane's vector
    // reads stay inside the allocation)
    for (int base = blockIdx.x * kDtTilesPerBlock; base < d.ntiles; base += gridDim.x * kDtTilesPerBlock) {
        const int t0 = base + 4 * tid;
        if (t0 < d.ntiles) {
            const uchar4 a4 = *reinterpret_cast<const uchar4*>(d.act + t0);
            const float4 x1 = *reinterpret_cast<const float4*>(d.mx1 + t0);
            const float4 x0 = *reinterpret_cast<const float4*>(d.mx + t0);
            const float4 y1 = *reinterpret_cast<const float4*>(d.my1 + t0);
            const float4 y0 = *reinterpret_cast<const float4*>(d.my + t0);
            const bool a[4] = {a4.x != 0, a4.y != 0, a4.z != 0, a4.w != 0};
            const float ax1[4] = {x1.x, x1.y, x1.z, x1.w}, ax0[4] = {x0.x, x0.y, x0.z, x0.w};
            const float ay1[4] = {y1.x, y1.y, y1.z, y1.w}, ay0[4] = {y0.x, y0.y, y0.z, y0.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int t = t0 + k;
                if (t >= d.ntiles) break;
                r.vx = fmaxf(r.vx, a[k] ? ax1[k] : ax0[k]);
                r.vy = fmaxf(r.vy, a[k] ? ay1[k] : ay0[k]);
                if (a[k]) {
                    r.cells += tile_cells(d, t);
                    ++r.nact;
                }
            }
        }
    }
    r = dt_block_reduce(r, red4);
    if (tid == 0) last = dt_slot_add(d, r);
    __syncthreads();
    if (!last) return;
    r = dt_slot_take(d);
    __syncthreads();   // red4 is reused
    r = dt_block_reduce(r, red4);
    dt_tail(d, r, ssc);
}

// dt formed by the stage-1 march itself (Dev::dt_in_march; single domain):
// the march's waves keep the largest maxima they flush into the active
// tiles; each workgroup takes the largest of its four waves into dt slot
// blockIdx % kDtSlots (atomic maxima), where grow left the stale maxima of
// the tiles it found inactive and the active tile and cell counts; the last
// workgroup to finish (dt_slot_add) reads and clears the slots, one per
// thread, and runs the dt kernel's tail.  The same maxima and counts as the
// dt kernel's pass over every tile: every active tile's stage-1 maxima are
// the largest of the values the waves flushed into it (grow cleared them),
// and an inactive tile contributes its stale maxima.  Saves the dt kernel's
// launch and its pass over the tile arrays.  Every value crossing
// workgroups travels in read-modify-write atomics (dt_slot_add), and every
// wave waits for its own memory operations (the budget atomics among them)
// before its workgroup takes the ticket.  The tail's loads (slots, scalars,
// the boundary outflow) are issued together.
__device__ __forceinline__ void march1_dt(const Dev& d, float2 wv)
{
    __shared__ float s_w[4][2];
    __shared__ bool s_last;
    __shared__ DtRed red4[4];
    __shared__ Scalars ssc;
    const int tid = threadIdx.x;
    __builtin_amdgcn_s_waitcnt(0);   // this wave's memory operations completed
    if ((tid & 63) == 0) {
        s_w[tid >> 6][0] = wv.x;
        s_w[tid >> 6][1] = wv.y;
    }
    __syncthreads();
    if (tid == 0) {
        DtRed r{s_w[0][0], s_w[0][1], 0ull, 0};
        for (int w = 1; w < 4; ++w) {
            r.vx = fmaxf(r.vx, s_w[w][0]);
            r.vy = fmaxf(r.vy, s_w[w][1]);
        }
        s_last = dt_slot_add(d, r);
    }
    __syncthreads();
    if (!s_last) return;
    DtRed r = dt_slot_take(d);
    // the scalars did not change in this launch but for the ticket and the
    // stage-1 boundary outflow (atomics of every workgroup, read here)
    double bf = 0.0;
    if (tid == 0 && (d.flags & RIFT_MASS_BUDGET)) bf = atomicAdd(&d.sc->b_flux1, 0.0);
    stage_scalars(d, ssc);
    r = dt_block_reduce(r, red4);   // its barrier also publishes ssc
    if (tid == 0 && (d.flags & RIFT_MASS_BUDGET)) ssc.b_flux1 = bf;
    dt_tail(d, r, ssc);
}

// Single domain, one thread per tile: grow; when the stage-1 march forms dt
// (Dev::dt_in_march), also each wave's share of the dt reduction over its
// tiles (stale maxima of the inactive ones, counts of the active ones and
// their interior cells), added by atomics into dt slot (wave index) %
// kDtSlots (one 128-byte line each, so that few waves meet on a line), which
// the march's dt tail reads and clears.  A wave with nothing to add (no
// active tile, no stale maxima: most of a sparse domain) adds nothing.
__global__ __launch_bounds__(256) void k_grow_tiles(Dev d)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool a = grow_tiles_body<false>(d, t);
    if (!d.dt_in_march) return;   // uniform
    float vx = 0.f, vy = 0.f;
    unsigned nact = 0u, cells = 0u;
    if (a) {
        cells = (unsigned)tile_cells(d, t);
        nact = 1u;
    } else if (t < d.ntiles) {
        vx = d.mx[t];
        vy = d.my[t];
    }
    if (__ballot(nact != 0u || vx > 0.f || vy > 0.f) == 0ull) return;   // uniform
    for (int off = 32; off > 0; off >>= 1) {
        vx = fmaxf(vx, __shfl_xor(vx, off));
        vy = fmaxf(vy, __shfl_xor(vy, off));
        nact += __shfl_xor(nact, off);
        cells += __shfl_xor(cells, off);
    }
    if ((threadIdx.x & 63) == 0) {
        unsigned* g = d.gs + 32 * ((t >> 6) % kDtSlots);
        if (vx > 0.f) atomicMax(g, __float_as_uint(vx));
        if (vy > 0.f) atomicMax(g + 1, __float_as_uint(vy));
        if (nact) {
            atomicAdd(g + 2, nact);
            atomicAdd(reinterpret_cast<unsigned long long*>(g + 4), (unsigned long long)cells);
        }
    }
}

__global__ __launch_bounds__(1024) void k_compact_quads(Dev d)
{
    compact_quads_body(d, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_timestep2(Dev d)
{
    timestep2_body(d);
}

void launch_timestep2(const Dev& d, hipStream_t s)
{
    hipLaunchKernelGGL(k_timestep2, dim3(dt_blocks(d.ntiles)), dim3(256), 0, s, d);
}

void launch_grow2(const Dev& d, hipStream_t s, bool dt_in_march)
{
    Dev arg = d;
    arg.dt_in_march = dt_in_march ? 1 : 0;
    hipLaunchKernelGGL(k_grow_tiles, dim3((d.ntiles + 255) / 256), dim3(256), 0, s, arg);
    hipLaunchKernelGGL(k_compact_quads, dim3((d.nunits + 1023) / 1024), dim3(1024), 0, s, d);
}

// Launch variants (all give identical results; they differ in register
// budget and work order): bit 0 = stage 1 launch bounds of 5 instead of 4
// waves per SIMD, bit 1 = stage 2 at 4 instead of 3 (the lean phased stage
// 2 needs 131 VGPRs, so 3 is the default; the lean stage 1 fits 87 and runs
// 5 resident either way); the dealing of work shares and their least size
// are Dev::band and Dev::min_rows.  The grid is
// the number of workgroups resident at once (march_resident_blocks).
// The lean instantiation serves the non-textbook domains that use none of
// the features outside kLeanMask, the coastal one those within kCoastMask
// whose surge (if any) comes from the device series.
bool march_lean(unsigned flags)
{
    return !(flags & RIFT_TEXTBOOK) && (flags & ~kLeanMask) == 0u;
}

int march_kind(const Dev& d)
{
    if (march_lean(d.flags)) return kKindLean;
    if (!(d.flags & RIFT_TEXTBOOK) && (d.flags & ~kCoastMask) == 0u &&
        (!(d.flags & RIFT_SURGE_GRID) || d.sg_series))
        return kKindCoast;
    return kKindFull;
}

template <int Stage, int Waves>
static const void* march_fn(int kind, bool tb, bool bud)
{
    if (kind == kKindLean)
        return bud ? (const void*)k_march<Stage, false, Waves, true, kKindLean>
                   : (const void*)k_march<Stage, false, Waves, false, kKindLean>;
    if (kind == kKindCoast)
        return bud ? (const void*)k_march<Stage, false, Waves, true, kKindCoast>
                   : (const void*)k_march<Stage, false, Waves, false, kKindCoast>;
    return tb ? (bud ? (const void*)k_march<Stage, true, Waves, true, kKindFull>
                     : (const void*)k_march<Stage, true, Waves, false, kKindFull>)
              : (bud ? (const void*)k_march<Stage, false, Waves, true, kKindFull>
                     : (const void*)k_march<Stage, false, Waves, false, kKindFull>);
}

static const void* march_kernel(int stage, int variant, int kind, unsigned flags)
{
    const bool tb = flags & RIFT_TEXTBOOK, bud = flags & RIFT_MASS_BUDGET;
    if (stage == 1) return (variant & 1) ? march_fn<1, 5>(kind, tb, bud) : march_fn<1, 4>(kind, tb, bud);
    return (variant & 2) ? march_fn<2, 4>(kind, tb, bud) : march_fn<2, 3>(kind, tb, bud);
}

void launch_march(const Dev& d, int stage, hipStream_t s, int variant, int blocks, bool dt_in_march)
{
    blocks = std::max(1, blocks);
    Dev arg = d;
    arg.dt_in_march = (stage == 1 && dt_in_march) ? 1 : 0;
    void* args[] = {&arg};
    (void)hipLaunchKernel(march_kernel(stage, variant, march_kind(d), d.flags), dim3(blocks), dim3(256), args,
                          0, s);
}

int march_resident_blocks(const Dev& d, int stage, int variant, int cus)
{
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, march_kernel(stage, variant, march_kind(d), d.flags), 256, 0);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    return per_cu * cus;
}

void launch_canonicalize(const Dev& d, hipStream_t s)
{
    const int tb = (d.ntiles + 255) / 256;
    hipLaunchKernelGGL(k_apply_flips, dim3(tb), dim3(256), 0, s, d);
    const int rows = kTileH * d.nby + 4;
    const long total = (long)rows * d.P;
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, d, rows);
    hipLaunchKernelGGL(k_clear_sel, dim3(tb), dim3(256), 0, s, d);
}

// ---------------------------------------------------------------------------
// Ensembles: many domains (scenarios) advanced together, one launch per
// phase.  devs[] holds every member's Dev in device memory; grow and dt run
// one grid row per member (blockIdx.y); the marching kernels deal the
// members' active rows, in member order, to the resident waves (bands across
// the strips of each fully active member, runs of units of the others), so
// small and large members share the device without a launch each.
// Every member keeps its own time, dt, forcing and tile sets; a member's
// cells see exactly the arithmetic of a single-domain step.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_grow_tiles_ens(const Dev* __restrict__ devs)
{
    grow_tiles_body<true>(devs[blockIdx.y], blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_compact_units_ens(const Dev* __restrict__ devs)
{
    compact_units_body(devs[blockIdx.y], blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_timestep2_ens(const Dev* __restrict__ devs)
{
    timestep2_body(devs[blockIdx.y]);
}

__device__ __forceinline__ StepScalars step_scalars(const Dev& d)
{
    StepScalars S;
    S.dt = (float)d.sc->dt;
    S.inv_dt = 1.f / S.dt;
    S.dtp = (float)d.sc->dt_prev;
    S.inv_dtp = 1.f / S.dtp;
    S.kappa = d.kappa;
    S.fl = d.flags;
    S.rain_rate = d.sc->rain_rate;
    S.drain_rate = d.sc->drain_rate;
    S.src_rate = d.sc->src_cell_rate;
    S.tnow = (float)d.sc->t;
    S.inv_dx = d.inv_dx;
    S.inv_dy = d.inv_dy;
    surge_scalars(d, S);
    return S;
}

template <int Stage>
__device__ __forceinline__ int member_units(const Dev& d)
{
    return min((Stage == 1) ? d.sc->n_units_acc : d.sc->n_units, d.nunits);
}

// Unit counts of members [c0, c0+64), one member per lane (loads in
// parallel), with their inclusive prefix sum across the wave.
template <int Stage>
__device__ __forceinline__ void member_chunk(const Dev* __restrict__ devs, int nmem, int c0, int lane,
                                             int& cnt, long long& incl)
{
    const int m = c0 + lane;
    cnt = (m < nmem) ? member_units<Stage>(devs[m]) : 0;
    incl = cnt;
    for (int off = 1; off < 64; off <<= 1) {
        const long long o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
    }
}

// Shares of one member for bands of hb rows.  A dense member is dealt like
// a dense domain in k_march: bands of hb rows across all strips (the last
// one shorter), one strip of one band per share, in band-major order, so
// that consecutive shares march the same rows of neighbouring strips at the
// same time and share their halo columns in L2 (dealt strip-major, the same
// member ran 40 % slower); bands alternate direction.  A sparse member's
// active units (strip-major) are cut into runs of hb / 14 units.
__device__ __forceinline__ int member_run_units(int hb)
{
    return max(1, hb / kTileH);
}
__device__ __forceinline__ int member_shares(const Dev& d, int cnt, int hb)
{
    if (cnt == 0) return 0;
    if (cnt == d.nunits) return ((d.ny - 4 + hb - 1) / hb) * d.nstrips;
    const int bu = member_run_units(hb);
    return (cnt + bu - 1) / bu;
}

// Unit and share counts of members [c0, c0+64), one member per lane, with
// the inclusive prefix sum of the shares across the wave.
template <int Stage>
__device__ __forceinline__ void member_chunk_sh(const Dev* __restrict__ devs, int nmem, int c0, int lane, int B,
                                                int& cnt, int& sh, long long& incl)
{
    const int m = c0 + lane;
    cnt = (m < nmem) ? member_units<Stage>(devs[m]) : 0;
    sh = (m < nmem) ? member_shares(devs[m], cnt, B) : 0;
    incl = sh;
    for (int off = 1; off < 64; off <<= 1) {
        const long long o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
    }
}

template <int Stage, bool Textbook, int Waves, bool Budget, bool Lean>
__global__ __launch_bounds__(256, Waves) void k_march_ens(const Dev* __restrict__ devs, int nmem, int min_rows)
{
    constexpr unsigned FM = Lean ? kLeanEnsMask : kFullMask;
    const int lane = threadIdx.x & 63;
    const int G = (int)gridDim.x;
    const int nb = min(G, 8), x = (int)blockIdx.x % nb;
    const long long W = 4LL * G;   // waves
    long long units = 0;            // active units over all members
    for (int c = 0; c < nmem; c += 64) {
        int cc;
        long long ic;
        member_chunk<Stage>(devs, nmem, c, lane, cc, ic);
        units += __shfl(ic, 63);
    }
    if (units == 0) return;
    // band height in rows: the total rows over the waves, raised in steps
    // of 1/16 while the shares (rounded up per member) outnumber the waves,
    // so that about one band of one strip falls to each wave.  Where that
    // does not fit (members a little taller than one band each), a quarter
    // of it: each wave then takes 4 or 5 shares, instead of some waves two
    // whole ones and the rest one.
    const int lo_rows = max(min_rows, 1);
    const long long rows_all = units * kTileH;
    int B = (int)max((long long)lo_rows, (rows_all + W - 1) / W);
    long long total = 0;
    for (int attempt = 0;; ++attempt) {
        total = 0;
        for (int c = 0; c < nmem; c += 64) {
            int cc, sc;
            long long ic;
            member_chunk_sh<Stage>(devs, nmem, c, lane, B, cc, sc, ic);
            total += __shfl(ic, 63);
        }
        if (total <= W) break;
        if (attempt == 8) {
            B = (int)max((long long)lo_rows, (rows_all + 4 * W - 1) / (4 * W));
            total = 0;
            for (int c = 0; c < nmem; c += 64) {
                int cc, sc;
                long long ic;
                member_chunk_sh<Stage>(devs, nmem, c, lane, B, cc, sc, ic);
                total += __shfl(ic, 63);
            }
            break;
        }
        B += max(1, B / 16);
    }
    // static dealing in band-major order: wave i takes shares [i T / W, (i + 1) T / W)
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int q = G / nb, r = G % nb;
    const int pos = x * q + min(x, r) + (int)blockIdx.x / nb;
    const long long wp = 4LL * pos + wid;
    const long long i0 = wp * total / W, i1 = (wp + 1) * total / W;
    if (i0 >= i1) return;

    int c0 = 0, cnt = 0, sh = 0;
    long long cbase = 0, incl = 0;
    member_chunk_sh<Stage>(devs, nmem, 0, lane, B, cnt, sh, incl);
    int smem = -1;
    StepScalars S{};
    for (long long i = i0; i < i1; ++i) {
        for (;;) {   // the 64-member chunk holding share i (i never decreases)
            const long long ctot = __shfl(incl, 63);
            if (i < cbase + ctot) break;
            cbase += ctot;
            c0 += 64;
            member_chunk_sh<Stage>(devs, nmem, c0, lane, B, cnt, sh, incl);
        }
        const int l = (int)__builtin_ctzll(__ballot(cbase + incl > i));
        const int mem = __builtin_amdgcn_readfirstlane(c0 + l);
        const int mcnt = __builtin_amdgcn_readfirstlane(__shfl(cnt, l));
        const int ls = (int)(i - (cbase + __shfl(incl, l) - __shfl(sh, l)));
        const Dev& d = devs[mem];
        if (mem != smem) {
            S = step_scalars(d);
            smem = mem;
        }
        // pacing: the wave's shares count B rows each
        const int stotal = (int)(i1 - i0) * B, sdone = (int)(i - i0) * B;
        if (mcnt == d.nunits) {
            const int band = ls / d.nstrips, s = ls - band * d.nstrips;
            const int ja = 2 + band * B, jb = min(ja + B, d.ny - 2);
            if (ja >= jb) continue;
            if (band & 1) march_run<Stage, Textbook, true, Budget, FM>(d, S, s, ja, jb, lane, sdone, stotal);
            else march_run<Stage, Textbook, false, Budget, FM>(d, S, s, ja, jb, lane, sdone, stotal);
        } else {
            // units [ls bu, (ls + 1) bu) of the list: runs of vertically
            // adjacent units of one strip
            const int bu = member_run_units(B);
            int kk = ls * bu;
            const int kend = min(kk + bu, mcnt);
            while (kk < kend) {
                const int kdone = (kk - ls * bu) * kTileH;
                const int u = __builtin_amdgcn_readfirstlane(d.ulist[kk]);
                const int s = u / d.nurows, rr = u - s * d.nurows;
                int last = u;
                while (kk + 1 < kend) {
                    const int un = __builtin_amdgcn_readfirstlane(d.ulist[kk + 1]);
                    if (un != last + 1 || un - s * d.nurows >= d.nurows) break;
                    last = un;
                    ++kk;
                }
                const int ja = rr * kTileH + 2;
                const int jb = min((last - s * d.nurows + 1) * kTileH + 2, d.ny - 2);
                if (ja < jb)
                    march_run<Stage, Textbook, false, Budget, FM>(d, S, s, ja, jb, lane, sdone + kdone, stotal);
                ++kk;
            }
        }
    }
}

void launch_grow2_ens(const Dev* devs, int nmem, int max_tiles, int max_units, hipStream_t s)
{
    hipLaunchKernelGGL(k_grow_tiles_ens, dim3((max_tiles + 255) / 256, nmem), dim3(256), 0, s, devs);
    hipLaunchKernelGGL(k_compact_units_ens, dim3((max_units + 255) / 256, nmem), dim3(256), 0, s, devs);
}

void launch_timestep2_ens(const Dev* devs, int nmem, int max_tiles, hipStream_t s)
{
    hipLaunchKernelGGL(k_timestep2_ens, dim3(dt_blocks(max_tiles), nmem), dim3(256), 0, s, devs);
}

template <int Stage, int Waves, bool Lean>
static const void* march_ens_kernel(bool textbook, bool budget)
{
    if constexpr (Lean)
        return budget ? (const void*)k_march_ens<Stage, false, Waves, true, true>
                      : (const void*)k_march_ens<Stage, false, Waves, false, true>;
    else
        return textbook ? (budget ? (const void*)k_march_ens<Stage, true, Waves, true, false>
                                  : (const void*)k_march_ens<Stage, true, Waves, false, false>)
                        : (budget ? (const void*)k_march_ens<Stage, false, Waves, true, false>
                                  : (const void*)k_march_ens<Stage, false, Waves, false, false>);
}

template <int Stage, int Waves, bool Lean>
static void launch_march_ens_w(const Dev* devs, int nmem, unsigned flags, int min_rows, int blocks,
                               hipStream_t s)
{
    const bool tb = flags & RIFT_TEXTBOOK, bud = flags & RIFT_MASS_BUDGET;
    const dim3 g(blocks), b(256);
    if constexpr (Lean) {
        if (bud) hipLaunchKernelGGL((k_march_ens<Stage, false, Waves, true, true>), g, b, 0, s, devs, nmem, min_rows);
        else hipLaunchKernelGGL((k_march_ens<Stage, false, Waves, false, true>), g, b, 0, s, devs, nmem, min_rows);
    } else if (tb && bud) {
        hipLaunchKernelGGL((k_march_ens<Stage, true, Waves, true, false>), g, b, 0, s, devs, nmem, min_rows);
    } else if (tb) {
        hipLaunchKernelGGL((k_march_ens<Stage, true, Waves, false, false>), g, b, 0, s, devs, nmem, min_rows);
    } else if (bud) {
        hipLaunchKernelGGL((k_march_ens<Stage, false, Waves, true, false>), g, b, 0, s, devs, nmem, min_rows);
    } else {
        hipLaunchKernelGGL((k_march_ens<Stage, false, Waves, false, false>), g, b, 0, s, devs, nmem, min_rows);
    }
}

// Ensembles use stage 1 at 4 waves per SIMD and stage 2 at 3: the kernels
// hold both marching directions and the member lookup, which spill at 5
// and 4.  Lean: every member's flags leave the lean instantiation's
// features unused.
void launch_march_ens(const Dev* devs, int nmem, unsigned flags, bool lean, int stage, int min_rows,
                      int blocks, hipStream_t s)
{
    blocks = std::max(1, blocks);
    if (lean) {
        if (stage == 1) launch_march_ens_w<1, 4, true>(devs, nmem, flags, min_rows, blocks, s);
        else launch_march_ens_w<2, 3, true>(devs, nmem, flags, min_rows, blocks, s);
    } else {
        if (stage == 1) launch_march_ens_w<1, 4, false>(devs, nmem, flags, min_rows, blocks, s);
        else launch_march_ens_w<2, 3, false>(devs, nmem, flags, min_rows, blocks, s);
    }
}

template <int Stage, int Waves, bool Lean>
static int resident_ens(bool textbook, bool budget, int cus)
{
    int per_cu = 0;
    hipError_t e =
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, march_ens_kernel<Stage, Waves, Lean>(textbook, budget), 256, 0);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    return per_cu * cus;
}

int march_ens_resident_blocks(int stage, bool textbook, bool budget, bool lean, int cus)
{
    if (stage == 1)
        return lean ? resident_ens<1, 4, true>(textbook, budget, cus) : resident_ens<1, 4, false>(textbook, budget, cus);
    return lean ? resident_ens<2, 3, true>(textbook, budget, cus) : resident_ens<2, 3, false>(textbook, budget, cus);
}

}  // namespace rift
