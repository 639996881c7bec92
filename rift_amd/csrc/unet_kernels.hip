// unet_kernels.hip -- gfx950 kernels of the U-Net flood surrogate.
//
// The surrogate has no counterpart in the reference.  Activations are NHWC
// bfloat16 with channel counts that are multiples of 32; weights are kept in
// FP32 (the optimizer's master copy) and in bfloat16 for the matrix cores.
//
// Every convolution, its input gradient and the transposed convolution are
// one implicit-GEMM kernel (k_gemm): GEMM row m is a pixel (b, i, j) of a row
// grid, column n an output channel, and the K dimension runs over taps x
// input channels.  A tap reads source pixel (stride*i + ty, stride*j + tx),
// zero outside the source; column n writes output pixel
// (ups*i + sub/ups, ups*j + sub%ups) with sub = n / cout.  The A and B
// fragments of v_mfma_f32_32x32x16_bf16 are 8 consecutive K values of one row
// or column, which in NHWC are 16 contiguous bytes of one pixel, so each lane
// loads its fragments straight from global memory (L1/L2 serve the 9-fold tap
// reuse).  Weight gradients (k_wgrad) reduce over pixels, so both operands
// are transposed through LDS.  Accumulation is FP32 throughout.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "unet_kernels.hpp"

namespace rift_nn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bf2f(bf16 b) { return __uint_as_float((uint32_t)b << 16); }

// Round to nearest even (NaN stays NaN).
__device__ __forceinline__ bf16 f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16)0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}

__device__ __forceinline__ bf16x8 load8(const bf16* p)
{
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0)); }

// ---------------------------------------------------------------------------
// Implicit GEMM: forward convolutions and input gradients.
//
// Block = 4 waves arranged WM x WN; each wave owns RM x RN tiles of 32 x 32.
// Lane (r = lane % 32, kh = lane / 32): A fragment = row r, K values
// [8 kh, 8 kh + 8) of the 16-wide step; B fragment = column r, same K; the
// 16 accumulators of tile (rm, rn) hold rows 8 q + 4 kh + e (acc[4 q + e]) of
// column r.

// Tap patterns: 3x3 (offsets -1..1), 3x3 mirrored (input gradients), one
// tap, and 2x2 at stride 2 (input gradient of the transposed convolution).
enum TapKind { kTap3 = 0, kTap3Mirror = 1, kTap1 = 2, kTap2x2 = 3 };

template <int TK>
struct TapGeo {
    static constexpr int NT = TK == kTap1 ? 1 : (TK == kTap2x2 ? 4 : 9);
    static constexpr int S = TK == kTap2x2 ? 2 : 1;
    __device__ static __forceinline__ void off(int t, int& ty, int& tx)
    {
        if constexpr (TK == kTap3) {
            ty = t / 3 - 1;
            tx = t - 3 * (t / 3) - 1;
        } else if constexpr (TK == kTap3Mirror) {
            ty = 1 - t / 3;
            tx = 1 - (t - 3 * (t / 3));
        } else if constexpr (TK == kTap2x2) {
            ty = t >> 1;
            tx = t & 1;
        } else {
            ty = 0;
            tx = 0;
        }
    }
};

// K advances in groups of 32 = one tap of one 32-channel block of one
// source; a stage holds KC groups.  A stage's A tile (BM pixels x 32 KC) and
// B tile (BN columns x 32 KC) are loaded from global memory into registers
// one stage ahead, stored to LDS (double-buffered), and the waves read their
// MFMA fragments from LDS.  The 16-byte chunks of an LDS row are
// XOR-swizzled by the row (gemm_chunk): the fragment reads of 8 consecutive
// rows and the stores of 2 consecutive rows each cover all 32 banks once.
constexpr int kGroup = 32;

// Element offset of 16-byte chunk c of LDS row r, rows of KC * 32 elements.
template <int KC>
__device__ __forceinline__ int gemm_chunk(int r, int c)
{
    if constexpr (KC == 1) return 8 * (c ^ ((r >> 1) & 3));
    else return 8 * (c ^ (((r & 1) << 2) | ((r >> 1) & 3)));
}

template <int TK, int WM, int WN, int RM, int RN, int KC>
__global__ __launch_bounds__(256) void k_gemm(Gemm g)
{
    using G = TapGeo<TK>;
    constexpr int BM = 32 * WM * RM, BN = 32 * WN * RN;
    static_assert(KC == 1 || KC == 2, "one or two K groups per stage");
    constexpr int LD = KC * kGroup;             // LDS row stride (elements)
    constexpr int PA = BM / 64;                 // A rows per thread and group
    constexpr int PB = (BN + 63) / 64;          // B rows per thread and group
    __shared__ __attribute__((aligned(16))) bf16 sA[2][BM * LD];
    __shared__ __attribute__((aligned(16))) bf16 sB[2][BN * LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int r = lane & 31, kh = lane >> 5;
    const int hw = g.hg * g.wg;
    const int M = g.batch * hw;
    const int bm0 = blockIdx.x * BM, bn0 = blockIdx.y * BN;
    const int m0 = bm0 + wm * 32 * RM;
    const int n0 = bn0 + wn * 32 * RN;
    const int lrow = tid >> 2, lq = (tid & 3) * 8;   // loader: row in a 64-row pass, channel octet

    // loader rows of the A tile: pixel of each
    int pb[PA], pi[PA], pj[PA];
    bool pv[PA];
#pragma unroll
    for (int a = 0; a < PA; ++a) {
        const int m = bm0 + a * 64 + lrow;
        pv[a] = m < M;
        const int mm = pv[a] ? m : 0;
        pb[a] = mm / hw;
        const int rem = mm - pb[a] * hw;
        pi[a] = G::S * (rem / g.wg);
        pj[a] = G::S * (rem - (rem / g.wg) * g.wg);
    }
    const int cin = g.c0 + g.c1;
    const int cpt = cin >> 5;                        // groups per tap
    const int QG = G::NT * cpt;                      // groups
    const int Q = (QG + KC - 1) / KC;                // stages

    uint4 ra[KC][PA], rb[KC][PB];
    auto gload = [&](int q) {
#pragma unroll
        for (int u = 0; u < KC; ++u) {
            const int gq = q * KC + u;
            const bool live = gq < QG;
            const int t = live ? gq / cpt : 0;
            const int k = live ? (gq - t * cpt) * kGroup : 0;   // channel offset in [0, cin)
            int ty, tx;
            G::off(t, ty, tx);
            const bool second = k >= g.c0;
            const bf16* src = second ? g.src1 : g.src0;
            const int cs = second ? g.c1 : g.c0;
            const int c = (second ? k - g.c0 : k) + lq;
#pragma unroll
            for (int a = 0; a < PA; ++a) {
                const int si = pi[a] + ty, sj = pj[a] + tx;
                const bool ok = live && pv[a] && (unsigned)si < (unsigned)g.hs && (unsigned)sj < (unsigned)g.ws;
                const size_t pix = (size_t)((pb[a] * g.hs + si) * g.ws + sj);
                ra[u][a] = ok ? *reinterpret_cast<const uint4*>(src + pix * cs + c) : make_uint4(0, 0, 0, 0);
            }
#pragma unroll
            for (int b = 0; b < PB; ++b) {
                const int n = b * 64 + lrow;
                rb[u][b] = (live && n < BN)
                               ? *reinterpret_cast<const uint4*>(g.w + (size_t)(bn0 + n) * g.K + t * cin + k + lq)
                               : make_uint4(0, 0, 0, 0);
            }
        }
    };
    auto swrite = [&](int buf) {
#pragma unroll
        for (int u = 0; u < KC; ++u) {
            const int c = u * 4 + (lq >> 3);
#pragma unroll
            for (int a = 0; a < PA; ++a) {
                const int row = a * 64 + lrow;
                *reinterpret_cast<uint4*>(&sA[buf][row * LD + gemm_chunk<KC>(row, c)]) = ra[u][a];
            }
#pragma unroll
            for (int b = 0; b < PB; ++b) {
                const int n = b * 64 + lrow;
                if (n < BN) *reinterpret_cast<uint4*>(&sB[buf][n * LD + gemm_chunk<KC>(n, c)]) = rb[u][b];
            }
        }
    };

    f32x16 acc[RM][RN];
#pragma unroll
    for (int a = 0; a < RM; ++a)
#pragma unroll
        for (int b = 0; b < RN; ++b)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[a][b][e] = 0.f;

    gload(0);
    swrite(0);
    __syncthreads();
    for (int q = 0; q < Q; ++q) {
        const int buf = q & 1;
        if (q + 1 < Q) gload(q + 1);
        // a zero-filled group past the end adds nothing
#pragma unroll
        for (int kk = 0; kk < KC * kGroup / 16; ++kk) {
            bf16x8 fa[RM], fb[RN];
#pragma unroll
            for (int a = 0; a < RM; ++a) {
                const int row = wm * 32 * RM + a * 32 + r;
                fa[a] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(
                                                       &sA[buf][row * LD + gemm_chunk<KC>(row, 2 * kk + kh)]));
            }
#pragma unroll
            for (int b = 0; b < RN; ++b) {
                const int row = wn * 32 * RN + b * 32 + r;
                fb[b] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(
                                                       &sB[buf][row * LD + gemm_chunk<KC>(row, 2 * kk + kh)]));
            }
#pragma unroll
            for (int a = 0; a < RM; ++a)
#pragma unroll
                for (int b = 0; b < RN; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
        }
        if (q + 1 < Q) swrite(buf ^ 1);
        __syncthreads();
    }

    // epilogue: bias, ReLU, mask, store
#pragma unroll
    for (int b = 0; b < RN; ++b) {
        const int n = n0 + b * 32 + r;
        if (n >= g.N) continue;
        const int sub = n / g.cout, co = n - sub * g.cout;
        const int su = sub / g.ups, sv = sub - su * g.ups;
        const float bias = g.bias ? g.bias[co] : 0.f;
        bf16* dst = g.dst0;
        int dstride = g.dst1 ? g.dsplit : g.cout, dco = co;
        if (g.dst1 && co >= g.dsplit) {
            dst = g.dst1;
            dstride = g.cout - g.dsplit;
            dco = co - g.dsplit;
        }
        // a lane's rows come in runs of 4 consecutive pixels starting at a
        // multiple of 4: with wg % 4 == 0 they share image and grid row, so
        // the pixel is decomposed once per run
        const bool quads = (g.wg & 3) == 0;
#pragma unroll
        for (int a = 0; a < RM; ++a)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m4 = m0 + a * 32 + 8 * q + 4 * kh;
                int ib4 = 0, rem4 = m4, i4 = 0, j4 = 0;
                if (quads && (g.ups != 1 || g.dstf)) {
                    ib4 = m4 / hw;
                    rem4 = m4 - ib4 * hw;
                    i4 = rem4 / g.wg;
                    j4 = rem4 - i4 * g.wg;
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int m = m4 + e;
                    if (m >= M) continue;
                    long long op = m;           // output pixel (same grid when ups == 1)
                    int ib = ib4, rem = rem4 + e, i = i4, j = j4 + e;
                    if (!quads && (g.ups != 1 || g.dstf)) {
                        ib = m / hw;
                        rem = m - ib * hw;
                        i = rem / g.wg;
                        j = rem - i * g.wg;
                    }
                    if (g.ups != 1) op = ((long long)ib * g.ho + (g.ups * i + su)) * g.wo + (g.ups * j + sv);
                    float v = acc[a][b][4 * q + e] + bias;
                    if (g.relu) v = fmaxf(v, 0.f);
                    if (g.mask && !(bf2f(g.mask[op * g.cout + co]) > 0.f)) v = 0.f;
                    if (g.dstf) {
                        if (co < g.nvalid) g.dstf[((long long)ib * g.nvalid + co) * hw + rem] = v;
                    } else {
                        dst[op * dstride + dco] = f2bf(v);
                    }
                }
            }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[n][k] += sum over pixels p of G(p, n) * A(p, k), with
// A the same tap gather as k_gemm and G the output gradient read through the
// same output mapping.  Block tile 32 (n) x 128 (k), one 32 x 32 tile per
// wave; pixels in chunks of 64, both operands staged transposed in LDS
// (rows of 64 pixels, padded to 72, columns swizzled per row group).  A
// variant storing packed pixel pairs (half the LDS instructions) ran slower.
// Partial sums go to dW with FP32
// atomics, so the summation order varies between runs.  The blocks of the
// first K tile also sum the staged output gradient over pixels: the bias
// gradient (for a transposed conv, every sub-pixel adds to its channel).
constexpr int kWgPix = 64, kWgPad = 72;

__global__ __launch_bounds__(256) void k_wgrad(WGrad w)
{
    __shared__ bf16 sG[32 * kWgPad];
    __shared__ bf16 sA[128 * kWgPad];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 31, kh = lane >> 5;
    const int k0 = blockIdx.x * 128, n0 = blockIdx.y * 32;
    const int hw = w.hg * w.wg, M = w.batch * hw;
    const int cin = w.c0 + w.c1;
    const int px = tid >> 2, cv = tid & 3;
    // pixel columns of LDS row R are XOR-swizzled by 16 x ((R / 8) mod 4):
    // the four channel octets a wave stores for one pixel land in different
    // banks, and 8-pixel fragments stay contiguous
    const int swz_w = 16 * cv, swz_r = 16 * (((threadIdx.x & 31) >> 3) & 3);
    // G column group: the 32 columns share one sub-pixel (cout % 32 == 0)
    const int sub = n0 / w.cout, co0 = n0 - sub * w.cout;
    const int su = sub / w.ups, sv = sub - su * w.ups;
    // K groups of this block: tap, source, channel block
    int kt[4], ksrc[4], kc[4];
    bool kvalid[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k = k0 + q * 32;
        kvalid[q] = k < w.K;
        const int t = kvalid[q] ? k / cin : 0;
        const int c = kvalid[q] ? k - t * cin : 0;
        kt[q] = t;
        ksrc[q] = c >= w.c0;
        kc[q] = ksrc[q] ? c - w.c0 : c;
    }
    const bool mine = k0 + wave * 32 < w.K;
    const bool bias = w.db != nullptr && blockIdx.x == 0;
    float bsum = 0.f;   // channel tid / 8, pixels (tid % 8) * 8 .. + 8 of each chunk
    f32x16 acc;
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.f;

    const int p_begin = blockIdx.z * w.chunk, p_end = min(M, p_begin + w.chunk);
    for (int pc = p_begin; pc < p_end; pc += kWgPix) {
        const int p = pc + px;
        const bool pvalid = p < p_end;
        const int pp = pvalid ? p : 0;
        const int ib = pp / hw, rem = pp - ib * hw;
        const int i = rem / w.wg, j = rem - i * w.wg;
        {
            bf16x8 v = zero8();
            if (pvalid) {
                const long long op = ((long long)ib * w.ho + (w.ups * i + su)) * w.wo + (w.ups * j + sv);
                v = load8(w.gy + op * w.cout + co0 + cv * 8);
            }
            const uint4 u = __builtin_bit_cast(uint4, v);
            const uint32_t wv[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int e = 0; e < 8; ++e)
                sG[(cv * 8 + e) * kWgPad + (px ^ swz_w)] = (bf16)(wv[e >> 1] >> ((e & 1) * 16));
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            bf16x8 v = zero8();
            if (pvalid && kvalid[q]) {
                const int si = w.stride * i + w.ty[kt[q]], sj = w.stride * j + w.tx[kt[q]];
                if (si >= 0 && si < w.hs && sj >= 0 && sj < w.ws) {
                    const bf16* src = ksrc[q] ? w.src1 : w.src0;
                    const int cs = ksrc[q] ? w.c1 : w.c0;
                    v = load8(src + (((long long)ib * w.hs + si) * w.ws + sj) * cs + kc[q] + cv * 8);
                }
            }
            const uint4 u = __builtin_bit_cast(uint4, v);
            const uint32_t wv[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int e = 0; e < 8; ++e)
                sA[(q * 32 + cv * 8 + e) * kWgPad + (px ^ swz_w)] = (bf16)(wv[e >> 1] >> ((e & 1) * 16));
        }
        __syncthreads();
        if (bias) {
            const uint4 u = *reinterpret_cast<const uint4*>(&sG[(tid >> 3) * kWgPad + (tid & 7) * 8]);
            const uint32_t wv[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) bsum += bf2f((bf16)(wv[e] & 0xffffu)) + bf2f((bf16)(wv[e] >> 16));
        }
        if (mine) {
#pragma unroll
            for (int st = 0; st < kWgPix / 16; ++st) {
                const int c = st * 16 + kh * 8;
                const int cs = c ^ swz_r;
                const bf16x8 fa = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&sG[r * kWgPad + cs]));
                const bf16x8 fb =
                    __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&sA[(wave * 32 + r) * kWgPad + cs]));
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb, acc, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    if (bias) {
        bsum += __shfl_xor(bsum, 1);
        bsum += __shfl_xor(bsum, 2);
        bsum += __shfl_xor(bsum, 4);
        if ((tid & 7) == 0 && bsum != 0.f) unsafeAtomicAdd(&w.db[co0 + (tid >> 3)], bsum);
    }
    if (!mine) return;
    const int k = k0 + wave * 32 + r;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int n = n0 + 8 * q + 4 * kh + e;
            const float v = acc[4 * q + e];
            if (v != 0.f) unsafeAtomicAdd(&w.dw[(size_t)n * w.K + k], v);
        }
}

// ---------------------------------------------------------------------------
// Element-wise kernels (8 channels per thread).

__global__ void k_pool2(const bf16* __restrict__ x, bf16* __restrict__ y, int B, int H, int W, int C)
{
    const int C8 = C / 8;
    const long long n = (long long)B * (H / 2) * (W / 2) * C8;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int c8 = (int)(t % C8);
    long long q = t / C8;
    const int j = (int)(q % (W / 2));
    q /= (W / 2);
    const int i = (int)(q % (H / 2));
    const int b = (int)(q / (H / 2));
    const bf16* p00 = x + (((long long)b * H + 2 * i) * W + 2 * j) * C + c8 * 8;
    const bf16* p10 = p00 + (long long)W * C;
    const bf16x8 a = load8(p00), bb = load8(p00 + C), cc = load8(p10), d = load8(p10 + C);
    uint32_t o[4];
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
        const float s0 = 0.25f * (((float)a[e] + (float)bb[e]) + ((float)cc[e] + (float)d[e]));
        const float s1 = 0.25f * (((float)a[e + 1] + (float)bb[e + 1]) + ((float)cc[e + 1] + (float)d[e + 1]));
        o[e / 2] = (uint32_t)f2bf(s0) | ((uint32_t)f2bf(s1) << 16);
    }
    *reinterpret_cast<uint4*>(y + t * 8) = make_uint4(o[0], o[1], o[2], o[3]);
}

// dE = (dskip + pool2 backward of dP) * (E > 0), E and dskip [B][H][W][C],
// dP [B][H/2][W/2][C].
__global__ void k_pool2_back(const bf16* __restrict__ dP, const bf16* __restrict__ dskip,
                             const bf16* __restrict__ E, bf16* __restrict__ dE, int B, int H, int W, int C)
{
    const int C8 = C / 8;
    const long long n = (long long)B * H * W * C8;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int c8 = (int)(t % C8);
    long long q = t / C8;
    const int j = (int)(q % W);
    q /= W;
    const int i = (int)(q % H);
    const int b = (int)(q / H);
    const bf16x8 p = load8(dP + (((long long)b * (H / 2) + i / 2) * (W / 2) + j / 2) * C + c8 * 8);
    const bf16x8 s = load8(dskip + t * 8);
    const bf16x8 e = load8(E + t * 8);
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 8; k += 2) {
        float v0 = (float)s[k] + 0.25f * (float)p[k];
        float v1 = (float)s[k + 1] + 0.25f * (float)p[k + 1];
        if (!((float)e[k] > 0.f)) v0 = 0.f;
        if (!((float)e[k + 1] > 0.f)) v1 = 0.f;
        o[k / 2] = (uint32_t)f2bf(v0) | ((uint32_t)f2bf(v1) << 16);
    }
    *reinterpret_cast<uint4*>(dE + t * 8) = make_uint4(o[0], o[1], o[2], o[3]);
}

// Planar FP32 [B][cin][H][W] -> NHWC bfloat16 [B][H][W][cpad], zero padded.
__global__ void k_to_nhwc(const float* __restrict__ x, bf16* __restrict__ y, int B, int cin, int H, int W, int cpad)
{
    const long long n = (long long)B * H * W;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const long long hw = (long long)H * W;
    const long long b = p / hw, s = p - b * hw;
    for (int c0 = 0; c0 < cpad; c0 += 8) {
        uint32_t o[4];
#pragma unroll
        for (int e = 0; e < 8; e += 2) {
            const int c = c0 + e;
            const float v0 = c < cin ? x[(b * cin + c) * hw + s] : 0.f;
            const float v1 = c + 1 < cin ? x[(b * cin + c + 1) * hw + s] : 0.f;
            o[e / 2] = (uint32_t)f2bf(v0) | ((uint32_t)f2bf(v1) << 16);
        }
        *reinterpret_cast<uint4*>(y + p * cpad + c0) = make_uint4(o[0], o[1], o[2], o[3]);
    }
}

// Weighted mean-square loss of planar y against planar target; writes the
// loss gradient NHWC bfloat16 [P][cpad] (zero beyond C) and adds the loss to
// *loss in FP64 (one atomic per block).
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ y, const float* __restrict__ tgt,
                                            const float* __restrict__ wts, int B, int C, int H, int W, int cpad,
                                            bf16* __restrict__ dy, double* __restrict__ loss)
{
    __shared__ double part[4];
    const long long hw = (long long)H * W, n = (long long)B * hw;
    const double inv = 1.0 / (double)n;
    double l = 0.0;
    // grid-stride, so that few blocks add to the one FP64 accumulator
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (long long)gridDim.x * blockDim.x) {
        const long long b = p / hw, s = p - b * hw;
        for (int c0 = 0; c0 < cpad; c0 += 8) {
            uint32_t o[4];
#pragma unroll
            for (int e = 0; e < 8; e += 2) {
                float g2[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int c = c0 + e + h;
                    g2[h] = 0.f;
                    if (c < C) {
                        const long long idx = (b * C + c) * hw + s;
                        const float d = y[idx] - tgt[idx];
                        l += (double)wts[c] * (double)d * (double)d;
                        g2[h] = (float)(2.0 * (double)wts[c] * (double)d * inv);
                    }
                }
                o[e / 2] = (uint32_t)f2bf(g2[0]) | ((uint32_t)f2bf(g2[1]) << 16);
            }
            *reinterpret_cast<uint4*>(dy + p * cpad + c0) = make_uint4(o[0], o[1], o[2], o[3]);
        }
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double s = (part[0] + part[1]) + (part[2] + part[3]);
        if (s != 0.0) unsafeAtomicAdd(loss, s * inv);
    }
}

// Adam with bias correction; writes the bfloat16 copy and clears the
// gradient for the next accumulation.
__global__ void k_adam(float* __restrict__ p, float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       bf16* __restrict__ pb, long long n, float lr, float b1, float b2, float eps, float c1,
                       float c2)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    p[i] = pi;
    pb[i] = f2bf(pi);
    g[i] = 0.f;
}

__global__ void k_to_bf16(const float* __restrict__ p, bf16* __restrict__ pb, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pb[i] = f2bf(p[i]);
}

__global__ void k_fill(float* __restrict__ p, long long n, float v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Input-gradient weights: out[ci][t * R + co] = in[co][t * cin + ci].
__global__ void k_wt_transpose(const bf16* __restrict__ in, bf16* __restrict__ out, int R, int T, int cin)
{
    const long long n = (long long)R * T * cin;
    const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int ci = (int)(x % cin);
    const long long q = x / cin;
    const int t = (int)(q % T), co = (int)(q / T);
    out[(long long)ci * T * R + (long long)t * R + co] = in[x];
}

// ---------------------------------------------------------------------------
// Launchers.

static inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

template <int TK, int WM, int WN, int RM, int RN, int KC>
static void launch_gemm_t(const Gemm& g, hipStream_t s)
{
    constexpr int BM = 32 * WM * RM, BN = 32 * WN * RN;
    const long long M = (long long)g.batch * g.hg * g.wg;
    dim3 grid(cdiv(M, BM), cdiv(g.N, BN));
    hipLaunchKernelGGL((k_gemm<TK, WM, WN, RM, RN, KC>), grid, dim3(256), 0, s, g);
}

template <int TK>
static void launch_gemm_tk(const Gemm& g, hipStream_t s)
{
    // two K groups per LDS stage for the 128 x 128 tiles; one for the
    // 256-row tiles (registers and LDS)
    if (g.N % 128 == 0) launch_gemm_t<TK, 2, 2, 2, 2, 2>(g, s);
    else if (g.N % 64 == 0) launch_gemm_t<TK, 4, 1, 2, 2, 1>(g, s);
    else launch_gemm_t<TK, 4, 1, 2, 1, 1>(g, s);
}

// The tap pattern is given by ntaps, stride and the first offset; the
// kernels derive the offsets at compile time.
bool launch_gemm(const Gemm& g, hipStream_t s)
{
    if (g.N % 32 || (g.c0 | g.c1) % 32) return false;
    if (g.ntaps == 9 && g.stride == 1 && g.ty[0] == -1 && g.tx[0] == -1) launch_gemm_tk<kTap3>(g, s);
    else if (g.ntaps == 9 && g.stride == 1 && g.ty[0] == 1 && g.tx[0] == 1) launch_gemm_tk<kTap3Mirror>(g, s);
    else if (g.ntaps == 1 && g.stride == 1 && g.ty[0] == 0 && g.tx[0] == 0) launch_gemm_tk<kTap1>(g, s);
    else if (g.ntaps == 4 && g.stride == 2 && g.ty[0] == 0 && g.tx[0] == 0) launch_gemm_tk<kTap2x2>(g, s);
    else return false;
    return true;
}

void launch_wgrad(WGrad w, hipStream_t s)
{
    const long long M = (long long)w.batch * w.hg * w.wg;
    const int kt = cdiv(w.K, 128), nt = w.N / 32;
    const int want = std::max(1, 2048 / std::max(1, kt * nt));
    const int splits = std::max(1, std::min(want, cdiv(M, kWgPix)));
    w.chunk = cdiv(cdiv(M, splits), kWgPix) * kWgPix;
    dim3 grid(kt, nt, cdiv(M, w.chunk));
    hipLaunchKernelGGL(k_wgrad, grid, dim3(256), 0, s, w);
}

void launch_pool2(const bf16* x, bf16* y, int B, int H, int W, int C, hipStream_t s)
{
    const long long n = (long long)B * (H / 2) * (W / 2) * (C / 8);
    hipLaunchKernelGGL(k_pool2, dim3(cdiv(n, 256)), dim3(256), 0, s, x, y, B, H, W, C);
}

void launch_pool2_back(const bf16* dP, const bf16* dskip, const bf16* E, bf16* dE, int B, int H, int W, int C,
                       hipStream_t s)
{
    const long long n = (long long)B * H * W * (C / 8);
    hipLaunchKernelGGL(k_pool2_back, dim3(cdiv(n, 256)), dim3(256), 0, s, dP, dskip, E, dE, B, H, W, C);
}

void launch_to_nhwc(const float* x, bf16* y, int B, int cin, int H, int W, int cpad, hipStream_t s)
{
    const long long n = (long long)B * H * W;
    hipLaunchKernelGGL(k_to_nhwc, dim3(cdiv(n, 256)), dim3(256), 0, s, x, y, B, cin, H, W, cpad);
}

void launch_mse(const float* y, const float* tgt, const float* wts, int B, int C, int H, int W, int cpad, bf16* dy,
                double* loss, hipStream_t s)
{
    const long long n = (long long)B * H * W;
    hipLaunchKernelGGL(k_mse, dim3(std::min(cdiv(n, 256), 512)), dim3(256), 0, s, y, tgt, wts, B, C, H, W, cpad, dy,
                       loss);
}

void launch_adam(float* p, float* g, float* m, float* v, bf16* pb, long long n, float lr, float b1, float b2,
                 float eps, float c1, float c2, hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, dim3(cdiv(n, 256)), dim3(256), 0, s, p, g, m, v, pb, n, lr, b1, b2, eps, c1, c2);
}

void launch_to_bf16(const float* p, bf16* pb, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_bf16, dim3(cdiv(n, 256)), dim3(256), 0, s, p, pb, n);
}

void launch_fill(float* p, long long n, float v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill, dim3(cdiv(n, 256)), dim3(256), 0, s, p, n, v);
}

void launch_wt_transpose(const bf16* in, bf16* out, int R, int T, int cin, hipStream_t s)
{
    const long long n = (long long)R * T * cin;
    hipLaunchKernelGGL(k_wt_transpose, dim3(cdiv(n, 256)), dim3(256), 0, s, in, out, R, T, cin);
}

}  // namespace rift_nn
