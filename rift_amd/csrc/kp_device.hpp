// kp_device.hpp -- per-face and per-cell arithmetic of the Kurganov-Petrova
// central-upwind scheme, FP32, in depth form, for gfx950.
//
// The reference (grid.cu:62-214) stores w = h + B in FP64 and reconstructs w.
// Here the prognostic field is the depth h; w never appears.  Every w
// difference the scheme needs is formed from depth and bed differences, which
// is the same arithmetic in exact numbers and keeps FP32 from losing the
// depth in the ulp of the bed elevation (3.8e-6 m at 60 m):
//   w(c+1) - w(c)  = (h(c+1) - h(c)) + (BC(c+1) - BC(c))
//   w_E - B_E      = h(c) + (BC(c) - B_E) + sigma        (face depth)
// The limiter is the reference's: slopes from a three-way minmod with
// theta = 1.3f (grid.cu:30, 78-80).  In the reference signf reads the low word
// of a double (grid.cu:62-71) and is never negative, so minmod reduces to
// min(|a|,|b|,|c|) (zero when any operand is zero); FP32 keeps that rule.  The
// textbook, sign-aware minmod is selected by a template flag.
#pragma once

#include <hip/hip_runtime.h>

namespace rift {

constexpr float kG = 9.80665f;         // grid.cu:28
constexpr float kEps = 1.19209e-07f;   // grid.cu:29
constexpr float kTheta = 1.3f;         // grid.cu:30
constexpr int kTileW = 16;             // BLOCK_COLS, grid.cu:19
constexpr int kTileH = 14;             // BLOCK_ROWS, grid.cu:18

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// x^p for x > 0 via the hardware log2/exp2 (v_log_f32, v_exp_f32).
__device__ __forceinline__ float fpow_pos(float x, float p)
{
    return __builtin_amdgcn_exp2f(p * __builtin_amdgcn_logf(x));
}

// minmod of grid.cu:73-76 (sign-blind) or the sign-aware textbook form.
template <bool Textbook>
__device__ __forceinline__ float minmod3(float a, float b, float c)
{
    float m = fminf(fminf(fabsf(a), fabsf(b)), fabsf(c));
    if constexpr (!Textbook) {
        return m;
    } else {
        bool pos = (a > 0.f) & (b > 0.f) & (c > 0.f);
        bool neg = (a < 0.f) & (b < 0.f) & (c < 0.f);
        return pos ? m : (neg ? -m : 0.f);
    }
}

// limiter(U, Up1, Um1) of grid.cu:78-80 from dr = Up1-U, dl = U-Um1, halved:
// 0.5 minmod(theta dr, (dr+dl)/2, theta dl).  Scaling by a power of two
// commutes with rounding, so the halving is folded into the operands
// (0.65f = 1.3f/2 exactly).
template <bool Textbook>
__device__ __forceinline__ float half_slope(float dr, float dl)
{
    return minmod3<Textbook>((0.5f * kTheta) * dr, 0.25f * (dr + dl), (0.5f * kTheta) * dl);
}

// Both face states of one cell (Construct, grid.cu:91-116).
struct FaceStates {
    float hp, qxp, qyp;   // east (x pass) / north (y pass) side
    float hm, qxm, qym;   // west / south side
};

// Differences across the face between a cell and its neighbour in the pass
// direction (n = plus side, s = minus side): of w = h + b, formed from depth
// and bed differences, and of the momenta.  A face's difference is the plus
// difference of the cell on its minus side and the minus difference of the
// cell on its plus side, so the marching kernels form it once.
struct FaceDiff {
    float w, qx, qy;
};
__device__ __forceinline__ FaceDiff face_diff(float hs, float hn, float bs, float bn, float qxs, float qxn,
                                              float qys, float qyn)
{
    return FaceDiff{(hn - hs) + (bn - bs), qxn - qxs, qyn - qys};
}

// Both face states of one cell from its cell values and the differences
// across its plus (dp) and minus (dm) faces.
template <bool Textbook>
__device__ __forceinline__ FaceStates construct_d(float h, float b, float qx, float qy, const FaceDiff& dp,
                                                  const FaceDiff& dm, float Bp, float Bm)
{
    FaceStates s;
    float sw = half_slope<Textbook>(dp.w, dm.w);
    const float hp0 = (h + (b - Bp)) + sw;
    const float hm0 = (h + (b - Bm)) - sw;
    const float sum = hp0 + hm0;
    const float hp = fminf(__builtin_amdgcn_fmed3f(hp0, 0.f, sum), sum);
    const float hmn = fmaxf(__builtin_amdgcn_fmed3f(hm0, 0.f, sum), 0.f);
    float sx = half_slope<Textbook>(dp.qx, dm.qx);
    float sy = half_slope<Textbook>(dp.qy, dm.qy);
    s.hp = hp;
    s.hm = hmn;
    s.qxp = qx + sx;
    s.qxm = qx - sx;
    s.qyp = qy + sy;
    s.qym = qy - sy;
    return s;
}

// h*, qx*, qy*: cell and neighbours along the pass (m = minus, p = plus);
// b*: cell-centre beds; Bp/Bm: the cell's east/west (north/south) face beds.
template <bool Textbook>
__device__ __forceinline__ FaceStates construct(float hm1, float h, float hp1,
                                                float bm1, float b, float bp1,
                                                float qxm1, float qx, float qxp1,
                                                float qym1, float qy, float qyp1,
                                                float Bp, float Bm)
{
    FaceStates s;
    float dr = (hp1 - h) + (bp1 - b);
    float dl = (h - hm1) + (b - bm1);
    float sw = half_slope<Textbook>(dr, dl);
    const float hp0 = (h + (b - Bp)) + sw;
    const float hm0 = (h + (b - Bm)) - sw;
    // positivity fix, both branches with the pre-fix sum s (wBar,
    // grid.cu:108-115):  hp < 0 -> (0, s);  then hm < 0 -> (s, 0).  As pure
    // selections: hp' = min(med3(hp, 0, s), s), hm' = max(med3(hm, 0, s), 0)
    // (for s >= 0 the median clips each side into [0, s]; for s < 0 the
    // result is (s, 0) whatever the signs were).
    const float sum = hp0 + hm0;
    const float hp = fminf(__builtin_amdgcn_fmed3f(hp0, 0.f, sum), sum);
    const float hmn = fmaxf(__builtin_amdgcn_fmed3f(hm0, 0.f, sum), 0.f);
    float sx = half_slope<Textbook>(qxp1 - qx, qx - qxm1);
    float sy = half_slope<Textbook>(qyp1 - qy, qy - qym1);
    s.hp = hp;
    s.hm = hmn;
    s.qxp = qx + sx;
    s.qxm = qx - sx;
    s.qyp = qy + sy;
    s.qym = qy - sy;
    return s;
}

// Desingularised velocity numerator factor: u = 2 h q / (h^2 + max(h^2, kappa))
// (getVelocity, grid.cu:82-84); returns 2h/(h^2+max(h^2,kappa)).
__device__ __forceinline__ float vel_factor_h2(float h, float h2, float kappa)
{
    return (2.f * h) * frcp(h2 + fmaxf(h2, kappa));
}
__device__ __forceinline__ float vel_factor(float h, float kappa)
{
    return vel_factor_h2(h, h * h, kappa);
}

struct Flux {
    float f0, f1, f2, speed;
};

// KP numerical flux through one face (getHX/getHY, grid.cu:118-214).
// L = state on the west/south side of the face, R = east/north side.
// Dir 0 = x (speeds from u), 1 = y (speeds from v).
template <int Dir>
__device__ __forceinline__ Flux face_flux(float hL, float qxL, float qyL,
                                          float hR, float qxR, float qyR, float kappa)
{
    const float h2L = hL * hL, h2R = hR * hR;
    const float fL = vel_factor_h2(hL, h2L, kappa), fR = vel_factor_h2(hR, h2R, kappa);
    const float uL = fL * qxL, vL = fL * qyL, uR = fR * qxR, vR = fR * qyR;
    // momenta recomputed from the desingularised velocities (grid.cu:133-136)
    const float mxL = hL * uL, myL = hL * vL, mxR = hR * uR, myR = hR * vR;
    const float cL = (Dir == 0) ? uL : vL;
    const float cR = (Dir == 0) ? uR : vR;
    // sqrtf of a negative face depth is NaN, as in the reference, and the
    // IEEE maxNum/minNum of v_max_f32/v_min_f32 drop it exactly as fmaxf and
    // fminf do there: a side with h < 0 does not bound the speeds.
    const float sL = fsqrt(kG * hL), sR = fsqrt(kG * hR);
    const float ap = fmaxf(fmaxf(cL + sL, cR + sR), 0.f);
    const float am = fminf(fminf(cL - sL, cR - sR), 0.f);

    // physical fluxes (hu^2/h = h u^2, hu hv / h = h u v)
    const float hg = 0.5f * kG;
    float F0L, F1L, F2L, F0R, F1R, F2R;
    if (Dir == 0) {
        F0L = mxL; F1L = fmaf(mxL, uL, hg * h2L); F2L = mxL * vL;
        F0R = mxR; F1R = fmaf(mxR, uR, hg * h2R); F2R = mxR * vR;
    } else {
        F0L = myL; F1L = mxL * vL; F2L = fmaf(myL, vL, hg * h2L);
        F0R = myR; F1R = mxR * vR; F2R = fmaf(myR, vR, hg * h2R);
    }

    // computeH (grid.cu:86-89): (ap FL - am FR)/(ap-am) + ap am/(ap-am) (UR-UL),
    // zero when ap - am = 0 (then ap = am = 0 and every term vanishes).  A dry
    // side's physical flux is zero: its weight is zeroed instead of its three
    // flux components (the fluxes are finite, so the products are the same).
    const float den = ap - am;
    const float r = (den > 0.f) ? frcp(den) : 0.f;
    const float a = ap * r, b = am * r, c = a * am;
    const float aL = (hL > 0.f) ? a : 0.f;
    const float bR = (hR > 0.f) ? b : 0.f;
    Flux out;
    out.f0 = fmaf(-bR, F0R, fmaf(aL, F0L, c * (hR - hL)));
    out.f1 = fmaf(-bR, F1R, fmaf(aL, F1L, c * (mxR - mxL)));
    out.f2 = fmaf(-bR, F2R, fmaf(aL, F2L, c * (myR - myL)));
    out.speed = fmaxf(ap, -am);
    return out;
}

// Flux part of dU for one cell: -(E - W)/dx - (N - S)/dy.
__device__ __forceinline__ float flux_divergence(float e, float w, float n, float s, float inv_dx,
                                                 float inv_dy)
{
    return fmaf(s - n, inv_dy, (w - e) * inv_dx);
}

// Manning coefficient G (grid.cu:753-758):
//   G = -4 g n^2 |q| h^(5/3) / (h^2 + max(h^2, kappa))^2
__device__ __forceinline__ float manning_g(float n, float hu, float hv, float h, float kappa)
{
    if (!(h > 0.f) || n == 0.f) return 0.f;
    float h2 = h * h;
    float D = h2 + fmaxf(h2, kappa);
    float q = fsqrt(hu * hu + hv * hv);
    return -4.f * kG * n * n * q * fpow_pos(h, 5.f / 3.f) * frcp(D * D);
}

// Green-Ampt infiltration rate (grid.cu:792-807) clamped to [0, h/dt].  The
// root (p1 + sqrt(p1^2 + q))/(2 dt) is evaluated as q/(sqrt(p1^2+q) - p1) when
// p1 < 0, which is the same number without the cancellation.
__device__ __forceinline__ float green_ampt(float K, float F, float psi_dtheta, float h,
                                            float dt, float inv_dt)
{
    float p1 = K * dt - 2.f * F;
    float p2 = K * (F + psi_dtheta);
    float q = 8.f * p2 * dt;
    float root = fsqrt(fmaxf(p1 * p1 + q, 0.f));
    float num = (p1 >= 0.f) ? (p1 + root) : q * frcp(root - p1);
    float rate = num * (0.5f * inv_dt);
    float cap = h * inv_dt;
    rate = (rate > cap) ? cap : rate;
    return (rate > 0.f) ? rate : 0.f;
}

// Beds from the corner elevations (the DEM points, simulator.cu:384-414):
// with p(j) = D(c, j) + D(c+1, j) the pair sum of corner row j, the cell
// bed is (p(j) + p(j+1))/4 (BC, the mean of the four face beds, grid.cu:
// 531-537), the south face (p(j))/2 (BY) and the west face
// (D(c, j) + D(c, j+1))/2 (BX).  Every kernel forms them with these
// expressions, so that all paths see the same FP32 beds.
__device__ __forceinline__ float bed_cell(float p_lo, float p_hi) { return 0.25f * (p_lo + p_hi); }
__device__ __forceinline__ float bed_south(float p_lo) { return 0.5f * p_lo; }
__device__ __forceinline__ float bed_west(float d_lo, float d_hi) { return 0.5f * (d_lo + d_hi); }

// Critical depth (q^2/g)^(1/3) of the outflow boundary (grid.cu:1342).
__device__ __forceinline__ float critical_depth(float q)
{
    float x = q * q * (1.f / kG);
    return (x > 0.f) ? fpow_pos(x, 1.f / 3.f) : 0.f;
}

}  // namespace rift
