// unet_kernels.hpp -- kernel arguments and launchers of the U-Net surrogate
// (unet_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace rift_nn {

typedef uint16_t bf16;  // bfloat16 storage

// Implicit GEMM (see k_gemm): rows = pixels of a batch x hg x wg grid,
// columns n = sub * cout + co, K = ntaps x (c0 + c1).
struct Gemm {
    const bf16* src0 = nullptr;
    const bf16* src1 = nullptr;  // second source of a channel concatenation
    int c0 = 0, c1 = 0;          // channels, multiples of 32
    int hs = 0, ws = 0;          // source image size
    int hg = 0, wg = 0;          // row grid per image
    int batch = 0;
    int ntaps = 1, stride = 1;   // tap t reads (stride*i + ty[t], stride*j + tx[t])
    int ty[9] = {}, tx[9] = {};
    const bf16* w = nullptr;     // [N][K]
    int K = 0, N = 0;
    const float* bias = nullptr; // [cout] or null
    int cout = 0, ups = 1;       // output pixel (ups*i + sub/ups, ups*j + sub%ups)
    int ho = 0, wo = 0;
    int relu = 0;
    const bf16* mask = nullptr;  // zero where mask <= 0 ([out pixel][cout])
    bf16* dst0 = nullptr;        // channels [0, dsplit) (all, without dst1)
    bf16* dst1 = nullptr;        // channels [dsplit, cout)
    int dsplit = 0;
    float* dstf = nullptr;       // planar FP32 [batch][nvalid][ho][wo] instead
    int nvalid = 0;
};

// Weight gradient (see k_wgrad): the source gather of Gemm, the output
// gradient gy read through Gemm's output mapping, dw [N][K] accumulated.
struct WGrad {
    const bf16* src0 = nullptr;
    const bf16* src1 = nullptr;
    int c0 = 0, c1 = 0;
    int hs = 0, ws = 0, hg = 0, wg = 0, batch = 0;
    int ntaps = 1, stride = 1;
    int ty[9] = {}, tx[9] = {};
    int K = 0, N = 0;
    const bf16* gy = nullptr;
    int cout = 0, ups = 1, ho = 0, wo = 0;
    float* dw = nullptr;
    float* db = nullptr;         // bias gradient [cout] accumulated too, or null
    int chunk = 0;               // pixels per block (set by the launcher)
};

// False (nothing launched) for a tap pattern or shape the kernels lack.
bool launch_gemm(const Gemm& g, hipStream_t s);
void launch_wgrad(WGrad w, hipStream_t s);
void launch_pool2(const bf16* x, bf16* y, int B, int H, int W, int C, hipStream_t s);
void launch_pool2_back(const bf16* dP, const bf16* dskip, const bf16* E, bf16* dE, int B, int H, int W, int C,
                       hipStream_t s);
void launch_to_nhwc(const float* x, bf16* y, int B, int cin, int H, int W, int cpad, hipStream_t s);
void launch_mse(const float* y, const float* tgt, const float* wts, int B, int C, int H, int W, int cpad, bf16* dy,
                double* loss, hipStream_t s);
void launch_adam(float* p, float* g, float* m, float* v, bf16* pb, long long n, float lr, float b1, float b2,
                 float eps, float c1, float c2, hipStream_t s);
void launch_to_bf16(const float* p, bf16* pb, long long n, hipStream_t s);
void launch_fill(float* p, long long n, float v, hipStream_t s);
void launch_wt_transpose(const bf16* in, bf16* out, int R, int T, int cin, hipStream_t s);

}  // namespace rift_nn
