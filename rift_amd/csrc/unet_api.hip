// unet_api.hip -- the rift_unet_* C ABI (include/rift_unet.h): parameter
// layout, buffers, and the launch sequences of the forward pass, the
// backward pass and the optimizer step.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rift_unet.h"
#include "unet_kernels.hpp"

namespace rift {
int set_last_error(int code, const std::string& msg);
}

using namespace rift_nn;

namespace {

int fail(int code, const std::string& msg) { return rift::set_last_error(code, msg); }

#define UCHK(expr)                                                                     \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(RIFT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define GEMM(desc, stream)                                                         \
    do {                                                                           \
        if (!launch_gemm((desc), (stream)))                                        \
            return fail(RIFT_E_UNSUPPORTED, "rift_unet: layer shape not supported"); \
    } while (0)

enum Kind { kConv3 = 0, kUp2 = 1, kHead = 2 };

struct Layer {
    Kind kind;
    int cin, cout;      // canonical channels
    int dcin;           // device input channels (the network input is padded to 32)
    int rows, K;        // device weight [rows][K]
    int nbias;          // device bias entries
    long long w_off, b_off, wd_off;  // element offsets (params / input-gradient weights)
    long long canon_off;
};

long long round64(long long x) { return (x + 63) / 64 * 64; }

// SplitMix64 + Box-Muller for the initial weights.
struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        const double u = uniform(), v = uniform();
        return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
    }
};

}  // namespace

struct rift_unet {
    rift_unet_desc desc{};
    int L = 0;
    std::vector<int> ch;          // channels per level
    std::vector<Layer> layers;    // canonical order
    long long n_dev = 0, n_wd = 0, n_canon = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    bf16 *pb = nullptr, *wd = nullptr;
    float* lossw = nullptr;
    double* loss = nullptr;
    float* yb = nullptr;          // network output of the training calls
    // activations and their gradients, [B][H_l][W_l][C]
    bf16* x32 = nullptr;
    std::vector<bf16*> enc_a, enc_b, pool, up, dec_a, dec_b;
    std::vector<bf16*> genc_a, genc_b, gpool, gup, gdec_a, gdec_b, gskip;
    bf16* ghead = nullptr;
    float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    long long adam_t = 0;

    int H(int l) const { return desc.height >> l; }
    int W(int l) const { return desc.width >> l; }
    const Layer& enc(int l, int s) const { return layers[2 * l + s]; }
    const Layer& upl(int l) const { return layers[2 * L + 3 * (L - 2 - l)]; }
    const Layer& dec(int l, int s) const { return layers[2 * L + 3 * (L - 2 - l) + 1 + s]; }
    const Layer& head() const { return layers.back(); }
};

namespace {

template <class T>
int dalloc(rift_unet* n, T** out, long long count)
{
    void* ptr = nullptr;
    UCHK(hipMalloc(&ptr, (size_t)std::max<long long>(count, 1) * sizeof(T)));
    n->allocs.push_back(ptr);
    *out = static_cast<T*>(ptr);
    return RIFT_OK;
}

void free_all(rift_unet* n)
{
    for (void* a : n->allocs) (void)hipFree(a);
    n->allocs.clear();
    if (n->stream) (void)hipStreamDestroy(n->stream);
    n->stream = nullptr;
}

// canonical <-> device parameter layouts
void canon_to_dev(const rift_unet* n, const float* c, std::vector<float>& d)
{
    d.assign((size_t)n->n_dev, 0.f);
    for (const Layer& ly : n->layers) {
        const float* cw = c + ly.canon_off;
        long long nw = 0;
        if (ly.kind == kConv3) {
            for (int co = 0; co < ly.cout; ++co)
                for (int t = 0; t < 9; ++t)
                    for (int ci = 0; ci < ly.cin; ++ci)
                        d[ly.w_off + (long long)co * ly.K + t * ly.dcin + ci] = cw[((long long)co * 9 + t) * ly.cin + ci];
            nw = 9LL * ly.cout * ly.cin;
        } else {   // up2 rows (s*cout + co) and head rows co share the canonical order
            const int rows = ly.kind == kUp2 ? 4 * ly.cout : ly.cout;
            for (int rr = 0; rr < rows; ++rr)
                for (int ci = 0; ci < ly.cin; ++ci) d[ly.w_off + (long long)rr * ly.K + ci] = cw[(long long)rr * ly.cin + ci];
            nw = (long long)rows * ly.cin;
        }
        for (int co = 0; co < ly.cout; ++co) d[ly.b_off + co] = cw[nw + co];
    }
}

void dev_to_canon(const rift_unet* n, const float* d, float* c)
{
    for (const Layer& ly : n->layers) {
        float* cw = c + ly.canon_off;
        long long nw = 0;
        if (ly.kind == kConv3) {
            for (int co = 0; co < ly.cout; ++co)
                for (int t = 0; t < 9; ++t)
                    for (int ci = 0; ci < ly.cin; ++ci)
                        cw[((long long)co * 9 + t) * ly.cin + ci] = d[ly.w_off + (long long)co * ly.K + t * ly.dcin + ci];
            nw = 9LL * ly.cout * ly.cin;
        } else {
            const int rows = ly.kind == kUp2 ? 4 * ly.cout : ly.cout;
            for (int rr = 0; rr < rows; ++rr)
                for (int ci = 0; ci < ly.cin; ++ci) cw[(long long)rr * ly.cin + ci] = d[ly.w_off + (long long)rr * ly.K + ci];
            nw = (long long)rows * ly.cin;
        }
        for (int co = 0; co < ly.cout; ++co) cw[nw + co] = d[ly.b_off + co];
    }
}

// Input-gradient weights from the bfloat16 copy of the weights.
void refresh_wd(rift_unet* n)
{
    for (size_t i = 0; i < n->layers.size(); ++i) {
        const Layer& ly = n->layers[i];
        if (i == 0) continue;   // no gradient with respect to the network input
        if (ly.kind == kConv3) launch_wt_transpose(n->pb + ly.w_off, n->wd + ly.wd_off, ly.rows, 9, ly.dcin, n->stream);
        else launch_wt_transpose(n->pb + ly.w_off, n->wd + ly.wd_off, ly.rows, 1, ly.K, n->stream);
    }
}

// --- GEMM descriptions of the layers ---------------------------------------

Gemm conv3_fwd(const rift_unet* n, const Layer& ly, int B, int l, const bf16* s0, int c0, const bf16* s1, int c1,
               bf16* dst)
{
    Gemm g;
    g.src0 = s0;
    g.src1 = s1;
    g.c0 = c0;
    g.c1 = c1;
    g.hs = g.hg = g.ho = n->H(l);
    g.ws = g.wg = g.wo = n->W(l);
    g.batch = B;
    g.ntaps = 9;
    for (int t = 0; t < 9; ++t) {
        g.ty[t] = t / 3 - 1;
        g.tx[t] = t % 3 - 1;
    }
    g.w = n->pb + ly.w_off;
    g.K = ly.K;
    g.N = g.cout = ly.cout;
    g.bias = n->p + ly.b_off;
    g.relu = 1;
    g.dst0 = dst;
    return g;
}

// Input gradient of a conv3x3 layer at level l: source = output gradient
// (cout channels), taps mirrored, weights [dcin][9 cout].
Gemm conv3_dgrad(const rift_unet* n, const Layer& ly, int B, int l, const bf16* gy, const bf16* mask, bf16* d0,
                 bf16* d1, int split)
{
    Gemm g;
    g.src0 = gy;
    g.c0 = ly.cout;
    g.hs = g.hg = g.ho = n->H(l);
    g.ws = g.wg = g.wo = n->W(l);
    g.batch = B;
    g.ntaps = 9;
    for (int t = 0; t < 9; ++t) {
        g.ty[t] = 1 - t / 3;
        g.tx[t] = 1 - t % 3;
    }
    g.w = n->wd + ly.wd_off;
    g.K = 9 * ly.cout;
    g.N = g.cout = ly.dcin;
    g.mask = mask;
    g.dst0 = d0;
    g.dst1 = d1;
    g.dsplit = split;
    return g;
}

WGrad conv3_wgrad(const rift_unet* n, const Layer& ly, int B, int l, const bf16* s0, int c0, const bf16* s1, int c1,
                  const bf16* gy)
{
    WGrad w;
    w.src0 = s0;
    w.src1 = s1;
    w.c0 = c0;
    w.c1 = c1;
    w.hs = w.hg = w.ho = n->H(l);
    w.ws = w.wg = w.wo = n->W(l);
    w.batch = B;
    w.ntaps = 9;
    for (int t = 0; t < 9; ++t) {
        w.ty[t] = t / 3 - 1;
        w.tx[t] = t % 3 - 1;
    }
    w.K = ly.K;
    w.N = w.cout = ly.cout;
    w.gy = gy;
    w.dw = n->g + ly.w_off;
    w.db = n->g + ly.b_off;
    return w;
}

// Transposed conv from level l+1 (cin channels) to level l (cout channels).
Gemm up_fwd(const rift_unet* n, const Layer& ly, int B, int l, const bf16* src, bf16* dst)
{
    Gemm g;
    g.src0 = src;
    g.c0 = ly.cin;
    g.hs = g.hg = n->H(l + 1);
    g.ws = g.wg = n->W(l + 1);
    g.ho = n->H(l);
    g.wo = n->W(l);
    g.batch = B;
    g.w = n->pb + ly.w_off;
    g.K = ly.K;
    g.N = 4 * ly.cout;
    g.cout = ly.cout;
    g.ups = 2;
    g.bias = n->p + ly.b_off;
    g.dst0 = dst;
    return g;
}

Gemm up_dgrad(const rift_unet* n, const Layer& ly, int B, int l, const bf16* gy, const bf16* mask, bf16* dst)
{
    Gemm g;
    g.src0 = gy;
    g.c0 = ly.cout;
    g.hs = n->H(l);
    g.ws = n->W(l);
    g.hg = g.ho = n->H(l + 1);
    g.wg = g.wo = n->W(l + 1);
    g.batch = B;
    g.ntaps = 4;
    g.stride = 2;
    for (int t = 0; t < 4; ++t) {
        g.ty[t] = t / 2;
        g.tx[t] = t % 2;
    }
    g.w = n->wd + ly.wd_off;
    g.K = 4 * ly.cout;
    g.N = g.cout = ly.cin;
    g.mask = mask;
    g.dst0 = dst;
    return g;
}

WGrad up_wgrad(const rift_unet* n, const Layer& ly, int B, int l, const bf16* src, const bf16* gy)
{
    WGrad w;
    w.src0 = src;
    w.c0 = ly.cin;
    w.hs = w.hg = n->H(l + 1);
    w.ws = w.wg = n->W(l + 1);
    w.batch = B;
    w.K = ly.K;
    w.N = 4 * ly.cout;
    w.gy = gy;
    w.cout = ly.cout;
    w.ups = 2;
    w.ho = n->H(l);
    w.wo = n->W(l);
    w.dw = n->g + ly.w_off;
    w.db = n->g + ly.b_off;
    return w;
}

int forward(rift_unet* n, const float* x, int B, float* y)
{
    hipStream_t s = n->stream;
    const int L = n->L;
    launch_to_nhwc(x, n->x32, B, n->desc.in_channels, n->H(0), n->W(0), 32, s);
    for (int l = 0; l < L; ++l) {
        const bf16* in = l == 0 ? n->x32 : n->pool[l - 1];
        const int cin = l == 0 ? 32 : n->ch[l - 1];
        GEMM(conv3_fwd(n, n->enc(l, 0), B, l, in, cin, nullptr, 0, n->enc_a[l]), s);
        GEMM(conv3_fwd(n, n->enc(l, 1), B, l, n->enc_a[l], n->ch[l], nullptr, 0, n->enc_b[l]), s);
        if (l < L - 1) launch_pool2(n->enc_b[l], n->pool[l], B, n->H(l), n->W(l), n->ch[l], s);
    }
    const bf16* a = n->enc_b[L - 1];
    for (int l = L - 2; l >= 0; --l) {
        GEMM(up_fwd(n, n->upl(l), B, l, a, n->up[l]), s);
        GEMM(conv3_fwd(n, n->dec(l, 0), B, l, n->up[l], n->ch[l], n->enc_b[l], n->ch[l], n->dec_a[l]), s);
        GEMM(conv3_fwd(n, n->dec(l, 1), B, l, n->dec_a[l], n->ch[l], nullptr, 0, n->dec_b[l]), s);
        a = n->dec_b[l];
    }
    const Layer& hd = n->head();
    Gemm g;
    g.src0 = a;
    g.c0 = n->ch[0];
    g.hs = g.hg = g.ho = n->H(0);
    g.ws = g.wg = g.wo = n->W(0);
    g.batch = B;
    g.w = n->pb + hd.w_off;
    g.K = hd.K;
    g.N = g.cout = 32;
    g.bias = n->p + hd.b_off;
    g.dstf = y;
    g.nvalid = hd.cout;
    GEMM(g, s);
    UCHK(hipGetLastError());
    return RIFT_OK;
}

// Gradient of the loss into n->g (which must hold zeros); loss into n->loss.
int backward(rift_unet* n, const float* y, const float* tgt, int B)
{
    hipStream_t s = n->stream;
    const int L = n->L;
    const int C = n->desc.out_channels;
    launch_mse(y, tgt, n->lossw, B, C, n->H(0), n->W(0), 32, n->ghead, n->loss, s);
    {   // head
        const Layer& hd = n->head();
        WGrad w;
        w.src0 = n->dec_b[0];
        w.c0 = n->ch[0];
        w.hs = w.hg = w.ho = n->H(0);
        w.ws = w.wg = w.wo = n->W(0);
        w.batch = B;
        w.K = hd.K;
        w.N = w.cout = 32;
        w.gy = n->ghead;
        w.dw = n->g + hd.w_off;
        w.db = n->g + hd.b_off;
        launch_wgrad(w, s);
        Gemm g;
        g.src0 = n->ghead;
        g.c0 = 32;
        g.hs = g.hg = g.ho = n->H(0);
        g.ws = g.wg = g.wo = n->W(0);
        g.batch = B;
        g.w = n->wd + hd.wd_off;
        g.K = 32;
        g.N = g.cout = n->ch[0];
        g.mask = n->dec_b[0];
        g.dst0 = n->gdec_b[0];
        GEMM(g, s);
    }
    for (int l = 0; l <= L - 2; ++l) {
        const int c = n->ch[l];
        const Layer& lb = n->dec(l, 1);
        launch_wgrad(conv3_wgrad(n, lb, B, l, n->dec_a[l], c, nullptr, 0, n->gdec_b[l]), s);
        GEMM(conv3_dgrad(n, lb, B, l, n->gdec_b[l], n->dec_a[l], n->gdec_a[l], nullptr, 0), s);
        const Layer& la = n->dec(l, 0);
        launch_wgrad(conv3_wgrad(n, la, B, l, n->up[l], c, n->enc_b[l], c, n->gdec_a[l]), s);
        GEMM(conv3_dgrad(n, la, B, l, n->gdec_a[l], nullptr, n->gup[l], n->gskip[l], c), s);
        const Layer& lu = n->upl(l);
        const bool top = l + 1 == L - 1;
        const bf16* src = top ? n->enc_b[L - 1] : n->dec_b[l + 1];
        bf16* gsrc = top ? n->genc_b[L - 1] : n->gdec_b[l + 1];
        launch_wgrad(up_wgrad(n, lu, B, l, src, n->gup[l]), s);
        GEMM(up_dgrad(n, lu, B, l, n->gup[l], src, gsrc), s);
    }
    for (int l = L - 1; l >= 0; --l) {
        const int c = n->ch[l];
        if (l < L - 1)
            launch_pool2_back(n->gpool[l], n->gskip[l], n->enc_b[l], n->genc_b[l], B, n->H(l), n->W(l), c, s);
        const Layer& lb = n->enc(l, 1);
        launch_wgrad(conv3_wgrad(n, lb, B, l, n->enc_a[l], c, nullptr, 0, n->genc_b[l]), s);
        GEMM(conv3_dgrad(n, lb, B, l, n->genc_b[l], n->enc_a[l], n->genc_a[l], nullptr, 0), s);
        const Layer& la = n->enc(l, 0);
        const bf16* in = l == 0 ? n->x32 : n->pool[l - 1];
        launch_wgrad(conv3_wgrad(n, la, B, l, in, la.dcin, nullptr, 0, n->genc_a[l]), s);
        if (l > 0) GEMM(conv3_dgrad(n, la, B, l, n->genc_a[l], nullptr, n->gpool[l - 1], nullptr, 0), s);
    }
    UCHK(hipGetLastError());
    return RIFT_OK;
}

int check_io(rift_unet* n, const void* x, int batch, const void* y, const char* what)
{
    if (!n || !x || !y) return fail(RIFT_E_ARG, std::string(what) + ": null argument");
    if (batch < 1 || batch > n->desc.max_batch)
        return fail(RIFT_E_ARG, std::string(what) + ": batch must lie in [1, max_batch]");
    UCHK(hipSetDevice(n->desc.device));
    return RIFT_OK;
}

}  // namespace

extern "C" {

int rift_unet_create(const rift_unet_desc* desc, rift_unet** out)
{
    if (!desc || !out) return fail(RIFT_E_ARG, "rift_unet_create: null argument");
    const rift_unet_desc& d = *desc;
    if (d.levels < 2 || d.levels > 5) return fail(RIFT_E_ARG, "rift_unet_create: levels must lie in [2, 5]");
    if (d.base_channels < 32 || d.base_channels % 32 || (d.base_channels << (d.levels - 1)) > 512)
        return fail(RIFT_E_ARG, "rift_unet_create: base_channels must be a multiple of 32 with base << (levels-1) <= 512");
    if (d.in_channels < 1 || d.in_channels > 32 || d.out_channels < 1 || d.out_channels > 32)
        return fail(RIFT_E_ARG, "rift_unet_create: in/out channels must lie in [1, 32]");
    const int q = 1 << (d.levels - 1);
    if (d.height < q || d.width < q || d.height % q || d.width % q)
        return fail(RIFT_E_ARG, "rift_unet_create: height and width must be positive multiples of 2^(levels-1)");
    if (d.max_batch < 1) return fail(RIFT_E_ARG, "rift_unet_create: max_batch must be >= 1");
    if ((long long)d.max_batch * d.height * d.width * (d.base_channels * 2) >= (1LL << 31))
        return fail(RIFT_E_ARG, "rift_unet_create: max_batch x height x width too large");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || d.device < 0 || d.device >= ndev)
        return fail(RIFT_E_ARG, "rift_unet_create: no such HIP device");
    UCHK(hipSetDevice(d.device));

    rift_unet* n = new rift_unet();
    n->desc = d;
    n->L = d.levels;
    for (int l = 0; l < n->L; ++l) n->ch.push_back(d.base_channels << l);
    auto add = [&](Kind k, int cin, int cout) {
        Layer ly{};
        ly.kind = k;
        ly.cin = cin;
        ly.cout = cout;
        ly.dcin = (k == kConv3 && n->layers.empty()) ? 32 : cin;
        ly.rows = k == kConv3 ? cout : (k == kUp2 ? 4 * cout : 32);
        ly.K = k == kConv3 ? 9 * ly.dcin : cin;
        ly.nbias = k == kHead ? 32 : cout;
        n->layers.push_back(ly);
    };
    {
        int prev = d.in_channels;
        for (int l = 0; l < n->L; ++l) {
            add(kConv3, prev, n->ch[l]);
            add(kConv3, n->ch[l], n->ch[l]);
            prev = n->ch[l];
        }
        for (int l = n->L - 2; l >= 0; --l) {
            add(kUp2, n->ch[l + 1], n->ch[l]);
            add(kConv3, 2 * n->ch[l], n->ch[l]);
            add(kConv3, n->ch[l], n->ch[l]);
        }
        add(kHead, n->ch[0], d.out_channels);
    }
    long long off = 0, offd = 0, offc = 0;
    for (Layer& ly : n->layers) {
        ly.w_off = off;
        off = round64(off + (long long)ly.rows * ly.K);
        ly.b_off = off;
        off = round64(off + ly.nbias);
        ly.wd_off = offd;
        offd = round64(offd + (long long)ly.rows * ly.K);
        ly.canon_off = offc;
        const long long nw = ly.kind == kConv3 ? 9LL * ly.cout * ly.cin
                                               : (ly.kind == kUp2 ? 4LL * ly.cout * ly.cin : (long long)ly.cout * ly.cin);
        offc += nw + ly.cout;
    }
    n->n_dev = off;
    n->n_wd = offd;
    n->n_canon = offc;

    auto bail = [&](int rc) {
        free_all(n);
        delete n;
        return rc;
    };
    if (hipStreamCreate(&n->stream) != hipSuccess) return bail(fail(RIFT_E_HIP, "rift_unet_create: hipStreamCreate"));
    const long long B = d.max_batch;
    int rc = RIFT_OK;
    rc |= dalloc(n, &n->p, n->n_dev);
    rc |= dalloc(n, &n->g, n->n_dev);
    rc |= dalloc(n, &n->m, n->n_dev);
    rc |= dalloc(n, &n->v, n->n_dev);
    rc |= dalloc(n, &n->pb, n->n_dev);
    rc |= dalloc(n, &n->wd, n->n_wd);
    rc |= dalloc(n, &n->lossw, 32);
    rc |= dalloc(n, &n->loss, 1);
    rc |= dalloc(n, &n->yb, B * d.out_channels * n->H(0) * n->W(0));
    rc |= dalloc(n, &n->x32, B * n->H(0) * n->W(0) * 32);
    rc |= dalloc(n, &n->ghead, B * n->H(0) * n->W(0) * 32);
    for (int l = 0; l < n->L && rc == RIFT_OK; ++l) {
        const long long sz = B * n->H(l) * n->W(l) * n->ch[l];
        bf16* t[13] = {};
        for (int k = 0; k < 13 && rc == RIFT_OK; ++k) {
            const bool want = k < 4 || l < n->L - 1;   // enc_a, enc_b and their gradients at every level
            if (want) rc |= dalloc(n, &t[k], (k == 4 || k == 5) ? sz / 4 : sz);
        }
        n->enc_a.push_back(t[0]);
        n->enc_b.push_back(t[1]);
        n->genc_a.push_back(t[2]);
        n->genc_b.push_back(t[3]);
        n->pool.push_back(t[4]);
        n->gpool.push_back(t[5]);
        n->up.push_back(t[6]);
        n->dec_a.push_back(t[7]);
        n->dec_b.push_back(t[8]);
        n->gup.push_back(t[9]);
        n->gdec_a.push_back(t[10]);
        n->gdec_b.push_back(t[11]);
        n->gskip.push_back(t[12]);
    }
    if (rc != RIFT_OK) return bail(fail(RIFT_E_HIP, "rift_unet_create: device allocation failed"));

    // He-normal initial weights, zero biases
    std::vector<float> canon((size_t)n->n_canon, 0.f);
    Rng rng{d.seed};
    for (const Layer& ly : n->layers) {
        const long long nw = ly.kind == kConv3 ? 9LL * ly.cout * ly.cin
                                               : (ly.kind == kUp2 ? 4LL * ly.cout * ly.cin : (long long)ly.cout * ly.cin);
        const double fan = ly.kind == kConv3 ? 9.0 * ly.cin : (double)ly.cin;
        const double sd = std::sqrt(2.0 / fan);
        for (long long i = 0; i < nw; ++i) canon[ly.canon_off + i] = (float)(sd * rng.normal());
    }
    *out = n;
    rc = rift_unet_set_params(n, canon.data());
    if (rc != RIFT_OK) {
        *out = nullptr;
        return bail(rc);
    }
    std::vector<float> ones(32, 1.f);
    rc = rift_unet_set_loss_weights(n, ones.data());
    if (rc != RIFT_OK) {
        *out = nullptr;
        return bail(rc);
    }
    return RIFT_OK;
}

int rift_unet_destroy(rift_unet* n)
{
    if (!n) return RIFT_OK;
    (void)hipSetDevice(n->desc.device);
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    free_all(n);
    delete n;
    return RIFT_OK;
}

int rift_unet_param_count(const rift_unet* n, long* count)
{
    if (!n || !count) return fail(RIFT_E_ARG, "rift_unet_param_count: null argument");
    *count = (long)n->n_canon;
    return RIFT_OK;
}

int rift_unet_set_params(rift_unet* n, const float* host)
{
    if (!n || !host) return fail(RIFT_E_ARG, "rift_unet_set_params: null argument");
    UCHK(hipSetDevice(n->desc.device));
    std::vector<float> dev;
    canon_to_dev(n, host, dev);
    UCHK(hipMemcpyAsync(n->p, dev.data(), dev.size() * sizeof(float), hipMemcpyHostToDevice, n->stream));
    launch_to_bf16(n->p, n->pb, n->n_dev, n->stream);
    launch_fill(n->g, n->n_dev, 0.f, n->stream);
    launch_fill(n->m, n->n_dev, 0.f, n->stream);
    launch_fill(n->v, n->n_dev, 0.f, n->stream);
    refresh_wd(n);
    n->adam_t = 0;
    UCHK(hipGetLastError());
    UCHK(hipStreamSynchronize(n->stream));
    return RIFT_OK;
}

int rift_unet_get_params(rift_unet* n, float* host)
{
    if (!n || !host) return fail(RIFT_E_ARG, "rift_unet_get_params: null argument");
    UCHK(hipSetDevice(n->desc.device));
    std::vector<float> dev((size_t)n->n_dev);
    UCHK(hipMemcpyAsync(dev.data(), n->p, dev.size() * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    UCHK(hipStreamSynchronize(n->stream));
    dev_to_canon(n, dev.data(), host);
    return RIFT_OK;
}

int rift_unet_set_loss_weights(rift_unet* n, const float* w)
{
    if (!n || !w) return fail(RIFT_E_ARG, "rift_unet_set_loss_weights: null argument");
    UCHK(hipSetDevice(n->desc.device));
    float buf[32] = {};
    for (int c = 0; c < n->desc.out_channels; ++c) buf[c] = w[c];
    UCHK(hipMemcpyAsync(n->lossw, buf, sizeof(buf), hipMemcpyHostToDevice, n->stream));
    UCHK(hipStreamSynchronize(n->stream));
    return RIFT_OK;
}

int rift_unet_set_adam(rift_unet* n, float beta1, float beta2, float eps)
{
    if (!n) return fail(RIFT_E_ARG, "rift_unet_set_adam: null handle");
    if (!(beta1 >= 0.f && beta1 < 1.f) || !(beta2 >= 0.f && beta2 < 1.f) || !(eps > 0.f))
        return fail(RIFT_E_ARG, "rift_unet_set_adam: need 0 <= beta < 1 and eps > 0");
    UCHK(hipSetDevice(n->desc.device));
    n->b1 = beta1;
    n->b2 = beta2;
    n->eps = eps;
    n->adam_t = 0;
    launch_fill(n->m, n->n_dev, 0.f, n->stream);
    launch_fill(n->v, n->n_dev, 0.f, n->stream);
    UCHK(hipGetLastError());
    UCHK(hipStreamSynchronize(n->stream));
    return RIFT_OK;
}

int rift_unet_forward(rift_unet* n, const float* x, int batch, float* y)
{
    int rc = check_io(n, x, batch, y, "rift_unet_forward");
    if (rc != RIFT_OK) return rc;
    rc = forward(n, x, batch, y);
    if (rc != RIFT_OK) return rc;
    UCHK(hipStreamSynchronize(n->stream));
    return RIFT_OK;
}

int rift_unet_gradient(rift_unet* n, const float* x, const float* target, int batch, double* loss, float* grad_host)
{
    int rc = check_io(n, x, batch, target, "rift_unet_gradient");
    if (rc != RIFT_OK) return rc;
    if (!grad_host) return fail(RIFT_E_ARG, "rift_unet_gradient: null gradient array");
    float* y = n->yb;
    launch_fill(n->g, n->n_dev, 0.f, n->stream);
    UCHK(hipMemsetAsync(n->loss, 0, sizeof(double), n->stream));
    rc = forward(n, x, batch, y);
    if (rc == RIFT_OK) rc = backward(n, y, target, batch);
    std::vector<float> dev((size_t)n->n_dev);
    double l = 0;
    if (rc == RIFT_OK) {
        UCHK(hipMemcpyAsync(dev.data(), n->g, dev.size() * sizeof(float), hipMemcpyDeviceToHost, n->stream));
        UCHK(hipMemcpyAsync(&l, n->loss, sizeof(double), hipMemcpyDeviceToHost, n->stream));
    }
    UCHK(hipStreamSynchronize(n->stream));
    if (rc != RIFT_OK) return rc;
    launch_fill(n->g, n->n_dev, 0.f, n->stream);
    dev_to_canon(n, dev.data(), grad_host);
    if (loss) *loss = l;
    UCHK(hipStreamSynchronize(n->stream));
    return RIFT_OK;
}

int rift_unet_train_step(rift_unet* n, const float* x, const float* target, int batch, float lr, double* loss)
{
    int rc = check_io(n, x, batch, target, "rift_unet_train_step");
    if (rc != RIFT_OK) return rc;
    if (!(lr > 0.f)) return fail(RIFT_E_ARG, "rift_unet_train_step: lr must be > 0");
    float* y = n->yb;
    UCHK(hipMemsetAsync(n->loss, 0, sizeof(double), n->stream));
    rc = forward(n, x, batch, y);
    if (rc == RIFT_OK) rc = backward(n, y, target, batch);
    if (rc == RIFT_OK) {
        ++n->adam_t;
        const float c1 = (float)(1.0 / (1.0 - std::pow((double)n->b1, (double)n->adam_t)));
        const float c2 = (float)(1.0 / (1.0 - std::pow((double)n->b2, (double)n->adam_t)));
        launch_adam(n->p, n->g, n->m, n->v, n->pb, n->n_dev, lr, n->b1, n->b2, n->eps, c1, c2, n->stream);
        refresh_wd(n);
    }
    double l = 0;
    UCHK(hipMemcpyAsync(&l, n->loss, sizeof(double), hipMemcpyDeviceToHost, n->stream));
    UCHK(hipStreamSynchronize(n->stream));
    if (rc != RIFT_OK) return rc;
    UCHK(hipGetLastError());
    if (loss) *loss = l;
    return RIFT_OK;
}

int rift_unet_flops(const rift_unet* n, int batch, double* fwd, double* train)
{
    if (!n || batch < 1) return fail(RIFT_E_ARG, "rift_unet_flops: bad argument");
    double f = 0, first = 0;
    for (size_t i = 0; i < n->layers.size(); ++i) {
        const Layer& ly = n->layers[i];
        double x = 0;
        // each layer's output pixels at its level
        if (ly.kind == kConv3 || ly.kind == kHead) {
            // level of the layer: find by channel count at the output
            int l = 0;
            while (l < n->L - 1 && n->ch[l] != (ly.kind == kHead ? n->ch[0] : ly.cout)) ++l;
            const double P = (double)batch * n->H(l) * n->W(l);
            x = 2.0 * P * (ly.kind == kConv3 ? 9.0 : 1.0) * ly.cin * ly.cout;
        } else {
            int l = 0;
            while (l < n->L - 1 && n->ch[l] != ly.cout) ++l;
            const double P = (double)batch * n->H(l + 1) * n->W(l + 1);
            x = 2.0 * P * 4.0 * ly.cout * ly.cin;
        }
        if (i == 0) first = x;
        f += x;
    }
    if (fwd) *fwd = f;
    if (train) *train = 3.0 * f - first;   // no input gradient for the first layer
    return RIFT_OK;
}

}  // extern "C"
