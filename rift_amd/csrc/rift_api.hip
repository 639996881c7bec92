// rift_api.hip -- the C ABI of include/rift_kp.h over the gfx950 kernels.
//
// A handle owns one stream and every device buffer of one domain.  Host
// arrays cross the boundary in the reference's double layout and are
// converted to the FP32 device layout described in kp_kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/rift_host.h"
#include "../../include/rift_kp.h"
#include "kp_device.hpp"
#include "kp_kernels.hpp"

using namespace rift;

namespace {

thread_local std::string g_last_error;
// set while an ensemble creates its members: they share its stream
thread_local hipStream_t g_member_stream = nullptr;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

}  // namespace

namespace rift {
int set_last_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace rift

namespace {

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(RIFT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kProfClasses = 6;

}  // namespace

struct rift_grid {
    rift_desc desc{};
    Dev d{};
    hipStream_t stream = nullptr;
    int R = 0;                 // padded rows
    size_t plane = 0;          // P*R floats
    std::vector<void*> allocs;
    std::vector<size_t> alloc_sizes;   // bytes of each allocation (same order)
    std::vector<double> bc_ref;   // reference BC (InitGrid_k rule), host copy
    std::vector<double> n_host;
    double* dsum = nullptr;       // device scratch for FP64 sums
    int grid_tiles = 0;
    bool static_loaded = false, inited = false;
    bool scattered = false;       // fused steps left per-tile buffer selects
    // the device holds this step's forcing (or the next step's, evaluated by
    // the last dt kernel) and zero unit counters: a fused step needs no
    // begin_step launch; anything that changes time, series, sources or the
    // counters clears it
    bool primed = false;
    int variant = 0;              // launch variant of the marching kernels
    bool own_stream = true;       // false for ensemble members (the ensemble's stream)
    double budget_v0 = 0.0;       // stored + infiltrated volume at the budget reset, m^3
    int cus = 0;                  // compute units of the device
    int march_blocks[2][kMarchVariants] = {};   // resident workgroups per stage/variant
    bool per_phase = false;       // rift_step uses the one-kernel-per-phase path
    // the face beds are means of one corner field (d.DC holds it): the
    // fused step, which forms the beds from the corners, is available
    bool corner_beds = false;
    // captured fused steps (1 and kGraphSteps steps); the kernels take the
    // handle's Dev by value, so any change of it invalidates the graphs
    hipGraphExec_t graph1 = nullptr, graphN = nullptr;
    bool use_graphs = true;
    bool dt_in_march = true;      // the stage-1 march forms dt (else the dt kernel)
    bool prof = false;
    int sg_dt_int = 0;            // device surge series: map interval, seconds
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_events;
    std::vector<hipEvent_t> event_pool;
};

void drop_graphs(rift_grid* g)
{
    if (g->graph1) (void)hipGraphExecDestroy(g->graph1);
    if (g->graphN) (void)hipGraphExecDestroy(g->graphN);
    g->graph1 = g->graphN = nullptr;
}

namespace {

// Cell fields start kPlaneShift floats into their allocation (which has
// kPlaneSlack floats more than the fields need): column 2, the first
// interior column and tile column 0's first, then falls on a 64-byte
// boundary in every row (the pitch is a multiple of 64 floats), and each
// tile's 16 columns of a row are exactly one 64-byte sector.  Neighbouring
// tiles keep their state in different buffers whenever they became active
// at steps of different parity, so with tiles straddling sectors (column 2
// at byte 8) every sector of a row was written half by one tile and half by
// the other in another plane: stage 2 wrote 1.56x the bytes of U^{n+1}.
constexpr size_t kPlaneShift = 14, kPlaneSlack = 16;

// n floats of cell fields (one or more consecutive planes), zeroed; returns
// the shifted start.
float* alloc_cells(rift_grid* g, size_t n)
{
    void* p = nullptr;
    const size_t bytes = (n + kPlaneSlack) * sizeof(float);
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    if (hipMemset(p, 0, bytes) != hipSuccess) return nullptr;
    g->allocs.push_back(p);
    g->alloc_sizes.push_back(bytes);
    return static_cast<float*>(p) + kPlaneShift;
}

float* alloc_plane(rift_grid* g) { return alloc_cells(g, g->plane); }

template <class T>
T* alloc_bytes(rift_grid* g, size_t n)
{
    void* p = nullptr;
    if (hipMalloc(&p, n * sizeof(T) + 64) != hipSuccess) return nullptr;
    if (hipMemset(p, 0, n * sizeof(T) + 64) != hipSuccess) return nullptr;
    g->allocs.push_back(p);
    g->alloc_sizes.push_back(n * sizeof(T) + 64);
    return static_cast<T*>(p);
}

hipEvent_t take_event(rift_grid* g)
{
    if (!g->event_pool.empty()) {
        hipEvent_t e = g->event_pool.back();
        g->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

struct ProfScope {
    rift_grid* g;
    int cls;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(rift_grid* g_, int c) : g(g_), cls(c)
    {
        if (g->prof) {
            a = take_event(g);
            b = take_event(g);
            (void)hipEventRecord(a, g->stream);
        }
    }
    ~ProfScope()
    {
        if (g->prof) {
            (void)hipEventRecord(b, g->stream);
            g->prof_events.push_back({cls, {a, b}});
        }
    }
};

// Host -> device field conversion into the padded layout.
int put_field(rift_grid* g, float* dst, const double* src, int cols, int rows, int ld)
{
    std::vector<float> tmp(g->plane, 0.f);
    for (int j = 0; j < rows; ++j)
        for (int i = 0; i < cols; ++i) tmp[(size_t)j * g->d.P + i] = (float)src[(size_t)j * ld + i];
    HIPCHK(hipMemcpyAsync(dst, tmp.data(), g->plane * sizeof(float), hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    return RIFT_OK;
}

int get_plane(rift_grid* g, const float* src, std::vector<float>& tmp)
{
    tmp.resize(g->plane);
    HIPCHK(hipMemcpyAsync(tmp.data(), src, g->plane * sizeof(float), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    return RIFT_OK;
}

int check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(RIFT_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return RIFT_OK;
}

float* map_plane(rift_grid* g, int which, int plane)
{
    if (which < 0 || which > 2 || plane < 0 || plane > 2) return nullptr;
    return g->d.map[which][plane];
}

}  // namespace

// ---------------------------------------------------------------------------
// FP64 interior sum without library allocation (rift_sum_interior_noalloc).
// Scratch lives in one file-scope device array per GPU: per-block FP64
// partials, the result, and the completion ticket.  Host code reaches it
// through the address hipGetSymbolAddress returns for the current device
// (the array's name on the host is only a stand-in).  Every call on a GPU
// shares that GPU's array: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
namespace {

constexpr int kSumCellsPerBlock = 256 * 64;     // interior cells per workgroup
constexpr int kSumMaxBlocks = 4096;             // partials: up to 8192^2 interior cells
constexpr int kSumResult = kSumMaxBlocks;       // index of the result
constexpr int kSumTicket = kSumMaxBlocks + 8;   // index of the ticket (as unsigned bits)
constexpr int kSumScratch = kSumMaxBlocks + 16; // doubles (32.9 KB per GPU)

static __device__ __align__(256) double scratch[kSumScratch];

__global__ __launch_bounds__(64) void k_sum_scratch_clear(double* sc)
{
    if (threadIdx.x == 0) reinterpret_cast<unsigned*>(sc + kSumTicket)[0] = 0u;
}

// Each workgroup sums its kSumCellsPerBlock consecutive interior cells
// (row-major) in FP64 into its partial; the last to finish sums the partials
// in a fixed order (strided per lane, then a fixed tree), so the result does
// not depend on completion order.
__global__ __launch_bounds__(256) void k_sum_partials(Dev d, const float* x, double* sc)
{
    __shared__ double red[256];
    __shared__ bool last;
    const int ni = d.nx - 4, nj = d.ny - 4;
    const long total = (long)ni * nj;
    const long k0 = (long)blockIdx.x * kSumCellsPerBlock;
    const long k1 = min(k0 + kSumCellsPerBlock, total);
    double acc = 0.0;
    for (long k = k0 + threadIdx.x; k < k1; k += 256) {
        const int j = (int)(k / ni) + 2, i = (int)(k % ni) + 2;
        acc += (double)x[cell(d, i, j)];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sc[blockIdx.x] = red[0];
        __threadfence();
        last = atomicAdd(reinterpret_cast<unsigned*>(sc + kSumTicket), 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double a = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += 256) a += __hip_atomic_load(sc + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    red[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sc[kSumResult] = red[0];
        reinterpret_cast<unsigned*>(sc + kSumTicket)[0] = 0u;
    }
}

// Device address of `scratch` per device, resolved once with the device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_for(int device, double** out)
{
    if (device < 0 || device >= kMaxDevices) return fail(RIFT_E_ARG, "scratch: device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        HIPCHK(hipGetDevice(&cur));
        HIPCHK(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(RIFT_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[device];
    return RIFT_OK;
}

}  // namespace

extern "C" {

const char* rift_last_error(void) { return g_last_error.c_str(); }

int rift_create(const rift_desc* desc, rift_grid** out)
{
    if (!desc || !out) return fail(RIFT_E_ARG, "rift_create: null argument");
    if (desc->nx < 6 || desc->ny < 6) return fail(RIFT_E_ARG, "rift_create: nx, ny must be >= 6");
    if (!(desc->dx > 0) || !(desc->dy > 0)) return fail(RIFT_E_ARG, "rift_create: dx, dy must be > 0");
    if (!(desc->courant > 0) || !(desc->courant < 1))
        return fail(RIFT_E_ARG, "rift_create: courant must lie in (0, 1)");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (desc->device < 0 || desc->device >= ndev)
        return fail(RIFT_E_ARG, "rift_create: no such HIP device");
    HIPCHK(hipSetDevice(desc->device));

    rift_grid* g = new rift_grid();
    g->desc = *desc;
    Dev& d = g->d;
    d.nx = desc->nx;
    d.ny = desc->ny;
    d.nbx = (int)std::ceil((double)d.nx / (double)kTileW);   // grid.cu:336-337
    d.nby = (int)std::ceil((double)d.ny / (double)kTileH);
    d.ntiles = d.nbx * d.nby;
    d.nstrips = (d.nx - 4 + kStrip - 1) / kStrip;
    d.nurows = (d.ny - 4 + kTileH - 1) / kTileH;
    d.nunits = d.nstrips * d.nurows;
    d.nquads = (d.nstrips + 3) / 4;
    // the pitch covers the 20-column window of the last tile and the 64
    // lanes of the last strip
    d.P = ((std::max(kTileW * d.nbx + 4, kStrip * d.nstrips + 4) + 63) / 64) * 64;
    g->R = kTileH * d.nby + 4;
    g->plane = (size_t)d.P * (size_t)g->R;
    // the marching kernels address a cell of either state buffer by a 32-bit
    // byte offset from its row (column + one plane): a plane stays < 2^30
    if (g->plane + (size_t)d.P >= ((size_t)1 << 30)) {
        delete g;
        return fail(RIFT_E_UNSUPPORTED, "rift_create: domain too large (a field plane must hold < 2^30 cells)");
    }
    d.dx = desc->dx;
    d.dy = desc->dy;
    d.inv_dx = (float)(1.0 / desc->dx);
    d.inv_dy = (float)(1.0 / desc->dy);
    d.kappa = (float)desc->kappa;
    d.kappa_d = desc->kappa;
    d.courant = desc->courant;
    d.psi_dtheta = (float)(desc->psi * desc->dtheta);
    d.nodata = (float)desc->nodata;
    d.flags = desc->flags;
    d.si0 = d.sj0 = 1;
    d.si1 = d.sj1 = 0;
    d.nsrc = 0;
    d.band = 1;
    d.min_rows = kTileH;
    d.quads = 1;

    if (g_member_stream) {
        g->stream = g_member_stream;
        g->own_stream = false;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete g;
            return fail(RIFT_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        }
    }
    const unsigned f = desc->flags;
    bool ok = true;
    auto P = [&](float*& p) { p = alloc_plane(g); ok = ok && p; };
    // the two state buffers are six consecutive planes, each field followed
    // by its second buffer (h, h0, hu, hu0, hv, hv0), so a tile's buffer
    // select is one plane (d.soff elements) added to the h/hu/hv indices and
    // fits the 32-bit lane offsets of the marching kernels
    {
        float* b = alloc_cells(g, 6 * g->plane);
        ok = b != nullptr;
        if (ok) {
            d.h = b; d.h0 = b + g->plane;
            d.hu = b + 2 * g->plane; d.hu0 = b + 3 * g->plane;
            d.hv = b + 4 * g->plane; d.hv0 = b + 5 * g->plane;
            d.soff = (long long)g->plane;
        }
    }
    P(d.dh); P(d.dhu); P(d.dhv); P(d.G); P(d.n); P(d.BC); P(d.BX); P(d.BY);
    // corners: one row more than a plane's padding needs is never read;
    // the marching kernels' last lane reads one column past its own
    d.DC = alloc_cells(g, g->plane + (size_t)d.P);
    ok = ok && d.DC;
    if (f & RIFT_INFILTRATION) { P(d.F); P(d.F0); P(d.dF); P(d.K); }
    if (f & RIFT_SAVE_MAX) { P(d.hmax); P(d.qmax); P(d.tpeak); P(d.tdry); }
    if (f & RIFT_SAVE_ARRIVAL) P(d.twet);
    if (f & RIFT_RAIN_GRID) for (int k = 0; k < 3; ++k) P(d.map[RIFT_MAP_RAIN][k]);
    if (f & RIFT_DRAIN_GRID) for (int k = 0; k < 3; ++k) P(d.map[RIFT_MAP_DRAIN][k]);
    if (f & RIFT_SURGE_GRID) for (int k = 0; k < 3; ++k) P(d.map[RIFT_MAP_SURGE][k]);
    d.wet = alloc_bytes<unsigned char>(g, d.ntiles);
    d.act = alloc_bytes<unsigned char>(g, d.ntiles);
    d.list = alloc_bytes<int>(g, d.ntiles);
    d.mx = alloc_bytes<float>(g, d.ntiles);
    d.my = alloc_bytes<float>(g, d.ntiles);
    d.mx1 = alloc_bytes<float>(g, d.ntiles);
    d.my1 = alloc_bytes<float>(g, d.ntiles);
    d.sel = alloc_bytes<unsigned char>(g, d.ntiles);
    d.flip = alloc_bytes<unsigned char>(g, d.ntiles);
    d.ulist = alloc_bytes<int>(g, d.nunits);
    d.umark = alloc_bytes<unsigned char>(g, d.nunits);
    d.qlist = alloc_bytes<int>(g, (size_t)d.nquads * d.nurows);
    d.utag = alloc_bytes<int>(g, d.nunits);
    d.sc = alloc_bytes<Scalars>(g, 1);
    // 8 ticket counters a 128-byte line apart
    d.dtk = alloc_bytes<unsigned>(g, 8 * 32);
    // dt formed by the stage-1 march: the slots grow and the march add into
    d.gs = alloc_bytes<unsigned>(g, 32 * (size_t)kDtSlots);
    d.ser = alloc_bytes<SeriesTable>(g, 3);
    g->dsum = alloc_bytes<double>(g, 4);
    ok = ok && d.wet && d.act && d.list && d.mx && d.my && d.mx1 && d.my1 && d.sc && d.dtk && d.gs && d.ser &&
         g->dsum && d.sel &&
         d.flip && d.ulist && d.umark && d.qlist && d.utag;
    if (!ok) {
        rift_destroy(g);
        return fail(RIFT_E_HIP, "rift_create: device allocation failed");
    }
    g->grid_tiles = std::min(d.ntiles, 256 * 8);
    {
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, desc->device));
        g->cus = prop.multiProcessorCount;
        for (int st = 0; st < 2; ++st)
            for (int v = 0; v < kMarchVariants; ++v)
                g->march_blocks[st][v] = march_resident_blocks(g->d, st + 1, v, g->cus);
    }
    *out = g;
    return RIFT_OK;
}

int rift_destroy(rift_grid* g)
{
    if (!g) return RIFT_OK;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    drop_graphs(g);
    for (auto& pe : g->prof_events) {
        (void)hipEventDestroy(pe.second.first);
        (void)hipEventDestroy(pe.second.second);
    }
    for (hipEvent_t e : g->event_pool) (void)hipEventDestroy(e);
    for (void* p : g->allocs) (void)hipFree(p);
    if (g->stream && g->own_stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return RIFT_OK;
}

// The corner plane in the field layout, extended with period 2 beyond the
// corners (D(1) = D(3), D(0) = D(2), D(nx-1) = D(nx-3), D(nx) = D(nx-2), and
// so on; rows alike).  Formed from it, every ghost face is the face two
// cells in, and every ghost cell's bed is that of the interior cell its
// state comes from, as in the reference's arrays.
static int put_corners(rift_grid* g, const std::vector<double>& D)
{
    const Dev& d = g->d;
    const int W = d.nx - 3, H = d.ny - 3;
    auto ext = [](int c, int n) {
        if (c < 0) return c & 1;
        if (c > n - 1) return n - 1 - ((n - 1 - c) & 1);
        return c;
    };
    std::vector<float> tmp(g->plane + (size_t)d.P, 0.f);
    for (int j = 0; j < g->R; ++j) {
        const int cj = ext(j - 2, H);
        for (int i = 0; i < d.P; ++i) tmp[(size_t)j * d.P + i] = (float)D[(size_t)cj * W + ext(i - 2, W)];
    }
    HIPCHK(hipMemcpyAsync(d.DC, tmp.data(), tmp.size() * sizeof(float), hipMemcpyHostToDevice, g->stream));
    launch_beds_from_corners(d, g->R, g->stream);
    HIPCHK(hipStreamSynchronize(g->stream));
    return check_launch("beds from corners");
}

int rift_upload_static(rift_grid* g, const double* BX, const double* BY, const double* n,
                       const double* K, const double* h0)
{
    if (!g || !BX || !BY) return fail(RIFT_E_ARG, "rift_upload_static: null argument");
    Dev& d = g->d;
    const int nx = d.nx, ny = d.ny;
    int rc;
    if ((rc = put_field(g, d.BX, BX, nx + 1, ny, nx + 1))) return rc;
    if ((rc = put_field(g, d.BY, BY, nx, ny + 1, nx))) return rc;
    // InitGrid_k (grid.cu:531-537): BC = 0.25f*(BE+BW+BN+BS) on i, j >= 2.
    g->bc_ref.assign((size_t)nx * ny, 0.0);
    for (int j = 2; j < ny; ++j) {
        for (int i = 2; i < nx; ++i) {
            double BE = BX[(size_t)j * (nx + 1) + i + 1], BW = BX[(size_t)j * (nx + 1) + i];
            double BN = BY[(size_t)(j + 1) * nx + i], BS = BY[(size_t)j * nx + i];
            g->bc_ref[(size_t)j * nx + i] = 0.25f * (BE + BW + BN + BS);
        }
    }
    // Device BC: interior as above; every other cell takes the BC of the
    // nearest interior cell, the cell its ghost value is copied from.
    std::vector<double> bc(g->plane, 0.0);
    for (int j = 0; j < g->R; ++j) {
        int jc = std::min(std::max(j, 2), ny - 3);
        for (int i = 0; i < d.P; ++i) {
            int ic = std::min(std::max(i, 2), nx - 3);
            bc[(size_t)j * d.P + i] = g->bc_ref[(size_t)jc * nx + ic];
        }
    }
    if ((rc = put_field(g, d.BC, bc.data(), d.P, g->R, d.P))) return rc;
    // beds that are means of one corner field (the reference's geometry
    // always gives them) are formed on the device from the corners, as the
    // fused step forms them; any other beds keep the uploaded planes and run
    // the per-phase path
    {
        std::vector<double> corners((size_t)(nx - 3) * (ny - 3));
        g->corner_beds = rift_corners_from_faces(nx, ny, BX, BY, corners.data()) == RIFT_OK;
        if (g->corner_beds && (rc = put_corners(g, corners))) return rc;
    }
    std::vector<double> zero((size_t)nx * ny, 0.0);
    g->n_host.assign(n ? n : zero.data(), (n ? n : zero.data()) + (size_t)nx * ny);
    if ((rc = put_field(g, d.n, g->n_host.data(), nx, ny, nx))) return rc;
    // a Manning field that is constant over the interior is read as a scalar
    d.n_uniform = 1;
    d.n_const = (float)g->n_host[(size_t)2 * nx + 2];
    for (int j = 2; j <= ny - 3 && d.n_uniform; ++j)
        for (int i = 2; i <= nx - 3; ++i)
            if ((float)g->n_host[(size_t)j * nx + i] != d.n_const) { d.n_uniform = 0; break; }
    if (d.K) {
        if (!K) return fail(RIFT_E_ARG, "rift_upload_static: K required with infiltration");
        if ((rc = put_field(g, d.K, K, nx, ny, nx))) return rc;
    }
    // the initial depth is held in h0 until rift_init
    if ((rc = put_field(g, d.h0, h0 ? h0 : zero.data(), nx, ny, nx))) return rc;
    g->static_loaded = true;
    drop_graphs(g);
    return RIFT_OK;
}

int rift_init(rift_grid* g, double t0)
{
    if (!g) return fail(RIFT_E_ARG, "rift_init: null handle");
    if (!g->static_loaded) return fail(RIFT_E_STATE, "rift_init: call rift_upload_static first");
    Dev& d = g->d;
    const size_t bytes = g->plane * sizeof(float);
    // w = BC + h0 on the interior (InitGrid_k, grid.cu:539-547); h0 = 0 elsewhere
    if (d.flags & RIFT_H_INIT) HIPCHK(hipMemcpyAsync(d.h, d.h0, bytes, hipMemcpyDeviceToDevice, g->stream));
    else HIPCHK(hipMemsetAsync(d.h, 0, bytes, g->stream));
    HIPCHK(hipMemcpyAsync(d.h0, d.h, bytes, hipMemcpyDeviceToDevice, g->stream));
    float* zeros[] = {d.hu, d.hv, d.hu0, d.hv0, d.dh, d.dhu, d.dhv, d.G,
                      d.F, d.F0, d.dF, d.hmax, d.qmax, d.tpeak, d.tdry};
    for (float* p : zeros)
        if (p) HIPCHK(hipMemsetAsync(p, 0, bytes, g->stream));
    if (d.twet) {
        std::vector<float> tw(g->plane, 0.f);
        for (int j = 2; j < d.ny; ++j)
            for (int i = 2; i < d.nx; ++i) tw[(size_t)j * d.P + i] = -9999.f;
        HIPCHK(hipMemcpyAsync(d.twet, tw.data(), bytes, hipMemcpyHostToDevice, g->stream));
        HIPCHK(hipStreamSynchronize(g->stream));
    }
    // tiles whose origin cell lies inside nx x ny start wet and active
    std::vector<unsigned char> wet(d.ntiles, 0);
    for (int by = 0; by < d.nby; ++by)
        for (int bx = 0; bx < d.nbx; ++bx)
            if (bx * kTileW + 2 < d.nx && by * kTileH + 2 < d.ny) wet[by * d.nbx + bx] = 1;
    HIPCHK(hipMemcpyAsync(d.wet, wet.data(), d.ntiles, hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipMemcpyAsync(d.act, wet.data(), d.ntiles, hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipMemsetAsync(d.mx, 0, d.ntiles * sizeof(float), g->stream));
    HIPCHK(hipMemsetAsync(d.my, 0, d.ntiles * sizeof(float), g->stream));
    HIPCHK(hipMemsetAsync(d.sel, 0, d.ntiles, g->stream));
    HIPCHK(hipMemsetAsync(d.flip, 0, d.ntiles, g->stream));
    g->scattered = false;
    Scalars sc{};
    HIPCHK(hipMemcpyAsync(&sc, d.sc, sizeof(Scalars), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    int ext = sc.hydro_external;
    double hv = sc.hydro_value;
    sc = Scalars{};
    sc.t = t0;
    sc.dt = (double)0.0000001f;   // simulator.cu:35
    sc.dt_prev = sc.dt;
    sc.hydro_external = ext;
    sc.hydro_value = hv;
    HIPCHK(hipMemcpyAsync(d.sc, &sc, sizeof(Scalars), hipMemcpyHostToDevice, g->stream));
    launch_boundaries(d, g->stream);
    int rc = check_launch("rift_init");
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(g->stream));
    g->inited = true;
    g->primed = false;
    return rift_budget_reset(g);
}

int rift_set_series(rift_grid* g, int which, const float* time_s, const float* value, int n)
{
    if (!g || which < 0 || which > 2 || n < 1 || !time_s || !value)
        return fail(RIFT_E_ARG, "rift_set_series: bad argument");
    if (n > kMaxSeries) return fail(RIFT_E_ARG, "rift_set_series: too many points");
    std::vector<SeriesTable> tab(1);
    tab[0].n = n;
    std::memcpy(tab[0].time, time_s, n * sizeof(float));
    std::memcpy(tab[0].rate, value, n * sizeof(float));
    HIPCHK(hipMemcpyAsync(g->d.ser + which, tab.data(), sizeof(SeriesTable), hipMemcpyHostToDevice,
                          g->stream));
    g->primed = false;
    if (which == RIFT_SER_HYDROGRAPH) {
        int zero = 0;
        HIPCHK(hipMemcpyAsync(&g->d.sc->hydro_external, &zero, sizeof(int), hipMemcpyHostToDevice,
                              g->stream));
    }
    HIPCHK(hipStreamSynchronize(g->stream));
    return RIFT_OK;
}

int rift_set_source_cells(rift_grid* g, const int* idx, long n)
{
    if (!g || (n > 0 && !idx)) return fail(RIFT_E_ARG, "rift_set_source_cells: bad argument");
    Dev& d = g->d;
    g->primed = false;
    if (n == 0) {
        d.si0 = d.sj0 = 1;
        d.si1 = d.sj1 = 0;
        d.nsrc = 0;
        drop_graphs(g);
    return RIFT_OK;
    }
    int i0 = 1 << 30, i1 = -1, j0 = 1 << 30, j1 = -1;
    std::vector<long> seen;
    for (long k = 0; k < n; ++k) {
        int i = idx[k] % d.nx, j = idx[k] / d.nx;
        if (idx[k] < 0 || j >= d.ny)
            return fail(RIFT_E_UNSUPPORTED, "rift_set_source_cells: source cell outside the domain");
        i0 = std::min(i0, i); i1 = std::max(i1, i);
        j0 = std::min(j0, j); j1 = std::max(j1, j);
        seen.push_back(idx[k]);
    }
    std::sort(seen.begin(), seen.end());
    seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
    if ((long)seen.size() != n || (long)(i1 - i0 + 1) * (j1 - j0 + 1) != n)
        return fail(RIFT_E_UNSUPPORTED, "rift_set_source_cells: cells must form a rectangle");
    d.si0 = i0; d.si1 = i1; d.sj0 = j0; d.sj1 = j1;
    d.nsrc = n;
    drop_graphs(g);
    return RIFT_OK;
}

int rift_set_source_value(rift_grid* g, double value)
{
    if (!g) return fail(RIFT_E_ARG, "rift_set_source_value: null handle");
    struct { int ext; int pad; double v; } s{1, 0, value};
    g->primed = false;
    HIPCHK(hipMemcpyAsync(&g->d.sc->hydro_external, &s.ext, sizeof(int), hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipMemcpyAsync(&g->d.sc->hydro_value, &s.v, sizeof(double), hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    return RIFT_OK;
}

int rift_set_map(rift_grid* g, int which, int plane, const double* host)
{
    if (!g || !host) return fail(RIFT_E_ARG, "rift_set_map: null argument");
    float* p = map_plane(g, which, plane);
    if (!p) return fail(RIFT_E_STATE, "rift_set_map: map not enabled by the flags");
    return put_field(g, p, host, g->d.nx, g->d.ny, g->d.nx);
}

int rift_interp_map(rift_grid* g, int which, double factor)
{
    if (!g || !map_plane(g, which, 0)) return fail(RIFT_E_STATE, "rift_interp_map: map not enabled");
    launch_interp_map(g->d, which, (float)factor, g->stream);
    return check_launch("rift_interp_map");
}

int rift_copy_map(rift_grid* g, int which, int dst, int src)
{
    float* a = g ? map_plane(g, which, dst) : nullptr;
    float* b = g ? map_plane(g, which, src) : nullptr;
    if (!a || !b) return fail(RIFT_E_STATE, "rift_copy_map: map not enabled");
    HIPCHK(hipMemcpyAsync(a, b, g->plane * sizeof(float), hipMemcpyDeviceToDevice, g->stream));
    return RIFT_OK;
}

// Per-phase kernels and transfers work on the canonical layout: every
// tile's state in h/hu/hv with its ghost cells filled.
static int ensure_canonical(rift_grid* g)
{
    if (!g->scattered) return RIFT_OK;
    launch_canonicalize(g->d, g->stream);
    launch_boundaries(g->d, g->stream);
    int rc = check_launch("canonicalize");
    if (rc) return rc;
    g->scattered = false;
    return RIFT_OK;
}

#define NEED_INIT(g)                                                        \
    do {                                                                    \
        if (!(g)) return fail(RIFT_E_ARG, "null handle");                   \
        if (!(g)->inited) return fail(RIFT_E_STATE, "rift_init not called"); \
        (g)->primed = false;                                                \
        int rc_ = ensure_canonical(g);                                      \
        if (rc_) return rc_;                                                \
    } while (0)

int rift_update_source(rift_grid* g)
{
    NEED_INIT(g);
    launch_begin_step(g->d, g->stream);
    return check_launch("begin_step");
}

int rift_grow(rift_grid* g)
{
    NEED_INIT(g);
    ProfScope ps(g, 4);
    if (g->d.flags & (RIFT_RAIN_GRID | RIFT_SURGE_GRID)) launch_track(g->d, g->stream);
    launch_grow(g->d, g->stream);
    return check_launch("grow");
}

int rift_compute_fluxes(rift_grid* g)
{
    NEED_INIT(g);
    ProfScope ps(g, 0);
    launch_flux(g->d, 1, g->stream, g->grid_tiles);
    return check_launch("flux");
}

int rift_compute_timestep(rift_grid* g)
{
    NEED_INIT(g);
    ProfScope ps(g, 5);
    launch_timestep(g->d, g->stream);
    return check_launch("timestep");
}

int rift_integrate_1(rift_grid* g)
{
    NEED_INIT(g);
    ProfScope ps(g, 2);
    launch_integrate1(g->d, g->stream, g->grid_tiles);
    return check_launch("integrate1");
}

int rift_integrate_2(rift_grid* g)
{
    NEED_INIT(g);
    ProfScope ps(g, 2);
    launch_integrate2(g->d, g->stream, g->grid_tiles);
    return check_launch("integrate2");
}

int rift_apply_boundaries(rift_grid* g)
{
    NEED_INIT(g);
    ProfScope ps(g, 3);
    launch_boundaries(g->d, g->stream);
    return check_launch("boundaries");
}

constexpr int kGraphSteps = 16;

static int step_fused(rift_grid* g);

// rift_step runs the fused step unless asked for the per-phase path, in
// Euler mode (the fused step is Heun's), or when the face beds are not means
// of one corner field (the fused step forms every bed from the corners).
static bool fused_path(const rift_grid* g)
{
    return !g->per_phase && !(g->d.flags & RIFT_EULER) && g->corner_beds;
}

static int capture_steps(rift_grid* g, int n, hipGraphExec_t* out)
{
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    int rc = RIFT_OK;
    for (int k = 0; k < n && rc == RIFT_OK; ++k) rc = step_fused(g);
    hipError_t e = hipStreamEndCapture(g->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(RIFT_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(RIFT_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    // upload now, so that the first replay does not pay for it
    e = hipGraphUpload(*out, g->stream);
    if (e != hipSuccess) return fail(RIFT_E_HIP, std::string("hipGraphUpload: ") + hipGetErrorString(e));
    return RIFT_OK;
}

// Forcing and counters for a fused step when the last dt kernel did not
// prepare them.
static void prime(rift_grid* g)
{
    if (g->primed) return;
    launch_begin_step(g->d, g->stream);
    g->primed = true;
}

// The fused step: tile/unit compaction (installing the forcing), stage-1
// march, dt (+ the next step's forcing), stage-2 march (Integrate_1 + flux +
// Heun + diagnostics in one pass).  Runs primed, so that it can be captured.
static int step_fused(rift_grid* g)
{
    Dev& d = g->d;
    prime(g);
    const bool dt_in_march = g->dt_in_march;
    {
        ProfScope ps(g, 4);
        if (d.flags & (RIFT_RAIN_GRID | RIFT_SURGE_GRID)) launch_track(d, g->stream);
        launch_grow2(d, g->stream, dt_in_march);
    }
    {
        ProfScope ps(g, 0);
        launch_march(d, 1, g->stream, g->variant, g->march_blocks[0][g->variant], dt_in_march);
    }
    if (!dt_in_march) {
        ProfScope ps(g, 5);
        launch_timestep2(d, g->stream);
    }
    {
        ProfScope ps(g, 1);
        launch_march(d, 2, g->stream, g->variant, g->march_blocks[1][g->variant]);
    }
    g->scattered = true;
    return check_launch("fused step");
}

// Both step graphs captured now (capture records the launches without
// running them, so the state does not move): a later rift_step then only
// replays graphs, whatever its step count.
static int prepare_graphs(rift_grid* g)
{
    if (!g->inited || g->prof || !g->use_graphs || !fused_path(g)) return RIFT_OK;
    prime(g);
    int rc;
    if (!g->graphN && (rc = capture_steps(g, kGraphSteps, &g->graphN))) return rc;
    if (!g->graph1 && (rc = capture_steps(g, 1, &g->graph1))) return rc;
    return RIFT_OK;
}

int rift_step(rift_grid* g, int nsteps)
{
    if (!g) return fail(RIFT_E_ARG, "null handle");
    if (!g->inited) return fail(RIFT_E_STATE, "rift_init not called");
    if (nsteps < 0) return fail(RIFT_E_ARG, "rift_step: negative step count");
    const bool euler = g->d.flags & RIFT_EULER;
    if (fused_path(g)) {
        if (g->prof || !g->use_graphs) {
            for (int s = 0; s < nsteps; ++s) {
                int rc = step_fused(g);
                if (rc) return rc;
            }
            return RIFT_OK;
        }
        int rc;
        int left = nsteps;
        prime(g);   // before capture: graphs hold the primed step only
        if (left >= kGraphSteps && !g->graphN && (rc = capture_steps(g, kGraphSteps, &g->graphN)))
            return rc;
        for (; left >= kGraphSteps; left -= kGraphSteps) HIPCHK(hipGraphLaunch(g->graphN, g->stream));
        if (left > 0 && !g->graph1 && (rc = capture_steps(g, 1, &g->graph1))) return rc;
        for (; left > 0; --left) HIPCHK(hipGraphLaunch(g->graph1, g->stream));
        g->scattered = true;
        return RIFT_OK;
    }
    NEED_INIT(g);
    for (int s = 0; s < nsteps; ++s) {
        int rc;
        if (g->d.flags & (RIFT_HYDROGRAPH | RIFT_RAIN_AVG | RIFT_DRAIN_AVG))
            launch_begin_step(g->d, g->stream);
        // the per-phase kernels read the surge of this step from map plane 0
        if ((g->d.flags & RIFT_SURGE_GRID) && g->d.sg_series) launch_sg_materialize(g->d, g->stream);
        if ((rc = rift_grow(g))) return rc;
        if ((rc = rift_compute_fluxes(g))) return rc;
        if ((rc = rift_compute_timestep(g))) return rc;
        if ((rc = rift_integrate_1(g))) return rc;
        if ((rc = rift_apply_boundaries(g))) return rc;
        if (!euler) {
            {
                ProfScope ps(g, 1);
                launch_flux(g->d, 2, g->stream, g->grid_tiles);
                if ((rc = check_launch("flux2"))) return rc;
            }
            if ((rc = rift_integrate_2(g))) return rc;
            if ((rc = rift_apply_boundaries(g))) return rc;
        }
    }
    return RIFT_OK;
}

int rift_set_option(rift_grid* g, int option, int value)
{
    if (!g) return fail(RIFT_E_ARG, "rift_set_option: null handle");
    switch (option) {
    case RIFT_OPT_STEP_PATH:
        if (value != 0 && value != 1) return fail(RIFT_E_ARG, "rift_set_option: step path is 0 or 1");
        g->per_phase = value == 1;
        if (g->d.sg_series) {
            // the per-phase kernels do not keep the fused step's per-unit
            // surge flags: the next fused step recomputes them
            const int one = 1;
            HIPCHK(hipMemcpyAsync(&g->d.sc->sg_udirty, &one, sizeof(int), hipMemcpyHostToDevice, g->stream));
            HIPCHK(hipMemcpyAsync(&g->d.sc->sg_udirty_load, &one, sizeof(int), hipMemcpyHostToDevice, g->stream));
            HIPCHK(hipStreamSynchronize(g->stream));
        }
        return RIFT_OK;
    case RIFT_OPT_GRAPHS:
        g->use_graphs = value != 0;
        return RIFT_OK;
    case RIFT_OPT_DT_KERNEL:
        if (value != 0 && value != 1) return fail(RIFT_E_ARG, "rift_set_option: dt kernel is 0 or 1");
        if (g->dt_in_march == (value == 0)) return RIFT_OK;
        HIPCHK(hipStreamSynchronize(g->stream));
        drop_graphs(g);   // the graphs hold the launch sequence
        g->dt_in_march = value == 0;
        return prepare_graphs(g);
    case RIFT_OPT_MASS_BUDGET: {
        // the budget instantiation of the marching kernels on or off; the
        // accumulators restart from the current state when it comes on
        const bool on = value != 0;
        if (on == ((g->d.flags & RIFT_MASS_BUDGET) != 0)) return RIFT_OK;
        HIPCHK(hipStreamSynchronize(g->stream));
        drop_graphs(g);
        if (on) g->d.flags |= RIFT_MASS_BUDGET;
        else g->d.flags &= ~RIFT_MASS_BUDGET;
        g->desc.flags = g->d.flags;
        for (int st = 0; st < 2; ++st)
            for (int v = 0; v < kMarchVariants; ++v)
                g->march_blocks[st][v] = march_resident_blocks(g->d, st + 1, v, g->cus);
        if (on && g->inited) {
            int rc = rift_budget_reset(g);
            if (rc) return rc;
        }
        return prepare_graphs(g);
    }
    case RIFT_OPT_VARIANT:
        if (value < 0 || value >= RIFT_NVARIANTS) return fail(RIFT_E_ARG, "rift_set_option: no such variant");
        drop_graphs(g);   // the graphs hold the launch configuration
        g->variant = value & (kMarchVariants - 1);
        g->d.band = (value & kMarchVariants) ? 0 : 1;
        {
            static const int least_rows[4] = {kTileH, kTileH / 2, 2 * kTileH, 3 * kTileH};
            g->d.min_rows = least_rows[(value >> 3) & 3];
        }
        g->d.quads = (value & 32) ? 0 : 1;
        return prepare_graphs(g);
    default:
        return fail(RIFT_E_ARG, "rift_set_option: unknown option");
    }
}

// Device state of g copied into c, a handle created from the same
// descriptor (same allocations in the same order), with the host-side state
// that goes with it.
static int copy_state(rift_grid* g, rift_grid* c)
{
    if (c->allocs.size() != g->allocs.size()) return fail(RIFT_E_STATE, "copy_state: layouts differ");
    HIPCHK(hipStreamSynchronize(g->stream));
    for (size_t k = 0; k < g->allocs.size(); ++k) {
        if (c->alloc_sizes[k] != g->alloc_sizes[k]) return fail(RIFT_E_STATE, "copy_state: layouts differ");
        HIPCHK(hipMemcpyAsync(c->allocs[k], g->allocs[k], g->alloc_sizes[k], hipMemcpyDeviceToDevice,
                              c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    Dev& cd = c->d;
    const Dev& gd = g->d;
    cd.si0 = gd.si0; cd.si1 = gd.si1; cd.sj0 = gd.sj0; cd.sj1 = gd.sj1;
    cd.nsrc = gd.nsrc;
    cd.n_uniform = gd.n_uniform;
    cd.n_const = gd.n_const;
    c->bc_ref = g->bc_ref;
    c->n_host = g->n_host;
    c->static_loaded = g->static_loaded;
    c->corner_beds = g->corner_beds;
    c->inited = g->inited;
    c->scattered = g->scattered;
    c->primed = g->primed;
    return RIFT_OK;
}

int rift_autotune(rift_grid* g, int nsteps, int* best, double* ms_per_step)
{
    if (!g || !best || nsteps < 1) return fail(RIFT_E_ARG, "rift_autotune: bad argument");
    if (!g->inited) return fail(RIFT_E_STATE, "rift_init not called");
    if (!g->corner_beds)
        return fail(RIFT_E_UNSUPPORTED, "rift_autotune: the face beds are not means of one corner field (no fused step)");
    rift_grid* c = nullptr;
    int rc = rift_create(&g->desc, &c);
    if (rc) return rc;
    if (g->d.sg_series &&
        (rc = rift_map_series_create(c, RIFT_MAP_SURGE, g->d.sg_nslots, g->sg_dt_int, g->d.sg_tmax))) {
        rift_destroy(c);
        return rc;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail(RIFT_E_HIP, "rift_autotune: hipEventCreate failed");
    }
    // one variant timed from the snapshot: a warm-up run (captures the
    // variant's graphs), then the timed run from the same snapshot
    auto time_variant = [&](int v, double& per) -> int {
        c->per_phase = g->per_phase;
        c->use_graphs = g->use_graphs;   // timed the way the handle steps
        c->dt_in_march = g->dt_in_march;
        int r;
        if ((r = copy_state(g, c)) || (r = rift_set_option(c, RIFT_OPT_VARIANT, v))) return r;
        if ((r = rift_step(c, nsteps)) || (r = copy_state(g, c))) return r;
        if ((r = rift_step(c, 2))) return r;
        if (hipEventRecord(e0, c->stream) != hipSuccess) return fail(RIFT_E_HIP, "hipEventRecord");
        if ((r = rift_step(c, nsteps))) return r;
        float ms = 0.f;
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            return fail(RIFT_E_HIP, "rift_autotune: event timing failed");
        per = (double)ms / nsteps;
        return RIFT_OK;
    };
    std::vector<double> t(RIFT_NVARIANTS, 0.0);
    for (int v = 0; v < RIFT_NVARIANTS && rc == RIFT_OK; ++v) rc = time_variant(v, t[v]);
    int v_best = -1;
    if (rc == RIFT_OK) {
        if (ms_per_step)
            for (int v = 0; v < RIFT_NVARIANTS; ++v) ms_per_step[v] = t[v];
        // one timing of a variant varies by about 1 %, as much as the
        // fastest variants differ: the fastest few are timed twice more, in
        // turn, and the one with the least total is kept
        constexpr int kTop = 6;
        std::vector<int> order(RIFT_NVARIANTS);
        for (int v = 0; v < RIFT_NVARIANTS; ++v) order[v] = v;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return t[a] < t[b]; });
        std::vector<double> total(kTop);
        for (int k = 0; k < kTop; ++k) total[k] = t[order[k]];
        for (int rep = 0; rep < 2 && rc == RIFT_OK; ++rep)
            for (int k = 0; k < kTop && rc == RIFT_OK; ++k) {
                double per = 0.0;
                rc = time_variant(order[k], per);
                total[k] += per;
            }
        int kb = 0;
        for (int k = 1; k < kTop; ++k)
            if (total[k] < total[kb]) kb = k;
        v_best = order[kb];
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    rift_destroy(c);
    if (rc) return rc;
    *best = v_best;
    return rift_set_option(g, RIFT_OPT_VARIANT, v_best);
}

int rift_map_series_create(rift_grid* g, int which, int nslots, int deltat, double tmax)
{
    if (!g) return fail(RIFT_E_ARG, "rift_map_series_create: null handle");
    if (which != RIFT_MAP_SURGE)
        return fail(RIFT_E_UNSUPPORTED, "rift_map_series_create: only the surge series lives on the device");
    if (!(g->d.flags & RIFT_SURGE_GRID))
        return fail(RIFT_E_STATE, "rift_map_series_create: handle created without RIFT_SURGE_GRID");
    if (g->d.sg_series) return fail(RIFT_E_STATE, "rift_map_series_create: series exists");
    if (nslots < 2 || deltat <= 0) return fail(RIFT_E_ARG, "rift_map_series_create: nslots >= 2, deltat > 0");
    Dev& d = g->d;
    d.sgs = alloc_cells(g, (size_t)nslots * g->plane);
    if (!d.sgs) return fail(RIFT_E_HIP, "rift_map_series_create: device allocation failed");
    d.sgflag = alloc_bytes<unsigned char>(g, (size_t)nslots * d.ntiles);
    d.sgtag = alloc_bytes<int>(g, nslots);
    d.sgu = alloc_bytes<unsigned char>(g, d.nunits);
    if (!d.sgflag || !d.sgtag || !d.sgu) return fail(RIFT_E_HIP, "rift_map_series_create: device allocation failed");
    std::vector<int> tags(nslots, -1);
    HIPCHK(hipMemcpyAsync(d.sgtag, tags.data(), nslots * sizeof(int), hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    // plane 0 (what downloads and the per-phase kernels read) starts at
    // nodata, as p_init_dev fills the series' planes (GridSeries.cu:226-231)
    {
        std::vector<double> fill((size_t)d.nx * d.ny, d.nodata);
        int rc = put_field(g, d.map[RIFT_MAP_SURGE][0], fill.data(), d.nx, d.ny, d.nx);
        if (rc) return rc;
    }
    d.sg_nslots = nslots;
    d.sg_dt = (double)deltat;
    d.sg_tmax = tmax;
    d.sg_series = 1;
    g->sg_dt_int = deltat;
    // the surge series may select another instantiation of the marching kernels
    for (int st = 0; st < 2; ++st)
        for (int v = 0; v < kMarchVariants; ++v) g->march_blocks[st][v] = march_resident_blocks(d, st + 1, v, g->cus);
    g->primed = false;
    drop_graphs(g);
    return RIFT_OK;
}

int rift_map_series_load(rift_grid* g, int which, int index, const double* host)
{
    if (!g || !host) return fail(RIFT_E_ARG, "rift_map_series_load: null argument");
    if (which != RIFT_MAP_SURGE || !g->d.sg_series)
        return fail(RIFT_E_STATE, "rift_map_series_load: no device series for this map");
    if (index < 0) return fail(RIFT_E_ARG, "rift_map_series_load: negative map index");
    Dev& d = g->d;
    const int slot = sg_slot(d, index);
    int rc = put_field(g, d.sgs + (size_t)slot * g->plane, host, d.nx, d.ny, d.nx);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(d.sgtag + slot, &index, sizeof(int), hipMemcpyHostToDevice, g->stream));
    launch_sg_flags(d, slot, g->stream);
    // the next step recomputes the per-unit surge flags (dt kernels carry
    // the request over one step, so a load between steps is seen)
    const int one = 1;
    HIPCHK(hipMemcpyAsync(&d.sc->sg_udirty, &one, sizeof(int), hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipMemcpyAsync(&d.sc->sg_udirty_load, &one, sizeof(int), hipMemcpyHostToDevice, g->stream));
    if ((rc = check_launch("rift_map_series_load"))) return rc;
    HIPCHK(hipStreamSynchronize(g->stream));
    return RIFT_OK;
}

int rift_map_series_start(rift_grid* g, int which, double t)
{
    if (!g) return fail(RIFT_E_ARG, "rift_map_series_start: null handle");
    if (which != RIFT_MAP_SURGE || !g->d.sg_series)
        return fail(RIFT_E_STATE, "rift_map_series_start: no device series for this map");
    Scalars sc;
    HIPCHK(hipStreamSynchronize(g->stream));
    HIPCHK(hipMemcpy(&sc, g->d.sc, sizeof(Scalars), hipMemcpyDeviceToHost));
    // first InterpolatedGridSeries::p_update (GridSeries.cu:333-352)
    const int index = (int)std::trunc(t / (double)g->sg_dt_int);
    sc.sg_in = (double)(index * g->sg_dt_int);
    sc.sg_next = sc.sg_in + g->sg_dt_int;
    sc.sg_done = 0;
    sc.sg_nk0 = sc.sg_k0 = index;
    sc.sg_nk1 = sc.sg_k1 = index + 1;
    const double factor = (t - sc.sg_in) / (sc.sg_next - sc.sg_in);
    sc.sg_nf = sc.sg_f = (float)factor;
    sc.sg_next_valid = 1;
    sc.sg_missing = -1;
    sc.sg_udirty = 1;
    HIPCHK(hipMemcpy(g->d.sc, &sc, sizeof(Scalars), hipMemcpyHostToDevice));
    return RIFT_OK;
}

int rift_map_series_state(rift_grid* g, int which, int* k0, int* k1, double* factor, int* missing)
{
    if (!g) return fail(RIFT_E_ARG, "rift_map_series_state: null handle");
    if (which != RIFT_MAP_SURGE || !g->d.sg_series)
        return fail(RIFT_E_STATE, "rift_map_series_state: no device series for this map");
    Scalars sc;
    HIPCHK(hipMemcpyAsync(&sc, g->d.sc, sizeof(Scalars), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    const bool nv = sc.sg_next_valid != 0;
    if (k0) *k0 = nv ? sc.sg_nk0 : sc.sg_k0;
    if (k1) *k1 = nv ? sc.sg_nk1 : sc.sg_k1;
    if (factor) *factor = nv ? sc.sg_nf : sc.sg_f;
    if (missing) *missing = sc.sg_missing;
    return RIFT_OK;
}

int rift_sync(rift_grid* g)
{
    if (!g) return fail(RIFT_E_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(g->stream));
    return RIFT_OK;
}

int rift_get_time(rift_grid* g, double* t, double* dt)
{
    if (!g || !t || !dt) return fail(RIFT_E_ARG, "rift_get_time: null argument");
    double v[2];
    HIPCHK(hipMemcpyAsync(v, g->d.sc, 2 * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    *t = v[0];
    *dt = v[1];
    return RIFT_OK;
}

int rift_set_time(rift_grid* g, double t, double dt)
{
    if (!g) return fail(RIFT_E_ARG, "rift_set_time: null handle");
    g->primed = false;
    double v[2] = {t, dt};
    HIPCHK(hipMemcpyAsync(g->d.sc, v, 2 * sizeof(double), hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    return RIFT_OK;
}

static float* field_plane(rift_grid* g, int field)
{
    Dev& d = g->d;
    switch (field) {
    case RIFT_HU: return d.hu;
    case RIFT_HV: return d.hv;
    case RIFT_DEPTH: return d.h;
    case RIFT_N: return d.n;
    case RIFT_K: return d.K;
    case RIFT_F: return d.F;
    case RIFT_DW: return d.dh;
    case RIFT_DHU: return d.dhu;
    case RIFT_DHV: return d.dhv;
    case RIFT_G: return d.G;
    case RIFT_DF: return d.dF;
    case RIFT_HMAX: return d.hmax;
    case RIFT_QMAX: return d.qmax;
    case RIFT_TWET: return d.twet;
    case RIFT_TPEAK: return d.tpeak;
    case RIFT_TDRY: return d.tdry;
    case RIFT_SURGE: return d.map[RIFT_MAP_SURGE][0];
    case RIFT_RAINMAP: return d.map[RIFT_MAP_RAIN][0];
    case RIFT_DRAINMAP: return d.map[RIFT_MAP_DRAIN][0];
    default: return nullptr;
    }
}

// The cell whose bed carries the water surface of cell (i, j).  The outflow
// boundary stores a ghost cell's depth as the critical depth H over the
// adjacent interior cell, whose bed the reference adds (ghost w = BC_int + H,
// grid.cu:1325-1411); corner ghosts are never written and keep their own bed.
static inline void surface_bed_cell(int nx, int ny, int i, int j, int& bi, int& bj)
{
    const bool gx = i < 2 || i > nx - 3, gy = j < 2 || j > ny - 3;
    bi = i;
    bj = j;
    if (gx && !gy) bi = i < 2 ? 2 : nx - 3;
    if (gy && !gx) bj = j < 2 ? 2 : ny - 3;
}

int rift_download(rift_grid* g, int field, double* host)
{
    if (!g || !host) return fail(RIFT_E_ARG, "rift_download: null argument");
    {
        int rc0 = ensure_canonical(g);
        if (rc0) return rc0;
    }
    Dev& d = g->d;
    const int nx = d.nx, ny = d.ny;
    std::vector<float> a, b, c;
    int rc;
    auto at = [&](const std::vector<float>& v, int i, int j) { return v[(size_t)j * d.P + i]; };
    if (field == RIFT_BC) {
        std::copy(g->bc_ref.begin(), g->bc_ref.end(), host);
        return RIFT_OK;
    }
    if (field == RIFT_W) {
        // over the reference's FP64 cell bed (InitGrid_k's BC), so a dry
        // cell's surface is exactly its bed and w - BC is never negative for
        // a reader of this w (a hotstart read by the reference computes
        // powf(w - BC, 5/3) for Manning, NaN below zero)
        if ((rc = get_plane(g, d.h, a))) return rc;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int bi, bj;
                surface_bed_cell(nx, ny, i, j, bi, bj);
                host[(size_t)j * nx + i] = (double)at(a, i, j) + g->bc_ref[(size_t)bj * nx + bi];
            }
        return RIFT_OK;
    }
    if (field == RIFT_Q) {
        if ((rc = get_plane(g, d.h, a)) || (rc = get_plane(g, d.hu, b)) || (rc = get_plane(g, d.hv, c)))
            return rc;
        const float kap = d.kappa;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                float h = at(a, i, j), f = 2.f * h / (h * h + std::max(h * h, kap));
                float u = f * at(b, i, j), v = f * at(c, i, j);
                host[(size_t)j * nx + i] = (double)(h * std::sqrt(v * v + u * u));
            }
        return RIFT_OK;
    }
    if (field == RIFT_SURGE && d.sg_series) {
        // the series' window in force, as the reference's current surge map
        launch_sg_materialize(d, g->stream);
        if ((rc = check_launch("rift_download"))) return rc;
    }
    float* p = field_plane(g, field);
    if (!p) return fail(RIFT_E_ARG, "rift_download: field not available with these flags");
    if ((rc = get_plane(g, p, a))) return rc;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) host[(size_t)j * nx + i] = (double)at(a, i, j);
    return RIFT_OK;
}

int rift_copy_field(rift_grid* g, int field, int i0, int j0, int w, int h, float* dst)
{
    if (!g || !dst) return fail(RIFT_E_ARG, "rift_copy_field: null argument");
    {
        int rc0 = ensure_canonical(g);
        if (rc0) return rc0;
    }
    Dev& d = g->d;
    if (i0 < 0 || j0 < 0 || w < 1 || h < 1 || i0 + w > d.nx || j0 + h > d.ny)
        return fail(RIFT_E_ARG, "rift_copy_field: window outside the grid");
    float* p = field_plane(g, field);
    if (!p) return fail(RIFT_E_ARG, "rift_copy_field: field not stored with these flags");
    HIPCHK(hipSetDevice(g->desc.device));
    HIPCHK(hipMemcpy2DAsync(dst, (size_t)w * sizeof(float), p + (size_t)j0 * d.P + i0, (size_t)d.P * sizeof(float),
                            (size_t)w * sizeof(float), (size_t)h, hipMemcpyDeviceToDevice, g->stream));
    return RIFT_OK;
}

int rift_device_alloc(int device, size_t bytes, void** ptr)
{
    if (!ptr || bytes == 0) return fail(RIFT_E_ARG, "rift_device_alloc: bad argument");
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipMalloc(ptr, bytes));
    return RIFT_OK;
}

int rift_device_free(void* ptr)
{
    if (ptr) HIPCHK(hipFree(ptr));
    return RIFT_OK;
}

int rift_memcpy(void* dst, const void* src, size_t bytes, int kind)
{
    if ((!dst || !src) && bytes) return fail(RIFT_E_ARG, "rift_memcpy: null argument");
    if (kind < 0 || kind > 2) return fail(RIFT_E_ARG, "rift_memcpy: kind must be 0, 1 or 2");
    const hipMemcpyKind k[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice};
    if (bytes) HIPCHK(hipMemcpy(dst, src, bytes, k[kind]));
    return RIFT_OK;
}

int rift_upload(rift_grid* g, int field, const double* host)
{
    if (!g || !host) return fail(RIFT_E_ARG, "rift_upload: null argument");
    {
        int rc0 = ensure_canonical(g);
        if (rc0) return rc0;
    }
    Dev& d = g->d;
    const int nx = d.nx, ny = d.ny;
    if (field == RIFT_W) {
        // depth = w - BC over the reference's FP64 cell beds, as downloads
        // form w (a ghost's over the bed of the interior cell its surface is
        // defined on)
        std::vector<double> hh((size_t)nx * ny);
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int bi, bj;
                surface_bed_cell(nx, ny, i, j, bi, bj);
                hh[(size_t)j * nx + i] = host[(size_t)j * nx + i] - g->bc_ref[(size_t)bj * nx + bi];
            }
        return put_field(g, d.h, hh.data(), nx, ny, nx);
    }
    float* p = field_plane(g, field);
    if (!p || field == RIFT_Q) return fail(RIFT_E_ARG, "rift_upload: field not available");
    return put_field(g, p, host, nx, ny, nx);
}

int rift_get_tiles(rift_grid* g, unsigned char* wet, unsigned char* active, int* n_active)
{
    if (!g) return fail(RIFT_E_ARG, "rift_get_tiles: null handle");
    if (wet) HIPCHK(hipMemcpyAsync(wet, g->d.wet, g->d.ntiles, hipMemcpyDeviceToHost, g->stream));
    if (active) HIPCHK(hipMemcpyAsync(active, g->d.act, g->d.ntiles, hipMemcpyDeviceToHost, g->stream));
    int n = 0;
    HIPCHK(hipMemcpyAsync(&n, &g->d.sc->n_active, sizeof(int), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    if (n_active) *n_active = n;
    return RIFT_OK;
}

int rift_set_tiles(rift_grid* g, const unsigned char* wet, const unsigned char* active)
{
    if (!g || !wet || !active) return fail(RIFT_E_ARG, "rift_set_tiles: null argument");
    HIPCHK(hipMemcpyAsync(g->d.wet, wet, g->d.ntiles, hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipMemcpyAsync(g->d.act, active, g->d.ntiles, hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    return RIFT_OK;
}

int rift_get_tile_speeds(rift_grid* g, float* mx, float* my)
{
    if (!g || !mx || !my) return fail(RIFT_E_ARG, "rift_get_tile_speeds: null argument");
    HIPCHK(hipMemcpyAsync(mx, g->d.mx, g->d.ntiles * sizeof(float), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipMemcpyAsync(my, g->d.my, g->d.ntiles * sizeof(float), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    return RIFT_OK;
}

int rift_get_tile_dims(rift_grid* g, int* nbx, int* nby)
{
    if (!g || !nbx || !nby) return fail(RIFT_E_ARG, "rift_get_tile_dims: null argument");
    *nbx = g->d.nbx;
    *nby = g->d.nby;
    return RIFT_OK;
}

int rift_get_desc(rift_grid* g, rift_desc* out)
{
    if (!g || !out) return fail(RIFT_E_ARG, "rift_get_desc: null argument");
    *out = g->desc;
    return RIFT_OK;
}

int rift_get_rates(rift_grid* g, float* out3)
{
    if (!g || !out3) return fail(RIFT_E_ARG, "rift_get_rates: null argument");
    Scalars sc;
    HIPCHK(hipMemcpyAsync(&sc, g->d.sc, sizeof(Scalars), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    out3[0] = sc.rain_rate;
    out3[1] = sc.drain_rate;
    out3[2] = sc.src_cell_rate;
    return RIFT_OK;
}

int rift_sum_interior(rift_grid* g, int field, double* out)
{
    if (!g || !out) return fail(RIFT_E_ARG, "rift_sum_interior: null argument");
    {
        int rc0 = ensure_canonical(g);
        if (rc0) return rc0;
    }
    float* p = field_plane(g, field);
    if (!p) return fail(RIFT_E_ARG, "rift_sum_interior: field not available");
    HIPCHK(hipMemsetAsync(g->dsum, 0, sizeof(double), g->stream));
    launch_sum_interior(g->d, p, 0, g->dsum, g->stream);
    int rc = check_launch("sum_interior");
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out, g->dsum, sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    return RIFT_OK;
}

int rift_sum_interior_noalloc(rift_grid* g, int field, double* out)
{
    if (!g || !out) return fail(RIFT_E_ARG, "rift_sum_interior_noalloc: null argument");
    const long cells = (long)(g->d.nx - 4) * (long)(g->d.ny - 4);
    const long blocks = (cells + kSumCellsPerBlock - 1) / kSumCellsPerBlock;
    if (blocks > kSumMaxBlocks)
        return fail(RIFT_E_UNSUPPORTED, "rift_sum_interior_noalloc: " + std::to_string(cells) +
                                            " interior cells exceed the static scratch (at most " +
                                            std::to_string((long)kSumMaxBlocks * kSumCellsPerBlock) + ")");
    if (blocks < 1) {
        *out = 0.0;
        return RIFT_OK;
    }
    {
        int rc0 = ensure_canonical(g);
        if (rc0) return rc0;
    }
    float* p = field_plane(g, field);
    if (!p) return fail(RIFT_E_ARG, "rift_sum_interior_noalloc: field not available");
    double* sc = nullptr;
    int rc = scratch_for(g->desc.device, &sc);
    if (rc) return rc;
    HIPCHK(hipSetDevice(g->desc.device));
    hipLaunchKernelGGL(k_sum_scratch_clear, dim3(1), dim3(64), 0, g->stream, sc);
    hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)blocks), dim3(256), 0, g->stream, g->d, (const float*)p, sc);
    if ((rc = check_launch("sum_interior_noalloc"))) return rc;
    HIPCHK(hipMemcpyAsync(out, sc + kSumResult, sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    return RIFT_OK;
}

int rift_mass_balance(rift_grid* g, double* out6)
{
    if (!g || !out6) return fail(RIFT_E_ARG, "rift_mass_balance: null argument");
    const double area = g->d.dx * g->d.dy;
    double v = 0.0, vf = 0.0;
    int rc = rift_sum_interior(g, RIFT_DEPTH, &v);
    if (rc) return rc;
    if (g->d.F && (rc = rift_sum_interior(g, RIFT_F, &vf))) return rc;
    Scalars sc;
    HIPCHK(hipMemcpyAsync(&sc, g->d.sc, sizeof(Scalars), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    out6[0] = v * area;
    out6[1] = vf * area;
    out6[2] = sc.v_rain * area;
    out6[3] = sc.v_src * area;
    out6[4] = sc.b_drain * area;
    out6[5] = sc.b_out * area;
    return RIFT_OK;
}

int rift_budget_reset(rift_grid* g)
{
    if (!g) return fail(RIFT_E_ARG, "rift_budget_reset: null handle");
    double m[6];
    int rc = rift_mass_balance(g, m);
    if (rc) return rc;
    g->budget_v0 = m[0] + m[1];
    const double zeros[5] = {0, 0, 0, 0, 0};
    Scalars* sc = g->d.sc;
    HIPCHK(hipMemcpyAsync(&sc->v_rain, zeros, 4 * sizeof(double), hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipMemcpyAsync(&sc->b_rain_grid, zeros, 5 * sizeof(double), hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    return RIFT_OK;
}

int rift_budget(rift_grid* g, double* out9)
{
    if (!g || !out9) return fail(RIFT_E_ARG, "rift_budget: null argument");
    double m[6];
    int rc = rift_mass_balance(g, m);
    if (rc) return rc;
    Scalars sc;
    HIPCHK(hipMemcpyAsync(&sc, g->d.sc, sizeof(Scalars), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    const double area = g->d.dx * g->d.dy;
    out9[0] = m[0];
    out9[1] = m[1];
    out9[2] = g->budget_v0;
    out9[3] = m[2] + sc.b_rain_grid * area;
    out9[4] = m[3];
    out9[5] = sc.b_drain * area;
    out9[6] = sc.b_out * area;
    out9[7] = sc.b_exch * area;
    out9[8] = (m[0] + m[1] - out9[2]) - (out9[3] + out9[4] - out9[5] - out9[6] + out9[7]);
    return RIFT_OK;
}

int rift_get_stream(rift_grid* g, void** stream)
{
    if (!g || !stream) return fail(RIFT_E_ARG, "rift_get_stream: null argument");
    *stream = (void*)g->stream;
    return RIFT_OK;
}

int rift_profile_enable(rift_grid* g, int on)
{
    if (!g) return fail(RIFT_E_ARG, "rift_profile_enable: null handle");
    HIPCHK(hipStreamSynchronize(g->stream));
    for (auto& pe : g->prof_events) {
        g->event_pool.push_back(pe.second.first);
        g->event_pool.push_back(pe.second.second);
    }
    g->prof_events.clear();
    g->prof = on != 0;
    return RIFT_OK;
}

int rift_kernel_stats(rift_grid* g, int kernel, long* launches, double* total_ms)
{
    if (!g || !launches || !total_ms || kernel < 0 || kernel >= kProfClasses)
        return fail(RIFT_E_ARG, "rift_kernel_stats: bad argument");
    HIPCHK(hipStreamSynchronize(g->stream));
    long n = 0;
    double ms = 0.0;
    for (auto& pe : g->prof_events) {
        if (pe.first != kernel) continue;
        float e = 0.f;
        HIPCHK(hipEventElapsedTime(&e, pe.second.first, pe.second.second));
        ms += e;
        ++n;
    }
    *launches = n;
    *total_ms = ms;
    return RIFT_OK;
}

int rift_wet_cell_updates(rift_grid* g, long* last_step, long* total)
{
    if (!g || !last_step || !total) return fail(RIFT_E_ARG, "rift_wet_cell_updates: null argument");
    unsigned long long c[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(c, &g->d.sc->wet_cells, sizeof(c), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    *last_step = (long)c[0];
    *total = (long)c[1];
    return RIFT_OK;
}

// ---------------------------------------------------------------------------
// Ensembles
// ---------------------------------------------------------------------------
}  // extern "C"

struct rift_ensemble {
    std::vector<rift_grid*> members;
    hipStream_t stream = nullptr;
    Dev* devs = nullptr;            // device copy of the stepped members' Dev
    std::vector<Dev> host_devs;
    unsigned flags = 0;             // TEXTBOOK / MASS_BUDGET (common to all members)
    int march_blocks[2][2] = {{0, 0}, {0, 0}};   // [full|lean][stage]
    int max_units = 0, max_tiles = 0;             // over the stepped members
    int min_rows = kTileH;
    std::vector<Dev> uploaded;      // what devs holds (the graphs read it)
    hipGraphExec_t graph = nullptr; // kEnsGraphSteps captured steps
    // members seen at their end time by rift_ensemble_finished: no longer
    // stepped (a step leaves an ended member as it is), until rift_ensemble_set_tf
    std::vector<char> ended;
    std::vector<rift_grid*> stepped;   // the members the launches cover
    double* times = nullptr;        // pinned: the members' times, read together
};

constexpr int kEnsGraphSteps = 16;

static void ens_step_launches(rift_ensemble* e)
{
    const int n = (int)e->stepped.size();
    for (rift_grid* g : e->stepped)
        if (g->d.flags & (RIFT_RAIN_GRID | RIFT_SURGE_GRID)) launch_track(g->d, e->stream);
    // the lean kernels when no member uses a feature outside them (a member
    // whose flags change changes its Dev, which drops the captured graph)
    bool lean = true;
    for (rift_grid* g : e->stepped) lean = lean && march_lean(g->d.flags);
    const int* blocks = e->march_blocks[lean ? 1 : 0];
    launch_grow2_ens(e->devs, n, e->max_tiles, e->max_units, e->stream);
    launch_march_ens(e->devs, n, e->flags, lean, 1, e->min_rows, blocks[0], e->stream);
    launch_timestep2_ens(e->devs, n, e->max_tiles, e->stream);
    launch_march_ens(e->devs, n, e->flags, lean, 2, e->min_rows, blocks[1], e->stream);
}

extern "C" {

int rift_ensemble_create(int n, const rift_desc* descs, rift_ensemble** out)
{
    if (n < 1 || !descs || !out) return fail(RIFT_E_ARG, "rift_ensemble_create: bad argument");
    *out = nullptr;
    const unsigned common = RIFT_TEXTBOOK | RIFT_MASS_BUDGET | RIFT_EULER;
    for (int i = 0; i < n; ++i) {
        if (descs[i].device != descs[0].device)
            return fail(RIFT_E_UNSUPPORTED, "rift_ensemble_create: members must share one device");
        if ((descs[i].flags & common) != (descs[0].flags & common))
            return fail(RIFT_E_UNSUPPORTED,
                        "rift_ensemble_create: members must agree on TEXTBOOK, MASS_BUDGET and EULER");
    }
    if (descs[0].flags & RIFT_EULER)
        return fail(RIFT_E_UNSUPPORTED, "rift_ensemble_create: Euler members run one at a time");
    HIPCHK(hipSetDevice(descs[0].device));
    rift_ensemble* e = new rift_ensemble();
    hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        delete e;
        return fail(RIFT_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
    }
    e->flags = descs[0].flags & common;
    int rc = RIFT_OK;
    g_member_stream = e->stream;
    for (int i = 0; i < n && rc == RIFT_OK; ++i) {
        rift_grid* g = nullptr;
        rc = rift_create(&descs[i], &g);
        if (rc == RIFT_OK) e->members.push_back(g);
    }
    g_member_stream = nullptr;
    if (rc == RIFT_OK && (hipMalloc(&e->devs, sizeof(Dev) * n) != hipSuccess ||
                          hipHostMalloc(&e->times, sizeof(double) * n) != hipSuccess))
        rc = fail(RIFT_E_HIP, "rift_ensemble_create: allocation failed");
    if (rc != RIFT_OK) {
        rift_ensemble_destroy(e);
        return rc;
    }
    e->ended.assign(n, 0);
    const int cus = e->members[0]->cus;
    for (int lean = 0; lean < 2; ++lean)
        for (int st = 0; st < 2; ++st)
            e->march_blocks[lean][st] = march_ens_resident_blocks(
                st + 1, (e->flags & RIFT_TEXTBOOK) != 0, (e->flags & RIFT_MASS_BUDGET) != 0, lean != 0, cus);
    *out = e;
    return RIFT_OK;
}

rift_grid* rift_ensemble_member(rift_ensemble* e, int i)
{
    if (!e || i < 0 || i >= (int)e->members.size()) return nullptr;
    return e->members[i];
}

int rift_ensemble_size(rift_ensemble* e) { return e ? (int)e->members.size() : 0; }

int rift_ensemble_set_tf(rift_ensemble* e, int i, double tf)
{
    if (!e || i < 0 || i >= (int)e->members.size())
        return fail(RIFT_E_ARG, "rift_ensemble_set_tf: bad argument");
    e->members[i]->d.tf = tf;
    e->ended[i] = 0;   // stepped again until seen at its (new) end time
    return RIFT_OK;
}

int rift_ensemble_step(rift_ensemble* e, int nsteps)
{
    if (!e || nsteps < 0) return fail(RIFT_E_ARG, "rift_ensemble_step: bad argument");
    // the members not yet seen at their end time: an ended member's step
    // only applies its pending buffer flips (a read of its fields applies
    // them too) and clears its per-tile maxima and step counters, so the
    // launches leave it out and every member's fields, time and tiles are
    // what stepping it would leave
    e->stepped.clear();
    e->host_devs.clear();
    e->max_units = e->max_tiles = 0;
    for (size_t i = 0; i < e->members.size(); ++i) {
        rift_grid* g = e->members[i];
        if (!g->inited) return fail(RIFT_E_STATE, "rift_ensemble_step: member not initialised");
        if (g->per_phase) return fail(RIFT_E_UNSUPPORTED, "rift_ensemble_step: member on the per-phase path");
        if (!g->corner_beds)
            return fail(RIFT_E_UNSUPPORTED,
                        "rift_ensemble_step: member's face beds are not means of one corner field");
        if (e->ended[i]) continue;
        prime(g);
        e->stepped.push_back(g);
        e->host_devs.push_back(g->d);
        e->max_units = std::max(e->max_units, g->d.nunits);
        e->max_tiles = std::max(e->max_tiles, g->d.ntiles);
    }
    const int n = (int)e->stepped.size();
    if (n == 0 || nsteps == 0) return RIFT_OK;
    const bool same = e->uploaded.size() == (size_t)n &&
                      std::memcmp(e->uploaded.data(), e->host_devs.data(), sizeof(Dev) * n) == 0;
    if (!same) {
        if (e->graph) {
            (void)hipGraphExecDestroy(e->graph);
            e->graph = nullptr;
        }
        HIPCHK(hipMemcpyAsync(e->devs, e->host_devs.data(), sizeof(Dev) * n, hipMemcpyHostToDevice,
                              e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        e->uploaded = e->host_devs;
    }
    int left = nsteps;
    // graph replays unless a member has graphs switched off (RIFT_OPT_GRAPHS)
    bool graphs = true;
    for (rift_grid* g : e->members) graphs = graphs && g->use_graphs;
    if (!graphs) {
        for (; left > 0; --left) ens_step_launches(e);
        for (rift_grid* g : e->stepped) g->scattered = true;
        return check_launch("rift_ensemble_step");
    }
    if (left >= kEnsGraphSteps && !e->graph) {
        // the ensemble's step loop as one graph (no memset nodes)
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < kEnsGraphSteps; ++k) ens_step_launches(e);
        hipError_t ce = hipStreamEndCapture(e->stream, &graph);
        if (ce != hipSuccess) return fail(RIFT_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        ce = hipGraphInstantiate(&e->graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ce != hipSuccess) return fail(RIFT_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ce));
    }
    for (; left >= kEnsGraphSteps; left -= kEnsGraphSteps) HIPCHK(hipGraphLaunch(e->graph, e->stream));
    for (; left > 0; --left) ens_step_launches(e);
    for (rift_grid* g : e->stepped) g->scattered = true;
    return check_launch("rift_ensemble_step");
}

int rift_ensemble_finished(rift_ensemble* e, int* n_finished)
{
    if (!e || !n_finished) return fail(RIFT_E_ARG, "rift_ensemble_finished: null argument");
    const int n = (int)e->members.size();
    for (int i = 0; i < n; ++i)
        HIPCHK(hipMemcpyAsync(e->times + i, &e->members[i]->d.sc->t, sizeof(double), hipMemcpyDeviceToHost,
                              e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    int done = 0;
    for (int i = 0; i < n; ++i) {
        const double tf = e->members[i]->d.tf;
        if (tf > 0.0 && e->times[i] >= tf) {
            ++done;
            e->ended[i] = 1;
        }
    }
    *n_finished = done;
    return RIFT_OK;
}

int rift_ensemble_run(rift_ensemble* e, int check_every, long max_steps, long* steps_done)
{
    if (!e || check_every < 1) return fail(RIFT_E_ARG, "rift_ensemble_run: bad argument");
    for (rift_grid* g : e->members)
        if (!(g->d.tf > 0.0)) return fail(RIFT_E_STATE, "rift_ensemble_run: every member needs an end time");
    long done_steps = 0;
    int nfin = 0;
    int rc = rift_ensemble_finished(e, &nfin);
    while (rc == RIFT_OK && nfin < (int)e->members.size() && (max_steps < 0 || done_steps < max_steps)) {
        const long k = (max_steps < 0) ? check_every : std::min<long>(check_every, max_steps - done_steps);
        if ((rc = rift_ensemble_step(e, (int)k))) break;
        done_steps += k;
        rc = rift_ensemble_finished(e, &nfin);
    }
    if (steps_done) *steps_done = done_steps;
    return rc;
}

int rift_ensemble_destroy(rift_ensemble* e)
{
    if (!e) return RIFT_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->graph) (void)hipGraphExecDestroy(e->graph);
    for (rift_grid* g : e->members) rift_destroy(g);
    if (e->devs) (void)hipFree(e->devs);
    if (e->times) (void)hipHostFree(e->times);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return RIFT_OK;
}

}  // extern "C"
