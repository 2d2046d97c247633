// C-ABI of the lr3d library: contexts, the PatchMatch stereo set and the TSDF volume.
//
// Host semantics follow the reference's DenseStereo (frontend/src/app/vision/mvs/dense-stereo.ts,
// DS:<line>) and TsdfVolume (frontend/src/app/vision/fusion/tsdf-volume.ts, TV:<line>): the same
// parameter packing (DS:314-355, TV:244-253), seed derivation (DS:333), schedule (DS:185-239) and
// culling (TV:220-234).  The reference slices work into row bands and awaits every submit so a
// phone GPU stays responsive; bands do not change results, so here each dispatch covers whole
// images and a call synchronises once, at its end.
#include "lr3d_internal.h"

#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lr3d.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

using namespace lr3d;

// ---- errors ---------------------------------------------------------------------------------------

static thread_local std::string g_error;

static int32_t fail(int32_t code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

// Fault injection, for tests of the error paths: with LR3D_FAULT_INJECT=k (k >= 1) in the
// environment, the k-th HIP runtime call the library checks (counted over the process from 1)
// reports a device failure, once, whatever its own status; the device is synchronised first, so
// no work of the failing call is still in flight when it returns.  Unset or 0: no injection.
static int64_t fault_inject_at() {
    static const int64_t k = [] {
        const char *v = getenv("LR3D_FAULT_INJECT");
        return v ? (int64_t)atoll(v) : (int64_t)0;
    }();
    return k;
}
static std::atomic<int64_t> g_checked_calls{0};
static bool fault_injected() {
    const int64_t k = fault_inject_at();
    if (k <= 0 || g_checked_calls.fetch_add(1) + 1 != k) return false;
    (void)hipDeviceSynchronize();
    return true;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (fault_injected())                                                                      \
            return fail(LR3D_ERR_DEVICE, "%s: injected fault (LR3D_FAULT_INJECT=%lld) (%s:%d)", #expr, \
                        (long long)fault_inject_at(), __FILE__, __LINE__);                          \
        if (e_ != hipSuccess)                                                                      \
            return fail(e_ == hipErrorOutOfMemory ? LR3D_ERR_NOMEM : LR3D_ERR_DEVICE, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                         \
    } while (0)

extern "C" const char *lr3d_last_error(void) { return g_error.c_str(); }

// ---- device buffers ----------------------------------------------------------------------------------

// Owns its allocation: moves hand it over, copies are not allowed (a copy's destructor would free
// memory the original still uses).
template <typename T>
struct DevBuf {
    T *ptr = nullptr;
    size_t count = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), count(o.count) {
        o.ptr = nullptr;
        o.count = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            ptr = o.ptr;
            count = o.count;
            o.ptr = nullptr;
            o.count = 0;
        }
        return *this;
    }
    hipError_t alloc(size_t n) {
        release();
        if (n == 0) return hipSuccess;
        hipError_t e = hipMalloc((void **)&ptr, n * sizeof(T));
        if (e == hipSuccess) count = n; else ptr = nullptr;
        return e;
    }
    // Grow-only reallocation (contents are not kept).
    hipError_t reserve(size_t n) { return n <= count ? hipSuccess : alloc(n); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    ~DevBuf() { release(); }
};

// ---- plan scratch of the allocation-free batched solve ------------------------------------------------
//
// lr3d_stereo_solve_batched_noalloc keeps its launch plan (solve descriptors and block tables) in
// this fixed array instead of a buffer it allocates: every GPU the library runs on holds one copy
// until the process ends, and every call on that GPU reuses it (calls that overlap on one GPU race
// on it).  In host code the array's name stands for a host-side stand-in, not device memory: host
// code only ever uses the address that scratch_on_device() resolved with the device current.
constexpr size_t kNoAllocScratchBytes = (size_t)1 << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static std::map<int32_t, uint8_t *> g_scratch_ptr;  // device -> its copy of `scratch`

// Device address of `scratch` on `device`, which must be current; resolved once per device.
static hipError_t scratch_on_device(int32_t device, uint8_t **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = (uint8_t *)p;
    *out = (uint8_t *)p;
    return hipSuccess;
}

// ---- context ------------------------------------------------------------------------------------------

struct KernelStats {
    int64_t launches = 0;
    double total_ms = 0.0;
};

struct PendingEvent {
    std::string name;
    hipEvent_t start, stop;
};

struct lr3d_ctx {
    int32_t device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    bool profiling = false;
    std::map<std::string, KernelStats> stats;
    std::vector<hipEvent_t> event_pool;
    std::vector<PendingEvent> pending;
    int64_t graph_launches = 0;
    int64_t users = 0;      // live stereo sets, volumes, meshes and clouds made on this context
    bool closing = false;   // lr3d_ctx_destroy was called while users remained
    uint8_t *scratch = nullptr;  // this device's copy of `scratch` (device address)
    DevBuf<float> wtab;          // bilateral weight table of the PatchMatch window
    DevBuf<float4> ntab;         // decoded normal of every packed octahedral code (TSDF pass)

    hipEvent_t take_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    // Fold finished launch timings into the statistics (stream must be idle).
    void collect() {
        for (auto &p : pending) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
                auto &s = stats[p.name];
                s.launches++;
                s.total_ms += ms;
            } else {
                (void)hipGetLastError();
            }
            event_pool.push_back(p.start);
            event_pool.push_back(p.stop);
        }
        pending.clear();
    }
    ~lr3d_ctx() {
        for (auto &p : pending) {
            (void)hipEventDestroy(p.start);
            (void)hipEventDestroy(p.stop);
        }
        for (auto e : event_pool) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Object lifetimes: every stereo set, volume, mesh and cloud is a user of its context.  A context
// destroyed while users remain is released by the destroy of its last user, so no object ever
// outlives the stream and mutex it uses.  Creation counts a user with ctx->mu held; the destroy of
// an object calls ctx_release with ctx->mu held through `lock`.
static void ctx_release(lr3d_ctx *ctx, std::unique_lock<std::mutex> &lock) {
    ctx->users--;
    if (ctx->closing && ctx->users == 0) {
        lock.unlock();
        delete ctx;
    }
}

// Records start/stop events around one launch when profiling is on.
template <typename F>
static hipError_t timed(lr3d_ctx *ctx, const char *name, F &&launch) {
    (void)hipGetLastError();  // launchers report hipGetLastError(): start from a clean slate
    if (!ctx->profiling) return launch();
    hipEvent_t a = ctx->take_event(), b = ctx->take_event();
    if (!a || !b) return hipErrorOutOfMemory;
    hipError_t e = hipEventRecord(a, ctx->stream);
    if (e != hipSuccess) return e;
    e = launch();
    if (e != hipSuccess) return e;
    e = hipEventRecord(b, ctx->stream);
    if (e != hipSuccess) return e;
    ctx->pending.push_back({name, a, b});
    return hipSuccess;
}

static int32_t finish(lr3d_ctx *ctx) {
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->profiling) ctx->collect();
    return LR3D_OK;
}

// PM:90-110 bilateral weights of every (offset class, |dc|): the f32 argument -(dc dc) kColorK -
// |o|^2 kSpaceK as the window writes it, exp in double, rounded to f32.
static void window_weight_table(std::vector<float> &out) {
    out.resize((size_t)kWindowClasses * 256);
    for (int r = 0; r < kWindowClasses; r++) {
        const float ds = (float)kWindowClassDist[r];
        for (int a = 0; a < 256; a++) {
            const float dc = (float)a;
            const float arg = -(dc * dc) * kColorK - ds * kSpaceK;
            out[(size_t)r * 256 + a] = (float)exp((double)arg);
        }
    }
}

extern "C" int32_t lr3d_ctx_create(int32_t device, lr3d_ctx **out) {
    if (!out) return fail(LR3D_ERR_INVALID, "lr3d_ctx_create: out is NULL");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (fault_injected())
        return fail(LR3D_ERR_DEVICE, "lr3d_ctx_create: hipGetDeviceCount: injected fault (LR3D_FAULT_INJECT=%lld)",
                    (long long)fault_inject_at());
    if (e != hipSuccess || count == 0)
        return fail(LR3D_ERR_DEVICE, "lr3d_ctx_create: no HIP device (%s)",
                    e != hipSuccess ? hipGetErrorString(e) : "0 devices");
    if (device < 0 || device >= count)
        return fail(LR3D_ERR_INVALID, "lr3d_ctx_create: device %d out of range [0, %d)", device, count);
    HIP_TRY(hipSetDevice(device));
    uint8_t *plan_scratch = nullptr;
    e = scratch_on_device(device, &plan_scratch);
    if (e != hipSuccess)
        return fail(LR3D_ERR_DEVICE, "lr3d_ctx_create: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    lr3d_ctx *ctx = new lr3d_ctx();
    ctx->device = device;
    ctx->scratch = plan_scratch;
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(LR3D_ERR_DEVICE, "lr3d_ctx_create: hipStreamCreate: %s", hipGetErrorString(e));
    }
    std::vector<float> table;
    window_weight_table(table);
    e = ctx->wtab.alloc(table.size());
    if (e == hipSuccess)
        e = hipMemcpy(ctx->wtab.ptr, table.data(), table.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = ctx->ntab.alloc(4096);
    if (e == hipSuccess) e = tsdf_launch_normal_table(ctx->stream, ctx->ntab.ptr);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        delete ctx;
        return fail(e == hipErrorOutOfMemory ? LR3D_ERR_NOMEM : LR3D_ERR_DEVICE,
                    "lr3d_ctx_create: weight and normal tables: %s", hipGetErrorString(e));
    }
    *out = ctx;
    return LR3D_OK;
}

extern "C" int32_t lr3d_ctx_destroy(lr3d_ctx *ctx) {
    if (!ctx) return LR3D_OK;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        if (ctx->users > 0) {  // released by the last user's destroy
            ctx->closing = true;
            return LR3D_OK;
        }
    }
    delete ctx;
    return LR3D_OK;
}

extern "C" int32_t lr3d_ctx_synchronize(lr3d_ctx *ctx) {
    if (!ctx) return fail(LR3D_ERR_INVALID, "lr3d_ctx_synchronize: ctx is NULL");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    return finish(ctx);
}

extern "C" int32_t lr3d_ctx_set_profiling(lr3d_ctx *ctx, int32_t enable) {
    if (!ctx) return fail(LR3D_ERR_INVALID, "lr3d_ctx_set_profiling: ctx is NULL");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    int32_t rc = finish(ctx);
    if (rc != LR3D_OK) return rc;
    ctx->profiling = enable != 0;
    if (ctx->profiling) ctx->stats.clear();
    return LR3D_OK;
}

extern "C" int32_t lr3d_ctx_kernel_stats(lr3d_ctx *ctx, const char *name, int64_t *launches,
                                         double *total_ms) {
    if (!ctx || !name) return fail(LR3D_ERR_INVALID, "lr3d_ctx_kernel_stats: NULL argument");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (strcmp(name, "graph_launch") == 0) {  // replays of captured launch sequences (lifetime count)
        if (launches) *launches = ctx->graph_launches;
        if (total_ms) *total_ms = 0.0;
        return LR3D_OK;
    }
    auto it = ctx->stats.find(name);
    if (launches) *launches = it == ctx->stats.end() ? 0 : it->second.launches;
    if (total_ms) *total_ms = it == ctx->stats.end() ? 0.0 : it->second.total_ms;
    return LR3D_OK;
}

// ---- stereo -----------------------------------------------------------------------------------------

// A captured launch sequence of one batch layout, replayed while the layout is unchanged.
struct GraphEntry {
    std::vector<int64_t> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static void destroy_graph(GraphEntry &g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
}

struct lr3d_stereo {
    lr3d_ctx *ctx = nullptr;
    std::vector<lr3d_view> views;
    std::vector<uint32_t> off;  // pixel offset per layer
    size_t total_px = 0;
    std::vector<uint32_t> qoff;  // texel offset per layer in the quad pool (quad_texels per layer)
    size_t total_q = 0;
    std::vector<uint8_t> has_image;
    std::vector<uint8_t> solved;
    DevBuf<uint32_t> quad, depth, filt;
    DevBuf<float4> hyp;
    DevBuf<float> cost, prior;
    DevBuf<uint8_t> staging;  // u8 upload of one layer
    DevBuf<uint8_t> plan_buf; // descriptors + block tables of the current launch plan
    DevBuf<float> adj_weights, adj_cost, adj_gd, adj_gn;  // adjoint inputs / outputs
    DevBuf<double> adj_partials;
    DevBuf<AdjGeom> adj_geom;
    lr3d_stereo *coarse = nullptr;
    std::vector<lr3d_stereo *> fine_levels;  // stereos using this one as their coarse level
    std::vector<GraphEntry> graphs;
    // launch shapes (threads x, y) of init, propagate and pack/consistency; set by the autotuner
    TileShape tiles[3] = {pm_tile_for("init", 0, 0, 0), pm_tile_for("propagate", 0, 0, 0),
                          pm_tile_for("pack", 0, 0, 0)};

    ~lr3d_stereo() {
        for (auto &g : graphs) destroy_graph(g);
    }
};

static bool valid_view(const lr3d_view &v) {
    return v.width > 0 && v.height > 0 && v.width <= 16384 && v.height <= 16384 && v.fx > 0 &&
           v.fy > 0 && std::isfinite(v.cx) && std::isfinite(v.cy);
}

extern "C" int32_t lr3d_stereo_create(lr3d_ctx *ctx, const lr3d_view *views, int32_t n_views,
                                      uint32_t flags, lr3d_stereo **out) {
    if (!ctx || !views || !out) return fail(LR3D_ERR_INVALID, "lr3d_stereo_create: NULL argument");
    *out = nullptr;
    if (n_views <= 0) return fail(LR3D_ERR_INVALID, "lr3d_stereo_create: n_views must be > 0");
    if (flags & ~(uint32_t)LR3D_STEREO_PRIOR_POOL)
        return fail(LR3D_ERR_INVALID, "lr3d_stereo_create: unknown flags 0x%x", flags);
    size_t total = 0, total_q = 0;
    std::vector<uint32_t> off(n_views), qoff(n_views);
    for (int32_t i = 0; i < n_views; i++) {
        if (!valid_view(views[i]))
            return fail(LR3D_ERR_INVALID, "lr3d_stereo_create: view %d has an invalid size or intrinsics", i);
        off[i] = (uint32_t)total;
        qoff[i] = (uint32_t)total_q;
        total += (size_t)views[i].width * (size_t)views[i].height;
        total_q += quad_texels(views[i].width, views[i].height);
        if (total_q >= (1ull << 32))
            return fail(LR3D_ERR_INVALID, "lr3d_stereo_create: more than 2^32 pixels in one set");
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (ctx->closing) return fail(LR3D_ERR_STATE, "lr3d_stereo_create: the context has been destroyed");
    HIP_TRY(hipSetDevice(ctx->device));
    lr3d_stereo *st = new lr3d_stereo();
    st->ctx = ctx;
    st->views.assign(views, views + n_views);
    st->off = off;
    st->total_px = total;
    st->qoff = qoff;
    st->total_q = total_q;
    st->has_image.assign(n_views, 0);
    st->solved.assign(n_views, 0);
    hipError_t e = st->quad.alloc(total_q);
    if (e == hipSuccess) e = st->depth.alloc(total);
    if (e == hipSuccess) e = st->filt.alloc(total);
    if (e == hipSuccess) e = st->hyp.alloc(total);
    if (e == hipSuccess) e = st->cost.alloc(total);
    if (e == hipSuccess && (flags & LR3D_STEREO_PRIOR_POOL)) e = st->prior.alloc(total);
    if (e == hipSuccess) e = hipMemsetAsync(st->quad.ptr, 0, total_q * 4, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(st->depth.ptr, 0, total * 4, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(st->filt.ptr, 0, total * 4, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(st->hyp.ptr, 0, total * 16, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(st->cost.ptr, 0, total * 4, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        delete st;
        return fail(e == hipErrorOutOfMemory ? LR3D_ERR_NOMEM : LR3D_ERR_DEVICE,
                    "lr3d_stereo_create: %s", hipGetErrorString(e));
    }
    ctx->users++;
    *out = st;
    return LR3D_OK;
}

static void detach_coarse(lr3d_stereo *st) {
    if (!st->coarse) return;
    auto &users = st->coarse->fine_levels;
    users.erase(std::remove(users.begin(), users.end(), st), users.end());
    st->coarse = nullptr;
}

extern "C" int32_t lr3d_stereo_destroy(lr3d_stereo *st) {
    if (!st) return LR3D_OK;
    lr3d_ctx *ctx = st->ctx;
    std::unique_lock<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    detach_coarse(st);
    // stereos that read this one as their coarse level lose it (a later from_coarse solve fails
    // with LR3D_ERR_STATE instead of reading freed memory); their graphs hold its planes' address
    for (lr3d_stereo *fine : st->fine_levels) {
        fine->coarse = nullptr;
        for (auto &g : fine->graphs) destroy_graph(g);
        fine->graphs.clear();
    }
    delete st;
    ctx_release(ctx, lock);
    return LR3D_OK;
}

static int32_t check_layer(const lr3d_stereo *st, int32_t layer, const char *what) {
    if (layer < 0 || layer >= (int32_t)st->views.size())
        return fail(LR3D_ERR_INVALID, "%s: layer %d out of range [0, %d)", what, layer,
                    (int32_t)st->views.size());
    return LR3D_OK;
}

extern "C" int32_t lr3d_stereo_set_image(lr3d_stereo *st, int32_t layer, const uint8_t *gray,
                                         int32_t width, int32_t height) {
    if (!st || !gray) return fail(LR3D_ERR_INVALID, "lr3d_stereo_set_image: NULL argument");
    int32_t rc = check_layer(st, layer, "lr3d_stereo_set_image");
    if (rc) return rc;
    const lr3d_view &v = st->views[layer];
    if (width != v.width || height != v.height)
        return fail(LR3D_ERR_INVALID, "lr3d_stereo_set_image: layer %d is %dx%d, got %dx%d", layer,
                    v.width, v.height, width, height);
    lr3d_ctx *ctx = st->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t px = (size_t)width * height;
    HIP_TRY(st->staging.reserve(px));
    HIP_TRY(hipMemcpyAsync(st->staging.ptr, gray, px, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(timed(ctx, "pm_build_quad", [&] {
        return pm_build_quad(ctx->stream, st->staging.ptr, width, height, quad_rows(height),
                             st->quad.ptr + st->qoff[layer]);
    }));
    st->has_image[layer] = 1;
    return finish(ctx);
}

// Launch shapes of the PatchMatch kernels (threads x, y).  The per-pixel kernels use 64-wide
// rows of one wavefront; propagate uses 8 x 8 threads = 16 x 8 pixels, the reference's tile.
TileShape lr3d::pm_tile_for(const char *kernel, int32_t W, int32_t H, int32_t n_src) {
    (void)W;
    (void)H;
    (void)n_src;
    if (strcmp(kernel, "propagate") == 0) return {8, 8};
    if (strcmp(kernel, "init") == 0) return {8, 8};
    return {64, 4};
}

// DS:314-355 for one solve, in double then rounded to f32 like Float32Array.set.
// CAM:73-77 relativePose(ref, src) with math/mat3.ts multiply3 / apply3, in double:
// R = R_s R_r^T, t = t_s - R t_r.
static void relative_pose(const lr3d_view &rv, const lr3d_view &sv, double R[9], double t[3]) {
    const double *a = rv.R, *b = sv.R;
    const double aT[9] = {a[0], a[3], a[6], a[1], a[4], a[7], a[2], a[5], a[8]};
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++)
            R[r * 3 + c] = b[r * 3] * aT[c] + b[r * 3 + 1] * aT[3 + c] + b[r * 3 + 2] * aT[6 + c];
    for (int r = 0; r < 3; r++) {
        const double rt = R[r * 3] * rv.t[0] + R[r * 3 + 1] * rv.t[1] + R[r * 3 + 2] * rv.t[2];
        t[r] = sv.t[r] - rt;
    }
}

// The plane-independent parts of the homography H = Ks (R + t n^T / rho) Kr^-1 (PM:116-128) of
// reference view rv and source view sv, in double, rounded to f32: G = Ks R Kr^-1 and e = Ks t, so
// that H = G + (e / rho) (Kr^-T n)^T.  R, t: the relative pose ref -> src (CAM:73-77).
static void homography_parts(const lr3d_view &rv, const lr3d_view &sv, const double R[9], const double t[3],
                             float G[9], float e[3]) {
    double A[9];  // Ks R
    for (int j = 0; j < 3; j++) {
        A[j] = sv.fx * R[j] + sv.cx * R[6 + j];
        A[3 + j] = sv.fy * R[3 + j] + sv.cy * R[6 + j];
        A[6 + j] = R[6 + j];
    }
    for (int i = 0; i < 3; i++) {  // (Ks R) Kr^-1
        G[i * 3 + 0] = (float)(A[i * 3 + 0] / rv.fx);
        G[i * 3 + 1] = (float)(A[i * 3 + 1] / rv.fy);
        G[i * 3 + 2] = (float)((A[i * 3 + 2] - A[i * 3 + 0] * rv.cx / rv.fx) - A[i * 3 + 1] * rv.cy / rv.fy);
    }
    e[0] = (float)(sv.fx * t[0] + sv.cx * t[2]);
    e[1] = (float)(sv.fy * t[1] + sv.cy * t[2]);
    e[2] = (float)t[2];
}

static int32_t build_desc(const lr3d_stereo *st, const lr3d_solve_desc &sd, uint32_t prior_off,
                          bool require_images, PmDesc &d) {
    const char *what = "lr3d_stereo_solve";
    int32_t rc = check_layer(st, sd.ref, what);
    if (rc) return rc;
    if (sd.n_sources < 0) return fail(LR3D_ERR_INVALID, "%s: n_sources < 0", what);
    const int32_t n_src = std::min(sd.n_sources, (int32_t)kMaxSources);  // DS:186
    if (!(sd.min_depth > 0.0) || !(sd.max_depth > sd.min_depth))
        return fail(LR3D_ERR_INVALID, "%s: depth range must satisfy 0 < min_depth < max_depth", what);
    if (sd.iterations < 0 || sd.iterations > 64)
        return fail(LR3D_ERR_INVALID, "%s: iterations must be in [0, 64]", what);
    if (sd.topk < 0 || sd.topk > kMaxSources)
        return fail(LR3D_ERR_INVALID, "%s: topk must be in [0, %d]", what, kMaxSources);
    if (require_images && !st->has_image[sd.ref])
        return fail(LR3D_ERR_STATE, "%s: layer %d has no image", what, sd.ref);
    const lr3d_view &rv = st->views[sd.ref];
    memset(&d, 0, sizeof(d));
    d.W = rv.width;
    d.H = rv.height;
    d.ref = sd.ref;
    d.n_src = n_src;
    d.k[0] = (float)rv.fx;
    d.k[1] = (float)rv.fy;
    d.k[2] = (float)rv.cx;
    d.k[3] = (float)rv.cy;
    d.ikx = 1.0f / d.k[0];
    d.iky = 1.0f / d.k[1];
    d.kx2 = -d.k[2] / d.k[0];
    d.ky2 = -d.k[3] / d.k[1];
    d.inv_min = (float)(1.0 / sd.min_depth);
    d.inv_max = (float)(1.0 / sd.max_depth);
    d.geo_weight = 0.2f;  // DS:46
    d.flags = sd.geometric ? 1u : 0u;
    int32_t k = sd.topk > 0 ? std::min(sd.topk, n_src) : std::max(1, std::min(3, n_src - 1));  // PM:218
    d.topk = std::max(1, k);
    d.iterations = sd.iterations;
    d.seed_base = (uint32_t)(sd.ref + 1) * 0x9e3779b1u + d.flags * 7u;  // DS:333
    d.off = st->off[sd.ref];
    d.qrows = quad_rows(rv.height);
    d.qoff = st->qoff[sd.ref];
    d.has_prior = sd.prior != nullptr;
    d.prior_off = prior_off;
    d.from_coarse = sd.from_coarse && st->coarse != nullptr;
    if (d.from_coarse) {
        const lr3d_view &cv = st->coarse->views[sd.ref];
        // CAM:28-36 scaleIntrinsics(k, cw / W, ch / H), in double
        const double sx = (double)cv.width / (double)rv.width, sy = (double)cv.height / (double)rv.height;
        d.ck[0] = (float)(rv.fx * sx);
        d.ck[1] = (float)(rv.fy * sy);
        d.ck[2] = (float)((rv.cx + 0.5) * sx - 0.5);
        d.ck[3] = (float)((rv.cy + 0.5) * sy - 0.5);
        d.coarse_W = cv.width;
        d.coarse_H = cv.height;
        d.coarse_off = st->coarse->off[sd.ref];
    }
    for (int32_t s = 0; s < n_src; s++) {
        const int32_t layer = sd.sources[s];
        rc = check_layer(st, layer, what);
        if (rc) return rc;
        if (require_images && !st->has_image[layer])
            return fail(LR3D_ERR_STATE, "%s: source layer %d has no image", what, layer);
        const lr3d_view &sv = st->views[layer];
        double R[9], t[3];
        relative_pose(rv, sv, R, t);
        PmSrc &o = d.src[s];
        o.k[0] = (float)sv.fx;
        o.k[1] = (float)sv.fy;
        o.k[2] = (float)sv.cx;
        o.k[3] = (float)sv.cy;
        for (int i = 0; i < 9; i++) o.R[i] = (float)R[i];
        for (int i = 0; i < 3; i++) o.t[i] = (float)t[i];
        homography_parts(rv, sv, R, t, o.G, o.e);
        o.W = sv.width;
        o.H = sv.height;
        o.off = st->off[layer];
        o.layer = layer;
        o.qrows = quad_rows(sv.height);
        o.qoff = st->qoff[layer];
    }
    return LR3D_OK;
}

// Device copy of descriptors and the three block tables (init, propagate, pack) of one launch
// group.
struct Plans {
    PmPlan init, prop, pack;
    int32_t max_iterations = 0;
    bool any_pack = false;
};

// Launch plans of consecutive groups of `descs` (group g = descs[groups[g], groups[g + 1])), all
// in one device buffer: the descriptors, then each group's three block tables.  The buffer is the
// set's plan buffer (grown as needed), or `dest` (dest_bytes long) when given.
// tiles: launch shapes of init, propagate, pack (the stereo's own when NULL).
static int32_t upload_groups(lr3d_stereo *st, const std::vector<PmDesc> &descs, const std::vector<uint8_t> &pack,
                             const std::vector<int32_t> &groups, std::vector<Plans> &plans,
                             std::vector<int64_t> &key, const TileShape *tiles = nullptr,
                             uint8_t *dest = nullptr, size_t dest_bytes = 0) {
    lr3d_ctx *ctx = st->ctx;
    const int32_t n = (int32_t)descs.size();
    const int32_t n_groups = (int32_t)groups.size() - 1;
    if (!tiles) tiles = st->tiles;
    const TileShape ti = tiles[0], tp = tiles[1], tk = tiles[2];
    const size_t desc_bytes = sizeof(PmDesc) * n;
    // per group: 3 tables of (group size + 1) entries
    const size_t table_words = 3 * ((size_t)n + (size_t)n_groups);
    const size_t total = desc_bytes + 4 * table_words;
    std::vector<uint8_t> host(total);
    memcpy(host.data(), descs.data(), desc_bytes);
    uint32_t *tab = (uint32_t *)(host.data() + desc_bytes);
    key.clear();
    key.push_back(n);
    key.push_back(n_groups);
    plans.assign(n_groups, Plans());
    std::vector<size_t> table_at(n_groups);
    size_t w = 0;
    for (int32_t g = 0; g < n_groups; g++) {
        const int32_t b = groups[g], m = groups[g + 1] - b;
        key.push_back(b);
        uint32_t *bi = tab + w, *bp = bi + (m + 1), *bk = bp + (m + 1);
        table_at[g] = w;
        w += 3 * (size_t)(m + 1);
        int32_t kmax = 1, geometric = 0;
        bi[0] = bp[0] = bk[0] = 0;
        Plans &pl = plans[g];
        for (int32_t v = 0; v < m; v++) {
            const PmDesc &d = descs[b + v];
            const uint64_t rows_i = (d.H + ti.ty - 1) / ti.ty, cols_i = (d.W + ti.tx - 1) / ti.tx;
            const uint64_t rows_p = (d.H + tp.ty - 1) / tp.ty, cols_p = (d.W + 2 * tp.tx - 1) / (2 * tp.tx);
            const uint64_t rows_k = (d.H + tk.ty - 1) / tk.ty, cols_k = (d.W + tk.tx - 1) / tk.tx;
            bi[v + 1] = bi[v] + (uint32_t)(rows_i * cols_i);
            bp[v + 1] = bp[v] + (d.iterations > 0 ? (uint32_t)(rows_p * cols_p) : 0u);
            bk[v + 1] = bk[v] + (pack[b + v] ? (uint32_t)(rows_k * cols_k) : 0u);
            kmax = std::max(kmax, d.topk);
            geometric |= (int32_t)(d.flags & 1u);
            pl.max_iterations = std::max(pl.max_iterations, d.iterations);
            pl.any_pack |= pack[b + v] != 0;
            key.push_back(d.W);
            key.push_back(d.H);
            key.push_back(d.iterations);
            key.push_back(pack[b + v]);
        }
        pl.init = {nullptr, nullptr, m, bi[m], ti.tx, ti.ty, kmax, geometric};
        pl.prop = {nullptr, nullptr, m, bp[m], tp.tx, tp.ty, kmax, geometric};
        pl.pack = {nullptr, nullptr, m, bk[m], tk.tx, tk.ty, kmax, geometric};
        key.push_back(kmax);
        key.push_back(geometric);
    }
    uint8_t *buf = dest;
    if (!buf) {
        HIP_TRY(st->plan_buf.reserve(total));
        buf = st->plan_buf.ptr;
    } else if (total > dest_bytes) {
        return fail(LR3D_ERR_INVALID, "launch plan of %zu bytes exceeds the %zu-byte plan scratch", total, dest_bytes);
    }
    HIP_TRY(hipMemcpyAsync(buf, host.data(), total, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));  // host staging goes out of scope
    const PmDesc *dd = (const PmDesc *)buf;
    const uint32_t *tables = (const uint32_t *)(buf + desc_bytes);
    for (int32_t g = 0; g < n_groups; g++) {
        const int32_t m = groups[g + 1] - groups[g];
        const uint32_t *t0 = tables + table_at[g];
        plans[g].init.descs = plans[g].prop.descs = plans[g].pack.descs = dd + groups[g];
        plans[g].init.block_start = t0;
        plans[g].prop.block_start = t0 + (m + 1);
        plans[g].pack.block_start = t0 + 2 * (m + 1);
    }
    key.push_back((int64_t)(uintptr_t)buf);
    for (int i = 0; i < 3; i++) {
        key.push_back(tiles[i].tx);
        key.push_back(tiles[i].ty);
    }
    return LR3D_OK;
}

// One launch group over all of `descs`.
static int32_t upload_plans(lr3d_stereo *st, const std::vector<PmDesc> &descs,
                            const std::vector<uint8_t> &pack, Plans &plans, std::vector<int64_t> &key,
                            const TileShape *tiles = nullptr) {
    std::vector<Plans> groups;
    int32_t rc = upload_groups(st, descs, pack, {0, (int32_t)descs.size()}, groups, key, tiles);
    if (rc) return rc;
    plans = groups[0];
    return LR3D_OK;
}

// The geometric pass runs layer after layer, each layer reading the packed maps its predecessors
// have just written (…/pipeline/reconstructor.ts:96-108, Gauss-Seidel over the views).  A batch is
// cut into groups that run one after another, each ending with its pack: a solve starts a new
// group when it is a geometric pass whose sources include the layer of an earlier solve of the
// current group that packs.  Every other read sees the same map as in a sequence of single
// solves: photometric passes read no packed map; a geometric pass reads its own layer's map (no
// other solve writes it) and its sources' maps, which a later solve of its group only overwrites
// at the group's end.  Photometric batches stay one group.
static std::vector<int32_t> gauss_seidel_groups(const std::vector<PmDesc> &dv, const std::vector<uint8_t> &packs) {
    std::vector<int32_t> groups(1, 0);
    std::vector<int32_t> written;  // layers packed by the current group
    for (int32_t i = 0; i < (int32_t)dv.size(); i++) {
        const PmDesc &d = dv[i];
        bool reads_group = false;
        if (d.flags & 1u)
            for (int32_t s = 0; s < d.n_src && !reads_group; s++)
                reads_group = std::find(written.begin(), written.end(), d.src[s].layer) != written.end();
        if (reads_group) {
            groups.push_back(i);
            written.clear();
        }
        if (packs[i]) written.push_back(d.ref);
    }
    groups.push_back((int32_t)dv.size());
    return groups;
}

static PmPools pools_of(lr3d_stereo *st) {
    PmPools p;
    p.quad = st->quad.ptr;
    p.hyp = st->hyp.ptr;
    p.cost = st->cost.ptr;
    p.depth = st->depth.ptr;
    p.filt = st->filt.ptr;
    p.prior = st->prior.ptr;
    p.coarse_hyp = st->coarse ? st->coarse->hyp.ptr : nullptr;
    p.wtab = st->ctx->wtab.ptr;
    return p;
}

// The schedule of DS:185-239 for one launch group: init, iterations x {phase 0, 1}, pack.
// with_timing: event pairs around the launches when profiling is on (events are created on
// demand, so the allocation-free entry point launches without them).
static hipError_t enqueue_schedule(lr3d_stereo *st, const Plans &plans, bool with_timing) {
    lr3d_ctx *ctx = st->ctx;
    const PmPools pools = pools_of(st);
    auto launch = [&](const char *name, auto &&f) -> hipError_t {
        if (with_timing) return timed(ctx, name, f);
        (void)hipGetLastError();
        return f();
    };
    hipError_t e = launch("pm_init", [&] { return pm_launch_init(ctx->stream, plans.init, pools); });
    if (e != hipSuccess) return e;
    for (int32_t it = 0; it < plans.max_iterations; it++) {
        const float step = (float)pow(0.5, (double)it);  // DS:203
        for (uint32_t phase = 0; phase < 2; phase++) {
            e = launch("pm_propagate", [&] {
                return pm_launch_propagate(ctx->stream, plans.prop, pools, (uint32_t)it, phase, step);
            });
            if (e != hipSuccess) return e;
        }
    }
    if (plans.any_pack) {
        e = launch("pm_pack", [&] { return pm_launch_pack(ctx->stream, plans.pack, pools); });
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

static hipError_t enqueue_groups(lr3d_stereo *st, const std::vector<Plans> &groups, bool with_timing = true) {
    for (const Plans &g : groups) {
        hipError_t e = enqueue_schedule(st, g, with_timing);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

static int32_t stage_priors(lr3d_stereo *st, const lr3d_solve_desc *descs, int32_t n,
                            std::vector<uint32_t> &prior_off, bool may_allocate = true) {
    lr3d_ctx *ctx = st->ctx;
    prior_off.assign(n, 0);
    bool any = false;
    for (int32_t i = 0; i < n; i++) any |= descs[i].prior != nullptr;
    if (!any) return LR3D_OK;
    if (!st->prior.ptr && !may_allocate)
        return fail(LR3D_ERR_STATE, "lr3d_stereo_solve_batched_noalloc: a solve passes a prior but the set has "
                                    "no prior pool (create it with LR3D_STEREO_PRIOR_POOL)");
    if (!st->prior.ptr) {
        HIP_TRY(st->prior.alloc(st->total_px));
        // graphs captured so far hold the pools as they were (no prior buffer): capture again
        for (auto &g : st->graphs) destroy_graph(g);
        st->graphs.clear();
    }
    for (int32_t i = 0; i < n; i++) {
        if (!descs[i].prior) continue;
        const int32_t ref = descs[i].ref;
        const lr3d_view &v = st->views[ref];
        prior_off[i] = st->off[ref];
        HIP_TRY(hipMemcpyAsync(st->prior.ptr + st->off[ref], descs[i].prior,
                               (size_t)v.width * v.height * 4, hipMemcpyHostToDevice, ctx->stream));
    }
    return LR3D_OK;
}

// batched: shared launches replayed from a captured graph; noalloc: shared launches with the plan in
// the device's `scratch` array, launched eagerly, no device memory allocated.
static int32_t solve_impl(lr3d_stereo *st, const lr3d_solve_desc *descs, int32_t n, bool batched,
                          bool noalloc = false) {
    lr3d_ctx *ctx = st->ctx;
    const char *what = noalloc ? "lr3d_stereo_solve_batched_noalloc"
                               : batched ? "lr3d_stereo_solve_batched" : "lr3d_stereo_solve";
    for (int32_t i = 0; i < n; i++) {
        int32_t rc = check_layer(st, descs[i].ref, what);
        if (rc) return rc;
        for (int32_t j = 0; j < i; j++)
            if (descs[j].ref == descs[i].ref)
                return fail(LR3D_ERR_INVALID, "%s: layer %d appears twice", what, descs[i].ref);
    }
    std::vector<uint32_t> prior_off;
    int32_t rc = stage_priors(st, descs, n, prior_off, !noalloc);
    if (rc) return rc;
    std::vector<PmDesc> dv;
    dv.reserve(n);
    std::vector<int32_t> refs;
    std::vector<uint8_t> packs;
    for (int32_t i = 0; i < n; i++) {
        if (std::min(descs[i].n_sources, (int32_t)kMaxSources) <= 0) continue;  // DS:187: no-op
        if (descs[i].from_coarse && !st->coarse)
            return fail(LR3D_ERR_STATE, "lr3d_stereo_solve: from_coarse without a coarse level");
        PmDesc d;
        rc = build_desc(st, descs[i], prior_off[i], true, d);
        if (rc) return rc;
        dv.push_back(d);
        packs.push_back(descs[i].pack ? 1 : 0);
        refs.push_back(descs[i].ref);
    }
    if (dv.empty()) return finish(ctx);
    std::vector<Plans> plans;
    std::vector<int64_t> key;
    rc = upload_groups(st, dv, packs, gauss_seidel_groups(dv, packs), plans, key, nullptr,
                       noalloc ? ctx->scratch : nullptr, noalloc ? kNoAllocScratchBytes : 0);
    if (rc) return rc;
    hipError_t e = hipSuccess;
    // HIP cannot time event records captured into a graph (hipEventElapsedTime reports an
    // invalid handle), so profiled calls launch eagerly with an event pair around each kernel.
    const bool use_graph = batched && !noalloc && !ctx->profiling;
    GraphEntry *entry = nullptr;
    if (use_graph) {
        for (auto &g : st->graphs)
            if (g.key == key) entry = &g;
        if (!entry) {
            GraphEntry g;
            g.key = key;
            HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            e = enqueue_groups(st, plans);
            hipGraph_t graph = nullptr;
            hipError_t e2 = hipStreamEndCapture(ctx->stream, &graph);
            g.graph = graph;
            if (e == hipSuccess && e2 == hipSuccess) e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            if (e != hipSuccess || e2 != hipSuccess) {
                destroy_graph(g);
                return fail(LR3D_ERR_DEVICE, "lr3d_stereo_solve_batched: graph capture failed: %s",
                            hipGetErrorString(e != hipSuccess ? e : e2));
            }
            if (st->graphs.size() >= 8) {
                destroy_graph(st->graphs.front());
                st->graphs.erase(st->graphs.begin());
            }
            st->graphs.push_back(std::move(g));
            entry = &st->graphs.back();
        }
        HIP_TRY(hipGraphLaunch(entry->exec, ctx->stream));
        ctx->graph_launches++;
    } else {
        e = enqueue_groups(st, plans, !noalloc);
        if (e != hipSuccess) return fail(LR3D_ERR_DEVICE, "%s: launch failed: %s", what, hipGetErrorString(e));
    }
    rc = finish(ctx);
    if (rc) return rc;
    for (int32_t r : refs) st->solved[r] = 1;
    return LR3D_OK;
}

extern "C" int32_t lr3d_stereo_solve(lr3d_stereo *st, const lr3d_solve_desc *desc) {
    if (!st || !desc) return fail(LR3D_ERR_INVALID, "lr3d_stereo_solve: NULL argument");
    std::lock_guard<std::mutex> lock(st->ctx->mu);
    HIP_TRY(hipSetDevice(st->ctx->device));
    return solve_impl(st, desc, 1, false);
}

extern "C" int32_t lr3d_stereo_solve_batched(lr3d_stereo *st, const lr3d_solve_desc *descs, int32_t n) {
    if (!st || (!descs && n > 0)) return fail(LR3D_ERR_INVALID, "lr3d_stereo_solve_batched: NULL argument");
    if (n < 0) return fail(LR3D_ERR_INVALID, "lr3d_stereo_solve_batched: n < 0");
    std::lock_guard<std::mutex> lock(st->ctx->mu);
    HIP_TRY(hipSetDevice(st->ctx->device));
    return solve_impl(st, descs, n, true);
}

extern "C" int32_t lr3d_stereo_solve_batched_noalloc(lr3d_stereo *st, const lr3d_solve_desc *descs, int32_t n) {
    if (!st || (!descs && n > 0)) return fail(LR3D_ERR_INVALID, "lr3d_stereo_solve_batched_noalloc: NULL argument");
    if (n < 0 || n > LR3D_NOALLOC_MAX_SOLVES)
        return fail(LR3D_ERR_INVALID, "lr3d_stereo_solve_batched_noalloc: n = %d, must be in [0, %d]", n,
                    LR3D_NOALLOC_MAX_SOLVES);
    std::lock_guard<std::mutex> lock(st->ctx->mu);
    HIP_TRY(hipSetDevice(st->ctx->device));
    return solve_impl(st, descs, n, true, true);
}

// ---- launch-shape autotuner (no counterpart in the reference) ----------------------------------------
//
// Candidate block shapes (threads x, y; whole wavefronts, at most 256 threads).  Every pixel's
// result is independent of the tiling, so any choice gives identical maps.

static const TileShape kTileCandidates[3][10] = {
    {{8, 8}, {16, 4}, {32, 2}, {64, 1}, {16, 8}, {64, 4}},                                   // init
    {{8, 8}, {16, 4}, {4, 16}, {32, 2}, {64, 1}, {16, 8}, {8, 16}, {32, 4}, {16, 16}, {64, 4}},  // propagate
    {{64, 4}, {64, 1}, {64, 2}, {32, 8}, {256, 1}},                                          // pack, consistency
};
static const int kTileCandidateCount[3] = {6, 10, 5};

static int tile_slot(const char *kernel) {
    if (!kernel) return -1;
    if (strcmp(kernel, "init") == 0) return 0;
    if (strcmp(kernel, "propagate") == 0) return 1;
    if (strcmp(kernel, "pack") == 0) return 2;
    return -1;
}

static bool valid_tile(int32_t tx, int32_t ty) {
    return tx > 0 && ty > 0 && tx * ty <= 256 && (tx * ty) % 64 == 0;
}

extern "C" int32_t lr3d_stereo_get_tile(lr3d_stereo *st, const char *kernel, int32_t *tx, int32_t *ty) {
    if (!st || !tx || !ty) return fail(LR3D_ERR_INVALID, "lr3d_stereo_get_tile: NULL argument");
    const int slot = tile_slot(kernel);
    if (slot < 0) return fail(LR3D_ERR_INVALID, "lr3d_stereo_get_tile: kernel must be init, propagate or pack");
    std::lock_guard<std::mutex> lock(st->ctx->mu);
    *tx = st->tiles[slot].tx;
    *ty = st->tiles[slot].ty;
    return LR3D_OK;
}

extern "C" int32_t lr3d_stereo_set_tile(lr3d_stereo *st, const char *kernel, int32_t tx, int32_t ty) {
    if (!st) return fail(LR3D_ERR_INVALID, "lr3d_stereo_set_tile: NULL argument");
    const int slot = tile_slot(kernel);
    if (slot < 0) return fail(LR3D_ERR_INVALID, "lr3d_stereo_set_tile: kernel must be init, propagate or pack");
    if (!valid_tile(tx, ty))
        return fail(LR3D_ERR_INVALID, "lr3d_stereo_set_tile: %d x %d is not 64, 128, 192 or 256 threads", tx, ty);
    std::lock_guard<std::mutex> lock(st->ctx->mu);
    st->tiles[slot] = {tx, ty};
    return LR3D_OK;
}

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

extern "C" int32_t lr3d_stereo_autotune(lr3d_stereo *st, const lr3d_solve_desc *descs, int32_t n, int32_t repeats) {
    if (!st || !descs || n <= 0) return fail(LR3D_ERR_INVALID, "lr3d_stereo_autotune: NULL argument or n <= 0");
    lr3d_ctx *ctx = st->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    repeats = std::max(1, std::min(repeats, 16));
    for (int32_t i = 0; i < n; i++) {
        int32_t rc = check_layer(st, descs[i].ref, "lr3d_stereo_autotune");
        if (rc) return rc;
        for (int32_t j = 0; j < i; j++)
            if (descs[j].ref == descs[i].ref)
                return fail(LR3D_ERR_INVALID, "lr3d_stereo_autotune: layer %d appears twice", descs[i].ref);
        if (descs[i].from_coarse && !st->coarse)
            return fail(LR3D_ERR_STATE, "lr3d_stereo_autotune: from_coarse without a coarse level");
    }
    std::vector<uint32_t> prior_off;
    int32_t rc = stage_priors(st, descs, n, prior_off);
    if (rc) return rc;
    std::vector<PmDesc> dv;
    std::vector<uint8_t> packs;
    bool any_iteration = false;
    int32_t max_it = 0;
    for (int32_t i = 0; i < n; i++) {
        if (std::min(descs[i].n_sources, (int32_t)kMaxSources) <= 0) continue;
        PmDesc d;
        rc = build_desc(st, descs[i], prior_off[i], true, d);
        if (rc) return rc;
        d.iterations = std::min(d.iterations, 2);  // propagate is timed on the second iteration
        any_iteration |= d.iterations > 0;
        max_it = std::max(max_it, d.iterations);
        dv.push_back(d);
        packs.push_back(1);
    }
    if (dv.empty()) return LR3D_OK;
    // Tuning leaves the set as it found it: planes, costs and packed maps are restored.
    DevBuf<uint32_t> sdepth;
    DevBuf<float4> shyp;
    DevBuf<float> scost;
    HIP_TRY(sdepth.alloc(st->total_px));
    HIP_TRY(shyp.alloc(st->total_px));
    HIP_TRY(scost.alloc(st->total_px));
    HIP_TRY(hipMemcpyAsync(sdepth.ptr, st->depth.ptr, st->total_px * 4, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(shyp.ptr, st->hyp.ptr, st->total_px * 16, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(scost.ptr, st->cost.ptr, st->total_px * 4, hipMemcpyDeviceToDevice, ctx->stream));
    EventPair ev;
    HIP_TRY(hipEventCreate(&ev.a));
    HIP_TRY(hipEventCreate(&ev.b));
    const PmPools pools = pools_of(st);
    // Each candidate starts from the same state: the packed maps restored (geometric passes read
    // them) and, for propagate and pack, a fresh init.
    auto time_slot = [&](int slot, TileShape cand, float &avg_ms) -> int32_t {
        TileShape tl[3] = {st->tiles[0], st->tiles[1], st->tiles[2]};
        tl[slot] = cand;
        Plans plans;
        std::vector<int64_t> key;
        int32_t r = upload_plans(st, dv, packs, plans, key, tl);
        if (r) return r;
        float total = 0.0f;
        for (int32_t k = 0; k < repeats; k++) {
            HIP_TRY(hipMemcpyAsync(st->depth.ptr, sdepth.ptr, st->total_px * 4, hipMemcpyDeviceToDevice, ctx->stream));
            if (slot != 0) HIP_TRY(pm_launch_init(ctx->stream, plans.init, pools));
            // propagate: when the batch has a second iteration, the first runs untimed so that the
            // timed one sees settled planes (solves with a single iteration then sit it out)
            const uint32_t it = max_it >= 2 ? 1u : 0u;
            if (slot == 1 && it == 1u)
                for (uint32_t phase = 0; phase < 2; phase++)
                    HIP_TRY(pm_launch_propagate(ctx->stream, plans.prop, pools, 0u, phase, 1.0f));
            HIP_TRY(hipEventRecord(ev.a, ctx->stream));
            if (slot == 0) HIP_TRY(pm_launch_init(ctx->stream, plans.init, pools));
            if (slot == 1)
                for (uint32_t phase = 0; phase < 2; phase++)
                    HIP_TRY(pm_launch_propagate(ctx->stream, plans.prop, pools, it, phase, (float)pow(0.5, (double)it)));
            if (slot == 2) HIP_TRY(pm_launch_pack(ctx->stream, plans.pack, pools));
            HIP_TRY(hipEventRecord(ev.b, ctx->stream));
            HIP_TRY(hipEventSynchronize(ev.b));
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
            total += ms;
        }
        avg_ms = total / (float)repeats;
        return LR3D_OK;
    };
    const int order[3] = {1, 0, 2};
    for (int o = 0; o < 3; o++) {
        const int slot = order[o];
        if (slot == 1 && !any_iteration) continue;
        float best_ms = 0.0f;
        TileShape best = st->tiles[slot];
        for (int c = 0; c < kTileCandidateCount[slot]; c++) {
            float ms = 0.0f;
            rc = time_slot(slot, kTileCandidates[slot][c], ms);
            if (rc) return rc;
            if (c == 0 || ms < best_ms) {
                best_ms = ms;
                best = kTileCandidates[slot][c];
            }
        }
        st->tiles[slot] = best;
    }
    HIP_TRY(hipMemcpyAsync(st->depth.ptr, sdepth.ptr, st->total_px * 4, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(st->hyp.ptr, shyp.ptr, st->total_px * 16, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(st->cost.ptr, scost.ptr, st->total_px * 4, hipMemcpyDeviceToDevice, ctx->stream));
    return finish(ctx);
}

extern "C" int32_t lr3d_stereo_filter(lr3d_stereo *st, int32_t ref, const int32_t *checks,
                                      int32_t n_checks, int32_t needed) {
    if (!st || (!checks && n_checks > 0)) return fail(LR3D_ERR_INVALID, "lr3d_stereo_filter: NULL argument");
    int32_t rc = check_layer(st, ref, "lr3d_stereo_filter");
    if (rc) return rc;
    if (n_checks < 0) return fail(LR3D_ERR_INVALID, "lr3d_stereo_filter: n_checks < 0");
    lr3d_ctx *ctx = st->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    // DS:250-263: the filter packs params with depth range 0.01..100 and step 0.
    lr3d_solve_desc sd;
    memset(&sd, 0, sizeof(sd));
    sd.ref = ref;
    sd.n_sources = std::min(n_checks, (int32_t)kMaxSources);
    for (int32_t i = 0; i < sd.n_sources; i++) sd.sources[i] = checks[i];
    sd.min_depth = 0.01;
    sd.max_depth = 100.0;
    PmDesc d;
    // the filter reads packed maps only: images are not required
    rc = build_desc(st, sd, 0, false, d);
    if (rc) return rc;
    std::vector<PmDesc> dv(1, d);
    std::vector<uint8_t> packs(1, 1);
    Plans plans;
    std::vector<int64_t> key;
    rc = upload_plans(st, dv, packs, plans, key);
    if (rc) return rc;
    const PmPools pools = pools_of(st);
    HIP_TRY(timed(ctx, "pm_consistency", [&] {
        return pm_launch_consistency(ctx->stream, plans.pack, pools, (uint32_t)std::max(needed, 0));
    }));
    return finish(ctx);
}

static int32_t read_layer(lr3d_stereo *st, int32_t layer, const void *src, void *dst, size_t elem,
                          const char *what) {
    int32_t rc = check_layer(st, layer, what);
    if (rc) return rc;
    lr3d_ctx *ctx = st->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const lr3d_view &v = st->views[layer];
    HIP_TRY(hipMemcpyAsync(dst, (const uint8_t *)src + (size_t)st->off[layer] * elem,
                           (size_t)v.width * v.height * elem, hipMemcpyDeviceToHost, ctx->stream));
    return finish(ctx);
}

extern "C" int32_t lr3d_stereo_read_depth(lr3d_stereo *st, int32_t layer, uint32_t *out) {
    if (!st || !out) return fail(LR3D_ERR_INVALID, "lr3d_stereo_read_depth: NULL argument");
    return read_layer(st, layer, st->depth.ptr, out, 4, "lr3d_stereo_read_depth");
}

extern "C" int32_t lr3d_stereo_read_filtered(lr3d_stereo *st, int32_t layer, uint32_t *out) {
    if (!st || !out) return fail(LR3D_ERR_INVALID, "lr3d_stereo_read_filtered: NULL argument");
    return read_layer(st, layer, st->filt.ptr, out, 4, "lr3d_stereo_read_filtered");
}

extern "C" int32_t lr3d_stereo_read_hyp(lr3d_stereo *st, int32_t layer, float *hyp, float *cost) {
    if (!st) return fail(LR3D_ERR_INVALID, "lr3d_stereo_read_hyp: NULL argument");
    if (hyp) {
        int32_t rc = read_layer(st, layer, st->hyp.ptr, hyp, 16, "lr3d_stereo_read_hyp");
        if (rc) return rc;
    }
    if (cost) return read_layer(st, layer, st->cost.ptr, cost, 4, "lr3d_stereo_read_hyp");
    return LR3D_OK;
}

extern "C" int32_t lr3d_stereo_write_depth(lr3d_stereo *st, int32_t layer, const uint32_t *packed) {
    if (!st || !packed) return fail(LR3D_ERR_INVALID, "lr3d_stereo_write_depth: NULL argument");
    int32_t rc = check_layer(st, layer, "lr3d_stereo_write_depth");
    if (rc) return rc;
    lr3d_ctx *ctx = st->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const lr3d_view &v = st->views[layer];
    HIP_TRY(hipMemcpyAsync(st->depth.ptr + st->off[layer], packed, (size_t)v.width * v.height * 4,
                           hipMemcpyHostToDevice, ctx->stream));
    return finish(ctx);
}

extern "C" int32_t lr3d_stereo_set_coarse(lr3d_stereo *st, lr3d_stereo *coarse) {
    if (!st) return fail(LR3D_ERR_INVALID, "lr3d_stereo_set_coarse: NULL argument");
    if (coarse) {
        if (coarse == st) return fail(LR3D_ERR_INVALID, "lr3d_stereo_set_coarse: a set cannot be its own coarse level");
        if (coarse->ctx != st->ctx)
            return fail(LR3D_ERR_INVALID, "lr3d_stereo_set_coarse: levels belong to different contexts");
        if (coarse->views.size() != st->views.size())
            return fail(LR3D_ERR_INVALID, "lr3d_stereo_set_coarse: layer counts differ (%d vs %d)",
                        (int32_t)coarse->views.size(), (int32_t)st->views.size());
    }
    std::lock_guard<std::mutex> lock(st->ctx->mu);
    detach_coarse(st);
    st->coarse = coarse;
    if (coarse) coarse->fine_levels.push_back(st);
    for (auto &g : st->graphs) destroy_graph(g);
    st->graphs.clear();
    return LR3D_OK;
}

// ---- pose / plane updates and the NCC adjoint -------------------------------------------------------

extern "C" int32_t lr3d_stereo_set_pose(lr3d_stereo *st, int32_t layer, const double R[9], const double t[3]) {
    if (!st || !R || !t) return fail(LR3D_ERR_INVALID, "lr3d_stereo_set_pose: NULL argument");
    int32_t rc = check_layer(st, layer, "lr3d_stereo_set_pose");
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(st->ctx->mu);
    for (int i = 0; i < 9; i++) st->views[layer].R[i] = R[i];
    for (int i = 0; i < 3; i++) st->views[layer].t[i] = t[i];
    return LR3D_OK;
}

extern "C" int32_t lr3d_stereo_write_hyp(lr3d_stereo *st, int32_t layer, const float *hyp) {
    if (!st || !hyp) return fail(LR3D_ERR_INVALID, "lr3d_stereo_write_hyp: NULL argument");
    int32_t rc = check_layer(st, layer, "lr3d_stereo_write_hyp");
    if (rc) return rc;
    lr3d_ctx *ctx = st->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const lr3d_view &v = st->views[layer];
    HIP_TRY(hipMemcpyAsync(st->hyp.ptr + st->off[layer], hyp, (size_t)v.width * v.height * 16,
                           hipMemcpyHostToDevice, ctx->stream));
    return finish(ctx);
}

// Shared part of the adjoint entry points: plans over the given solves (8 x 8 tiles, one
// wavefront per block), one launch, per-view pose gradients and losses summed from the block
// partials in block order.
static int32_t adjoint_impl(lr3d_stereo *st, const lr3d_solve_desc *descs, int32_t n, const float *weights,
                            bool want_maps, float step_depth, float step_normal, int32_t apply,
                            double *grad_pose, double *loss_before, double *loss_after) {
    lr3d_ctx *ctx = st->ctx;
    std::vector<PmDesc> dv(n);
    std::vector<AdjGeom> geo(n);
    std::vector<uint8_t> packs(n, 1);
    for (int32_t i = 0; i < n; i++) {
        int32_t rc = build_desc(st, descs[i], 0, true, dv[i]);
        if (rc) return rc;
        if (dv[i].n_src <= 0) return fail(LR3D_ERR_INVALID, "lr3d_ncc: solve %d has no sources", i);
        for (int32_t j = 0; j < i; j++)
            if (descs[j].ref == descs[i].ref) return fail(LR3D_ERR_INVALID, "lr3d_ncc: layer %d appears twice", descs[i].ref);
        const lr3d_view &rv = st->views[descs[i].ref];
        AdjGeom &g = geo[i];
        memset(&g, 0, sizeof(g));
        g.kr[0] = rv.fx;
        g.kr[1] = rv.fy;
        g.kr[2] = rv.cx;
        g.kr[3] = rv.cy;
        for (int32_t s = 0; s < dv[i].n_src; s++) {
            const lr3d_view &sv = st->views[descs[i].sources[s]];
            relative_pose(rv, sv, g.R[s], g.t[s]);
            g.k[s][0] = sv.fx;
            g.k[s][1] = sv.fy;
            g.k[s][2] = sv.cx;
            g.k[s][3] = sv.cy;
        }
        packs[i] = descs[i].pack ? 1 : 0;
    }
    Plans plans;
    std::vector<int64_t> key;
    // the backward kernel runs one 8 x 8 wavefront per block
    const TileShape tiles[3] = {{8, 8}, st->tiles[1], st->tiles[2]};
    int32_t rc = upload_plans(st, dv, packs, plans, key, tiles);
    if (rc) return rc;
    const PmPlan &plan = plans.init;
    HIP_TRY(st->adj_geom.reserve((size_t)n));
    HIP_TRY(hipMemcpyAsync(st->adj_geom.ptr, geo.data(), sizeof(AdjGeom) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(st->adj_partials.reserve((size_t)plan.blocks * kAdjPartial));
    HIP_TRY(hipMemsetAsync(st->adj_partials.ptr, 0, (size_t)plan.blocks * kAdjPartial * 8, ctx->stream));
    const float *wdev = nullptr;
    if (weights) {
        HIP_TRY(st->adj_weights.reserve(st->total_px));
        const lr3d_view &v = st->views[descs[0].ref];
        HIP_TRY(hipMemcpyAsync(st->adj_weights.ptr + st->off[descs[0].ref], weights, (size_t)v.width * v.height * 4,
                               hipMemcpyHostToDevice, ctx->stream));
        wdev = st->adj_weights.ptr;
    }
    if (want_maps) {
        HIP_TRY(st->adj_cost.reserve(st->total_px));
        HIP_TRY(st->adj_gd.reserve(st->total_px));
        HIP_TRY(st->adj_gn.reserve(st->total_px * 3));
    }
    const PmPools pools = pools_of(st);
    HIP_TRY(timed(ctx, "ncc_backward", [&] {
        return ncc_launch_backward(ctx->stream, plan, pools, st->adj_geom.ptr, wdev,
                                   want_maps ? st->adj_cost.ptr : nullptr, want_maps ? st->adj_gd.ptr : nullptr,
                                   want_maps ? st->adj_gn.ptr : nullptr, st->adj_partials.ptr, step_depth,
                                   step_normal, apply);
    }));
    if (apply && plans.pack.blocks > 0) {
        // refined planes -> packed depth maps of the descs that pack (DS:233-236)
        HIP_TRY(timed(ctx, "pm_pack", [&] { return pm_launch_pack(ctx->stream, plans.pack, pools); }));
    }
    std::vector<double> partials((size_t)plan.blocks * kAdjPartial);
    HIP_TRY(hipMemcpyAsync(partials.data(), st->adj_partials.ptr, partials.size() * 8, hipMemcpyDeviceToHost,
                           ctx->stream));
    rc = finish(ctx);
    if (rc) return rc;
    std::vector<uint32_t> starts(n + 1, 0);
    for (int32_t v = 0; v < n; v++) {
        const PmDesc &d = dv[v];
        starts[v + 1] = starts[v] + (uint32_t)(((d.H + plan.ty - 1) / plan.ty) * ((d.W + plan.tx - 1) / plan.tx));
    }
    for (int32_t v = 0; v < n; v++) {
        double *gp = grad_pose ? grad_pose + (size_t)v * kMaxSources * 12 : nullptr;
        if (gp) std::fill(gp, gp + kMaxSources * 12, 0.0);
        double l0 = 0.0, l1 = 0.0;
        for (uint32_t b = starts[v]; b < starts[v + 1]; b++) {
            const double *p = &partials[(size_t)b * kAdjPartial];
            l0 += p[0];
            l1 += p[1];
            if (gp)
                for (int j = 0; j < kMaxSources * 12; j++) gp[j] += p[2 + j];
        }
        if (loss_before) loss_before[v] = l0;
        if (loss_after) loss_after[v] = l1;
    }
    return LR3D_OK;
}

extern "C" int32_t lr3d_ncc_backward(lr3d_stereo *st, const lr3d_solve_desc *desc, const float *weights,
                                     float *cost, float *grad_depth, float *grad_normal, double *grad_pose,
                                     double *loss) {
    if (!st || !desc) return fail(LR3D_ERR_INVALID, "lr3d_ncc_backward: NULL argument");
    int32_t rc = check_layer(st, desc->ref, "lr3d_ncc_backward");
    if (rc) return rc;
    lr3d_ctx *ctx = st->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    rc = adjoint_impl(st, desc, 1, weights, true, 0.0f, 0.0f, 0, grad_pose, loss, nullptr);
    if (rc) return rc;
    const lr3d_view &v = st->views[desc->ref];
    const size_t px = (size_t)v.width * v.height, off = st->off[desc->ref];
    if (cost) HIP_TRY(hipMemcpyAsync(cost, st->adj_cost.ptr + off, px * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (grad_depth) HIP_TRY(hipMemcpyAsync(grad_depth, st->adj_gd.ptr + off, px * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (grad_normal)
        HIP_TRY(hipMemcpyAsync(grad_normal, st->adj_gn.ptr + off * 3, px * 12, hipMemcpyDeviceToHost, ctx->stream));
    return finish(ctx);
}

extern "C" int32_t lr3d_ncc_refine_step(lr3d_stereo *st, const lr3d_solve_desc *descs, int32_t n,
                                        float step_depth, float step_normal, double *grad_pose,
                                        double *loss_before, double *loss_after) {
    if (!st || !descs || n <= 0) return fail(LR3D_ERR_INVALID, "lr3d_ncc_refine_step: NULL argument or n <= 0");
    lr3d_ctx *ctx = st->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    if (!(step_depth >= 0.0f) || !(step_normal >= 0.0f))
        return fail(LR3D_ERR_INVALID, "lr3d_ncc_refine_step: steps must be >= 0");
    // zero steps evaluate only: no plane is rewritten (a renormalised copy of a plane can score a
    // last-bit lower cost) and the packed maps are left alone
    const int32_t apply = (step_depth > 0.0f || step_normal > 0.0f) ? 1 : 0;
    return adjoint_impl(st, descs, n, nullptr, false, step_depth, step_normal, apply, grad_pose, loss_before,
                        loss_after);
}

// ---- TSDF ---------------------------------------------------------------------------------------------

struct lr3d_tsdf {
    lr3d_ctx *ctx = nullptr;
    TsdfGridDev grid{};
    int32_t dims[3] = {0, 0, 0};
    DevBuf<int4> coords;
    DevBuf<float> tsdf, weight;
    DevBuf<uint32_t> vis;
    DevBuf<uint32_t> brick_list, brick_ctrs;  // bricks with a kept view, and the pass's counters
    DevBuf<float> tiles;               // depth tiles of the views of the last pass
    bool depth_culling = true;         // drop (brick, view) pairs that cannot update a voxel
    DevBuf<TsdfView> views;
    DevBuf<float> priors;
    DevBuf<uint32_t> host_map;
    DevBuf<unsigned long long> counter;
    bool counting = false;
    int32_t bricks_per_block = 4;  // integrate_kernel: wavefronts (bricks) per block; set by the autotuner
    int32_t xcd_chunk = 1;         // integrate_kernel: consecutive blocks per XCD (1: round robin); autotuned
    std::vector<int32_t> coords_host;  // brick grid coordinates (4 per brick), for the mesher's lookup
    DevBuf<int32_t> lookup;            // grid index -> pool brick or -1 (built on first use)
    // adjoint scratch
    DevBuf<uint32_t> adj_voxels, adj_pixel_off;
    DevBuf<float> adj_targets, adj_tsdf, adj_gdf;
    DevBuf<float2> adj_ckpt;
    DevBuf<double> adj_gd, adj_gp, adj_loss;
};

__global__ void fill_kernel(float *p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

extern "C" int32_t lr3d_tsdf_create(lr3d_ctx *ctx, const double origin[3], double voxel,
                                    const int32_t dims[3], const int32_t *coords, int32_t n_bricks,
                                    double truncation_voxels, lr3d_tsdf **out) {
    if (!ctx || !origin || !dims || !coords || !out) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_create: NULL argument");
    *out = nullptr;
    if (!(voxel > 0.0)) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_create: voxel must be > 0");
    if (!(truncation_voxels > 0.0)) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_create: truncation must be > 0");
    if (n_bricks <= 0) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_create: n_bricks must be > 0");
    if ((uint64_t)n_bricks * 512u >= (1ull << 32))
        return fail(LR3D_ERR_INVALID, "lr3d_tsdf_create: more than 2^32 voxels");
    for (int a = 0; a < 3; a++)
        if (dims[a] <= 0) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_create: dims must be > 0");
    for (int32_t b = 0; b < n_bricks; b++)
        for (int a = 0; a < 3; a++)
            if (coords[b * 4 + a] < 0 || coords[b * 4 + a] >= dims[a])
                return fail(LR3D_ERR_INVALID, "lr3d_tsdf_create: brick %d lies outside the grid", b);
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (ctx->closing) return fail(LR3D_ERR_STATE, "lr3d_tsdf_create: the context has been destroyed");
    HIP_TRY(hipSetDevice(ctx->device));
    lr3d_tsdf *ts = new lr3d_tsdf();
    ts->ctx = ctx;
    ts->coords_host.assign(coords, coords + (size_t)n_bricks * 4);
    for (int a = 0; a < 3; a++) {
        ts->dims[a] = dims[a];
        ts->grid.origin[a] = (float)origin[a];
        ts->grid.origin_d[a] = origin[a];
    }
    ts->grid.voxel = (float)voxel;
    ts->grid.voxel_d = voxel;
    ts->grid.mu_d = voxel * truncation_voxels;  // TV:179
    ts->grid.mu = (float)ts->grid.mu_d;
    ts->grid.inv_mu = 1.0f / ts->grid.mu;
    ts->grid.n_bricks = n_bricks;
    const size_t voxels = (size_t)n_bricks * 512u;
    hipError_t e = ts->coords.alloc(n_bricks);
    if (e == hipSuccess) e = ts->tsdf.alloc(voxels);
    if (e == hipSuccess) e = ts->weight.alloc(voxels);
    if (e == hipSuccess) e = ts->counter.alloc(4);
    if (e == hipSuccess) e = ts->brick_list.alloc(n_bricks);
    if (e == hipSuccess) e = ts->brick_ctrs.alloc(kTsdfCounters);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ts->coords.ptr, coords, (size_t)n_bricks * 16, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((voxels + 255) / 256)), dim3(256), 0, ctx->stream,
                           ts->tsdf.ptr, voxels, 1.0f);  // TV:188
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemsetAsync(ts->weight.ptr, 0, voxels * 4, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(ts->counter.ptr, 0, 32, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        delete ts;
        return fail(e == hipErrorOutOfMemory ? LR3D_ERR_NOMEM : LR3D_ERR_DEVICE, "lr3d_tsdf_create: %s",
                    hipGetErrorString(e));
    }
    ctx->users++;
    *out = ts;
    return LR3D_OK;
}

extern "C" int32_t lr3d_tsdf_destroy(lr3d_tsdf *ts) {
    if (!ts) return LR3D_OK;
    lr3d_ctx *ctx = ts->ctx;
    std::unique_lock<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete ts;
    ctx_release(ctx, lock);
    return LR3D_OK;
}

// TV:244-253 uniform + double copies for the culling.
static int32_t make_tsdf_view(const lr3d_integration_view &v, uint32_t map_off, uint32_t prior_off,
                              TsdfView &o) {
    if (v.width <= 0 || v.height <= 0 || !(v.fx > 0) || !(v.fy > 0))
        return fail(LR3D_ERR_INVALID, "lr3d_tsdf_integrate: invalid view size or intrinsics");
    memset(&o, 0, sizeof(o));
    o.k[0] = (float)v.fx;
    o.k[1] = (float)v.fy;
    o.k[2] = (float)v.cx;
    o.k[3] = (float)v.cy;
    for (int i = 0; i < 9; i++) o.R[i] = (float)v.R[i];
    for (int i = 0; i < 3; i++) o.t[i] = (float)v.t[i];
    o.W = v.width;
    o.H = v.height;
    o.confidence = (float)v.confidence;
    o.map_off = map_off;
    o.has_prior = v.prior != nullptr;
    o.prior_off = prior_off;
    o.kd[0] = v.fx;
    o.kd[1] = v.fy;
    o.kd[2] = v.cx;
    o.kd[3] = v.cy;
    for (int i = 0; i < 9; i++) o.Rd[i] = v.R[i];
    for (int i = 0; i < 3; i++) o.td[i] = v.t[i];
    o.max_depth = v.max_depth;
    return LR3D_OK;
}

// Priors, view table, depth tiles and culling bits of a pass (stream-ordered).
static int32_t stage_pass(lr3d_tsdf *ts, std::vector<TsdfView> &tv, const lr3d_integration_view *views,
                          const uint32_t *maps) {
    lr3d_ctx *ctx = ts->ctx;
    const int32_t n = (int32_t)tv.size();
    const int32_t words = (n + 31) / 32;
    size_t prior_total = 0;
    for (int32_t i = 0; i < n; i++)
        if (views[i].prior) prior_total += (size_t)views[i].width * views[i].height;
    if (prior_total) {
        HIP_TRY(ts->priors.reserve(prior_total));
        for (int32_t i = 0; i < n; i++)
            if (views[i].prior)
                HIP_TRY(hipMemcpyAsync(ts->priors.ptr + tv[i].prior_off, views[i].prior,
                                       (size_t)views[i].width * views[i].height * 4, hipMemcpyHostToDevice,
                                       ctx->stream));
    }
    size_t tiles = 0;
    int32_t max_tw = 0, max_th = 0;
    for (int32_t i = 0; i < n; i++) {
        TsdfView &v = tv[i];
        v.tw = (v.W + kTsdfTile - 1) / kTsdfTile;
        v.th = (v.H + kTsdfTile - 1) / kTsdfTile;
        v.tw2 = (v.tw + kTsdfTile2 - 1) / kTsdfTile2;
        v.th2 = (v.th + kTsdfTile2 - 1) / kTsdfTile2;
        v.tile_off = (uint32_t)tiles;
        tiles += (size_t)v.tw * v.th;
        v.tile2_off = (uint32_t)tiles;
        tiles += (size_t)v.tw2 * v.th2;
        max_tw = std::max(max_tw, v.tw);
        max_th = std::max(max_th, v.th);
    }
    if (tiles > 0xffffffffull) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_integrate: too many depth tiles");
    const bool cull = ts->depth_culling && n <= 65535;
    HIP_TRY(ts->views.reserve(n));
    HIP_TRY(hipMemcpyAsync(ts->views.ptr, tv.data(), sizeof(TsdfView) * n, hipMemcpyHostToDevice, ctx->stream));
    if (cull) {
        HIP_TRY(ts->tiles.reserve(tiles));
        HIP_TRY(timed(ctx, "tsdf_tiles", [&] {
            return tsdf_launch_depth_tiles(ctx->stream, ts->views.ptr, n, max_tw, max_th, maps, ts->priors.ptr,
                                           ts->tiles.ptr);
        }));
    }
    HIP_TRY(ts->vis.reserve((size_t)ts->grid.n_bricks * words));
    HIP_TRY(timed(ctx, "tsdf_visibility", [&] {
        return tsdf_launch_visibility(ctx->stream, ts->grid, ts->coords.ptr, ts->views.ptr, n, ts->vis.ptr, words,
                                      cull ? ts->tiles.ptr : nullptr, ts->brick_ctrs.ptr,
                                      n <= kListPassViews ? ts->brick_list.ptr : nullptr);
    }));
    return LR3D_OK;
}

// The prior pool when a view of the pass has a prior, else nullptr (the integrate kernel without
// prior lookups).
static const float *pass_priors(const lr3d_tsdf *ts, const std::vector<TsdfView> &tv) {
    for (const TsdfView &v : tv)
        if (v.has_prior) return ts->priors.ptr;
    return nullptr;
}

// Whether a pass lies in the ranges where integrate_kernel's correctly rounded short forms equal
// the IEEE divisions and square roots: every view's camera coordinates of the grid's voxel centres
// below 2^40 in magnitude (bounded over the box of all the grid's voxel centres, with a margin far
// above the FP32 rounding of cam = R world + t), |fx|, |fy| below 2^40, and mu in [2^-40, 2^40].
static bool integrate_fast_ok(const lr3d_tsdf *ts, const std::vector<TsdfView> &tv) {
    const double mu = ts->grid.mu;
    if (!(mu >= 0x1p-40 && mu <= 0x1p40)) return false;
    double m[3];
    for (int a = 0; a < 3; a++) {
        const double lo = (double)ts->grid.origin[a] + 0.5 * (double)ts->grid.voxel;
        const double hi = (double)ts->grid.origin[a] + ((double)ts->dims[a] * 8.0 - 0.5) * (double)ts->grid.voxel;
        m[a] = std::max(std::fabs(lo), std::fabs(hi));
    }
    for (const TsdfView &v : tv) {
        if (!(std::fabs((double)v.k[0]) < 0x1p40 && std::fabs((double)v.k[1]) < 0x1p40)) return false;
        for (int i = 0; i < 3; i++) {
            const double b = std::fabs((double)v.R[3 * i]) * m[0] + std::fabs((double)v.R[3 * i + 1]) * m[1] +
                             std::fabs((double)v.R[3 * i + 2]) * m[2] + std::fabs((double)v.t[i]);
            if (!(b * 1.001 < 0x1p40)) return false;
        }
    }
    return true;
}

static int32_t integrate_impl(lr3d_tsdf *ts, const uint32_t *maps, std::vector<TsdfView> &tv,
                              const lr3d_integration_view *views) {
    lr3d_ctx *ctx = ts->ctx;
    const int32_t n = (int32_t)tv.size();
    if (n == 0) return finish(ctx);
    const int32_t words = (n + 31) / 32;
    int32_t rc = stage_pass(ts, tv, views, maps);
    if (rc) return rc;
    HIP_TRY(timed(ctx, "tsdf_integrate", [&] {
        return tsdf_launch_integrate(ctx->stream, ts->grid, ts->views.ptr, n, ts->vis.ptr, words,
                                     ts->coords.ptr, maps, pass_priors(ts, tv), ctx->ntab.ptr,
                                     integrate_fast_ok(ts, tv), ts->tsdf.ptr, ts->weight.ptr,
                                     ts->counting ? ts->counter.ptr : nullptr, ts->brick_ctrs.ptr,
                                     n <= kListPassViews ? ts->brick_list.ptr : nullptr, ts->bricks_per_block,
                                     ts->xcd_chunk);
    }));
    return finish(ctx);
}

static int32_t stereo_views(lr3d_tsdf *ts, lr3d_stereo *st, const int32_t *layers, int32_t map,
                            const lr3d_integration_view *views, int32_t n, std::vector<TsdfView> &tv,
                            const uint32_t **maps) {
    if (!st) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_integrate: stereo is NULL");
    if (st->ctx != ts->ctx) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_integrate: stereo and volume use different contexts");
    if (map != LR3D_MAP_RAW && map != LR3D_MAP_FILTERED)
        return fail(LR3D_ERR_INVALID, "lr3d_tsdf_integrate: unknown map selector %d", map);
    *maps = map == LR3D_MAP_RAW ? st->depth.ptr : st->filt.ptr;
    tv.resize(n);
    uint32_t prior_off = 0;
    for (int32_t i = 0; i < n; i++) {
        int32_t rc = check_layer(st, layers[i], "lr3d_tsdf_integrate");
        if (rc) return rc;
        const lr3d_view &sv = st->views[layers[i]];
        if (views[i].width != sv.width || views[i].height != sv.height)
            return fail(LR3D_ERR_INVALID, "lr3d_tsdf_integrate: view %d is %dx%d but layer %d is %dx%d", i,
                        views[i].width, views[i].height, layers[i], sv.width, sv.height);
        rc = make_tsdf_view(views[i], st->off[layers[i]], prior_off, tv[i]);
        if (rc) return rc;
        if (views[i].prior) prior_off += (uint32_t)(views[i].width * views[i].height);
    }
    return LR3D_OK;
}

extern "C" int32_t lr3d_tsdf_integrate(lr3d_tsdf *ts, lr3d_stereo *st, int32_t layer, int32_t map,
                                       const lr3d_integration_view *view) {
    return lr3d_tsdf_integrate_batched(ts, st, &layer, map, view, 1);
}

extern "C" int32_t lr3d_tsdf_integrate_batched(lr3d_tsdf *ts, lr3d_stereo *st, const int32_t *layers,
                                               int32_t map, const lr3d_integration_view *views, int32_t n) {
    if (!ts || !layers || !views) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_integrate: NULL argument");
    if (n < 0) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_integrate: n < 0");
    std::lock_guard<std::mutex> lock(ts->ctx->mu);
    HIP_TRY(hipSetDevice(ts->ctx->device));
    std::vector<TsdfView> tv;
    const uint32_t *maps = nullptr;
    int32_t rc = stereo_views(ts, st, layers, map, views, n, tv, &maps);
    if (rc) return rc;
    return integrate_impl(ts, maps, tv, views);
}

extern "C" int32_t lr3d_tsdf_integrate_host(lr3d_tsdf *ts, const uint32_t *packed,
                                            const lr3d_integration_view *view) {
    if (!ts || !packed || !view) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_integrate_host: NULL argument");
    lr3d_ctx *ctx = ts->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<TsdfView> tv(1);
    int32_t rc = make_tsdf_view(*view, 0, 0, tv[0]);
    if (rc) return rc;
    const size_t px = (size_t)view->width * view->height;
    HIP_TRY(ts->host_map.reserve(px));
    HIP_TRY(hipMemcpyAsync(ts->host_map.ptr, packed, px * 4, hipMemcpyHostToDevice, ctx->stream));
    return integrate_impl(ts, ts->host_map.ptr, tv, view);
}

extern "C" int32_t lr3d_tsdf_backward(lr3d_tsdf *ts, lr3d_stereo *st, const int32_t *layers, int32_t map,
                                      const lr3d_integration_view *views, int32_t n, const uint32_t *voxels,
                                      const float *targets, int64_t n_voxels, float *tsdf_out, float *grad_depth,
                                      double *grad_pose, double *loss) {
    if (!ts || !layers || !views || (n_voxels > 0 && (!voxels || !targets)))
        return fail(LR3D_ERR_INVALID, "lr3d_tsdf_backward: NULL argument");
    if (n <= 0) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_backward: n must be > 0");
    if (n_voxels < 0 || n_voxels > (int64_t)INT32_MAX)
        return fail(LR3D_ERR_INVALID, "lr3d_tsdf_backward: n_voxels must be in [0, 2^31)");
    const uint64_t pool = (uint64_t)ts->grid.n_bricks * 512u;
    for (int64_t i = 0; i < n_voxels; i++)
        if (voxels[i] >= pool) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_backward: voxel %lld is outside the pool", (long long)i);
    lr3d_ctx *ctx = ts->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<TsdfView> tv;
    const uint32_t *maps = nullptr;
    int32_t rc = stereo_views(ts, st, layers, map, views, n, tv, &maps);
    if (rc) return rc;
    const int32_t words = (n + 31) / 32;
    std::vector<uint32_t> pixel_off(n + 1, 0);
    for (int32_t i = 0; i < n; i++) pixel_off[i + 1] = pixel_off[i] + (uint32_t)(views[i].width * views[i].height);
    const size_t total_px = pixel_off[n];
    rc = stage_pass(ts, tv, views, maps);
    if (rc) return rc;
    HIP_TRY(ts->adj_pixel_off.reserve(n + 1));
    HIP_TRY(hipMemcpyAsync(ts->adj_pixel_off.ptr, pixel_off.data(), 4 * (n + 1), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ts->adj_gd.reserve(std::max<size_t>(total_px, 1)));
    HIP_TRY(ts->adj_gp.reserve((size_t)n * 12));
    HIP_TRY(hipMemsetAsync(ts->adj_gd.ptr, 0, total_px * 8, ctx->stream));
    HIP_TRY(hipMemsetAsync(ts->adj_gp.ptr, 0, (size_t)n * 12 * 8, ctx->stream));
    const uint32_t nv = (uint32_t)n_voxels;
    const uint32_t waves = (nv + 63u) / 64u;
    std::vector<double> partials(waves);
    if (nv > 0) {
        const int32_t n_seg = (n + 15) / 16;
        HIP_TRY(ts->adj_voxels.reserve(nv));
        HIP_TRY(ts->adj_targets.reserve(nv));
        HIP_TRY(ts->adj_tsdf.reserve(nv));
        HIP_TRY(ts->adj_ckpt.reserve((size_t)n_seg * nv));
        HIP_TRY(ts->adj_loss.reserve(((nv + 255u) / 256u) * 4u));
        HIP_TRY(hipMemcpyAsync(ts->adj_voxels.ptr, voxels, (size_t)nv * 4, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(ts->adj_targets.ptr, targets, (size_t)nv * 4, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(timed(ctx, "tsdf_backward", [&] {
            return tsdf_launch_backward(ctx->stream, ts->grid, ts->views.ptr, n, ts->vis.ptr, words, ts->coords.ptr,
                                        maps, ts->priors.ptr, ts->adj_voxels.ptr, ts->adj_targets.ptr, nv,
                                        ts->adj_pixel_off.ptr, ts->adj_ckpt.ptr, ts->adj_tsdf.ptr, ts->adj_gd.ptr,
                                        ts->adj_gp.ptr, ts->adj_loss.ptr);
        }));
        HIP_TRY(hipMemcpyAsync(partials.data(), ts->adj_loss.ptr, (size_t)waves * 8, hipMemcpyDeviceToHost, ctx->stream));
        if (tsdf_out)
            HIP_TRY(hipMemcpyAsync(tsdf_out, ts->adj_tsdf.ptr, (size_t)nv * 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    if (grad_depth && total_px) {
        HIP_TRY(ts->adj_gdf.reserve(total_px));
        HIP_TRY(launch_to_float(ctx->stream, ts->adj_gd.ptr, ts->adj_gdf.ptr, total_px));
        HIP_TRY(hipMemcpyAsync(grad_depth, ts->adj_gdf.ptr, total_px * 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    if (grad_pose) HIP_TRY(hipMemcpyAsync(grad_pose, ts->adj_gp.ptr, (size_t)n * 12 * 8, hipMemcpyDeviceToHost, ctx->stream));
    rc = finish(ctx);
    if (rc) return rc;
    if (loss) {
        double l = 0.0;
        for (uint32_t i = 0; i < waves; i++) l += partials[i];
        *loss = l;
    }
    return LR3D_OK;
}

extern "C" int32_t lr3d_tsdf_set_depth_culling(lr3d_tsdf *ts, int32_t on) {
    if (!ts) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_set_depth_culling: NULL argument");
    std::lock_guard<std::mutex> lock(ts->ctx->mu);
    ts->depth_culling = on != 0;
    return LR3D_OK;
}

extern "C" int32_t lr3d_tsdf_get_bricks_per_block(lr3d_tsdf *ts, int32_t *out) {
    if (!ts || !out) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_get_bricks_per_block: NULL argument");
    std::lock_guard<std::mutex> lock(ts->ctx->mu);
    *out = ts->bricks_per_block;
    return LR3D_OK;
}

extern "C" int32_t lr3d_tsdf_set_bricks_per_block(lr3d_tsdf *ts, int32_t bricks) {
    if (!ts) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_set_bricks_per_block: NULL argument");
    if (bricks != 1 && bricks != 2 && bricks != 4)
        return fail(LR3D_ERR_INVALID, "lr3d_tsdf_set_bricks_per_block: %d is not 1, 2 or 4", bricks);
    std::lock_guard<std::mutex> lock(ts->ctx->mu);
    ts->bricks_per_block = bricks;
    return LR3D_OK;
}

extern "C" int32_t lr3d_tsdf_get_xcd_chunk(lr3d_tsdf *ts, int32_t *out) {
    if (!ts || !out) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_get_xcd_chunk: NULL argument");
    std::lock_guard<std::mutex> lock(ts->ctx->mu);
    *out = ts->xcd_chunk;
    return LR3D_OK;
}

extern "C" int32_t lr3d_tsdf_set_xcd_chunk(lr3d_tsdf *ts, int32_t chunk) {
    if (!ts) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_set_xcd_chunk: NULL argument");
    if (chunk < 1 || chunk > kMaxXcdChunk)
        return fail(LR3D_ERR_INVALID, "lr3d_tsdf_set_xcd_chunk: %d is outside [1, %d]", chunk, kMaxXcdChunk);
    std::lock_guard<std::mutex> lock(ts->ctx->mu);
    ts->xcd_chunk = chunk;
    return LR3D_OK;
}

extern "C" int32_t lr3d_tsdf_autotune(lr3d_tsdf *ts, lr3d_stereo *st, const int32_t *layers, int32_t map,
                                      const lr3d_integration_view *views, int32_t n, int32_t repeats) {
    if (!ts || !layers || !views || n <= 0) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_autotune: NULL argument or n <= 0");
    lr3d_ctx *ctx = ts->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    repeats = std::max(1, std::min(repeats, 16));
    std::vector<TsdfView> tv;
    const uint32_t *maps = nullptr;
    int32_t rc = stereo_views(ts, st, layers, map, views, n, tv, &maps);
    if (rc) return rc;
    rc = stage_pass(ts, tv, views, maps);
    if (rc) return rc;
    const int32_t words = (n + 31) / 32;
    const size_t voxels = (size_t)ts->grid.n_bricks * 512u;
    // every candidate integrates into the same starting volume; the volume is restored at the end
    DevBuf<float> s_tsdf, s_weight;
    HIP_TRY(s_tsdf.alloc(voxels));
    HIP_TRY(s_weight.alloc(voxels));
    HIP_TRY(hipMemcpyAsync(s_tsdf.ptr, ts->tsdf.ptr, voxels * 4, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(s_weight.ptr, ts->weight.ptr, voxels * 4, hipMemcpyDeviceToDevice, ctx->stream));
    EventPair ev;
    HIP_TRY(hipEventCreate(&ev.a));
    HIP_TRY(hipEventCreate(&ev.b));
    // candidates: bricks (wavefronts) per block x chunks of consecutive blocks per XCD (round robin,
    // a few sizes, and the length of a row of the grid's bricks)
    std::vector<int32_t> chunks = {1, 16, 32, 64, 128};
    if (ts->dims[0] > 1 && ts->dims[0] <= kMaxXcdChunk &&
        std::find(chunks.begin(), chunks.end(), ts->dims[0]) == chunks.end())
        chunks.push_back(ts->dims[0]);
    const bool fast = integrate_fast_ok(ts, tv);
    const float *priors = pass_priors(ts, tv);
    float best_ms = 0.0f;
    int32_t best_bricks = ts->bricks_per_block, best_chunk = ts->xcd_chunk;
    bool first = true;
    for (const int32_t bricks : {1, 2, 4}) {
        for (const int32_t chunk : chunks) {
            float total = 0.0f;
            for (int32_t k = 0; k < repeats; k++) {
                HIP_TRY(hipMemcpyAsync(ts->tsdf.ptr, s_tsdf.ptr, voxels * 4, hipMemcpyDeviceToDevice, ctx->stream));
                HIP_TRY(hipMemcpyAsync(ts->weight.ptr, s_weight.ptr, voxels * 4, hipMemcpyDeviceToDevice, ctx->stream));
                HIP_TRY(hipEventRecord(ev.a, ctx->stream));
                HIP_TRY(tsdf_launch_integrate(ctx->stream, ts->grid, ts->views.ptr, n, ts->vis.ptr, words,
                                              ts->coords.ptr, maps, priors, ctx->ntab.ptr, fast, ts->tsdf.ptr,
                                              ts->weight.ptr, nullptr, ts->brick_ctrs.ptr,
                                              n <= kListPassViews ? ts->brick_list.ptr : nullptr, bricks, chunk));
                HIP_TRY(hipEventRecord(ev.b, ctx->stream));
                HIP_TRY(hipEventSynchronize(ev.b));
                float ms = 0.0f;
                HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
                total += ms;
            }
            if (first || total < best_ms) {
                best_ms = total;
                best_bricks = bricks;
                best_chunk = chunk;
                first = false;
            }
        }
    }
    ts->bricks_per_block = best_bricks;
    ts->xcd_chunk = best_chunk;
    HIP_TRY(hipMemcpyAsync(ts->tsdf.ptr, s_tsdf.ptr, voxels * 4, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ts->weight.ptr, s_weight.ptr, voxels * 4, hipMemcpyDeviceToDevice, ctx->stream));
    return finish(ctx);
}

extern "C" int32_t lr3d_tsdf_write(lr3d_tsdf *ts, const float *tsdf, const float *weight) {
    if (!ts || !tsdf || !weight) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_write: NULL argument");
    lr3d_ctx *ctx = ts->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t bytes = (size_t)ts->grid.n_bricks * 512u * 4u;
    HIP_TRY(hipMemcpyAsync(ts->tsdf.ptr, tsdf, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ts->weight.ptr, weight, bytes, hipMemcpyHostToDevice, ctx->stream));
    return finish(ctx);
}

// ---- surface nets and mesh post-processing (SN = surface-nets.ts, MS = mesh.ts) ----------------------

struct lr3d_mesh {
    lr3d_ctx *ctx = nullptr;
    DevBuf<float> positions, normals;
    DevBuf<uint32_t> indices;
    uint32_t n_vertices = 0, n_indices = 0;
};

// Device-wide exclusive prefix sum of n uint32 values (rocPRIM); returns the total in *total.
struct Scanner {
    DevBuf<uint8_t> tmp;
    uint32_t last[2] = {0, 0};
    hipError_t run(hipStream_t s, const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *total) {
        *total = 0;
        if (n == 0) return hipSuccess;
        size_t bytes = 0;
        hipError_t e = rocprim::exclusive_scan(nullptr, bytes, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s);
        if (e != hipSuccess) return e;
        if ((e = tmp.reserve(std::max<size_t>(bytes, 1))) != hipSuccess) return e;
        e = rocprim::exclusive_scan(tmp.ptr, bytes, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s);
        if (e != hipSuccess) return e;
        if ((e = hipMemcpyAsync(&last[0], in + n - 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(&last[1], out + n - 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
        *total = last[0] + last[1];
        return hipSuccess;
    }
};

// computeNormals (MS:9-32) of a mesh whose positions and indices are set.
static int32_t mesh_normals(lr3d_mesh *m, Scanner &scan) {
    hipStream_t s = m->ctx->stream;
    const uint32_t V = m->n_vertices, N = m->n_indices;
    HIP_TRY(m->normals.alloc((size_t)V * 3));
    if (V == 0) return LR3D_OK;
    DevBuf<uint32_t> deg, off, cursor, adj;
    HIP_TRY(deg.alloc(V));
    HIP_TRY(off.alloc(V));
    HIP_TRY(cursor.alloc(V));
    HIP_TRY(adj.alloc(std::max<uint32_t>(N, 1)));
    HIP_TRY(hipMemsetAsync(deg.ptr, 0, (size_t)V * 4, s));
    HIP_TRY(hipMemsetAsync(cursor.ptr, 0, (size_t)V * 4, s));
    HIP_TRY(mesh_launch_degree(s, m->indices.ptr, N, deg.ptr));
    uint32_t total = 0;
    HIP_TRY(scan.run(s, deg.ptr, off.ptr, V, &total));
    HIP_TRY(mesh_launch_adjacency(s, m->indices.ptr, N, off.ptr, cursor.ptr, adj.ptr));
    HIP_TRY(mesh_launch_normal_sum(s, m->positions.ptr, m->indices.ptr, off.ptr, deg.ptr, adj.ptr, V, m->normals.ptr));
    return finish(m->ctx);
}

extern "C" int32_t lr3d_tsdf_extract_surface(lr3d_tsdf *ts, double min_weight, lr3d_mesh **out) {
    if (!ts || !out) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_extract_surface: NULL argument");
    *out = nullptr;
    lr3d_ctx *ctx = ts->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int64_t grid_cells = (int64_t)ts->dims[0] * ts->dims[1] * ts->dims[2];
    if (!ts->lookup.ptr) {
        // SN:67-74's inverse: grid index -> pool brick (BrickGrid.build, TV:54-64)
        std::vector<int32_t> lookup((size_t)grid_cells, -1);
        for (int32_t b = 0; b < ts->grid.n_bricks; b++) {
            const int32_t *c = &ts->coords_host[(size_t)b * 4];
            lookup[((size_t)c[2] * ts->dims[1] + c[1]) * ts->dims[0] + c[0]] = b;
        }
        HIP_TRY(ts->lookup.alloc((size_t)grid_cells));
        HIP_TRY(hipMemcpyAsync(ts->lookup.ptr, lookup.data(), lookup.size() * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    MeshVolume v;
    for (int a = 0; a < 3; a++) {
        v.origin[a] = ts->grid.origin_d[a];
        v.dims[a] = ts->dims[a];
    }
    v.voxel = ts->grid.voxel_d;
    v.min_weight = min_weight;
    v.lookup = ts->lookup.ptr;
    v.coords = ts->coords.ptr;
    v.tsdf = ts->tsdf.ptr;
    v.weight = ts->weight.ptr;
    const uint32_t cells = (uint32_t)((uint64_t)ts->grid.n_bricks * 512u);
    std::unique_ptr<lr3d_mesh> m(new lr3d_mesh());
    m->ctx = ctx;
    Scanner scan;
    DevBuf<uint32_t> flag, vidx, qcount, qoff, live, lpos, bricks;
    DevBuf<int32_t> cell_vertex;
    DevBuf<double> pos_d;
    // the cell kernels run over the bricks holding an observed voxel; the others' cells are empty
    const uint32_t n_bricks = (uint32_t)ts->grid.n_bricks;
    HIP_TRY(live.alloc(std::max<uint32_t>(n_bricks, 1)));
    HIP_TRY(lpos.alloc(std::max<uint32_t>(n_bricks, 1)));
    HIP_TRY(timed(ctx, "mesh_bricks", [&] { return mesh_launch_live_bricks(s, v, n_bricks, live.ptr); }));
    uint32_t L = 0;
    HIP_TRY(scan.run(s, live.ptr, lpos.ptr, n_bricks, &L));
    HIP_TRY(bricks.alloc(std::max<uint32_t>(L, 1)));
    HIP_TRY(mesh_launch_brick_list(s, live.ptr, lpos.ptr, n_bricks, bricks.ptr));
    v.bricks = bricks.ptr;
    const uint32_t threads = L * 512u;
    HIP_TRY(flag.alloc(cells));
    HIP_TRY(vidx.alloc(cells));
    HIP_TRY(cell_vertex.alloc(cells));
    HIP_TRY(hipMemsetAsync(flag.ptr, 0, (size_t)cells * 4, s));
    HIP_TRY(hipMemsetAsync(cell_vertex.ptr, 0xff, (size_t)cells * 4, s));  // -1: no vertex
    // SN:87-115: active cells, numbered in pool order
    HIP_TRY(timed(ctx, "mesh_active", [&] { return mesh_launch_active(s, v, threads, flag.ptr); }));
    uint32_t V = 0;
    HIP_TRY(scan.run(s, flag.ptr, vidx.ptr, cells, &V));
    m->n_vertices = V;
    HIP_TRY(pos_d.alloc((size_t)std::max<uint32_t>(V, 1) * 3));
    HIP_TRY(m->positions.alloc((size_t)std::max<uint32_t>(V, 1) * 3));
    HIP_TRY(timed(ctx, "mesh_vertices", [&] {
        return mesh_launch_vertices(s, v, threads, vidx.ptr, cell_vertex.ptr, pos_d.ptr, m->positions.ptr);
    }));
    flag.release();
    vidx.release();
    // SN:117-156: quads counted per voxel, written at their sequential offsets
    HIP_TRY(qcount.alloc(cells));
    HIP_TRY(qoff.alloc(cells));
    HIP_TRY(hipMemsetAsync(qcount.ptr, 0, (size_t)cells * 4, s));
    HIP_TRY(timed(ctx, "mesh_quads", [&] {
        return mesh_launch_quads(s, v, threads, cell_vertex.ptr, pos_d.ptr, nullptr, qcount.ptr, nullptr);
    }));
    uint32_t Q = 0;
    HIP_TRY(scan.run(s, qcount.ptr, qoff.ptr, cells, &Q));
    if ((uint64_t)Q * 6 >= (1ull << 32)) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_extract_surface: more than 2^32 indices");
    m->n_indices = Q * 6;
    HIP_TRY(m->indices.alloc((size_t)std::max<uint32_t>(m->n_indices, 1)));
    HIP_TRY(timed(ctx, "mesh_quads", [&] {
        return mesh_launch_quads(s, v, threads, cell_vertex.ptr, pos_d.ptr, qoff.ptr, nullptr, m->indices.ptr);
    }));
    int32_t rc = finish(ctx);
    if (rc) return rc;
    rc = mesh_normals(m.get(), scan);
    if (rc) return rc;
    ctx->users++;  // a mesh made from a live volume keeps the context (held: ctx->mu)
    *out = m.release();
    return LR3D_OK;
}

extern "C" int32_t lr3d_mesh_remove_small_components(lr3d_mesh *in, double fraction, int32_t min_faces,
                                                     lr3d_mesh **out) {
    if (!in || !out) return fail(LR3D_ERR_INVALID, "lr3d_mesh_remove_small_components: NULL argument");
    *out = nullptr;
    lr3d_ctx *ctx = in->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const uint32_t V = in->n_vertices, F = in->n_indices / 3;
    std::unique_ptr<lr3d_mesh> m(new lr3d_mesh());
    m->ctx = ctx;
    Scanner scan;
    DevBuf<uint32_t> parent, changed, faces, keep, koff, first, sflag, rank;
    HIP_TRY(parent.alloc(std::max<uint32_t>(V, 1)));
    HIP_TRY(changed.alloc(1));
    uint32_t changed_host = 0;
    HIP_TRY(mesh_launch_components(s, in->indices.ptr, F, V, parent.ptr, changed.ptr, &changed_host));
    HIP_TRY(faces.alloc(std::max<uint32_t>(V, 1)));
    HIP_TRY(keep.alloc(std::max<uint32_t>(F, 1)));
    HIP_TRY(koff.alloc(std::max<uint32_t>(F, 1)));
    HIP_TRY(hipMemsetAsync(faces.ptr, 0, (size_t)std::max<uint32_t>(V, 1) * 4, s));
    const double limit = std::max((double)min_faces, (double)F * fraction);  // MS:65
    HIP_TRY(mesh_launch_keep(s, in->indices.ptr, F, parent.ptr, faces.ptr, limit, keep.ptr));
    uint32_t K = 0;
    HIP_TRY(scan.run(s, keep.ptr, koff.ptr, F, &K));
    const uint32_t stream_len = K * 3;
    HIP_TRY(first.alloc(std::max<uint32_t>(V, 1)));
    HIP_TRY(sflag.alloc(std::max<uint32_t>(stream_len, 1)));
    HIP_TRY(rank.alloc(std::max<uint32_t>(stream_len, 1)));
    HIP_TRY(hipMemsetAsync(first.ptr, 0xff, (size_t)std::max<uint32_t>(V, 1) * 4, s));
    HIP_TRY(mesh_launch_first_use(s, in->indices.ptr, F, keep.ptr, koff.ptr, first.ptr, sflag.ptr));
    uint32_t NV = 0;
    HIP_TRY(scan.run(s, sflag.ptr, rank.ptr, stream_len, &NV));
    m->n_vertices = NV;
    m->n_indices = stream_len;
    HIP_TRY(m->positions.alloc((size_t)std::max<uint32_t>(NV, 1) * 3));
    HIP_TRY(m->normals.alloc((size_t)std::max<uint32_t>(NV, 1) * 3));
    HIP_TRY(m->indices.alloc(std::max<uint32_t>(stream_len, 1)));
    HIP_TRY(mesh_launch_compact(s, in->indices.ptr, F, keep.ptr, koff.ptr, first.ptr, rank.ptr, in->positions.ptr,
                                in->normals.ptr, m->indices.ptr, m->positions.ptr, m->normals.ptr));
    int32_t rc = finish(ctx);
    if (rc) return rc;
    ctx->users++;  // held: ctx->mu
    *out = m.release();
    return LR3D_OK;
}

// ---- dense-cloud fusion (DF = depth-fusion.ts, PA = point-accumulator.ts) ---------------------------

struct lr3d_cloud {
    lr3d_ctx *ctx = nullptr;
    DevBuf<float> positions, normals;
    DevBuf<uint8_t> colors;
    DevBuf<uint16_t> support;
    uint32_t n = 0;
};

// Stable LSD radix sort of (63-bit key, value) pairs.
static hipError_t sort_pairs(hipStream_t s, DevBuf<uint8_t> &tmp, const uint64_t *kin, uint64_t *kout,
                             const uint32_t *vin, uint32_t *vout, uint32_t n) {
    if (n == 0) return hipSuccess;
    size_t bytes = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, (size_t)n, 0, 63, s);
    if (e != hipSuccess) return e;
    if ((e = tmp.reserve(std::max<size_t>(bytes, 1))) != hipSuccess) return e;
    return rocprim::radix_sort_pairs(tmp.ptr, bytes, kin, kout, vin, vout, (size_t)n, 0, 63, s);
}

__global__ void iota_u32_kernel(uint32_t *a, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// The accumulator's cells (sorted by key): keys, 9 float sums, counts.
struct CellSet {
    DevBuf<uint64_t> keys;
    DevBuf<float> sums;
    DevBuf<uint32_t> counts;
    uint32_t n = 0;
    hipError_t alloc(uint32_t count) {
        hipError_t e = keys.alloc(std::max<uint32_t>(count, 1));
        if (e == hipSuccess) e = sums.alloc((size_t)std::max<uint32_t>(count, 1) * 9);
        if (e == hipSuccess) e = counts.alloc(std::max<uint32_t>(count, 1));
        n = count;
        return e;
    }
    // Room for `count` cells, grow-only (contents are not kept; n is left to the caller).
    hipError_t reserve(uint32_t count) {
        hipError_t e = keys.reserve(std::max<uint32_t>(count, 1));
        if (e == hipSuccess) e = sums.reserve((size_t)std::max<uint32_t>(count, 1) * 9);
        if (e == hipSuccess) e = counts.reserve(std::max<uint32_t>(count, 1));
        return e;
    }
};

extern "C" int32_t lr3d_fuse_depth_maps(lr3d_stereo *st, const int32_t *layers, int32_t map, int32_t n,
                                        const lr3d_thumb *thumbs, double cell, int64_t cell_budget,
                                        int32_t min_support, lr3d_cloud **out, double *final_cell) {
    if (!st || !layers || !thumbs || !out) return fail(LR3D_ERR_INVALID, "lr3d_fuse_depth_maps: NULL argument");
    *out = nullptr;
    if (n <= 0) return fail(LR3D_ERR_INVALID, "lr3d_fuse_depth_maps: n must be > 0");
    if (!(cell > 0.0)) return fail(LR3D_ERR_INVALID, "lr3d_fuse_depth_maps: cell must be > 0");
    if (cell_budget < 1) return fail(LR3D_ERR_INVALID, "lr3d_fuse_depth_maps: cell_budget must be >= 1");
    if (map != LR3D_MAP_RAW && map != LR3D_MAP_FILTERED)
        return fail(LR3D_ERR_INVALID, "lr3d_fuse_depth_maps: unknown map selector %d", map);
    size_t thumb_bytes = 0;
    for (int32_t i = 0; i < n; i++) {
        int32_t rc = check_layer(st, layers[i], "lr3d_fuse_depth_maps");
        if (rc) return rc;
        if (thumbs[i].width <= 0 || thumbs[i].height <= 0 || !thumbs[i].rgb)
            return fail(LR3D_ERR_INVALID, "lr3d_fuse_depth_maps: thumbnail %d is empty", i);
        thumb_bytes += (size_t)thumbs[i].width * thumbs[i].height * 3;
    }
    lr3d_ctx *ctx = st->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const uint32_t *maps = map == LR3D_MAP_RAW ? st->depth.ptr : st->filt.ptr;
    DevBuf<uint8_t> dthumbs, sort_tmp;
    HIP_TRY(dthumbs.alloc(thumb_bytes));
    std::vector<FusionView> fv(n);
    uint64_t toff = 0;
    for (int32_t i = 0; i < n; i++) {
        const lr3d_view &v = st->views[layers[i]];
        FusionView &f = fv[i];
        f.W = v.width;
        f.H = v.height;
        f.k[0] = v.fx;
        f.k[1] = v.fy;
        f.k[2] = v.cx;
        f.k[3] = v.cy;
        for (int j = 0; j < 9; j++) f.R[j] = v.R[j];
        for (int j = 0; j < 3; j++) f.t[j] = v.t[j];
        f.map_off = st->off[layers[i]];
        f.tw = thumbs[i].width;
        f.th = thumbs[i].height;
        f.thumb_off = toff;
        const size_t bytes = (size_t)f.tw * f.th * 3;
        HIP_TRY(hipMemcpyAsync(dthumbs.ptr + toff, thumbs[i].rgb, bytes, hipMemcpyHostToDevice, s));
        toff += bytes;
    }
    DevBuf<uint32_t> overflow;
    HIP_TRY(overflow.alloc(1));
    HIP_TRY(hipMemsetAsync(overflow.ptr, 0, 4, s));
    Scanner scan;
    CellSet state;
    HIP_TRY(state.alloc(0));
    // The samples go in chunks: a chunk's valid samples are keyed, sorted (stably, so each cell keeps
    // the arrival order: views in order, pixels row-major) and added in one pass.  The reference
    // changes its cells at two kinds of points, and a chunk ends at the first of them:
    //  - after a view, while the accumulator holds more than the budget's cells, it coarsens
    //    (DF:36): the new cells per view of first appearance give the size after each view;
    //  - before a sample, when its hash table is over 60 % full (PA:67), it doubles the table and
    //    re-bins every cell by its mean position at the same cell size (PA:87-110), which moves a
    //    cell whose float mean rounds outside it: the first sample of the cell that fills the
    //    table ends the chunk, and the next chunk starts with the re-binning.
    // Chunks grow from 4 views while nothing cuts them; LR3D_FUSION_CHUNK_VIEWS=k fixes them at k
    // views.
    DevBuf<FusionView> dviews;
    DevBuf<uint32_t> vstart, hist, flag, pos, idx, sidx, head, seg, start, is_new, new_pos, gflag, gscan, gsel;
    DevBuf<uint64_t> keys, skeys, tk;
    DevBuf<uint32_t> perm0, perm, big, n_big;
    DevBuf<int32_t> exist;
    CellSet cat, next;
    HIP_TRY(dviews.alloc(n));
    HIP_TRY(n_big.alloc(1));
    HIP_TRY(hipMemcpyAsync(dviews.ptr, fv.data(), sizeof(FusionView) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(vstart.alloc((size_t)n + 1));
    HIP_TRY(hist.alloc(n));
    HIP_TRY(gsel.alloc(1));
    int32_t fixed = 0;
    if (const char *e = getenv("LR3D_FUSION_CHUNK_VIEWS")) fixed = std::min(std::max(0, atoi(e)), kFusionMaxChunkViews);
    int32_t cap = fixed > 0 ? fixed : 4;
    int64_t capacity = 1 << 16;  // PA:17 the hash table's slots
    // PA:87-110 rebuild: every cell re-binned by its mean position at `to`, cells that fall
    // together merged (in key order; the reference merges in slot order)
    // (its buffers, like the chunks' below, grow with headroom and are kept: every allocation and
    // free of device memory costs more than the work of a re-binning)
    auto grown = [](size_t n) { return n + n / 2 + 1024; };
    DevBuf<uint64_t> rk, rsk;
    DevBuf<uint32_t> ri, rsi, rh, rs, rst;
    CellSet rnext;
    auto rebin = [&](double to) -> int32_t {
        const uint32_t S = state.n;
        if (S > rk.count) {
            HIP_TRY(rk.alloc(grown(S)));
            HIP_TRY(rsk.alloc(grown(S)));
            HIP_TRY(ri.alloc(grown(S)));
            HIP_TRY(rsi.alloc(grown(S)));
            HIP_TRY(rh.alloc(grown(S)));
            HIP_TRY(rs.alloc(grown(S)));
            HIP_TRY(rst.alloc(grown(S)));
        }
        DevBuf<uint64_t> &keys = rk, &skeys = rsk;
        DevBuf<uint32_t> &idx = ri, &sidx = rsi, &head = rh, &seg = rs, &start = rst;
        HIP_TRY(fusion_launch_rekey(s, state.sums.ptr, state.counts.ptr, S, to, keys.ptr, idx.ptr, overflow.ptr));
        HIP_TRY(sort_pairs(s, sort_tmp, keys.ptr, skeys.ptr, idx.ptr, sidx.ptr, S));
        HIP_TRY(fusion_launch_heads(s, skeys.ptr, S, head.ptr));
        uint32_t G = 0;
        HIP_TRY(scan.run(s, head.ptr, seg.ptr, S, &G));
        HIP_TRY(fusion_launch_starts(s, head.ptr, seg.ptr, S, start.ptr));
        CellSet &next = rnext;
        HIP_TRY(next.reserve(G));
        HIP_TRY(fusion_launch_merge(s, skeys.ptr, sidx.ptr, start.ptr, G, S, state.sums.ptr, state.counts.ptr,
                                    next.keys.ptr, next.sums.ptr, next.counts.ptr));
        std::swap(state.keys, next.keys);
        std::swap(state.sums, next.sums);
        std::swap(state.counts, next.counts);
        state.n = G;
        return LR3D_OK;
    };
    // the per-sample buffers sized once for the largest chunk this call can form
    uint64_t total_px = 0, max_px = 0;
    for (int32_t i = 0; i < n; i++) {
        total_px += (uint64_t)fv[i].W * fv[i].H;
        max_px = std::max<uint64_t>(max_px, (uint64_t)fv[i].W * fv[i].H);
    }
    const size_t most = (size_t)std::min<uint64_t>(total_px, kFusionChunkPixels + max_px);
    HIP_TRY(flag.reserve(most));
    HIP_TRY(pos.reserve(most));
    HIP_TRY(keys.reserve(most));
    HIP_TRY(skeys.reserve(most));
    HIP_TRY(idx.reserve(most));
    HIP_TRY(sidx.reserve(most));
    HIP_TRY(head.reserve(most));
    HIP_TRY(seg.reserve(most));
    HIP_TRY(gflag.reserve(max_px));
    HIP_TRY(gscan.reserve(max_px));
    std::vector<uint32_t> hstart, hhist;
    uint32_t from = 0;  // first pixel of view a still to add
    for (int32_t a = 0; a < n;) {
        // chunk [a, b): at most cap views and kFusionChunkPixels pixels, at least one view
        int32_t b = a;
        uint64_t P = 0;
        hstart.assign(1, 0u);
        while (b < n && b - a < cap) {
            const uint64_t px = (uint64_t)fv[b].W * fv[b].H;
            if (b > a && P + px > kFusionChunkPixels) break;
            P += px;
            hstart.push_back((uint32_t)P);
            b++;
        }
        const uint32_t nv = (uint32_t)(b - a);
        int32_t last = b - 1;     // the view holding the chunk's last sample
        uint32_t limit = (uint32_t)P;  // samples numbered below it are added
        bool view_end = true;     // the chunk ends at the end of view `last`
        HIP_TRY(flag.reserve(P));
        HIP_TRY(pos.reserve(P));
        HIP_TRY(hipMemcpyAsync(vstart.ptr, hstart.data(), (size_t)(nv + 1) * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(fusion_launch_chunk_valid(s, dviews.ptr + a, vstart.ptr, nv, (uint32_t)P, from, maps, flag.ptr));
        uint32_t M = 0;
        HIP_TRY(scan.run(s, flag.ptr, pos.ptr, (uint32_t)P, &M));
        if (M > 0) {
            // PA:67: before its next sample the table doubles if it is over 60 % full
            if ((int64_t)state.n * 10 > capacity * 6) {
                capacity *= 2;
                int32_t rc = rebin(cell);
                if (rc) return rc;
            }
            const uint64_t fill = (uint64_t)(capacity * 6 / 10) + 1;  // the cells that make it over 60 % full
            // DF:52-71 + PA:37-48: samples keyed by cell, grouped by a stable sort
            HIP_TRY(keys.reserve(M));
            HIP_TRY(skeys.reserve(M));
            HIP_TRY(idx.reserve(M));
            HIP_TRY(sidx.reserve(M));
            HIP_TRY(head.reserve(M));
            HIP_TRY(seg.reserve(M));
            HIP_TRY(fusion_launch_chunk_keys(s, dviews.ptr + a, vstart.ptr, nv, (uint32_t)P, maps, dthumbs.ptr, flag.ptr,
                                             pos.ptr, cell, keys.ptr, idx.ptr, overflow.ptr));
            HIP_TRY(sort_pairs(s, sort_tmp, keys.ptr, skeys.ptr, idx.ptr, sidx.ptr, M));
            HIP_TRY(fusion_launch_heads(s, skeys.ptr, M, head.ptr));
            uint32_t G = 0;
            HIP_TRY(scan.run(s, head.ptr, seg.ptr, M, &G));
            if (G > start.count) {  // per-cell buffers: grown with headroom
                HIP_TRY(start.alloc(grown(G)));
                HIP_TRY(exist.alloc(grown(G)));
                HIP_TRY(is_new.alloc(grown(G)));
                HIP_TRY(new_pos.alloc(grown(G)));
                HIP_TRY(big.alloc(grown(G)));
            }
            HIP_TRY(fusion_launch_starts(s, head.ptr, seg.ptr, M, start.ptr));
            HIP_TRY(fusion_launch_lookup(s, skeys.ptr, start.ptr, G, state.keys.ptr, state.n, exist.ptr, is_new.ptr));
            // new cells per view of first appearance
            HIP_TRY(hipMemsetAsync(hist.ptr, 0, (size_t)nv * 4, s));
            HIP_TRY(fusion_launch_first_views(s, sidx.ptr, start.ptr, is_new.ptr, G, vstart.ptr, nv, hist.ptr));
            hhist.resize(nv);
            HIP_TRY(hipMemcpyAsync(hhist.data(), hist.ptr, (size_t)nv * 4, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            // the first view after which the budget is exceeded (DF:36), and the view of the new
            // cell that fills the table (its rank among that view's new cells)
            int32_t over = -1, fills = -1;
            uint32_t rank = 0;
            uint64_t size = state.n;
            for (uint32_t c = 0; c < nv && over < 0; c++) {
                if (fills < 0 && size + hhist[c] >= fill) {
                    fills = a + (int32_t)c;
                    rank = (uint32_t)(fill - size - 1);
                }
                size += hhist[c];
                if ((int64_t)size > cell_budget) over = a + (int32_t)c;
            }
            if (fills >= 0 && (over < 0 || fills <= over)) {
                // the sample that creates the filling cell is the chunk's last
                const uint32_t lo = hstart[fills - a], hi = hstart[fills - a + 1];
                HIP_TRY(gflag.reserve(hi - lo));
                HIP_TRY(gscan.reserve(hi - lo));
                HIP_TRY(hipMemsetAsync(gflag.ptr, 0, (size_t)(hi - lo) * 4, s));
                HIP_TRY(fusion_launch_new_cell_flags(s, sidx.ptr, start.ptr, is_new.ptr, G, lo, hi, gflag.ptr));
                uint32_t flagged = 0;
                HIP_TRY(scan.run(s, gflag.ptr, gscan.ptr, hi - lo, &flagged));
                HIP_TRY(hipMemsetAsync(gsel.ptr, 0xff, 4, s));
                HIP_TRY(fusion_launch_select(s, gflag.ptr, gscan.ptr, hi - lo, rank, gsel.ptr));
                uint32_t at = 0, later[2] = {0, 0};
                HIP_TRY(hipMemcpyAsync(&at, gsel.ptr, 4, hipMemcpyDeviceToHost, s));
                HIP_TRY(hipStreamSynchronize(s));
                if (rank >= flagged || at >= hi - lo)
                    return fail(LR3D_ERR_DEVICE, "lr3d_fuse_depth_maps: new cell %u of view %d not found (%u flagged)",
                                rank, fills, flagged);
                limit = lo + at + 1;
                last = fills;
                // valid samples of that view after it: none means the cut is at the view's end
                HIP_TRY(hipMemcpyAsync(&later[0], pos.ptr + limit - 1, 4, hipMemcpyDeviceToHost, s));
                if (hi < P) HIP_TRY(hipMemcpyAsync(&later[1], pos.ptr + hi, 4, hipMemcpyDeviceToHost, s));
                HIP_TRY(hipStreamSynchronize(s));
                const uint32_t upto = hi < P ? later[1] : M;  // valid samples before the view's end
                view_end = upto == later[0] + 1;              // the cut sample is the view's last
            } else if (over >= 0) {
                last = over;
                limit = hstart[over - a + 1];
            }
            if (limit < P) HIP_TRY(fusion_launch_limit(s, sidx.ptr, start.ptr, G, limit, is_new.ptr));
            uint32_t Nn = 0;
            HIP_TRY(scan.run(s, is_new.ptr, new_pos.ptr, G, &Nn));
            // merged set: the old cells, then the new ones, re-sorted by key below
            HIP_TRY(cat.reserve(state.n + Nn > cat.keys.count ? grown(state.n + Nn) : state.n + Nn));
            if (state.n) {
                HIP_TRY(hipMemcpyAsync(cat.keys.ptr, state.keys.ptr, (size_t)state.n * 8, hipMemcpyDeviceToDevice, s));
                HIP_TRY(hipMemcpyAsync(cat.sums.ptr, state.sums.ptr, (size_t)state.n * 36, hipMemcpyDeviceToDevice, s));
                HIP_TRY(hipMemcpyAsync(cat.counts.ptr, state.counts.ptr, (size_t)state.n * 4, hipMemcpyDeviceToDevice, s));
            }
            // PA:49-62: running sums of existing cells updated in place, new cells appended
            HIP_TRY(hipMemsetAsync(n_big.ptr, 0, 4, s));
            HIP_TRY(fusion_launch_reduce(s, dviews.ptr + a, vstart.ptr, nv, limit, maps, dthumbs.ptr, skeys.ptr, sidx.ptr,
                                         start.ptr, G, M, exist.ptr, is_new.ptr, new_pos.ptr, cat.sums.ptr,
                                         cat.counts.ptr, cat.keys.ptr + state.n, cat.sums.ptr + (size_t)state.n * 9,
                                         cat.counts.ptr + state.n, big.ptr, n_big.ptr));
            const uint32_t T = state.n + Nn;
            HIP_TRY(tk.reserve(T > tk.count ? grown(T) : T));
            HIP_TRY(perm0.reserve(T > perm0.count ? grown(T) : T));
            HIP_TRY(perm.reserve(T > perm.count ? grown(T) : T));
            (void)hipGetLastError();
            hipLaunchKernelGGL(iota_u32_kernel, dim3((T + 255) / 256), dim3(256), 0, s, perm0.ptr, T);
            HIP_TRY(hipGetLastError());
            HIP_TRY(sort_pairs(s, sort_tmp, cat.keys.ptr, tk.ptr, perm0.ptr, perm.ptr, T));
            HIP_TRY(next.reserve(T > next.keys.count ? grown(T) : T));
            HIP_TRY(fusion_launch_gather(s, perm.ptr, T, cat.keys.ptr, cat.sums.ptr, cat.counts.ptr, next.keys.ptr,
                                         next.sums.ptr, next.counts.ptr));
            std::swap(state.keys, next.keys);
            std::swap(state.sums, next.sums);
            std::swap(state.counts, next.counts);
            state.n = T;
        }
        if (view_end) {
            // DF:36-37 / PA:98-131: coarsen by 1.4 while the cloud is over budget
            while ((int64_t)state.n > cell_budget) {
                const uint32_t before = state.n;
                cell = cell * 1.4;
                int32_t rc = rebin(cell);
                if (rc) return rc;
                if (state.n == before && cell > 1e300) break;  // cannot merge further
            }
            from = 0;
            if (fixed == 0) cap = last == b - 1 ? std::min(cap * 2, kFusionMaxChunkViews) : std::max(1, last - a + 1);
            a = last + 1;
        } else {
            // the next chunk resumes inside view `last`, after the cut sample
            from = limit - hstart[last - a];
            if (fixed == 0) cap = std::max(1, last - a + 1);
            a = last;
        }
    }
    uint32_t ovf = 0;
    HIP_TRY(hipMemcpyAsync(&ovf, overflow.ptr, 4, hipMemcpyDeviceToHost, s));
    int32_t rc = finish(ctx);
    if (rc) return rc;
    if (ovf) return fail(LR3D_ERR_INVALID, "lr3d_fuse_depth_maps: a point lies more than 2^20 cells from the origin");
    // PA:108-143 cloud(minSupport), in key order
    std::unique_ptr<lr3d_cloud> c(new lr3d_cloud());
    c->ctx = ctx;
    HIP_TRY(flag.reserve(std::max<uint32_t>(state.n, 1)));
    HIP_TRY(pos.reserve(std::max<uint32_t>(state.n, 1)));
    HIP_TRY(fusion_launch_support(s, state.counts.ptr, state.n, (uint32_t)std::max(min_support, 0), flag.ptr));
    uint32_t total = 0;
    HIP_TRY(scan.run(s, flag.ptr, pos.ptr, state.n, &total));
    c->n = total;
    HIP_TRY(c->positions.alloc((size_t)std::max<uint32_t>(total, 1) * 3));
    HIP_TRY(c->normals.alloc((size_t)std::max<uint32_t>(total, 1) * 3));
    HIP_TRY(c->colors.alloc((size_t)std::max<uint32_t>(total, 1) * 3));
    HIP_TRY(c->support.alloc(std::max<uint32_t>(total, 1)));
    HIP_TRY(fusion_launch_cloud(s, state.sums.ptr, state.counts.ptr, state.n, flag.ptr, pos.ptr, c->positions.ptr,
                                c->normals.ptr, c->colors.ptr, c->support.ptr));
    rc = finish(ctx);
    if (rc) return rc;
    if (final_cell) *final_cell = cell;
    ctx->users++;  // held: ctx->mu
    *out = c.release();
    return LR3D_OK;
}

extern "C" int32_t lr3d_cloud_size(lr3d_cloud *c, int64_t *points) {
    if (!c || !points) return fail(LR3D_ERR_INVALID, "lr3d_cloud_size: NULL argument");
    *points = c->n;
    return LR3D_OK;
}

extern "C" int32_t lr3d_cloud_read(lr3d_cloud *c, float *positions, float *normals, uint8_t *colors, uint16_t *support) {
    if (!c) return fail(LR3D_ERR_INVALID, "lr3d_cloud_read: NULL argument");
    lr3d_ctx *ctx = c->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t n = c->n;
    if (n && positions) HIP_TRY(hipMemcpyAsync(positions, c->positions.ptr, n * 12, hipMemcpyDeviceToHost, ctx->stream));
    if (n && normals) HIP_TRY(hipMemcpyAsync(normals, c->normals.ptr, n * 12, hipMemcpyDeviceToHost, ctx->stream));
    if (n && colors) HIP_TRY(hipMemcpyAsync(colors, c->colors.ptr, n * 3, hipMemcpyDeviceToHost, ctx->stream));
    if (n && support) HIP_TRY(hipMemcpyAsync(support, c->support.ptr, n * 2, hipMemcpyDeviceToHost, ctx->stream));
    return finish(ctx);
}

extern "C" int32_t lr3d_cloud_destroy(lr3d_cloud *c) {
    if (!c) return LR3D_OK;
    lr3d_ctx *ctx = c->ctx;
    std::unique_lock<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete c;
    ctx_release(ctx, lock);
    return LR3D_OK;
}

extern "C" int32_t lr3d_mesh_size(lr3d_mesh *m, int64_t *vertices, int64_t *indices) {
    if (!m || !vertices || !indices) return fail(LR3D_ERR_INVALID, "lr3d_mesh_size: NULL argument");
    *vertices = m->n_vertices;
    *indices = m->n_indices;
    return LR3D_OK;
}

extern "C" int32_t lr3d_mesh_read(lr3d_mesh *m, float *positions, float *normals, uint32_t *indices) {
    if (!m) return fail(LR3D_ERR_INVALID, "lr3d_mesh_read: NULL argument");
    lr3d_ctx *ctx = m->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    if (positions && m->n_vertices)
        HIP_TRY(hipMemcpyAsync(positions, m->positions.ptr, (size_t)m->n_vertices * 12, hipMemcpyDeviceToHost, ctx->stream));
    if (normals && m->n_vertices)
        HIP_TRY(hipMemcpyAsync(normals, m->normals.ptr, (size_t)m->n_vertices * 12, hipMemcpyDeviceToHost, ctx->stream));
    if (indices && m->n_indices)
        HIP_TRY(hipMemcpyAsync(indices, m->indices.ptr, (size_t)m->n_indices * 4, hipMemcpyDeviceToHost, ctx->stream));
    return finish(ctx);
}

extern "C" int32_t lr3d_mesh_destroy(lr3d_mesh *m) {
    if (!m) return LR3D_OK;
    lr3d_ctx *ctx = m->ctx;
    std::unique_lock<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete m;
    ctx_release(ctx, lock);
    return LR3D_OK;
}

extern "C" int32_t lr3d_tsdf_read(lr3d_tsdf *ts, float *tsdf, float *weight) {
    if (!ts) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_read: NULL argument");
    lr3d_ctx *ctx = ts->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t bytes = (size_t)ts->grid.n_bricks * 512u * 4u;
    if (tsdf) HIP_TRY(hipMemcpyAsync(tsdf, ts->tsdf.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    if (weight) HIP_TRY(hipMemcpyAsync(weight, ts->weight.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    return finish(ctx);
}

extern "C" int32_t lr3d_tsdf_set_counting(lr3d_tsdf *ts, int32_t enable) {
    if (!ts) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_set_counting: NULL argument");
    lr3d_ctx *ctx = ts->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    ts->counting = enable != 0;
    HIP_TRY(hipMemsetAsync(ts->counter.ptr, 0, 32, ctx->stream));
    return finish(ctx);
}

extern "C" int32_t lr3d_tsdf_update_count(lr3d_tsdf *ts, int64_t *out) {
    if (!ts || !out) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_update_count: NULL argument");
    lr3d_ctx *ctx = ts->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    unsigned long long v = 0;
    HIP_TRY(hipMemcpyAsync(&v, ts->counter.ptr, 8, hipMemcpyDeviceToHost, ctx->stream));
    int32_t rc = finish(ctx);
    if (rc) return rc;
    *out = (int64_t)v;
    return LR3D_OK;
}

extern "C" int32_t lr3d_tsdf_pass_counts(lr3d_tsdf *ts, int64_t out[4]) {
    if (!ts || !out) return fail(LR3D_ERR_INVALID, "lr3d_tsdf_pass_counts: NULL argument");
    lr3d_ctx *ctx = ts->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    unsigned long long v[4] = {0, 0, 0, 0};
    HIP_TRY(hipMemcpyAsync(v, ts->counter.ptr, 32, hipMemcpyDeviceToHost, ctx->stream));
    int32_t rc = finish(ctx);
    if (rc) return rc;
    for (int i = 0; i < 4; i++) out[i] = (int64_t)v[i];
    return LR3D_OK;
}
