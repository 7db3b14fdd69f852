// C-ABI of libcrayhip.so (include/crayhip.h): contexts, device buffers,
// BVH build, scene upload, ray queries, film and render dispatch.
//
// Replaces the reference's host layer: DeviceBuffer (Core/Buffer.cu:7-63),
// Scene::Scene (Scene/Scene.cu:61-78), TriangleMeshManager
// (Scene/Triangle.cu:98-165), MaterialManager (Core/Material.cu:15-30),
// Sensor (Core/Sensor.cu:45-112) and PathTraceIntegrator::dispatch
// (Core/Integrator.cu:233-262).  Differences by design: size_t buffers, the
// integrator configuration actually reaches the device, the film is zeroed
// on creation, every launch is checked, and spp passes are queued back to
// back on one stream instead of synchronising after each.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/crayhip.h"
#include "bvh_build.h"
#include "bvh_device.h"
#include "kernels.h"
#include "noalloc.h"
#include "spectral.h"
#include "wavefront.h"
#include "guided_mk.h"
#include "guiding_kernels.h"

static_assert(sizeof(crh_ray) == 32, "crh_ray layout");
static_assert(sizeof(crh_hit) == 16, "crh_hit layout");
static_assert(sizeof(crh_bvh_node_ref) == 36, "crh_bvh_node_ref layout");
static_assert(sizeof(crh_material) == 68, "crh_material layout");
static_assert(sizeof(crh_camera) == 96, "crh_camera layout");

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

int hipFail(const char* expr, hipError_t e, const char* file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof(buf),
                  "HIP call %s failed with error %d (%s).\n\n%s:%d", expr,
                  int(e), hipGetErrorString(e), file, line);
    g_err = buf;
    return e == hipErrorOutOfMemory ? CRH_ERR_OOM : CRH_ERR_HIP;
}

#define HIP_TRY(expr)                                                   \
    do {                                                                \
        hipError_t _e = (expr);                                         \
        if (_e != hipSuccess) return hipFail(#expr, _e, __FILE__, __LINE__); \
    } while (0)

#define ARG_CHECK(cond, msg)                                            \
    do {                                                                \
        if (!(cond)) return fail(CRH_ERR_INVALID_ARG, msg);             \
    } while (0)

#define GUARD_BEGIN try {
#define GUARD_END                                                       \
    }                                                                   \
    catch (const std::bad_alloc&) {                                     \
        return fail(CRH_ERR_OOM, "host allocation failed");             \
    }                                                                   \
    catch (const std::exception& ex) {                                  \
        return fail(CRH_ERR_INTERNAL, ex.what());                       \
    }

// Owned device allocation.
struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n) {
        reset();
        bytes = n;
        if (n == 0) return hipSuccess;
        return hipMalloc(&p, n);
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

template <class T>
hipError_t upload(DevMem& m, const std::vector<T>& v, hipStream_t st) {
    hipError_t e = m.alloc(v.size() * sizeof(T));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(m.p, v.data(), v.size() * sizeof(T),
                          hipMemcpyHostToDevice, st);
}

double nowMs() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

struct crh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    crh::Tuning tuning;
    uint32_t* counter = nullptr;  // work counter of the persistent kernels
    // One reference for the creator (dropped by crh_context_destroy) and one
    // per live buffer, scene, film and guide made on it: the stream and the
    // events outlive every object that uses them, whatever order the caller
    // destroys things in (a garbage-collected host may finalise a context
    // before the objects that point at it).
    std::atomic<int> refs{1};
};

namespace {

void contextRelease(crh_context* c) {
    if (c->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->counter) (void)hipFree(c->counter);
    delete c;
}

// A counted reference to the owning context.  Declared as the first member of
// each object so that it is released after the object's device memory.
struct CtxRef {
    crh_context* p = nullptr;
    CtxRef() = default;
    CtxRef(const CtxRef&) = delete;
    CtxRef& operator=(const CtxRef&) = delete;
    ~CtxRef() {
        if (p) contextRelease(p);
    }
    void set(crh_context* c) {
        c->refs.fetch_add(1, std::memory_order_relaxed);
        p = c;
    }
    crh_context* operator->() const { return p; }
    operator crh_context*() const { return p; }
};

// The private stack holds kPrivStack levels; deeper trees use the LDS stack.
crh::Tuning effective(const crh::Tuning& t, const crh::SceneView& v) {
    crh::Tuning e = t;
    if (v.stackCap > uint32_t(crh::kPrivStackLevels))
        e.privStack = e.renderPriv = e.wfPriv = e.batchPriv = e.guidedPriv = 0;
    return e;
}
// Core/CameraProxy.h:25-40 view of a camera for a w x h film: aspect in
// float, tan(radians(fovY * 0.5)) in double (fovY * 0.5 promotes), narrowed.
crh::CameraView makeCameraView(const crh_camera* cam, uint32_t w, uint32_t h) {
    crh::CameraView cv;
    for (int k = 0; k < 3; ++k) {
        cv.pos[k] = cam->pos[k];
        cv.u[k] = cam->u[k];
        cv.v[k] = cam->v[k];
        cv.w[k] = cam->w[k];
    }
    cv.hh = float(std::tan(double(cam->fov_y) * 0.5 * 0.01745329251994329576923690768489));
    cv.aspect = float(w) / float(h);
    cv.width = w;
    cv.height = h;
    return cv;
}
crh::WfLaunch wfLaunch(const crh::Tuning& t) {
    crh::WfLaunch w;
    w.traceGrid = t.wfTraceGrid;
    w.traceBlock = t.wfTraceBlock;
    w.shadeGrid = t.wfShadeGrid;
    w.privStack = t.wfPriv;
    w.persistent = t.wfPersistent;
    w.refill = t.wfRefill;
    w.finishDepth = t.wfFinish;
    w.finishWaves = t.renderWaves;
    w.finishGrid = t.renderGrid;
    return w;
}
}  // namespace

struct crh_buffer {
    CtxRef ctx;
    DevMem mem;
};

struct crh_bvh {
    crh::BvhResult r;
};

struct crh_scene {
    CtxRef ctx;
    crh::BvhResult bvh;
    DevMem nodes, tris, pos, nrm, uv, idx, primMat, mats, emissive, spec;
    DevMem spheres, sphereMat, texels, texInfo;
    crh::SceneView view{};
};

struct crh_film {
    CtxRef ctx;
    uint32_t w = 0, h = 0;
    DevMem sum, wsum, L, P;
    // Work counters of crh_render (rays, node visits, triangle tests),
    // allocated with the film so that a render allocates nothing itself.
    DevMem work;
    // Pass timing events of crh_render, created once and grown on demand.
    std::vector<hipEvent_t> passEv;
    // Wavefront path state, allocated on first wavefront render.
    DevMem wfArena;
    crh::WfState wf{};
    bool wfReady = false;
    // Guided-mode state, allocated on first guided render.
    DevMem gArena, gSamples;
    bool guidedReady = false;
    // Training samples waiting in gSamples (accumulated over passes until a
    // batch is full) and the buffer's capacity in samples.
    size_t gPending = 0, gCap = 0;
    // Fused guided rendering: per-wave network outputs, training records and
    // their counters (records written, samples produced).
    DevMem gScratch, gRec, gCnt;
    uint32_t gRecCap = 0;
    // Guided rendering of several views (crh_render_guided_views, kept by
    // the first view's film): the pass's sample buffers of all views and the
    // device view table.
    DevMem vL, vP, vTab;
    ~crh_film() {
        for (hipEvent_t e : passEv) (void)hipEventDestroy(e);
    }
};

namespace {

// Carves the wavefront structure-of-arrays state for `slots` samples per
// pixel out of one allocation (reallocated when more slots are needed).
// Path state of n paths of the wavefront in one arena (SoA arrays, queues,
// counters); sizes and guided buffers are the caller's.
hipError_t layoutWavefront(DevMem& arena, crh::WfState& w, size_t n) {
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t v4 = align(n * 16), u1 = align(n * 4), u2 = align(n * 8);
    const size_t total = 17 * v4 + u2 + 9 * u1 + align(64) + align(64);
    hipError_t e = arena.alloc(total);
    if (e != hipSuccess) return e;
    char* base = arena.as<char>();
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base + off;
        off += bytes;
        return p;
    };
    w.rayO = (float4*)take(v4);
    w.rayD = (float4*)take(v4);
    w.hit = (uint4*)take(v4);
    w.beta = (float4*)take(v4);
    w.rad = (float4*)take(v4);
    w.shO = (float4*)take(v4);
    w.shD = (float4*)take(v4);
    w.lgt = (float4*)take(v4);
    w.lgn = (float4*)take(v4);
    w.misO = (float4*)take(v4);
    w.misD = (float4*)take(v4);
    w.misHit = (uint4*)take(v4);
    w.misVal = (float4*)take(v4);
    w.misBw = (float4*)take(v4);
    w.misPos = (float4*)take(v4);
    w.misBeta = (float4*)take(v4);
    take(v4);  // spare
    w.P = (float2*)take(u2);
    w.rng = (uint32_t*)take(u1);
    w.flags = (uint32_t*)take(u1);
    w.lgId = (uint32_t*)take(u1);
    w.qExt[0] = (uint32_t*)take(u1);
    w.qExt[1] = (uint32_t*)take(u1);
    w.qMis[0] = (uint32_t*)take(u1);
    w.qMis[1] = (uint32_t*)take(u1);
    w.qShadow = (uint32_t*)take(u1);
    w.qCont = (uint32_t*)take(u1);
    w.counters = (uint32_t*)take(align(64));
    w.work = (unsigned long long*)take(align(64));
    w.gPos = w.gNrm = w.gWo = w.rec = nullptr;
    w.gMix = nullptr;
    w.recId = nullptr;
    w.recCap = 0;
    return hipSuccess;
}

hipError_t ensureWavefront(crh_film* f, uint32_t slots) {
    if (f->wfReady && f->wf.slots >= slots) return hipSuccess;
    f->wfReady = f->guidedReady = false;
    f->gArena.reset();
    crh::WfState& w = f->wf;
    hipError_t e = layoutWavefront(f->wfArena, w, size_t(f->w) * f->h * slots);
    if (e != hipSuccess) return e;
    w.width = f->w;
    w.height = f->h;
    w.npix = f->w * f->h;
    w.slots = slots;
    w.active = slots;
    f->wfReady = true;
    return hipSuccess;
}

// Guided-mode buffers: per-path shading frame and mixture, vertex records
// (4 per path) and the training samples made from them.
hipError_t ensureGuided(crh_film* f, uint32_t slots) {
    hipError_t e = ensureWavefront(f, slots);
    if (e != hipSuccess || f->guidedReady) return e;
    const size_t n = size_t(f->w) * f->h * f->wf.slots;
    const size_t cap = n * 4;
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t total = 3 * align(n * 16) + align(n * 40 * 4) + align(n * 4) +
                         align(cap * crh::kRecF4 * 16);
    e = f->gArena.alloc(total);
    if (e != hipSuccess) return e;
    char* base = f->gArena.as<char>();
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base + off;
        off += align(bytes);
        return p;
    };
    crh::WfState& w = f->wf;
    w.gPos = (float4*)take(n * 16);
    w.gNrm = (float4*)take(n * 16);
    w.gWo = (float4*)take(n * 16);
    w.gMix = (float*)take(n * 40 * 4);
    w.recId = (uint32_t*)take(n * 4);
    w.rec = (float4*)take(cap * crh::kRecF4 * 16);
    w.recCap = uint32_t(std::min<size_t>(cap, 0xFFFFFFF0u));
    f->guidedReady = true;
    return hipSuccess;
}

// Buffers of the fused guided megakernel: `scratchFloats` network outputs
// and `recCap` training records (none when not training).
hipError_t ensureGuidedFused(crh_film* f, size_t scratchFloats, uint32_t recCap) {
    hipError_t e = hipSuccess;
    if (f->gScratch.bytes < scratchFloats * sizeof(float))
        e = f->gScratch.alloc(scratchFloats * sizeof(float));
    if (e == hipSuccess && f->gCnt.bytes == 0) e = f->gCnt.alloc(256);
    if (e == hipSuccess && f->gRecCap < recCap) {
        e = f->gRec.alloc(size_t(recCap) * crh::kRecF4 * sizeof(float4));
        f->gRecCap = e == hipSuccess ? recCap : 0u;
    }
    return e;
}

// Samples per pixel rendered together: enough paths to fill `pool`, no
// more than the call renders.
uint32_t poolSlots(uint32_t pool, const crh_film* f, uint32_t spp) {
    const uint64_t npix = uint64_t(f->w) * f->h;
    uint64_t k = std::max<uint64_t>(1, pool / npix);
    k = std::min<uint64_t>(k, std::max<uint32_t>(spp, 1u));
    k = std::min<uint64_t>(k, std::max<uint64_t>(1, (uint64_t(1) << 31) / npix / 4));
    return uint32_t(k);
}
uint32_t wfSlots(const crh::Tuning& t, const crh_film* f, uint32_t spp) {
    return poolSlots(t.wfPool, f, spp);
}

// Per-pass sample buffers of the film for `slots` sample indices per launch.
hipError_t ensureFilmSlots(crh_film* f, uint32_t slots) {
    const size_t n = size_t(f->w) * f->h * slots;
    if (f->L.bytes >= n * sizeof(float4) && f->P.bytes >= n * sizeof(float2)) return hipSuccess;
    hipError_t e = f->L.alloc(n * sizeof(float4));
    if (e == hipSuccess) e = f->P.alloc(n * sizeof(float2));
    return e;
}

}  // namespace

extern "C" {

const char* crh_last_error(void) { return g_err.c_str(); }

int crh_version(uint32_t* version) {
    ARG_CHECK(version, "version is NULL");
    *version = 1u;
    return CRH_OK;
}

int crh_device_count(int* count) {
    ARG_CHECK(count, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return CRH_OK;
}

// ---- context ---------------------------------------------------------------
int crh_context_create(int device, crh_context** out) {
    ARG_CHECK(out, "out is NULL");
    *out = nullptr;
    GUARD_BEGIN
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    ARG_CHECK(device >= 0 && device < n, "device index out of range");
    HIP_TRY(hipSetDevice(device));
    auto* c = new crh_context();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = crh::guideSetLdsLimits();
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&c->counter), 256);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    // The allocation-free entry point's scratch: this device's address of it.
    if (e == hipSuccess) e = crh::noallocScratchInit(device);
    if (e != hipSuccess) {
        crh_context_destroy(c);
        return hipFail("context setup", e, __FILE__, __LINE__);
    }
    *out = c;
    return CRH_OK;
    GUARD_END
}

int crh_context_destroy(crh_context* ctx) {
    if (!ctx) return CRH_OK;
    contextRelease(ctx);
    return CRH_OK;
}

int crh_context_set_tuning(crh_context* ctx, const crh_tuning* t) {
    ARG_CHECK(ctx && t, "NULL argument");
    auto okBlock = [](uint32_t b) { return b == 0 || (b >= 64 && b <= 256 && b % 64 == 0); };
    auto okStack = [](uint32_t v) { return v <= 3; };
    ARG_CHECK(okStack(t->stack) && okStack(t->render_stack) && okStack(t->wf_stack),
              "stack must be 0 (keep), 1 (LDS), 2 (private) or 3 (short LDS + private)");
    ARG_CHECK(okBlock(t->trace_block) && okBlock(t->wf_trace_block),
              "block sizes must be multiples of 64 in [64, 256]");
    ARG_CHECK(t->wf_trace_grid <= (1u << 20) && t->wf_shade_grid <= (1u << 20) &&
                  t->pt_grid <= (1u << 20),
              "grid too large");
    ARG_CHECK(t->persistent <= 2 && t->wf_persistent <= 2,
              "persistent must be 0 (keep), 1 (on) or 2 (off)");
    ARG_CHECK(t->refill <= 64 && t->wf_refill <= 64, "refill must be in [1, 64]");
    ARG_CHECK(t->render_waves == 0 || t->render_waves == 1 ||
                  (t->render_waves >= 4 && t->render_waves <= 6) || t->render_waves == 8,
              "render_waves must be 0 (keep), 1, 4, 5, 6 or 8");
    ARG_CHECK(t->batch_waves == 0 || t->batch_waves == 1 ||
                  (t->batch_waves >= 4 && t->batch_waves <= 6) || t->batch_waves == 8,
              "batch_waves must be 0 (keep), 1, 4, 5, 6 or 8");
    ARG_CHECK(t->render_regen <= 2, "render_regen must be 0 (keep), 1 (on) or 2 (off)");
    ARG_CHECK(t->render_grid <= (1u << 20), "render_grid too large");
    ARG_CHECK(t->wf_finish <= 65, "wf_finish must be 0 (keep), 1..64 (bounces) or 65 (off)");
    ARG_CHECK(t->render_band <= 2, "render_band must be 0 (keep), 1 (on) or 2 (off)");
    ARG_CHECK(t->batch_slots <= 64, "batch_slots must be 0 (keep) or 1..64");
    ARG_CHECK(t->trace_tris <= 4, "trace_tris must be 0 (keep) or 1..4");
    ARG_CHECK(t->wf_pool <= (1u << 28) && t->render_pool <= (1u << 28),
              "wf_pool and render_pool must be at most 2^28 paths");
    ARG_CHECK(okStack(t->batch_stack),
              "batch_stack must be 0 (keep), 1 (LDS), 2 (private) or 3 (short LDS + private)");
    ARG_CHECK(t->batch_tile_w == 0 || t->batch_tile_w == 8 || t->batch_tile_w == 16 ||
                  t->batch_tile_w == 32 || t->batch_tile_w == 64,
              "batch_tile_w must be 8, 16, 32 or 64");
    ARG_CHECK(t->batch_grid <= (1u << 20), "grid too large");
    ARG_CHECK(t->batch_chunks <= crh::kBatchMaxChunks, "batch_chunks must be in [1, 8]");
    ARG_CHECK(t->batch_graph <= 2, "batch_graph must be 0 (keep), 1 (on) or 2 (off)");
    ARG_CHECK(t->guided_fused <= 2, "guided_fused must be 0 (keep), 1 (on) or 2 (off)");
    ARG_CHECK(okStack(t->guided_stack),
              "guided_stack must be 0 (keep), 1 (LDS), 2 (private) or 3 (short LDS + private)");
    ARG_CHECK(t->guided_waves == 0 || t->guided_waves == 1 ||
                  (t->guided_waves >= 4 && t->guided_waves <= 6) || t->guided_waves == 8,
              "guided_waves must be 0 (keep), 1, 4, 5, 6 or 8");
    ARG_CHECK(t->guided_grid <= (1u << 16), "guided_grid must be at most 65536");
    ARG_CHECK(t->guided_pool <= (1u << 28), "guided_pool must be at most 2^28 paths");
    crh::Tuning& o = ctx->tuning;
    if (t->stack) o.privStack = t->stack - 1u;
    if (t->trace_block) o.traceBlock = t->trace_block;
    if (t->persistent) o.persistent = t->persistent == 1 ? 1u : 0u;
    if (t->refill) o.refill = t->refill;
    if (t->pt_grid) o.ptGrid = t->pt_grid;
    if (t->render_stack) o.renderPriv = t->render_stack - 1u;
    if (t->wf_stack) o.wfPriv = t->wf_stack - 1u;
    if (t->wf_trace_grid) o.wfTraceGrid = t->wf_trace_grid;
    if (t->wf_trace_block) o.wfTraceBlock = t->wf_trace_block;
    if (t->wf_shade_grid) o.wfShadeGrid = t->wf_shade_grid;
    if (t->wf_persistent) o.wfPersistent = t->wf_persistent == 1 ? 1u : 0u;
    if (t->wf_refill) o.wfRefill = t->wf_refill;
    if (t->wf_pool) o.wfPool = t->wf_pool;
    if (t->render_pool) o.renderPool = t->render_pool;
    if (t->render_waves) o.renderWaves = t->render_waves;
    if (t->render_regen) o.renderRegen = t->render_regen == 1 ? 1u : 0u;
    if (t->render_grid) o.renderGrid = t->render_grid;
    if (t->wf_finish) o.wfFinish = t->wf_finish == 65 ? 0u : t->wf_finish;
    if (t->render_band) o.xcdBand = t->render_band == 1 ? 1u : 0u;
    if (t->batch_stack) o.batchPriv = t->batch_stack - 1u;
    if (t->batch_waves) o.batchWaves = t->batch_waves;
    if (t->batch_tile_w) o.batchTileW = t->batch_tile_w;
    if (t->batch_grid) o.batchGrid = t->batch_grid;
    if (t->batch_chunks) o.batchChunks = t->batch_chunks;
    if (t->batch_graph) o.batchGraph = t->batch_graph == 1 ? 1u : 0u;
    if (t->batch_slots) o.batchSlots = t->batch_slots;
    if (t->trace_tris) o.traceTris = t->trace_tris;
    if (t->guided_fused) o.guidedFused = t->guided_fused == 1 ? 1u : 0u;
    if (t->guided_stack) o.guidedPriv = t->guided_stack - 1u;
    if (t->guided_waves) o.guidedWaves = t->guided_waves;
    if (t->guided_grid) o.guidedGrid = t->guided_grid;
    if (t->guided_pool) o.guidedPool = t->guided_pool;
    return CRH_OK;
}

int crh_context_get_tuning(const crh_context* ctx, crh_tuning* t) {
    ARG_CHECK(ctx && t, "NULL argument");
    std::memset(t, 0, sizeof(*t));
    const crh::Tuning& o = ctx->tuning;
    t->stack = o.privStack + 1u;
    t->trace_block = o.traceBlock;
    t->persistent = o.persistent ? 1u : 2u;
    t->refill = o.refill;
    t->pt_grid = o.ptGrid;
    t->render_stack = o.renderPriv + 1u;
    t->wf_stack = o.wfPriv + 1u;
    t->wf_trace_grid = o.wfTraceGrid;
    t->wf_trace_block = o.wfTraceBlock;
    t->wf_shade_grid = o.wfShadeGrid;
    t->wf_persistent = o.wfPersistent ? 1u : 2u;
    t->wf_refill = o.wfRefill;
    t->wf_pool = o.wfPool;
    t->render_pool = o.renderPool;
    t->render_waves = o.renderWaves;
    t->render_regen = o.renderRegen ? 1u : 2u;
    t->render_grid = o.renderGrid;
    t->wf_finish = o.wfFinish ? o.wfFinish : 65u;
    t->render_band = o.xcdBand ? 1u : 2u;
    t->batch_stack = o.batchPriv + 1u;
    t->batch_waves = o.batchWaves;
    t->batch_tile_w = o.batchTileW;
    t->batch_grid = o.batchGrid;
    t->batch_chunks = o.batchChunks;
    t->batch_graph = o.batchGraph ? 1u : 2u;
    t->batch_slots = o.batchSlots;
    t->trace_tris = o.traceTris;
    t->guided_fused = o.guidedFused ? 1u : 2u;
    t->guided_stack = o.guidedPriv + 1u;
    t->guided_waves = o.guidedWaves;
    t->guided_grid = o.guidedGrid;
    t->guided_pool = o.guidedPool;
    return CRH_OK;
}

// ---- buffers ---------------------------------------------------------------
int crh_buffer_create(crh_context* ctx, size_t bytes, crh_buffer** out) {
    ARG_CHECK(ctx && out, "NULL argument");
    *out = nullptr;
    GUARD_BEGIN
    HIP_TRY(hipSetDevice(ctx->device));
    auto* b = new crh_buffer();
    b->ctx.set(ctx);
    hipError_t e = b->mem.alloc(bytes);
    if (e != hipSuccess) {
        delete b;
        return hipFail("hipMalloc", e, __FILE__, __LINE__);
    }
    *out = b;
    return CRH_OK;
    GUARD_END
}

int crh_buffer_upload(crh_buffer* buf, const void* src, size_t bytes) {
    ARG_CHECK(buf && (src || bytes == 0), "NULL argument");
    ARG_CHECK(bytes <= buf->mem.bytes, "upload larger than the buffer");
    if (bytes == 0) return CRH_OK;
    HIP_TRY(hipSetDevice(buf->ctx->device));
    HIP_TRY(hipMemcpyAsync(buf->mem.p, src, bytes, hipMemcpyHostToDevice,
                           buf->ctx->stream));
    HIP_TRY(hipStreamSynchronize(buf->ctx->stream));
    return CRH_OK;
}

int crh_buffer_download(const crh_buffer* buf, void* dst, size_t bytes) {
    ARG_CHECK(buf && (dst || bytes == 0), "NULL argument");
    ARG_CHECK(bytes <= buf->mem.bytes, "download larger than the buffer");
    if (bytes == 0) return CRH_OK;
    HIP_TRY(hipSetDevice(buf->ctx->device));
    HIP_TRY(hipMemcpyAsync(dst, buf->mem.p, bytes, hipMemcpyDeviceToHost,
                           buf->ctx->stream));
    HIP_TRY(hipStreamSynchronize(buf->ctx->stream));
    return CRH_OK;
}

int crh_buffer_memset(crh_buffer* buf, int value) {
    ARG_CHECK(buf, "NULL argument");
    if (buf->mem.bytes == 0) return CRH_OK;
    HIP_TRY(hipSetDevice(buf->ctx->device));
    HIP_TRY(hipMemsetAsync(buf->mem.p, value, buf->mem.bytes, buf->ctx->stream));
    HIP_TRY(hipStreamSynchronize(buf->ctx->stream));
    return CRH_OK;
}

int crh_buffer_copy(crh_buffer* dst, const crh_buffer* src, size_t bytes) {
    ARG_CHECK(dst && src, "NULL argument");
    ARG_CHECK(dst->ctx == src->ctx, "buffers belong to different contexts");
    ARG_CHECK(bytes <= dst->mem.bytes && bytes <= src->mem.bytes, "copy larger than a buffer");
    if (bytes == 0) return CRH_OK;
    HIP_TRY(hipSetDevice(dst->ctx->device));
    HIP_TRY(hipMemcpyAsync(dst->mem.p, src->mem.p, bytes, hipMemcpyDeviceToDevice,
                           dst->ctx->stream));
    HIP_TRY(hipStreamSynchronize(dst->ctx->stream));
    return CRH_OK;
}

int crh_buffer_size(const crh_buffer* buf, size_t* bytes) {
    ARG_CHECK(buf && bytes, "NULL argument");
    *bytes = buf->mem.bytes;
    return CRH_OK;
}

int crh_buffer_destroy(crh_buffer* buf) {
    if (!buf) return CRH_OK;
    (void)hipSetDevice(buf->ctx->device);
    delete buf;
    return CRH_OK;
}

// ---- BVH ---------------------------------------------------------------------
static void fillStats(const crh::BvhResult& r, crh_bvh_stats* st) {
    std::memset(st, 0, sizeof(*st));
    st->node_count = uint32_t(r.nodes.size());
    st->max_depth = r.maxDepth;
    st->max_leaf = r.maxLeaf;
    st->root = r.root;
    st->prim_count = uint32_t(r.primOrder.size());
    st->internal_count = r.internalCount;
    for (int k = 0; k < 3; ++k) {
        st->root_min[k] = r.rootMin[k];
        st->root_max[k] = r.rootMax[k];
    }
    st->build_ms = r.buildMs;
}

static void exportBvh(const crh::BvhResult& r, crh_bvh_node_ref* nodes,
                      uint32_t* prims) {
    if (nodes) std::memcpy(nodes, r.nodes.data(), r.nodes.size() * sizeof(*nodes));
    if (prims)
        std::memcpy(prims, r.primOrder.data(), r.primOrder.size() * sizeof(*prims));
}

int crh_bvh_build(const float* pos, uint32_t nvert, const uint32_t* idx,
                  uint32_t ntri, const crh_bvh_params* params, crh_bvh** out) {
    ARG_CHECK(out, "out is NULL");
    *out = nullptr;
    ARG_CHECK((pos || nvert == 0) && (idx || ntri == 0), "NULL geometry");
    GUARD_BEGIN
    for (size_t i = 0; i < size_t(ntri) * 3; ++i)
        ARG_CHECK(idx[i] < nvert, "vertex index out of range");
    crh::BuildParams bp = crh::resolveParams(params);
    ARG_CHECK(bp.bins >= 2 && bp.bins <= 1024, "bins must be in [2, 1024]");
    auto* b = new crh_bvh();
    b->r = crh::buildBvh(pos, nvert, idx, ntri, bp);
    *out = b;
    return CRH_OK;
    GUARD_END
}

namespace {
// The GPU builder from host arrays: uploads the geometry, builds, returns.
hipError_t buildOnDevice(crh_context* ctx, const float* pos, uint32_t nvert, const uint32_t* idx,
                         uint32_t ntri, const crh::BuildParams& bp, crh::BvhResult& out) {
    const double t0 = nowMs();
    DevMem dp, di;
    hipError_t e = dp.alloc(size_t(nvert) * 3 * sizeof(float));
    if (e == hipSuccess) e = di.alloc(size_t(ntri) * 3 * sizeof(uint32_t));
    if (e == hipSuccess && nvert)
        e = hipMemcpyAsync(dp.p, pos, dp.bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && ntri)
        e = hipMemcpyAsync(di.p, idx, di.bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = crh::buildBvhDevice(dp.as<float>(), di.as<uint32_t>(), ntri, bp, ctx->stream, out);
    out.buildMs = nowMs() - t0;
    return e;
}
}  // namespace

int crh_bvh_build_device(crh_context* ctx, const float* pos, uint32_t nvert,
                         const uint32_t* idx, uint32_t ntri, const crh_bvh_params* params,
                         crh_bvh** out) {
    ARG_CHECK(ctx && out, "NULL argument");
    *out = nullptr;
    ARG_CHECK((pos || nvert == 0) && (idx || ntri == 0), "NULL geometry");
    GUARD_BEGIN
    for (size_t i = 0; i < size_t(ntri) * 3; ++i)
        ARG_CHECK(idx[i] < nvert, "vertex index out of range");
    crh::BuildParams bp = crh::resolveParams(params);
    ARG_CHECK(bp.bins >= 2 && bp.bins <= crh::kDeviceMaxBins,
              "the GPU builder takes bins in [2, 64]");
    ARG_CHECK(bp.maxDepth <= crh::kDeviceMaxDepth, "the GPU builder takes max_depth <= 64");
    HIP_TRY(hipSetDevice(ctx->device));
    auto* b = new crh_bvh();
    hipError_t e = buildOnDevice(ctx, pos, nvert, idx, ntri, bp, b->r);
    if (e != hipSuccess) {
        delete b;
        return hipFail("GPU BVH build", e, __FILE__, __LINE__);
    }
    *out = b;
    return CRH_OK;
    GUARD_END
}

int crh_bvh_get_stats(const crh_bvh* bvh, crh_bvh_stats* stats) {
    ARG_CHECK(bvh && stats, "NULL argument");
    fillStats(bvh->r, stats);
    return CRH_OK;
}

int crh_bvh_export(const crh_bvh* bvh, crh_bvh_node_ref* nodes,
                   uint32_t* prim_order) {
    ARG_CHECK(bvh, "NULL argument");
    exportBvh(bvh->r, nodes, prim_order);
    return CRH_OK;
}

int crh_bvh_destroy(crh_bvh* bvh) {
    delete bvh;
    return CRH_OK;
}

// ---- materials ---------------------------------------------------------------
int crh_material_finalize(crh_material* m) {
    ARG_CHECK(m, "NULL argument");
    auto maxOf = [](const float* v) {
        float r = v[0];
        for (int i = 1; i < 3; ++i) r = (r < v[i]) ? v[i] : r;
        return r;
    };
    if (maxOf(m->emission) != 0.0f) m->flags |= CRH_MATF_EMISSIVE;
    if (maxOf(m->specular) == 0.0f) m->flags |= CRH_MATF_DIFFUSE;
    return CRH_OK;
}

// ---- scene ---------------------------------------------------------------------
int crh_scene_create(crh_context* ctx, const crh_scene_desc* d,
                     crh_scene** out) {
    ARG_CHECK(ctx && d && out, "NULL argument");
    *out = nullptr;
    ARG_CHECK(d->nmesh == 0 || d->meshes, "meshes is NULL");
    ARG_CHECK(d->nmaterial == 0 || d->materials, "materials is NULL");
    ARG_CHECK(d->nemissive == 0 || d->emissive_index, "emissive_index is NULL");
    GUARD_BEGIN
    // Concatenate meshes (TriangleMeshManager, Scene/Triangle.cu:98-165).
    std::vector<float> pos, nrm, uv;
    std::vector<uint32_t> idx, primMat;
    for (uint32_t m = 0; m < d->nmesh; ++m) {
        const crh_mesh& mesh = d->meshes[m];
        ARG_CHECK(mesh.nindex % 3 == 0, "mesh index count is not a multiple of 3");
        ARG_CHECK(mesh.nvert == 0 || (mesh.position && mesh.normal),
                  "mesh position/normal is NULL");
        ARG_CHECK(mesh.nindex == 0 || mesh.index, "mesh index is NULL");
        ARG_CHECK(mesh.material_id < d->nmaterial, "material_id out of range");
        uint32_t voff = uint32_t(pos.size() / 3);
        pos.insert(pos.end(), mesh.position, mesh.position + 3 * size_t(mesh.nvert));
        nrm.insert(nrm.end(), mesh.normal, mesh.normal + 3 * size_t(mesh.nvert));
        if (mesh.texcrd)
            uv.insert(uv.end(), mesh.texcrd, mesh.texcrd + 2 * size_t(mesh.nvert));
        else
            uv.insert(uv.end(), 2 * size_t(mesh.nvert), 0.0f);
        for (uint32_t i = 0; i < mesh.nindex; ++i) {
            ARG_CHECK(mesh.index[i] < mesh.nvert, "vertex index out of range");
            idx.push_back(mesh.index[i] + voff);
        }
        primMat.insert(primMat.end(), mesh.nindex / 3, mesh.material_id);
    }
    ARG_CHECK(idx.size() / 3 < crh::kSphereBit, "too many triangles (at most 2^30 - 1)");
    const uint32_t ntri = uint32_t(idx.size() / 3);
    for (uint32_t i = 0; i < d->nemissive; ++i)
        ARG_CHECK(d->emissive_index[i] < ntri, "emissive index out of range");
    // Spheres: centre + radius as 4 floats; their materials.
    ARG_CHECK(d->nsphere == 0 || (d->sphere_center && d->sphere_radius),
              "sphere_center / sphere_radius is NULL");
    ARG_CHECK(d->nsphere < crh::kSphereBit, "too many spheres");
    std::vector<float> sph(4 * size_t(d->nsphere));
    for (uint32_t i = 0; i < d->nsphere; ++i) {
        for (int k = 0; k < 3; ++k) sph[4 * size_t(i) + k] = d->sphere_center[3 * size_t(i) + k];
        sph[4 * size_t(i) + 3] = d->sphere_radius[i];
    }
    std::vector<uint32_t> sphMat;
    if (d->nsphere && d->sphere_material) {
        sphMat.assign(d->sphere_material, d->sphere_material + d->nsphere);
        for (uint32_t id : sphMat) ARG_CHECK(id < d->nmaterial, "sphere material out of range");
    } else {
        ARG_CHECK(d->nsphere <= ntri,
                  "without sphere_material a sphere takes the material of the triangle with "
                  "its index, so the sphere count must not exceed the triangle count");
    }
    // Textures: every image as 4-float texels, back to back.
    ARG_CHECK(d->ntexture == 0 || d->textures, "textures is NULL");
    std::vector<float> texels;
    std::vector<uint32_t> texInfo(4 * size_t(d->ntexture), 0u);
    for (uint32_t t = 0; t < d->ntexture; ++t) {
        const crh_texture& tx = d->textures[t];
        ARG_CHECK(tx.width > 0 && tx.height > 0 && tx.data, "empty texture");
        ARG_CHECK(tx.channels >= 1 && tx.channels <= 4, "texture channels must be 1..4");
        const size_t first = texels.size() / 4, n = size_t(tx.width) * tx.height;
        ARG_CHECK(first + n < (size_t(1) << 32), "texture data too large");
        texInfo[4 * size_t(t)] = uint32_t(first);
        texInfo[4 * size_t(t) + 1] = tx.width;
        texInfo[4 * size_t(t) + 2] = tx.height;
        texels.resize(4 * (first + n), 0.0f);
        for (size_t i = 0; i < n; ++i)
            for (uint32_t c = 0; c < tx.channels; ++c)
                texels[4 * (first + i) + c] = tx.data[i * tx.channels + c];
    }
    for (uint32_t i = 0; i < d->nmaterial; ++i)
        ARG_CHECK(!(d->materials[i].flags & CRH_MATF_DIFFUSE_TEXTURE) ||
                      d->materials[i].diffuse_tex < d->ntexture,
                  "diffuse_tex out of range");
    crh::BuildParams bp = crh::resolveParams(&d->bvh);
    ARG_CHECK(bp.bins >= 2 && bp.bins <= 1024, "bins must be in [2, 1024]");
    ARG_CHECK(d->bvh.device <= 1u, "bvh.device must be 0 (host) or 1 (GPU)");
    ARG_CHECK(!d->bvh.device || (bp.bins <= crh::kDeviceMaxBins && bp.maxDepth <= crh::kDeviceMaxDepth),
              "the GPU builder takes bins in [2, 64] and max_depth <= 64");

    auto* s = new crh_scene();
    s->ctx.set(ctx);
    if (d->bvh.device) {
        HIP_TRY(hipSetDevice(ctx->device));
        hipError_t e = buildOnDevice(ctx, pos.data(), uint32_t(pos.size() / 3), idx.data(), ntri,
                                     bp, s->bvh);
        if (e != hipSuccess) {
            delete s;
            return hipFail("GPU BVH build", e, __FILE__, __LINE__);
        }
    } else {
        s->bvh = crh::buildBvh(pos.data(), uint32_t(pos.size() / 3), idx.data(), ntri, bp);
    }
    crh::GpuBvh g = crh::toGpuLayout(s->bvh, pos.data(), idx.data());
    std::vector<crh::GMaterial> gm(d->nmaterial);
    for (uint32_t i = 0; i < d->nmaterial; ++i) {
        const crh_material& m = d->materials[i];
        crh::GMaterial& o = gm[i];
        for (int k = 0; k < 3; ++k) {
            o.kd[k] = m.diffuse[k];
            o.ks[k] = m.specular[k];
            o.emission[k] = m.emission[k];
        }
        o.type = m.type;
        o.flags = m.flags;
        o.shininess = m.shininess;
        o.tex = m.diffuse_tex;
    }
    std::vector<uint32_t> em(d->emissive_index, d->emissive_index + d->nemissive);

    hipStream_t st = ctx->stream;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = upload(s->nodes, g.nodes, st);
    if (e == hipSuccess) e = upload(s->tris, g.tris, st);
    if (e == hipSuccess) e = upload(s->pos, pos, st);
    if (e == hipSuccess) e = upload(s->nrm, nrm, st);
    if (e == hipSuccess) e = upload(s->uv, uv, st);
    if (e == hipSuccess) e = upload(s->idx, idx, st);
    if (e == hipSuccess) e = upload(s->primMat, primMat, st);
    if (e == hipSuccess) e = upload(s->mats, gm, st);
    if (e == hipSuccess) e = upload(s->emissive, em, st);
    // Spectral table: the given spectra, or the materials' and environment's
    // RGB uplifted to 95 bins (RGB-authored scenes render spectrally too).
    std::vector<float> specTab;
    if (d->spectra) {
        specTab = crh::buildSpectralTable(d->spectra, d->nmaterial, d->env_spectrum);
    } else {
        std::vector<float> rgb(size_t(d->nmaterial) * 9);
        for (uint32_t i = 0; i < d->nmaterial; ++i)
            for (int k = 0; k < 3; ++k) {
                rgb[9 * i + k] = d->materials[i].diffuse[k];
                rgb[9 * i + 3 + k] = d->materials[i].specular[k];
                rgb[9 * i + 6 + k] = d->materials[i].emission[k];
            }
        std::vector<float> sp = crh::upliftedSpectra(rgb.data(), d->nmaterial, 9);
        float env[crh::kCieBins];
        crh::rgbToSpectrum(d->env_radiance, env);
        specTab = crh::buildSpectralTable(sp.data(), d->nmaterial, env);
    }
    if (e == hipSuccess) e = upload(s->spec, specTab, st);
    if (e == hipSuccess) e = upload(s->spheres, sph, st);
    if (e == hipSuccess) e = upload(s->sphereMat, sphMat, st);
    if (e == hipSuccess) e = upload(s->texels, texels, st);
    if (e == hipSuccess) e = upload(s->texInfo, texInfo, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        delete s;
        return hipFail("scene upload", e, __FILE__, __LINE__);
    }
    crh::SceneView& v = s->view;
    v.nodes = s->nodes.as<crh::GNode>();
    v.tris = s->tris.as<crh::GTri>();
    v.rootRef = g.rootRef;
    v.hasGeom = g.hasGeom;
    v.stackCap = s->bvh.levels + 1;
    v.nEmissive = d->nemissive;
    for (int k = 0; k < 3; ++k) {
        v.rootMin[k] = s->bvh.rootMin[k];
        v.rootMax[k] = s->bvh.rootMax[k];
        v.env[k] = d->env_radiance[k];
    }
    v.pos = s->pos.as<float>();
    v.nrm = s->nrm.as<float>();
    v.uv = s->uv.as<float>();
    v.idx = s->idx.as<uint32_t>();
    v.primMat = s->primMat.as<uint32_t>();
    v.mats = s->mats.as<crh::GMaterial>();
    v.emissive = s->emissive.as<uint32_t>();
    v.spec = s->spec.as<float>();
    v.spheres = s->spheres.as<float>();
    v.sphereMat = sphMat.empty() ? nullptr : s->sphereMat.as<uint32_t>();
    v.nSphere = d->nsphere;
    v.texels = s->texels.as<float>();
    v.texInfo = s->texInfo.as<uint32_t>();
    v.nTexture = d->ntexture;
    *out = s;
    return CRH_OK;
    GUARD_END
}

int crh_scene_get_bvh_stats(const crh_scene* s, crh_bvh_stats* st) {
    ARG_CHECK(s && st, "NULL argument");
    fillStats(s->bvh, st);
    return CRH_OK;
}

int crh_scene_export_bvh(const crh_scene* s, crh_bvh_node_ref* nodes,
                         uint32_t* prims) {
    ARG_CHECK(s, "NULL argument");
    exportBvh(s->bvh, nodes, prims);
    return CRH_OK;
}

int crh_scene_destroy(crh_scene* s) {
    if (!s) return CRH_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    delete s;
    return CRH_OK;
}

// ---- camera --------------------------------------------------------------------
int crh_camera_update(crh_camera* c) {
    ARG_CHECK(c, "NULL argument");
    auto dot = [](const float* a, const float* b) {
        float x = a[0] * b[0], y = a[1] * b[1], z = a[2] * b[2];
        return (x + y) + z;
    };
    auto norm = [&](float* v) {
        float s = 1.0f / std::sqrt(dot(v, v));
        for (int k = 0; k < 3; ++k) v[k] = v[k] * s;
    };
    auto cross = [](const float* a, const float* b, float* o) {
        o[0] = a[1] * b[2] - b[1] * a[2];
        o[1] = a[2] * b[0] - b[2] * a[0];
        o[2] = a[0] * b[1] - b[0] * a[1];
    };
    float W[3], U[3], V[3];
    for (int k = 0; k < 3; ++k) W[k] = c->target[k] - c->pos[k];
    norm(W);
    cross(W, c->up, U);
    norm(U);
    cross(U, W, V);
    norm(V);
    for (int k = 0; k < 3; ++k) {
        c->w[k] = W[k];
        c->u[k] = U[k];
        c->v[k] = V[k];
    }
    return CRH_OK;
}

// ---- ray queries ---------------------------------------------------------------
static int traceClosestDev(crh_scene* s, const crh_ray* rays, uint64_t n,
                           crh_hit* hits, uint32_t* work, crh_stats* stats) {
    crh_context* ctx = s->ctx;
    double t0 = nowMs();
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(crh::launchTraceClosest(s->view, rays, n, hits, work,
                                    effective(ctx->tuning, s->view), ctx->counter, ctx->stream));
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    if (stats) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        stats->kernel_ms = ms;
        stats->main_kernel_ms = ms;
        stats->total_ms = nowMs() - t0;
        stats->rays = n;
        stats->launches = 1;
    }
    return CRH_OK;
}

int crh_trace_closest_device(crh_scene* s, const crh_buffer* rays, uint64_t n,
                             crh_buffer* hits, int count_work,
                             crh_stats* stats) {
    ARG_CHECK(s && rays && hits, "NULL argument");
    ARG_CHECK(rays->mem.bytes >= n * sizeof(crh_ray), "ray buffer too small");
    ARG_CHECK(hits->mem.bytes >= n * sizeof(crh_hit), "hit buffer too small");
    ARG_CHECK(n < (uint64_t(1) << 40), "ray count too large");
    GUARD_BEGIN
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (!count_work)
        return traceClosestDev(s, rays->mem.as<crh_ray>(), n,
                               hits->mem.as<crh_hit>(), nullptr, stats);
    DevMem work;
    HIP_TRY(hipSetDevice(s->ctx->device));
    HIP_TRY(work.alloc(n * 2 * sizeof(uint32_t)));
    int rc = traceClosestDev(s, rays->mem.as<crh_ray>(), n,
                             hits->mem.as<crh_hit>(), work.as<uint32_t>(), stats);
    if (rc != CRH_OK) return rc;
    std::vector<uint32_t> w(n * 2);
    HIP_TRY(hipMemcpy(w.data(), work.p, w.size() * 4, hipMemcpyDeviceToHost));
    uint64_t nodes = 0, tris = 0;
    for (uint64_t i = 0; i < n; ++i) {
        nodes += w[2 * i];
        tris += w[2 * i + 1];
    }
    if (stats) {
        stats->node_visits = nodes;
        stats->tri_tests = tris;
    }
    return CRH_OK;
    GUARD_END
}

int crh_trace_closest_device_noalloc(crh_scene* s, const crh_buffer* rays, uint64_t n,
                                     crh_buffer* hits, int count_work, crh_stats* stats) {
    ARG_CHECK(s && rays && hits, "NULL argument");
    ARG_CHECK(rays->mem.bytes >= n * sizeof(crh_ray), "ray buffer too small");
    ARG_CHECK(hits->mem.bytes >= n * sizeof(crh_hit), "hit buffer too small");
    ARG_CHECK(n < (uint64_t(1) << 40), "ray count too large");
    GUARD_BEGIN
    crh_context* ctx = s->ctx;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    double t0 = nowMs();
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    unsigned long long totals[2] = {0, 0};
    HIP_TRY(crh::traceClosestNoAlloc(ctx->device, s->view, rays->mem.as<crh_ray>(), n,
                                     hits->mem.as<crh_hit>(), count_work != 0,
                                     effective(ctx->tuning, s->view), ctx->stream, totals));
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    if (stats) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        stats->kernel_ms = ms;
        stats->main_kernel_ms = ms;
        stats->total_ms = nowMs() - t0;
        stats->rays = n;
        stats->node_visits = totals[0];
        stats->tri_tests = totals[1];
        stats->launches = 1;
    }
    return CRH_OK;
    GUARD_END
}

int crh_trace_any_device(crh_scene* s, const crh_buffer* rays, uint64_t n,
                         crh_buffer* occ, crh_stats* stats) {
    ARG_CHECK(s && rays && occ, "NULL argument");
    ARG_CHECK(rays->mem.bytes >= n * sizeof(crh_ray), "ray buffer too small");
    ARG_CHECK(occ->mem.bytes >= n, "occlusion buffer too small");
    GUARD_BEGIN
    crh_context* ctx = s->ctx;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    double t0 = nowMs();
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(crh::launchTraceAny(s->view, rays->mem.as<crh_ray>(), n,
                                occ->mem.as<uint8_t>(), effective(ctx->tuning, s->view),
                                ctx->counter, ctx->stream));
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    if (stats) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        stats->kernel_ms = ms;
        stats->main_kernel_ms = ms;
        stats->total_ms = nowMs() - t0;
        stats->rays = n;
        stats->launches = 1;
    }
    return CRH_OK;
    GUARD_END
}

int crh_trace_closest(crh_scene* s, const crh_ray* rays, uint64_t n,
                      crh_hit* hits) {
    ARG_CHECK(s && (rays || n == 0) && (hits || n == 0), "NULL argument");
    if (n == 0) return CRH_OK;
    GUARD_BEGIN
    HIP_TRY(hipSetDevice(s->ctx->device));
    DevMem dr, dh;
    HIP_TRY(dr.alloc(n * sizeof(crh_ray)));
    HIP_TRY(dh.alloc(n * sizeof(crh_hit)));
    HIP_TRY(hipMemcpy(dr.p, rays, n * sizeof(crh_ray), hipMemcpyHostToDevice));
    int rc = traceClosestDev(s, dr.as<crh_ray>(), n, dh.as<crh_hit>(), nullptr,
                             nullptr);
    if (rc != CRH_OK) return rc;
    HIP_TRY(hipMemcpy(hits, dh.p, n * sizeof(crh_hit), hipMemcpyDeviceToHost));
    return CRH_OK;
    GUARD_END
}

int crh_trace_any(crh_scene* s, const crh_ray* rays, uint64_t n,
                  uint8_t* occluded) {
    ARG_CHECK(s && (rays || n == 0) && (occluded || n == 0), "NULL argument");
    if (n == 0) return CRH_OK;
    GUARD_BEGIN
    crh_context* ctx = s->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    DevMem dr, dox;
    HIP_TRY(dr.alloc(n * sizeof(crh_ray)));
    HIP_TRY(dox.alloc(n));
    HIP_TRY(hipMemcpy(dr.p, rays, n * sizeof(crh_ray), hipMemcpyHostToDevice));
    HIP_TRY(crh::launchTraceAny(s->view, dr.as<crh_ray>(), n, dox.as<uint8_t>(),
                                effective(ctx->tuning, s->view), ctx->counter, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemcpy(occluded, dox.p, n, hipMemcpyDeviceToHost));
    return CRH_OK;
    GUARD_END
}

// ---- film ---------------------------------------------------------------------
int crh_film_create(crh_context* ctx, uint32_t w, uint32_t h, crh_film** out) {
    ARG_CHECK(ctx && out, "NULL argument");
    *out = nullptr;
    ARG_CHECK(w > 0 && h > 0, "film size must be positive");
    ARG_CHECK(uint64_t(w) * h <= (uint64_t(1) << 31), "film too large");
    GUARD_BEGIN
    HIP_TRY(hipSetDevice(ctx->device));
    auto* f = new crh_film();
    f->ctx.set(ctx);
    f->w = w;
    f->h = h;
    size_t n = size_t(w) * h;
    hipError_t e = f->sum.alloc(n * 3 * sizeof(float));
    if (e == hipSuccess) e = f->wsum.alloc(n * sizeof(float));
    if (e == hipSuccess) e = f->L.alloc(n * sizeof(float4));
    if (e == hipSuccess) e = f->P.alloc(n * sizeof(float2));
    if (e == hipSuccess) e = f->work.alloc(3 * sizeof(unsigned long long));
    if (e != hipSuccess) {
        delete f;
        return hipFail("film allocation", e, __FILE__, __LINE__);
    }
    *out = f;
    return crh_film_clear(f);
    GUARD_END
}

int crh_film_clear(crh_film* f) {
    ARG_CHECK(f, "NULL argument");
    HIP_TRY(hipSetDevice(f->ctx->device));
    HIP_TRY(hipMemsetAsync(f->sum.p, 0, f->sum.bytes, f->ctx->stream));
    HIP_TRY(hipMemsetAsync(f->wsum.p, 0, f->wsum.bytes, f->ctx->stream));
    HIP_TRY(hipStreamSynchronize(f->ctx->stream));
    return CRH_OK;
}

int crh_film_readback(const crh_film* f, float* sum, float* weight) {
    ARG_CHECK(f, "NULL argument");
    HIP_TRY(hipSetDevice(f->ctx->device));
    if (sum)
        HIP_TRY(hipMemcpy(sum, f->sum.p, f->sum.bytes, hipMemcpyDeviceToHost));
    if (weight)
        HIP_TRY(hipMemcpy(weight, f->wsum.p, f->wsum.bytes, hipMemcpyDeviceToHost));
    return CRH_OK;
}

int crh_film_destroy(crh_film* f) {
    if (!f) return CRH_OK;
    (void)hipSetDevice(f->ctx->device);
    (void)hipStreamSynchronize(f->ctx->stream);
    delete f;
    return CRH_OK;
}

// ---- render --------------------------------------------------------------------
int crh_render(crh_scene* s, const crh_camera* cam,
               const crh_integrator_cfg* cfg, uint32_t first_sample,
               uint32_t spp, crh_film* film, crh_stats* stats) {
    ARG_CHECK(s && cam && cfg && film, "NULL argument");
    ARG_CHECK(film->ctx == s->ctx, "film and scene belong to different contexts");
    ARG_CHECK(cfg->mode == CRH_MODE_MEGAKERNEL || cfg->mode == CRH_MODE_WAVEFRONT,
              "unknown integrator mode");
    ARG_CHECK(cfg->max_depth <= 1024, "max_depth too large");
    ARG_CHECK(!cfg->spectral || s->view.spec, "spectral transport needs a scene with spectra");
    ARG_CHECK(!cfg->spectral || !cfg->count_work, "spectral transport has no counting variant");
    GUARD_BEGIN
    crh_context* ctx = s->ctx;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    const crh::CameraView cv = makeCameraView(cam, film->w, film->h);
    crh::IntegratorView iv{cfg->max_depth, cfg->rr_threshold,
                           cfg->rr_start_depth, cfg->use_mis ? 1u : 0u};
    double t0 = nowMs();
    HIP_TRY(hipSetDevice(ctx->device));
    DevMem& counter = film->work;
    HIP_TRY(hipMemsetAsync(counter.p, 0, counter.bytes, ctx->stream));
    const bool wavefront = cfg->mode == CRH_MODE_WAVEFRONT;
    const crh::Tuning tun = effective(ctx->tuning, s->view);
    const crh::WfLaunch wl = wfLaunch(tun);
    const uint32_t slots = wavefront ? wfSlots(tun, film, spp)
                                     : poolSlots(tun.renderPool, film, spp);
    if (wavefront) {
        HIP_TRY(ensureWavefront(film, slots));
        HIP_TRY(hipMemsetAsync(film->wf.work, 0, 64, ctx->stream));
    } else {
        HIP_TRY(ensureFilmSlots(film, slots));
    }
    // One event pair per pass group times the main launches on their stream;
    // the film keeps the events for later calls.
    std::vector<hipEvent_t>& evs = film->passEv;
    const size_t nev = stats ? 2 * size_t((spp + slots - 1) / slots) : 0;
    while (evs.size() < nev) {
        hipEvent_t e = nullptr;
        HIP_TRY(hipEventCreate(&e));
        evs.push_back(e);
    }
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    // Pass groups: `slots` sample indices per megakernel launch or wavefront
    // pass, gathered into the film in sample order.
    for (uint32_t i = 0; i < spp; i += slots) {
        const uint32_t active = std::min(slots, spp - i);
        const size_t g = i / slots;
        if (stats) HIP_TRY(hipEventRecord(evs[2 * g], ctx->stream));
        if (wavefront)
            HIP_TRY(crh::wavefrontPass(s->view, cv, iv, first_sample + i, active, film->wf, wl,
                                       cfg->count_work != 0, ctx->stream, nullptr, 0.5f, nullptr,
                                       0u, 3u, 0u, 0u, cfg->spectral != 0));
        else
            HIP_TRY(crh::launchRenderPass(s->view, cv, iv, first_sample + i, active,
                                          film->L.as<float4>(), film->P.as<float2>(),
                                          counter.as<unsigned long long>(),
                                          cfg->count_work != 0, tun, ctx->stream,
                                          cfg->spectral != 0, ctx->counter));
        if (stats) HIP_TRY(hipEventRecord(evs[2 * g + 1], ctx->stream));
        HIP_TRY(crh::launchFilmGather(wavefront ? film->wf.rad : film->L.as<float4>(),
                                      wavefront ? film->wf.P : film->P.as<float2>(), film->w,
                                      film->h, film->sum.as<float>(), film->wsum.as<float>(),
                                      ctx->stream, active));
    }
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    if (stats) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        double mainMs = 0.0;
        for (size_t g = 0; 2 * g < nev; ++g) {
            float m = 0.f;
            HIP_TRY(hipEventElapsedTime(&m, evs[2 * g], evs[2 * g + 1]));
            mainMs += m;
        }
        unsigned long long c[3] = {0, 0, 0};
        HIP_TRY(hipMemcpy(c, wavefront ? (void*)film->wf.work : counter.p, sizeof(c),
                          hipMemcpyDeviceToHost));
        stats->kernel_ms = ms;
        stats->main_kernel_ms = mainMs;
        stats->total_ms = nowMs() - t0;
        stats->rays = c[0];
        stats->node_visits = c[1];
        stats->tri_tests = c[2];
        const uint64_t groups = (spp + slots - 1) / slots;
        stats->launches = wavefront ? groups * (6u * cfg->max_depth + 5u) + spp
                                    : groups + spp;
    }
    return CRH_OK;
    GUARD_END
}

// ---- learned path guiding ----------------------------------------------------
}  // extern "C"

struct crh_guide {
    CtxRef ctx;
    // loss: the losses of the last kLossHist Adam steps (step k in slot
    // k % kLossHist), then one slot for gradient-only steps.
    DevMem theta, m, v, grad, wq, wqT, loss, samples, hsave, dsave;
    crh::GuideView view{};
    uint32_t step = 0;
    uint32_t lastSamples = 0;
    size_t capSamples = 0;
    uint32_t width = 64;  // hidden width of the network
    uint32_t params() const { return uint32_t(crh::gParamsOf(int(width))); }
};

namespace {

constexpr uint32_t kLossHist = 1024;

float* lossSlot(crh_guide* g) { return g->loss.as<float>() + (g->step % kLossHist); }

// After a guided pass group: its records become training samples (fused
// kernel: from the final radiance in recL), which join the film's pending
// samples; then up to active * train_steps Adam steps, one per full batch.
int guidedTrainPass(crh_guide* g, crh_film* film, const crh_guided_cfg* gc, uint32_t active,
                    bool fused, uint32_t recCap, const float4* recL, crh::GuideSample* smpOut,
                    uint32_t smpRoom, uint32_t& lastSamples) {
    crh_context* ctx = film->ctx;
    const uint32_t batch = gc->train_batch;
    crh::GuideSample* smpBuf = film->gSamples.as<crh::GuideSample>();
    uint32_t* trainCounter = fused ? film->gCnt.as<uint32_t>() + 1
                                   : film->wf.counters + crh::kCTrain;
    if (fused)
        HIP_TRY(crh::launchGuidedRecordsToSamples(film->gRec.as<float4>(), film->gCnt.as<uint32_t>(),
                                                  recCap, recL, smpOut, smpRoom, trainCounter,
                                                  gc->target, ctx->stream));
    uint32_t produced = 0;
    HIP_TRY(hipMemcpyAsync(&produced, trainCounter, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    produced = uint32_t(std::min<size_t>(produced, film->gCap - film->gPending));
    lastSamples = produced;
    g->lastSamples = produced;
    const size_t pending = film->gPending + produced;
    uint32_t k = 0;
    for (; k < gc->train_steps * active && size_t(k + 1) * batch <= pending; ++k) {
        g->step++;
        float* slot = lossSlot(g);
        HIP_TRY(hipMemsetAsync(slot, 0, 4, ctx->stream));
        HIP_TRY(crh::launchGuideTrainStep(g->view, smpBuf + size_t(k) * batch, batch,
                                          g->hsave.as<uint16_t>(), g->dsave.as<uint16_t>(), slot,
                                          gc->lr, g->step, true, ctx->stream));
    }
    // Keep the newest unused samples (at most k batches' worth, so the move
    // never overlaps) at the front for the next pass.
    size_t rem = pending - size_t(k) * batch;
    if (k > 0) {
        rem = std::min(rem, size_t(k) * batch);
        if (rem)
            HIP_TRY(hipMemcpyAsync(smpBuf, smpBuf + (pending - rem),
                                   rem * sizeof(crh::GuideSample), hipMemcpyDeviceToDevice,
                                   ctx->stream));
    }
    film->gPending = rem;
    return CRH_OK;
}

// He-uniform init (limit sqrt(6 / fan-in), fan-in = the layer's inputs) from
// SplitMix64, layer by layer in the parameter order; biases zero.
template <int WD>
void heInitW(std::vector<float>& th, uint32_t seed) {
    using G = crh::GNet<WD>;
    uint64_t x = 0x9E3779B97F4A7C15ull * (uint64_t(seed) + 1);
    auto next = [&]() {
        x += 0x9E3779B97F4A7C15ull;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return double(z >> 11) * (1.0 / 9007199254740992.0);
    };
    th.assign(crh::kGParamsOf<WD>, 0.f);
    for (int l = 0; l < crh::kGLayers; ++l) {
        const double lim = std::sqrt(6.0 / double(G::K(l)));
        const int w0 = G::WOff(l), n = G::K(l) * G::N(l);
        for (int i = 0; i < n; ++i) th[w0 + i] = float((2.0 * next() - 1.0) * lim);
    }
}
void heInit(std::vector<float>& th, uint32_t seed, uint32_t width) {
    if (width == 128) heInitW<128>(th, seed);
    else heInitW<64>(th, seed);
}

int guideUpload(crh_guide* g, const float* theta) {
    crh_context* ctx = g->ctx;
    HIP_TRY(hipMemcpyAsync(g->theta.p, theta, size_t(g->params()) * 4, hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(hipMemsetAsync(g->m.p, 0, g->m.bytes, ctx->stream));
    HIP_TRY(hipMemsetAsync(g->v.p, 0, g->v.bytes, ctx->stream));
    HIP_TRY(hipMemsetAsync(g->grad.p, 0, g->grad.bytes, ctx->stream));
    HIP_TRY(crh::launchGuideQuantize(g->view, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    g->step = 0;
    return CRH_OK;
}

}  // namespace

extern "C" {

int crh_guide_create(crh_context* ctx, const float* bmin, const float* bmax, uint32_t seed,
                     crh_guide** out) {
    return crh_guide_create_width(ctx, bmin, bmax, seed, 64, out);
}

int crh_guide_create_width(crh_context* ctx, const float* bmin, const float* bmax,
                           uint32_t seed, uint32_t width, crh_guide** out) {
    ARG_CHECK(ctx && bmin && bmax && out, "NULL argument");
    ARG_CHECK(width == 64 || width == 128, "guide width must be 64 or 128");
    *out = nullptr;
    GUARD_BEGIN
    HIP_TRY(hipSetDevice(ctx->device));
    auto* g = new crh_guide();
    g->ctx.set(ctx);
    g->width = width;
    const size_t np = g->params(), nq = size_t(crh::gQElemsOf(int(width)));
    hipError_t e = g->theta.alloc(np * 4);
    if (e == hipSuccess) e = g->m.alloc(np * 4);
    if (e == hipSuccess) e = g->v.alloc(np * 4);
    if (e == hipSuccess) e = g->grad.alloc(np * 4);
    if (e == hipSuccess) e = g->wq.alloc(nq * 2);
    if (e == hipSuccess) e = g->wqT.alloc(nq * 2);
    if (e == hipSuccess) e = g->loss.alloc((kLossHist + 1) * sizeof(float));
    if (e != hipSuccess) {
        delete g;
        return hipFail("guide allocation", e, __FILE__, __LINE__);
    }
    crh::GuideView& v = g->view;
    v.theta = g->theta.as<float>();
    v.m = g->m.as<float>();
    v.v = g->v.as<float>();
    v.grad = g->grad.as<float>();
    v.wq = g->wq.as<uint16_t>();
    v.wqT = g->wqT.as<uint16_t>();
    v.width = width;
    for (int k = 0; k < 3; ++k) {
        v.bmin[k] = bmin[k];
        v.bext[k] = bmax[k] - bmin[k];
    }
    std::vector<float> th;
    heInit(th, seed, width);
    int rc = guideUpload(g, th.data());
    if (rc != CRH_OK) {
        delete g;
        return rc;
    }
    *out = g;
    return CRH_OK;
    GUARD_END
}

int crh_guide_set_params(crh_guide* g, const float* theta, uint32_t n) {
    ARG_CHECK(g && theta, "NULL argument");
    ARG_CHECK(n == g->params(), "parameter count must be the guide's (CRH_GUIDE_PARAMS at width 64)");
    HIP_TRY(hipSetDevice(g->ctx->device));
    return guideUpload(g, theta);
}

int crh_guide_get_params(const crh_guide* g, float* theta, uint32_t n) {
    ARG_CHECK(g && theta, "NULL argument");
    ARG_CHECK(n == g->params(), "parameter count must be the guide's (CRH_GUIDE_PARAMS at width 64)");
    HIP_TRY(hipSetDevice(g->ctx->device));
    HIP_TRY(hipMemcpy(theta, g->theta.p, size_t(n) * 4, hipMemcpyDeviceToHost));
    return CRH_OK;
}

int crh_guide_eval(crh_guide* g, const float* pos, const float* nrm, const float* wo,
                   uint32_t n, float* raw, float* mixture) {
    ARG_CHECK(g && pos && nrm && wo, "NULL argument");
    if (n == 0) return CRH_OK;
    GUARD_BEGIN
    crh_context* ctx = g->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<float4> hp(n), hn(n), hw(n);
    for (uint32_t i = 0; i < n; ++i) {
        hp[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.f);
        hn[i] = make_float4(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2], 0.f);
        hw[i] = make_float4(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2], 0.f);
    }
    DevMem dp, dn, dw, dmix, draw;
    HIP_TRY(dp.alloc(size_t(n) * 16));
    HIP_TRY(dn.alloc(size_t(n) * 16));
    HIP_TRY(dw.alloc(size_t(n) * 16));
    HIP_TRY(dmix.alloc(size_t(n) * 40 * 4));
    HIP_TRY(draw.alloc(size_t(n) * crh::kGOut * 4));
    HIP_TRY(hipMemcpy(dp.p, hp.data(), dp.bytes, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dn.p, hn.data(), dn.bytes, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dw.p, hw.data(), dw.bytes, hipMemcpyHostToDevice));
    uint32_t grid = std::min<uint32_t>((n + 63) / 64, 1024u);
    HIP_TRY(crh::launchGuideInfer(g->view, nullptr, nullptr, n, dp.as<float4>(), dn.as<float4>(),
                                  dw.as<float4>(), dmix.as<float>(), draw.as<float>(), grid,
                                  ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (raw) HIP_TRY(hipMemcpy(raw, draw.p, draw.bytes, hipMemcpyDeviceToHost));
    if (mixture) HIP_TRY(hipMemcpy(mixture, dmix.p, dmix.bytes, hipMemcpyDeviceToHost));
    return CRH_OK;
    GUARD_END
}

int crh_guide_train(crh_guide* g, const float* samples, uint32_t n, float lr, int apply_adam,
                    float* loss, float* grad) {
    ARG_CHECK(g && samples, "NULL argument");
    ARG_CHECK(n > 0 && n <= (1u << 24), "sample count must be in [1, 2^24]");
    GUARD_BEGIN
    crh_context* ctx = g->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t padN = (size_t(n) + 63) & ~size_t(63);
    if (padN > g->capSamples || padN * sizeof(crh::GuideSample) > g->samples.bytes) {
        HIP_TRY(g->samples.alloc(padN * sizeof(crh::GuideSample)));
        HIP_TRY(g->hsave.alloc(size_t(crh::kGLayers) * padN * g->width * 2));
        HIP_TRY(g->dsave.alloc(size_t(crh::kGLayers) * padN * g->width * 2));
        g->capSamples = padN;
    }
    HIP_TRY(hipMemcpyAsync(g->samples.p, samples, size_t(n) * sizeof(crh::GuideSample),
                           hipMemcpyHostToDevice, ctx->stream));
    if (apply_adam) g->step++;
    float* slot = apply_adam ? lossSlot(g) : g->loss.as<float>() + kLossHist;
    HIP_TRY(hipMemsetAsync(slot, 0, 4, ctx->stream));
    HIP_TRY(crh::launchGuideTrainStep(g->view, g->samples.as<crh::GuideSample>(), n,
                                      g->hsave.as<uint16_t>(), g->dsave.as<uint16_t>(),
                                      slot, lr, g->step, apply_adam != 0, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (loss) HIP_TRY(hipMemcpy(loss, slot, 4, hipMemcpyDeviceToHost));
    if (!apply_adam) {
        if (grad) HIP_TRY(hipMemcpy(grad, g->grad.p, size_t(g->params()) * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemset(g->grad.p, 0, g->grad.bytes));
    }
    return CRH_OK;
    GUARD_END
}

int crh_guide_get_info(const crh_guide* g, crh_guide_info* info) {
    ARG_CHECK(g && info, "NULL argument");
    GUARD_BEGIN
    std::memset(info, 0, sizeof(*info));
    info->steps = g->step;
    info->last_samples = g->lastSamples;
    info->width = g->width;
    if (g->step) {
        HIP_TRY(hipSetDevice(g->ctx->device));
        HIP_TRY(hipStreamSynchronize(g->ctx->stream));
        HIP_TRY(hipMemcpy(&info->last_loss, g->loss.as<float>() + (g->step % kLossHist), 4,
                          hipMemcpyDeviceToHost));
    }
    return CRH_OK;
    GUARD_END
}

int crh_guide_get_losses(const crh_guide* g, float* losses, uint32_t n) {
    ARG_CHECK(g && (losses || n == 0), "NULL argument");
    ARG_CHECK(n <= std::min(g->step, kLossHist), "n exceeds the recorded steps");
    GUARD_BEGIN
    if (n == 0) return CRH_OK;
    HIP_TRY(hipSetDevice(g->ctx->device));
    HIP_TRY(hipStreamSynchronize(g->ctx->stream));
    std::vector<float> all(kLossHist);
    HIP_TRY(hipMemcpy(all.data(), g->loss.p, kLossHist * sizeof(float), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i) losses[i] = all[(g->step - n + 1 + i) % kLossHist];
    return CRH_OK;
    GUARD_END
}

int crh_guide_destroy(crh_guide* g) {
    if (!g) return CRH_OK;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    delete g;
    return CRH_OK;
}

int crh_render_guided(crh_scene* s, const crh_camera* cam, const crh_integrator_cfg* cfg,
                      crh_guide* g, const crh_guided_cfg* gc, uint32_t first_sample,
                      uint32_t spp, crh_film* film, crh_stats* stats) {
    ARG_CHECK(s && cam && cfg && g && gc && film, "NULL argument");
    ARG_CHECK(film->ctx == s->ctx && g->ctx == s->ctx, "objects belong to different contexts");
    ARG_CHECK(gc->alpha >= 0.f && gc->alpha <= 1.f, "alpha must be in [0, 1]");
    ARG_CHECK(!cfg->spectral, "guided rendering transports RGB");
    ARG_CHECK(gc->materials <= 3u, "materials must be a mask of bits 0 and 1");
    ARG_CHECK(gc->target <= 1u, "target must be 0 (BSDF x radiance) or 1 (radiance)");
    ARG_CHECK(cfg->max_depth <= 1024, "max_depth too large");
    ARG_CHECK(!gc->train || (gc->train_batch >= 64 && gc->train_batch <= (1u << 22)),
              "train_batch must be in [64, 2^22]");
    GUARD_BEGIN
    crh_context* ctx = s->ctx;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    const crh::CameraView cv = makeCameraView(cam, film->w, film->h);
    crh::IntegratorView iv{cfg->max_depth, cfg->rr_threshold, cfg->rr_start_depth,
                           cfg->use_mis ? 1u : 0u};
    double t0 = nowMs();
    HIP_TRY(hipSetDevice(ctx->device));
    const crh::Tuning tun = effective(ctx->tuning, s->view);
    const bool fused = tun.guidedFused != 0u;
    const uint32_t gmask = gc->materials ? gc->materials : 3u;
    crh::WfLaunch wl = wfLaunch(tun);
    // The per-path finish never takes guided vertices away: with guiding at
    // every depth it is off, else it starts no earlier than max_depth.
    wl.finishDepth = (gc->max_depth == 0u || wl.finishDepth == 0u)
                         ? 0u
                         : std::max(wl.finishDepth, gc->max_depth);
    const uint32_t slots = fused ? poolSlots(tun.guidedPool, film, spp) : wfSlots(tun, film, spp);
    const size_t npix = size_t(film->w) * film->h;
    const uint32_t batch = gc->train_batch;
    uint32_t recCap = 0;
    if (fused) {
        HIP_TRY(ensureFilmSlots(film, slots));
        const uint64_t tiles = uint64_t((film->w + 15u) / 16u) * ((film->h + 15u) / 16u);
        const uint32_t grid = uint32_t(std::min<uint64_t>(tiles * slots, tun.guidedGrid));
        // Records for the training the pass group feeds (twice the samples
        // its steps consume, at least 2^20), at most four per path.
        if (gc->train)
            recCap = uint32_t(std::min<uint64_t>(
                {uint64_t(4) * npix * slots,
                 std::max<uint64_t>(uint64_t(2) * gc->train_steps * slots * batch, 1u << 20),
                 0xFFFFFFF0ull}));
        HIP_TRY(ensureGuidedFused(film, crh::guidedScratchFloats(grid), recCap));
        HIP_TRY(hipMemsetAsync(film->work.p, 0, film->work.bytes, ctx->stream));
    } else {
        HIP_TRY(ensureGuided(film, slots));
        HIP_TRY(hipMemsetAsync(film->wf.work, 0, 64, ctx->stream));
        recCap = film->wf.recCap;
    }
    const size_t padB = (size_t(batch) + 63) & ~size_t(63);
    if (gc->train && padB > g->capSamples) {
        HIP_TRY(g->hsave.alloc(size_t(crh::kGLayers) * padB * g->width * 2));
        HIP_TRY(g->dsave.alloc(size_t(crh::kGLayers) * padB * g->width * 2));
        g->capSamples = padB;
    }
    // Sample buffer: one pass group's records on top of a batch still filling.
    const size_t needCap = size_t(recCap) + batch;
    if (gc->train && film->gCap < needCap) {
        HIP_TRY(film->gSamples.alloc(needCap * sizeof(crh::GuideSample)));
        film->gCap = needCap;
        film->gPending = 0;
    }
    std::vector<hipEvent_t>& evs = film->passEv;
    const size_t nev = stats ? 2 * size_t((spp + slots - 1) / slots) : 0;
    while (evs.size() < nev) {
        hipEvent_t e = nullptr;
        HIP_TRY(hipEventCreate(&e));
        evs.push_back(e);
    }
    uint32_t lastSamples = 0;
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    for (uint32_t i = 0; i < spp; i += slots) {
        const uint32_t active = std::min(slots, spp - i);
        const size_t pg = i / slots;
        crh::GuideSample* smpBuf = film->gSamples.as<crh::GuideSample>();
        crh::GuideSample* smpOut = gc->train ? smpBuf + film->gPending : nullptr;
        const uint32_t smpRoom = uint32_t(film->gCap - film->gPending);
        if (stats) HIP_TRY(hipEventRecord(evs[2 * pg], ctx->stream));
        if (fused) {
            uint32_t* cnt = film->gCnt.as<uint32_t>();
            const crh::GuidedPassArgs ga{&g->view, gc->alpha, gmask, gc->max_depth,
                                         film->gScratch.as<float>(),
                                         gc->train ? film->gRec.as<float4>() : nullptr, cnt,
                                         recCap};
            HIP_TRY(crh::launchGuidedPass(s->view, cv, iv, first_sample + i, active,
                                          film->L.as<float4>(), film->P.as<float2>(),
                                          film->work.as<unsigned long long>(), tun, ctx->counter,
                                          ga, ctx->stream));
        } else {
            HIP_TRY(crh::wavefrontPass(s->view, cv, iv, first_sample + i, active, film->wf, wl,
                                       false, ctx->stream, &g->view, gc->alpha, smpOut, smpRoom,
                                       gmask, gc->max_depth, gc->target));
        }
        if (stats) HIP_TRY(hipEventRecord(evs[2 * pg + 1], ctx->stream));
        const float4* Lp = fused ? film->L.as<float4>() : film->wf.rad;
        const float2* Pp = fused ? film->P.as<float2>() : film->wf.P;
        HIP_TRY(crh::launchFilmGather(Lp, Pp, film->w, film->h, film->sum.as<float>(),
                                      film->wsum.as<float>(), ctx->stream, active));
        if (gc->train) {
            const int rc = guidedTrainPass(g, film, gc, active, fused, recCap,
                                           film->L.as<float4>(), smpOut, smpRoom, lastSamples);
            if (rc != CRH_OK) return rc;
        }
    }
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    if (stats) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        double mainMs = 0.0;
        for (size_t pg = 0; 2 * pg < nev; ++pg) {
            float m = 0.f;
            HIP_TRY(hipEventElapsedTime(&m, evs[2 * pg], evs[2 * pg + 1]));
            mainMs += m;
        }
        unsigned long long c[3] = {0, 0, 0};
        HIP_TRY(hipMemcpy(c, fused ? film->work.p : (void*)film->wf.work, sizeof(c),
                          hipMemcpyDeviceToHost));
        stats->kernel_ms = ms;
        stats->main_kernel_ms = mainMs;
        stats->total_ms = nowMs() - t0;
        stats->rays = c[0];
        stats->node_visits = lastSamples;
        stats->launches = (spp + slots - 1) / slots;
    }
    return CRH_OK;
    GUARD_END
}

int crh_render_guided_views(crh_scene* s, const crh_camera* cams, crh_film* const* films,
                            uint32_t nviews, const crh_integrator_cfg* cfg, crh_guide* g,
                            const crh_guided_cfg* gc, uint32_t first_sample, uint32_t spp,
                            crh_stats* stats) {
    ARG_CHECK(s && cams && films && cfg && g && gc, "NULL argument");
    ARG_CHECK(nviews >= 1 && nviews <= 4096, "view count must be in [1, 4096]");
    ARG_CHECK(g->ctx == s->ctx, "objects belong to different contexts");
    ARG_CHECK(gc->alpha >= 0.f && gc->alpha <= 1.f, "alpha must be in [0, 1]");
    ARG_CHECK(!cfg->spectral, "guided rendering transports RGB");
    ARG_CHECK(gc->materials <= 3u, "materials must be a mask of bits 0 and 1");
    ARG_CHECK(gc->target <= 1u, "target must be 0 (BSDF x radiance) or 1 (radiance)");
    ARG_CHECK(cfg->max_depth <= 1024, "max_depth too large");
    ARG_CHECK(!gc->train || (gc->train_batch >= 64 && gc->train_batch <= (1u << 22)),
              "train_batch must be in [64, 2^22]");
    for (uint32_t i = 0; i < nviews; ++i) {
        ARG_CHECK(films[i], "NULL film");
        ARG_CHECK(films[i]->ctx == s->ctx, "film belongs to a different context");
        for (uint32_t j = 0; j < i; ++j) ARG_CHECK(films[j] != films[i], "films must differ");
    }
    GUARD_BEGIN
    crh_context* ctx = s->ctx;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    const crh::IntegratorView iv{cfg->max_depth, cfg->rr_threshold, cfg->rr_start_depth,
                                 cfg->use_mis ? 1u : 0u};
    double t0 = nowMs();
    HIP_TRY(hipSetDevice(ctx->device));
    const crh::Tuning tun = effective(ctx->tuning, s->view);
    const uint32_t gmask = gc->materials ? gc->materials : 3u;
    // The view table: cameras, 16 x 16 tile and pixel prefixes, and the
    // films the gather updates.
    std::vector<crh::CameraView> cv(nviews);
    std::vector<uint32_t> pre(2 * (size_t(nviews) + 1), 0);
    uint32_t* tilePre = pre.data();
    uint32_t* pixPre = pre.data() + nviews + 1;
    std::vector<crh::BatchScene> tab(nviews);
    for (uint32_t i = 0; i < nviews; ++i) {
        crh_film* f = films[i];
        cv[i] = makeCameraView(&cams[i], f->w, f->h);
        const uint64_t tiles = uint64_t((f->w + 15u) / 16u) * ((f->h + 15u) / 16u);
        ARG_CHECK(uint64_t(tilePre[i]) + tiles < (uint64_t(1) << 23) &&
                      uint64_t(pixPre[i]) + uint64_t(f->w) * f->h < (uint64_t(1) << 31),
                  "views too large for one launch");
        tilePre[i + 1] = tilePre[i] + uint32_t(tiles);
        pixPre[i + 1] = pixPre[i] + f->w * f->h;
        std::memset(&tab[i], 0, sizeof(tab[i]));
        tab[i].cam = cv[i];
        tab[i].sum = f->sum.as<float>();
        tab[i].wsum = f->wsum.as<float>();
    }
    const uint32_t tiles = tilePre[nviews], totalPix = pixPre[nviews];
    uint64_t k = std::max<uint64_t>(1, tun.guidedPool / totalPix);
    k = std::min<uint64_t>(k, std::max<uint32_t>(spp, 1u));
    k = std::min<uint64_t>(k, std::max<uint64_t>(1, ((uint64_t(1) << 23) - 1) / tiles));
    k = std::min<uint64_t>(k, std::max<uint64_t>(1, (uint64_t(1) << 31) / totalPix / 4));
    const uint32_t slots = uint32_t(k);
    crh_film* f0 = films[0];
    const size_t np = size_t(totalPix) * slots;
    if (f0->vL.bytes < np * sizeof(float4)) HIP_TRY(f0->vL.alloc(np * sizeof(float4)));
    if (f0->vP.bytes < np * sizeof(float2)) HIP_TRY(f0->vP.alloc(np * sizeof(float2)));
    const size_t camB = nviews * sizeof(crh::CameraView), preB = pre.size() * sizeof(uint32_t),
                 tabB = nviews * sizeof(crh::BatchScene);
    const size_t preOff = (camB + 255) & ~size_t(255), tabOff = (preOff + preB + 255) & ~size_t(255);
    if (f0->vTab.bytes < tabOff + tabB) HIP_TRY(f0->vTab.alloc(tabOff + tabB));
    char* vt = f0->vTab.as<char>();
    HIP_TRY(hipMemcpyAsync(vt, cv.data(), camB, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(vt + preOff, pre.data(), preB, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(vt + tabOff, tab.data(), tabB, hipMemcpyHostToDevice, ctx->stream));
    const auto* dTilePre = reinterpret_cast<const uint32_t*>(vt + preOff);
    const auto* dPixPre = dTilePre + nviews + 1;
    const auto* dTab = reinterpret_cast<const crh::BatchScene*>(vt + tabOff);
    const crh::GuidedViews views{reinterpret_cast<const crh::CameraView*>(vt), dTilePre, dPixPre,
                                 nviews, tiles, totalPix};
    const uint32_t grid = uint32_t(std::min<uint64_t>(uint64_t(tiles) * slots, tun.guidedGrid));
    const uint32_t batch = gc->train_batch;
    uint32_t recCap = 0;
    if (gc->train)
        recCap = uint32_t(std::min<uint64_t>(
            {uint64_t(4) * np, std::max<uint64_t>(uint64_t(2) * gc->train_steps * slots * batch,
                                                   1u << 20),
             0xFFFFFFF0ull}));
    HIP_TRY(ensureGuidedFused(f0, crh::guidedScratchFloats(grid), recCap));
    HIP_TRY(hipMemsetAsync(f0->work.p, 0, f0->work.bytes, ctx->stream));
    const size_t padB = (size_t(batch) + 63) & ~size_t(63);
    if (gc->train && padB > g->capSamples) {
        HIP_TRY(g->hsave.alloc(size_t(crh::kGLayers) * padB * g->width * 2));
        HIP_TRY(g->dsave.alloc(size_t(crh::kGLayers) * padB * g->width * 2));
        g->capSamples = padB;
    }
    const size_t needCap = size_t(recCap) + batch;
    if (gc->train && f0->gCap < needCap) {
        HIP_TRY(f0->gSamples.alloc(needCap * sizeof(crh::GuideSample)));
        f0->gCap = needCap;
        f0->gPending = 0;
    }
    uint32_t lastSamples = 0;
    double mainMs = 0.0;
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    for (uint32_t i = 0; i < spp; i += slots) {
        const uint32_t active = std::min(slots, spp - i);
        crh::GuideSample* smpOut =
            gc->train ? f0->gSamples.as<crh::GuideSample>() + f0->gPending : nullptr;
        const uint32_t smpRoom = uint32_t(f0->gCap - f0->gPending);
        const crh::GuidedPassArgs ga{&g->view, gc->alpha, gmask, gc->max_depth,
                                     f0->gScratch.as<float>(),
                                     gc->train ? f0->gRec.as<float4>() : nullptr,
                                     f0->gCnt.as<uint32_t>(), recCap};
        HIP_TRY(crh::launchGuidedPass(s->view, crh::CameraView{}, iv, first_sample + i, active,
                                      f0->vL.as<float4>(), f0->vP.as<float2>(),
                                      f0->work.as<unsigned long long>(), tun, ctx->counter, ga,
                                      ctx->stream, &views));
        HIP_TRY(crh::launchBatchGatherPass(dTab, dPixPre, nviews, totalPix, active,
                                           f0->vL.as<float4>(), f0->vP.as<float2>(),
                                           ctx->stream));
        if (gc->train) {
            const int rc = guidedTrainPass(g, f0, gc, active, true, recCap, f0->vL.as<float4>(),
                                           smpOut, smpRoom, lastSamples);
            if (rc != CRH_OK) return rc;
        }
    }
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    if (stats) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        unsigned long long c[3] = {0, 0, 0};
        HIP_TRY(hipMemcpy(c, f0->work.p, sizeof(c), hipMemcpyDeviceToHost));
        stats->kernel_ms = ms;
        stats->main_kernel_ms = mainMs > 0.0 ? mainMs : ms;
        stats->total_ms = nowMs() - t0;
        stats->rays = c[0];
        stats->node_visits = lastSamples;
        stats->launches = (spp + slots - 1) / slots;
    }
    return CRH_OK;
    GUARD_END
}

// ---- batched rendering -----------------------------------------------------
struct crh_batch {
    CtxRef ctx;
    DevMem pool;  // every scene's nodes, then triangles, vertices, ... (256 B aligned slices)
    std::vector<crh::SceneView> views;
    uint64_t nodes = 0, tris = 0;
    uint32_t maxStack = 1;
    DevMem table, prefix, next, graphCounter;
    DevMem counter;  // work counters of the direct (non-graph) launches
    // Wavefront mode: path state of up to wfPaths paths (several sample
    // indices of every pixel of the batch per pass).
    DevMem wfArena;
    crh::WfState wf{};
    size_t wfPaths = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> sig;  // what the captured graph was built for
    ~crh_batch() { dropGraph(); }
    void dropGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        sig.clear();
    }
};

int crh_batch_create(crh_context* ctx, crh_scene* const* scenes, uint32_t n, crh_batch** out) {
    ARG_CHECK(ctx && out && (scenes || n == 0), "NULL argument");
    ARG_CHECK(n > 0 && n <= 4096, "scene count must be in [1, 4096]");
    *out = nullptr;
    for (uint32_t i = 0; i < n; ++i) {
        ARG_CHECK(scenes[i], "NULL scene");
        ARG_CHECK(scenes[i]->ctx == ctx, "scene belongs to a different context");
    }
    GUARD_BEGIN
    HIP_TRY(hipSetDevice(ctx->device));
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    // Region r of the pool holds member r of every distinct scene, scene
    // after scene; a scene listed again (several views of one scene) shares
    // the slices of its first listing.
    DevMem crh_scene::*members[] = {
        &crh_scene::nodes,    &crh_scene::tris,    &crh_scene::pos,     &crh_scene::nrm,
        &crh_scene::uv,       &crh_scene::idx,     &crh_scene::primMat, &crh_scene::mats,
        &crh_scene::emissive, &crh_scene::spheres, &crh_scene::sphereMat,
        &crh_scene::texels,   &crh_scene::texInfo, &crh_scene::spec};
    constexpr int kR = 14;
    std::vector<uint32_t> first(n);
    std::unordered_map<const crh_scene*, uint32_t> seen;
    for (uint32_t i = 0; i < n; ++i) first[i] = seen.emplace(scenes[i], i).first->second;
    std::vector<size_t> off(size_t(n) * kR);
    size_t total = 0;
    for (int r = 0; r < kR; ++r)
        for (uint32_t i = 0; i < n; ++i) {
            if (first[i] != i) continue;
            off[size_t(i) * kR + r] = total;
            total += align((scenes[i]->*members[r]).bytes);
        }
    for (uint32_t i = 0; i < n; ++i)
        for (int r = 0; r < kR; ++r) off[size_t(i) * kR + r] = off[size_t(first[i]) * kR + r];
    auto* b = new crh_batch();
    b->ctx.set(ctx);
    hipError_t e = b->pool.alloc(std::max<size_t>(total, 256));
    if (e == hipSuccess) e = b->next.alloc(crh::kBatchMaxChunks * sizeof(uint32_t));
    char* base = b->pool.as<char>();
    for (uint32_t i = 0; i < n && e == hipSuccess; ++i)
        for (int r = 0; r < kR && e == hipSuccess && first[i] == i; ++r) {
            const DevMem& src = scenes[i]->*members[r];
            if (src.bytes)
                e = hipMemcpyAsync(base + off[size_t(i) * kR + r], src.p, src.bytes,
                                   hipMemcpyDeviceToDevice, ctx->stream);
        }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        delete b;
        return hipFail("batch pool", e, __FILE__, __LINE__);
    }
    b->views.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const crh_scene* sc = scenes[i];
        crh::SceneView v = sc->view;
        auto at = [&](int r) { return base + off[size_t(i) * kR + r]; };
        v.nodes = reinterpret_cast<const crh::GNode*>(at(0));
        v.tris = reinterpret_cast<const crh::GTri*>(at(1));
        v.pos = reinterpret_cast<const float*>(at(2));
        v.nrm = reinterpret_cast<const float*>(at(3));
        v.uv = reinterpret_cast<const float*>(at(4));
        v.idx = reinterpret_cast<const uint32_t*>(at(5));
        v.primMat = reinterpret_cast<const uint32_t*>(at(6));
        v.mats = reinterpret_cast<const crh::GMaterial*>(at(7));
        v.emissive = reinterpret_cast<const uint32_t*>(at(8));
        v.spheres = reinterpret_cast<const float*>(at(9));
        v.sphereMat = sc->view.sphereMat ? reinterpret_cast<const uint32_t*>(at(10)) : nullptr;
        v.texels = reinterpret_cast<const float*>(at(11));
        v.texInfo = reinterpret_cast<const uint32_t*>(at(12));
        v.spec = reinterpret_cast<const float*>(at(13));
        b->views[i] = v;
        if (first[i] == i) {
            b->nodes += sc->nodes.bytes / sizeof(crh::GNode);
            b->tris += sc->tris.bytes / sizeof(crh::GTri);
        }
        b->maxStack = std::max(b->maxStack, v.stackCap);
    }
    *out = b;
    return CRH_OK;
    GUARD_END
}

int crh_cie_cmf(float* xyz) {
    ARG_CHECK(xyz, "NULL argument");
    for (int i = 0; i < crh::kCieBins; ++i) crh::cieTable(i, xyz + 3 * i);
    return CRH_OK;
}

int crh_rgb_to_spectrum(const float* rgb, float* spectrum) {
    GUARD_BEGIN
    ARG_CHECK(rgb && spectrum, "NULL argument");
    crh::rgbToSpectrum(rgb, spectrum);
    return CRH_OK;
    GUARD_END
}

int crh_spectrum_to_xyz(const float* spectrum, float* xyz) {
    ARG_CHECK(spectrum && xyz, "NULL argument");
    crh::spectrumToXYZ(spectrum, xyz);
    return CRH_OK;
}

int crh_batch_get_info(const crh_batch* b, crh_batch_info* info) {
    ARG_CHECK(b && info, "NULL argument");
    std::memset(info, 0, sizeof(*info));
    info->scenes = uint32_t(b->views.size());
    info->max_stack = b->maxStack;
    info->pool_bytes = b->pool.bytes;
    info->nodes = b->nodes;
    info->triangles = b->tris;
    return CRH_OK;
}

int crh_batch_render(crh_batch* b, const crh_camera* cams, crh_film* const* films,
                     const uint32_t* spp, const crh_integrator_cfg* cfg, uint32_t first_sample,
                     crh_stats* stats) {
    ARG_CHECK(b && cams && films && spp && cfg, "NULL argument");
    ARG_CHECK(cfg->mode == CRH_MODE_MEGAKERNEL || cfg->mode == CRH_MODE_WAVEFRONT,
              "unknown integrator mode");
    ARG_CHECK(!cfg->count_work, "batched rendering has no counting variant");
    ARG_CHECK(cfg->max_depth <= 1024, "max_depth too large");
    const uint32_t n = uint32_t(b->views.size());
    uint32_t maxSpp = 0;
    for (uint32_t i = 0; i < n; ++i) {
        ARG_CHECK(films[i], "NULL film");
        ARG_CHECK(films[i]->ctx == b->ctx, "film belongs to a different context");
        for (uint32_t j = 0; j < i; ++j) ARG_CHECK(films[j] != films[i], "films must differ");
        ARG_CHECK(spp[i] <= (1u << 20), "spp too large");
        maxSpp = std::max(maxSpp, spp[i]);
    }
    GUARD_BEGIN
    crh_context* ctx = b->ctx;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    double t0 = nowMs();
    HIP_TRY(hipSetDevice(ctx->device));
    crh::Tuning tun = ctx->tuning;
    if (b->maxStack > uint32_t(crh::kPrivStackLevels)) tun.batchPriv = tun.wfPriv = 0;
    const bool wavefront = cfg->mode == CRH_MODE_WAVEFRONT;
    const uint32_t maxSlots = std::max<uint32_t>(1u, tun.batchSlots);
    // Per-sample buffers for up to maxSlots sample indices per launch (a
    // scene takes part in launches of at most its own spp indices).
    for (uint32_t i = 0; i < n && !wavefront; ++i)
        HIP_TRY(ensureFilmSlots(films[i], std::max<uint32_t>(1u, std::min(maxSlots, spp[i]))));
    // Scenes by decreasing spp: those still sampling at pass k form a prefix.
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t c) { return spp[a] > spp[c]; });
    const uint32_t tw = tun.batchTileW, th = crh::kRenderBlock / tw;
    std::vector<crh::BatchScene> tab(n);
    std::vector<uint32_t> pre(2 * (size_t(n) + 1), 0);  // tile prefix, then pixel prefix
    uint32_t* tilePre = pre.data();
    uint32_t* pixPre = pre.data() + n + 1;
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t i = order[j];
        crh_film* f = films[i];
        crh::BatchScene& t = tab[j];
        std::memset(&t, 0, sizeof(t));
        t.view = b->views[i];
        t.cam = makeCameraView(&cams[i], f->w, f->h);
        t.L = f->L.as<float>();
        t.P = f->P.as<float>();
        t.sum = f->sum.as<float>();
        t.wsum = f->wsum.as<float>();
        t.tilesX = (f->w + tw - 1) / tw;
        const uint64_t tiles = uint64_t(t.tilesX) * ((f->h + th - 1) / th);
        ARG_CHECK(uint64_t(tilePre[j]) + tiles < (uint64_t(1) << 32) &&
                      uint64_t(pixPre[j]) + uint64_t(f->w) * f->h < (uint64_t(1) << 32),
                  "batch too large for 32-bit item indices");
        tilePre[j + 1] = tilePre[j] + uint32_t(tiles);
        pixPre[j + 1] = pixPre[j] + f->w * f->h;
    }
    if (b->table.bytes < tab.size() * sizeof(crh::BatchScene))
        HIP_TRY(b->table.alloc(tab.size() * sizeof(crh::BatchScene)));
    if (b->prefix.bytes < pre.size() * sizeof(uint32_t)) {
        HIP_TRY(b->prefix.alloc(pre.size() * sizeof(uint32_t)));
        b->dropGraph();  // the captured graph points at the old buffers
    }
    DevMem& counter = b->counter;
    if (!counter.p) HIP_TRY(counter.alloc(3 * sizeof(unsigned long long)));
    HIP_TRY(hipMemsetAsync(counter.p, 0, counter.bytes, ctx->stream));
    HIP_TRY(hipMemcpyAsync(b->table.p, tab.data(), tab.size() * sizeof(crh::BatchScene),
                           hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(b->prefix.p, pre.data(), pre.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice, ctx->stream));
    const crh::IntegratorView iv{cfg->max_depth, cfg->rr_threshold, cfg->rr_start_depth,
                                 cfg->use_mis ? 1u : 0u};
    const auto* dTab = b->table.as<crh::BatchScene>();
    const uint32_t* dTile = b->prefix.as<uint32_t>();
    const uint32_t* dPix = dTile + n + 1;
    auto* dCount = counter.as<unsigned long long>();
    // Wavefront mode: paths in flight per pass (wf_pool, at least every
    // pixel of the batch once), the same sample index grouping rule.
    const crh::WfLaunch wl = wfLaunch(tun);
    if (wavefront) {
        const size_t allPix = pixPre[n];
        const size_t want = std::max<size_t>(allPix, size_t(tun.wfPool) / allPix * allPix);
        const size_t paths = std::min<size_t>(want, allPix * std::max<uint32_t>(1u, maxSpp));
        if (b->wfPaths < paths) {
            b->dropGraph();  // the captured graph points at the old state
            HIP_TRY(layoutWavefront(b->wfArena, b->wf, paths));
            b->wfPaths = paths;
        }
    }
    // The pass sequence; launched directly or recorded into a graph.  Each
    // launch covers up to maxSlots sample indices (wavefront: as many as its
    // path state holds), as long as the set of scenes still sampling stays
    // the same.
    auto passes = [&](hipStream_t st, unsigned long long* cnt) -> hipError_t {
        uint32_t active = n;
        for (uint32_t k = 0; k < maxSpp;) {
            if (wavefront) {
                while (active > 0 && spp[order[active - 1]] <= k) --active;
                crh::WfState w = b->wf;
                w.npix = pixPre[active];
                w.width = w.height = 0;  // per scene, from the table
                w.slots = uint32_t(std::min<size_t>(b->wfPaths / w.npix, 0xFFFFFFFFu));
                w.work = cnt;
                const uint32_t g = std::min(w.slots, spp[order[active - 1]] - k);
                if (uint64_t(w.npix) * g >= (uint64_t(1) << 32)) return hipErrorInvalidValue;
                hipError_t e = crh::wavefrontBatchPass(dTab, dPix, active, b->maxStack, iv,
                                                       first_sample + k, g, w, wl, st,
                                                       cfg->spectral != 0);
                if (e == hipSuccess)
                    e = crh::launchBatchGatherPass(dTab, dPix, active, w.npix, g, w.rad, w.P, st);
                if (e != hipSuccess) return e;
                k += g;
                continue;
            }
            while (active > 0 && spp[order[active - 1]] <= k) --active;
            const uint32_t g = std::min(maxSlots, spp[order[active - 1]] - k);
            if (uint64_t(tilePre[active]) * g >= (uint64_t(1) << 32))
                return hipErrorInvalidValue;
            hipError_t e = crh::launchBatchPass(dTab, dTile, dPix, active, tilePre[active],
                                                pixPre[active], iv, first_sample + k, g,
                                                b->maxStack, tun, b->next.as<uint32_t>(), cnt,
                                                st, cfg->spectral != 0);
            if (e != hipSuccess) return e;
            k += g;
        }
        return hipSuccess;
    };
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    if (tun.batchGraph) {
        // The graph bakes in pass counts, item counts, launch shapes and the
        // counter address; cameras and films come from the table each call.
        std::vector<uint64_t> sig = {n, first_sample, maxSpp, iv.maxDepth,
                                     uint64_t(__builtin_bit_cast(uint32_t, iv.rrThreshold)),
                                     iv.rrStartDepth, iv.useMIS, tun.batchPriv, tw,
                                     tun.batchGrid, tun.batchChunks, tun.batchWaves, maxSlots,
                                     cfg->spectral != 0 ? 1u : 0u, wavefront ? 1u : 0u,
                                     wl.traceGrid, wl.traceBlock, wl.shadeGrid, wl.privStack,
                                     wl.finishDepth, wl.finishWaves, wl.finishGrid, b->wfPaths};
        for (uint32_t j = 0; j < n; ++j) sig.push_back(spp[order[j]]);
        sig.insert(sig.end(), pre.begin(), pre.end());
        if (!b->exec || sig != b->sig) {
            b->dropGraph();
            // counts of graph replays go to a counter the batch keeps
            if (!b->graphCounter.p) HIP_TRY(b->graphCounter.alloc(3 * sizeof(unsigned long long)));
            HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = passes(ctx->stream, b->graphCounter.as<unsigned long long>());
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
            if (e == hipSuccess) e = e2;
            if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
            b->graph = g;
            if (e != hipSuccess) {
                b->dropGraph();
                return hipFail("batch graph capture", e, __FILE__, __LINE__);
            }
            b->sig = sig;
        }
        HIP_TRY(hipMemsetAsync(b->graphCounter.p, 0, b->graphCounter.bytes, ctx->stream));
        HIP_TRY(hipGraphLaunch(b->exec, ctx->stream));
        HIP_TRY(hipMemcpyAsync(dCount, b->graphCounter.p, b->graphCounter.bytes,
                               hipMemcpyDeviceToDevice, ctx->stream));
    } else {
        HIP_TRY(passes(ctx->stream, dCount));
    }
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    if (stats) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        unsigned long long c[3] = {0, 0, 0};
        HIP_TRY(hipMemcpy(c, dCount, sizeof(c), hipMemcpyDeviceToHost));
        stats->kernel_ms = ms;
        stats->main_kernel_ms = ms;
        stats->total_ms = nowMs() - t0;
        stats->rays = c[0];
        // per launch group: counter reset, megakernel, one gather per index
        uint64_t groups = 0;
        for (uint32_t k = 0, active = n; k < maxSpp; ++groups) {
            while (active > 0 && spp[order[active - 1]] <= k) --active;
            k += std::min(maxSlots, spp[order[active - 1]] - k);
        }
        stats->launches = tun.batchGraph ? 1 : 2 * groups + maxSpp;
    }
    return CRH_OK;
    GUARD_END
}

int crh_batch_destroy(crh_batch* b) {
    if (!b) return CRH_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
    return CRH_OK;
}

}  // extern "C"
