// ctx.hpp — internal state behind the fnd_pt_* C ABI.
#pragma once
#include <new>
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <functional>
#include <map>
#include <stdexcept>
#include <vector>

#include "../../include/fnd_pt.h"
#include "bvh_build.hpp"
#include "scene_encode.hpp"
#include "trace_kernel.hpp"

namespace fnd {

enum HandleKind : uint64_t { kHandleScene = 1, kHandleBuffer = 2, kHandleFramebuffer = 3 };

inline fnd_handle make_handle(HandleKind k, uint64_t id) { return (uint64_t(k) << 56) | id; }
inline HandleKind handle_kind(fnd_handle h) { return HandleKind(h >> 56); }

struct Scene {
    WideNode* nodes = nullptr;
    float4* tris = nullptr;
    InstanceRec* insts = nullptr;
    DeviceMaterial* materials = nullptr;
    DeviceLight* lights = nullptr;
    uint32_t num_lights = 0;
    uint32_t num_materials = 0;
    float background[3] = {0, 0, 0};
    uint32_t root = 0;
    bool two_level = false;
    uint64_t sz_nodes = 0, sz_tris = 0, sz_insts = 0, sz_materials = 0, sz_lights = 0;
    fnd_scene_info info{};
};

struct Buffer {
    void* ptr = nullptr;
    uint64_t bytes = 0;
};

struct Framebuffer {
    uint32_t width = 0, height = 0, flags = 0;
    float* accum = nullptr;       // W*H*3 running sums (FP32)
    double* accum64 = nullptr;    // W*H*3 running sums (FP64, oracle images only)
    uint32_t* rgba8 = nullptr;    // W*H packed R8G8B8A8 (made by the first RGBA8 read)
    uint32_t samples = 0;
};

// Host memory of a context comes from the caller's allocator (malloc/free
// when none was given); every context-owned object is made and released here.
void* host_alloc(const fnd_host_allocator* a, size_t size, size_t align);
void host_free(const fnd_host_allocator* a, void* p);

// Standard-library allocator over the caller's fnd_host_allocator (null:
// malloc), so the context's tables and a call's temporaries take their
// memory where the caller said.  A refused allocation throws std::bad_alloc,
// which the entry point's guard turns into FND_E_OOM_HOST.
template <class T>
struct CtxAlloc {
    using value_type = T;
    const fnd_host_allocator* a = nullptr;
    CtxAlloc() = default;
    explicit CtxAlloc(const fnd_host_allocator* x) : a(x) {}
    template <class U>
    CtxAlloc(const CtxAlloc<U>& o) : a(o.a) {}
    T* allocate(size_t n)
    {
        if (n > size_t(-1) / sizeof(T)) throw std::bad_alloc();
        void* p = host_alloc(a, n * sizeof(T), alignof(T) < 16 ? 16 : alignof(T));
        if (!p) throw std::bad_alloc();
        return static_cast<T*>(p);
    }
    void deallocate(T* p, size_t) { host_free(a, p); }
    template <class U>
    bool operator==(const CtxAlloc<U>& o) const { return a == o.a; }
    template <class U>
    bool operator!=(const CtxAlloc<U>& o) const { return a != o.a; }
};

template <class T>
using CtxVector = std::vector<T, CtxAlloc<T>>;
template <class K, class V>
using CtxMap = std::map<K, V, std::less<K>, CtxAlloc<std::pair<const K, V>>>;

}  // namespace fnd

struct fnd_pt_ctx {
    // `a`: the caller's allocator, copied (null: malloc/free)
    explicit fnd_pt_ctx(const fnd_host_allocator* a)
        : alloc(a ? *a : fnd_host_allocator{}), has_alloc(a != nullptr), timing_events(host_alloc_ref()),
          scenes(host_alloc_ref()), buffers(host_alloc_ref()), framebuffers(host_alloc_ref())
    {
    }
    const fnd_host_allocator* host_alloc_ptr() const { return has_alloc ? &alloc : nullptr; }
    fnd::CtxAlloc<char> host_alloc_ref() const { return fnd::CtxAlloc<char>(host_alloc_ptr()); }

    fnd_host_allocator alloc{};
    bool has_alloc = false;
    fnd_log_fn log = nullptr;
    void* log_user = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;                // radiance-cache inference overlapped with a frame's last bounces
    // Radiance-cache training of a frame runs here, overlapped with the next
    // frame until that frame first reads the cache; train_ev marks its end
    // (train_pending: an event has been recorded since the last host join).
    hipStream_t train = nullptr;
    hipEvent_t train_ev = nullptr;
    bool train_pending = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    fnd::CtxVector<hipEvent_t> timing_events;  // pairs (start, end) per pending trace launch
    size_t timing_used = 0;                    // events in use (even)
    char last_error[512] = {};                 // fixed storage: recording a failure allocates nothing
    uint64_t budget = 0;
    uint64_t used = 0;
    uint32_t num_cus = 0;
    char arch[64] = {};
    fnd_tuning tuning{};
    fnd_pt_stats stats{};
    uint64_t next_id = 1;
    fnd::CtxMap<uint64_t, fnd::Scene*> scenes;
    fnd::CtxMap<uint64_t, fnd::Buffer*> buffers;
    fnd::CtxMap<uint64_t, fnd::Framebuffer*> framebuffers;
    // scratch owned by the context
    float* d_present = nullptr;                // 255 byte thresholds of the RGBA8 transfer
    double* d_relmse = nullptr;                // kRelMseParts partial sums + the result of fnd_pt_framebuffer_rel_mse
    uint32_t* d_counters = nullptr;            // traversal work counter (kCounterWords)
    unsigned long long* d_stats = nullptr;     // 8 counters
    uint2* d_spill = nullptr;
    uint64_t spill_bytes = 0;
    // Ray-query batches in spatial order (ray_order.hip), two slots used in
    // turn: a batch's order is made on the side stream while the previous
    // batch's traversal runs on the main stream.  Per slot: the ray indices
    // and, after them, each ray's cell (uint16); the cell counters
    // (ray_order_scratch_bytes()); the traversal's work counter
    // (kCounterWords); and events for "order made" and "traversal done".
    uint32_t* d_order[2] = {nullptr, nullptr};
    uint64_t order_bytes[2] = {0, 0};
    void* d_order_scratch[2] = {nullptr, nullptr};
    uint32_t* d_order_counter[2] = {nullptr, nullptr};
    hipEvent_t order_ready[2] = {nullptr, nullptr};
    hipEvent_t order_free[2] = {nullptr, nullptr};
    uint32_t order_slot = 0;
    void* d_tmp = nullptr;                     // host-array trace staging
    uint64_t tmp_bytes = 0;
    void* frame_memory = nullptr;              // caller's block for the per-frame queues, or null
    uint64_t frame_bytes = 0;
    void* render_state = nullptr;              // wavefront buffers (render.cpp)
    void* cache = nullptr;                     // radiance cache (cache_api.cpp)
    void* commands = nullptr;                  // command lists and fences made by this context (command.cpp)
    void* scratch = nullptr;                   // this device's fnd_pt_trace_noalloc scratch (noalloc.hip)
};

namespace fnd {

fnd_status set_error(fnd_pt_ctx* ctx, fnd_status st, const char* fmt, ...);
fnd_status check_hip(fnd_pt_ctx* ctx, hipError_t e, const char* what);
fnd_status dev_alloc(fnd_pt_ctx* ctx, void** ptr, uint64_t bytes);
// hipMemcpyAsync on the context's stream, then wait for it.
fnd_status copy_sync(fnd_pt_ctx* ctx, void* dst, const void* src, uint64_t bytes, hipMemcpyKind kind, const char* what);
void dev_free(fnd_pt_ctx* ctx, void* ptr, uint64_t bytes);
fnd_status ensure_spill(fnd_pt_ctx* ctx, uint32_t stack_entries, uint32_t stack_lds, uint32_t threads);
// ray_order.hip: the order in which a ray-query launch takes its rays, by the
// cells of a coarse Morton grid over [lo, hi] that hold their origins.
constexpr uint32_t kOrderMinRays = 131072;  // smaller batches are traced in the caller's order
uint64_t ray_order_scratch_bytes();
fnd_status ensure_order_slot(fnd_pt_ctx* ctx, uint32_t k, uint64_t n);
// The word in the scratch that says whether the order is used (1) or the
// caller's (0): set on the device by launch_ray_order.
const uint32_t* ray_order_flag(const void* scratch);
// It also clears the traversal's work counter (kCounterWords) for the trace
// launch that follows.
hipError_t launch_ray_order(const Ray* rays, uint32_t n, const float* lo, const float* hi, void* scratch,
                            uint32_t* order, uint16_t* cells, uint32_t* trace_counter, uint32_t blocks, hipStream_t s);
Scene* find_scene(fnd_pt_ctx* ctx, fnd_handle h);
Framebuffer* find_framebuffer(fnd_pt_ctx* ctx, fnd_handle h);
Buffer* find_buffer(fnd_pt_ctx* ctx, fnd_handle h);
// One trace launch over device-resident rays and hits (FND_TRACE_* flags;
// with FND_ASYNC it returns once the launch is queued).
// staged: the rays were uploaded on the context's stream within this call.
fnd_status trace_device(fnd_pt_ctx* ctx, Scene* s, const Ray* rays, float4* hits, uint32_t n, uint32_t flags,
                        bool staged = false);
// Queues the R8G8B8A8 conversion of a framebuffer's mean into device memory
// (`pitch` in pixels) on the context's stream; no wait.
fnd_status present_async(fnd_pt_ctx* ctx, Framebuffer* f, float exposure, uint32_t* out, uint32_t pitch);
// Queues the clearing of a framebuffer's sums and zeroes its sample count; no wait.
fnd_status clear_async(fnd_pt_ctx* ctx, Framebuffer* f);
// Why a camera has no basis (null when it has one).
const char* camera_problem(const fnd_camera& c);
// Command lists and fences still alive when the context is destroyed.
void commands_free(fnd_pt_ctx* ctx);

// Fixed device scratch of fnd_pt_trace_noalloc (noalloc.hip): 256 bytes of
// work counter, then the traversal's stack spill area.
// The host waits for the radiance-cache training a render queued (before
// anything outside a frame reads or replaces the cache state).
void finish_training(fnd_pt_ctx* ctx);

constexpr uint64_t kNoAllocScratchBytes = 16ull << 20;
constexpr uint64_t kNoAllocSpillOffset = kCounterWords * sizeof(uint32_t);
// Device address of the scratch array on the current device.
hipError_t noalloc_scratch_address(void** out);
// Zeroes the 32-bit counter with a one-thread kernel on `s` (a kernel, not a
// memset: callers may capture the stream into a graph).
void launch_zero_counter(uint32_t* counter, hipStream_t s);
uint32_t trace_grid(const fnd_pt_ctx* ctx, const TraceLaunch& l);
void resolve_timings(fnd_pt_ctx* ctx);
fnd_status timing_pair(fnd_pt_ctx* ctx, hipEvent_t* start, hipEvent_t* end);
void render_state_free(fnd_pt_ctx* ctx);
void cache_free(fnd_pt_ctx* ctx);

// Runs the body of an entry point.  No exception crosses the C boundary: a
// host allocation failure anywhere below (a container on the caller's
// allocator, a worker thread of a builder) becomes FND_E_OOM_HOST with a
// message, anything else FND_E_DEVICE.
template <class F>
fnd_status guard(fnd_pt_ctx* ctx, F&& body) noexcept
{
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return set_error(ctx, FND_E_OOM_HOST, "out of host memory");
    } catch (const std::length_error&) {
        return set_error(ctx, FND_E_OOM_HOST, "out of host memory (container size limit)");
    } catch (const std::exception& e) {
        return set_error(ctx, FND_E_DEVICE, "internal error: %s", e.what());
    } catch (...) {
        return set_error(ctx, FND_E_DEVICE, "internal error");
    }
}

template <class T>
T* ctx_new(fnd_pt_ctx* ctx)
{
    void* p = host_alloc(ctx->has_alloc ? &ctx->alloc : nullptr, sizeof(T), alignof(T) < 16 ? 16 : alignof(T));
    return p ? new (p) T() : nullptr;
}

template <class T>
void ctx_delete(fnd_pt_ctx* ctx, T* p)
{
    if (!p) return;
    p->~T();
    host_free(ctx->has_alloc ? &ctx->alloc : nullptr, p);
}

}  // namespace fnd
