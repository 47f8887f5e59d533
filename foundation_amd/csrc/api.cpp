// api.cpp — the fnd_pt_* C ABI: contexts, scenes, buffers, ray queries.
//
// Conventions kept from the reference (mos9527/Foundation):
//  * the context is the factory: it owns every scene, buffer and framebuffer
//    it hands out and destroys them with itself
//    (src/Platform/RHI/Details/Details.hpp:39-46,125-141);
//  * handles are uint64_t, ~0 invalid (src/Platform/RHI/Details/Details.hpp:11-12);
//  * host memory of the context comes from the caller's allocator, the C form
//    of the Core::Allocator injected into Renderer
//    (src/Renderer/Renderer.hpp:41, src/Platform/RHI/Vulkan/Application.hpp:13-25);
//  * one context = one queue (here a HIP stream); the host waits on an event
//    where the reference waits on the frame fence (src/Renderer/Renderer.cpp:394).
// Unlike the reference's CHECK -> BugCheck -> exit (src/Core/Core.hpp:17), no
// failure aborts: every entry point returns a status and records a message.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <thread>

#include "ctx.hpp"
#include "wavefront.hpp"

using namespace fnd;

namespace fnd {

fnd_status set_error(fnd_pt_ctx* ctx, fnd_status st, const char* fmt, ...)
{
    if (!ctx) return st;
    char buf[sizeof ctx->last_error];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    std::memcpy(ctx->last_error, buf, sizeof buf);
    if (ctx->log) ctx->log(ctx->log_user, 3, buf);
    return st;
}

fnd_status check_hip(fnd_pt_ctx* ctx, hipError_t e, const char* what)
{
    if (e == hipSuccess) return FND_OK;
    if (e == hipErrorOutOfMemory) return set_error(ctx, FND_E_OOM_DEVICE, "%s: %s", what, hipGetErrorString(e));
    return set_error(ctx, FND_E_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

fnd_status dev_alloc(fnd_pt_ctx* ctx, void** ptr, uint64_t bytes)
{
    *ptr = nullptr;
    if (bytes == 0) bytes = 256;
    if (ctx->budget && ctx->used + bytes > ctx->budget)
        return set_error(ctx, FND_E_OOM_DEVICE, "device memory budget exceeded (%llu + %llu > %llu bytes)",
                         (unsigned long long)ctx->used, (unsigned long long)bytes, (unsigned long long)ctx->budget);
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        (void)hipGetLastError();
        return check_hip(ctx, e, "hipMalloc");
    }
    ctx->used += bytes;
    return FND_OK;
}

void dev_free(fnd_pt_ctx* ctx, void* ptr, uint64_t bytes)
{
    if (!ptr) return;
    (void)hipFree(ptr);
    if (bytes == 0) bytes = 256;
    ctx->used -= std::min(ctx->used, bytes);
}

fnd_status ensure_spill(fnd_pt_ctx* ctx, uint32_t stack_entries, uint32_t stack_lds, uint32_t threads)
{
    const uint64_t per = stack_entries > stack_lds ? stack_entries - stack_lds : 0;
    const uint64_t need = std::max<uint64_t>(per * threads * sizeof(uint2), 256);
    if (need <= ctx->spill_bytes) return FND_OK;
    // an FND_ASYNC trace still in flight may use the old area
    (void)hipStreamSynchronize(ctx->stream);
    dev_free(ctx, ctx->d_spill, ctx->spill_bytes);
    ctx->d_spill = nullptr;
    ctx->spill_bytes = 0;
    void* p = nullptr;
    fnd_status st = dev_alloc(ctx, &p, need);
    if (st != FND_OK) return st;
    ctx->d_spill = static_cast<uint2*>(p);
    ctx->spill_bytes = need;
    return FND_OK;
}

// Buffers and events of ray-order slot k for a batch of n rays.
fnd_status ensure_order_slot(fnd_pt_ctx* ctx, uint32_t k, uint64_t n)
{
    fnd_status st = FND_OK;
    if (!ctx->order_ready[k]) {
        st = check_hip(ctx, hipEventCreateWithFlags(&ctx->order_ready[k], hipEventDisableTiming), "hipEventCreate");
        if (st == FND_OK)
            st = check_hip(ctx, hipEventCreateWithFlags(&ctx->order_free[k], hipEventDisableTiming), "hipEventCreate");
        if (st != FND_OK) return st;
    }
    if (!ctx->d_order_scratch[k]) {
        void* p = nullptr;
        st = dev_alloc(ctx, &p, ray_order_scratch_bytes());
        if (st != FND_OK) return st;
        ctx->d_order_scratch[k] = p;
        // the counters start at zero; each batch's scan zeroes them again
        st = check_hip(ctx, hipMemsetAsync(p, 0, ray_order_scratch_bytes(), ctx->stream), "ray order");
        if (st != FND_OK) return st;
        st = dev_alloc(ctx, &p, kCounterWords * sizeof(uint32_t));
        if (st != FND_OK) return st;
        ctx->d_order_counter[k] = static_cast<uint32_t*>(p);
        // the side stream uses the scratch next
        st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "ray order");
        if (st != FND_OK) return st;
    }
    const uint64_t need = n * (sizeof(uint32_t) + sizeof(uint16_t));
    if (need <= ctx->order_bytes[k]) return FND_OK;
    // a traversal still in flight may read the old order
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->side);
    dev_free(ctx, ctx->d_order[k], ctx->order_bytes[k]);
    ctx->d_order[k] = nullptr;
    ctx->order_bytes[k] = 0;
    void* p = nullptr;
    st = dev_alloc(ctx, &p, need);
    if (st != FND_OK) return st;
    ctx->d_order[k] = static_cast<uint32_t*>(p);
    ctx->order_bytes[k] = need;
    return FND_OK;
}

Scene* find_scene(fnd_pt_ctx* ctx, fnd_handle h)
{
    if (handle_kind(h) != kHandleScene) return nullptr;
    auto it = ctx->scenes.find(h);
    return it == ctx->scenes.end() ? nullptr : it->second;
}

Framebuffer* find_framebuffer(fnd_pt_ctx* ctx, fnd_handle h)
{
    if (handle_kind(h) != kHandleFramebuffer) return nullptr;
    auto it = ctx->framebuffers.find(h);
    return it == ctx->framebuffers.end() ? nullptr : it->second;
}

// Resolves the pending (start, end) event pairs of trace launches into the
// cumulative kernel time.  Call only after the stream has drained.
void resolve_timings(fnd_pt_ctx* ctx)
{
    for (size_t i = 0; i + 1 < ctx->timing_used; i += 2) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, ctx->timing_events[i], ctx->timing_events[i + 1]) == hipSuccess) {
            ctx->stats.trace_kernel_ms_total += ms;
            ctx->stats.trace_launches += 1;
            ctx->stats.last_kernel_ms = ms;
        }
    }
    ctx->timing_used = 0;
}

// A (start, end) pair for the next launch; drains the stream when the ring is full.
fnd_status timing_pair(fnd_pt_ctx* ctx, hipEvent_t* start, hipEvent_t* end)
{
    constexpr size_t kMaxPending = 2048;
    if (ctx->timing_used >= 2 * kMaxPending) {
        fnd_status st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
        if (st != FND_OK) return st;
        resolve_timings(ctx);
    }
    while (ctx->timing_events.size() < ctx->timing_used + 2) {
        hipEvent_t e;
        fnd_status st = check_hip(ctx, hipEventCreate(&e), "hipEventCreate");
        if (st != FND_OK) return st;
        ctx->timing_events.push_back(e);
    }
    *start = ctx->timing_events[ctx->timing_used];
    *end = ctx->timing_events[ctx->timing_used + 1];
    ctx->timing_used += 2;
    return FND_OK;
}

// Persistent grid: waves_per_cu waves per CU, never more blocks than fit on
// the chip at once (a persistent wave that cannot start until others finish
// only adds spill memory and a late start).
uint32_t trace_grid(const fnd_pt_ctx* ctx, const TraceLaunch& l)
{
    const uint32_t waves = std::max(1u, ctx->num_cus * ctx->tuning.waves_per_cu);
    uint32_t grid = std::max(1u, waves * 64u / l.block);
    const uint32_t per_cu = trace_blocks_per_cu(l);
    if (per_cu) grid = std::min(grid, per_cu * ctx->num_cus);
    return grid;
}

void* host_alloc(const fnd_host_allocator* a, size_t size, size_t align)
{
    if (a && a->alloc) return a->alloc(a->user, size, align);
    return std::aligned_alloc(align, (size + align - 1) / align * align);
}

void host_free(const fnd_host_allocator* a, void* p)
{
    if (!p) return;
    if (a && a->free)
        a->free(a->user, p);
    else
        std::free(p);
}

// Copies on the context stream and waits: the stream is non-blocking, so a
// plain hipMemcpy on the null stream would not order against its kernels.
fnd_status copy_sync(fnd_pt_ctx* ctx, void* dst, const void* src, uint64_t bytes, hipMemcpyKind kind, const char* what)
{
    if (!bytes) return FND_OK;
    fnd_status st = check_hip(ctx, hipMemcpyAsync(dst, src, bytes, kind, ctx->stream), what);
    if (st != FND_OK) return st;
    return check_hip(ctx, hipStreamSynchronize(ctx->stream), what);
}

}  // namespace fnd

namespace {


void default_tuning(fnd_tuning& t)
{
    t.block_size = 128;
    t.stack_lds = 12;
    t.waves_per_cu = 24;
    // refill 16 with the sharded work counter: incoherent rays at 10.08M
    // triangles 1875-1975 Mrays/s over refills 8-24 (16 best), against
    // 1760-1795 at 48 with one counter address
    t.refill = 16;
    t.queue_capacity = 1u << 21;
    t.tile = 8;
}

void free_scene(fnd_pt_ctx* ctx, Scene* s)
{
    dev_free(ctx, s->nodes, s->sz_nodes);
    dev_free(ctx, s->tris, s->sz_tris);
    dev_free(ctx, s->insts, s->sz_insts);
    dev_free(ctx, s->materials, s->sz_materials);
    dev_free(ctx, s->lights, s->sz_lights);
    ctx_delete(ctx, s);
}

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

uint32_t fnd_pt_abi_version(void) { return FND_PT_ABI_VERSION; }

const char* fnd_pt_status_string(fnd_status s)
{
    switch (s) {
    case FND_OK: return "ok";
    case FND_E_INVALID_ARG: return "invalid argument";
    case FND_E_OOM_HOST: return "out of host memory";
    case FND_E_OOM_DEVICE: return "out of device memory";
    case FND_E_DEVICE: return "device error";
    case FND_E_QUEUE_OVERFLOW: return "queue overflow";
    case FND_E_UNSUPPORTED: return "unsupported";
    case FND_E_NOT_FOUND: return "not found";
    }
    return "unknown status";
}

fnd_status fnd_pt_create(const fnd_pt_create_desc* desc, const fnd_host_allocator* allocator, fnd_pt_ctx** out)
{
    if (!out) return FND_E_INVALID_ARG;
    *out = nullptr;
    if (allocator && (!allocator->alloc || !allocator->free)) return FND_E_INVALID_ARG;
    void* mem = host_alloc(allocator, sizeof(fnd_pt_ctx), 64);
    if (!mem) return FND_E_OOM_HOST;
    fnd_pt_ctx* ctx = new (mem) fnd_pt_ctx(allocator);
    if (desc) {
        ctx->device = desc->device;
        ctx->budget = desc->device_memory_budget;
        ctx->log = desc->log;
        ctx->log_user = desc->log_user;
    }
    default_tuning(ctx->tuning);
    auto fail = [&](fnd_status st) {
        // keep the message for the caller: there is no context to ask
        fprintf(stderr, "fnd_pt_create: %s\n", ctx->last_error);
        fnd_pt_destroy(ctx);
        return st;
    };
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        set_error(ctx, FND_E_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
        return fail(FND_E_DEVICE);
    }
    if (ctx->device < 0 || ctx->device >= ndev) {
        set_error(ctx, FND_E_INVALID_ARG, "device %d out of range (%d devices)", ctx->device, ndev);
        return fail(FND_E_INVALID_ARG);
    }
    fnd_status st = check_hip(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    if (st != FND_OK) return fail(st);
    hipDeviceProp_t prop;
    st = check_hip(ctx, hipGetDeviceProperties(&prop, ctx->device), "hipGetDeviceProperties");
    if (st != FND_OK) return fail(st);
    // the allocation-free trace's scratch array, resolved once with this device current
    st = check_hip(ctx, noalloc_scratch_address(&ctx->scratch), "hipGetSymbolAddress(scratch)");
    if (st != FND_OK) return fail(st);
    ctx->num_cus = uint32_t(prop.multiProcessorCount);
    snprintf(ctx->arch, sizeof ctx->arch, "%s", prop.gcnArchName);
    ctx->stats.num_cus = ctx->num_cus;
    st = check_hip(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
    if (st != FND_OK) return fail(st);
    st = check_hip(ctx, hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking), "hipStreamCreate");
    if (st != FND_OK) return fail(st);
    st = check_hip(ctx, hipStreamCreateWithFlags(&ctx->train, hipStreamNonBlocking), "hipStreamCreate");
    if (st != FND_OK) return fail(st);
    st = check_hip(ctx, hipEventCreateWithFlags(&ctx->train_ev, hipEventDisableTiming), "hipEventCreate");
    if (st != FND_OK) return fail(st);
    for (auto& ev : ctx->ev) {
        st = check_hip(ctx, hipEventCreate(&ev), "hipEventCreate");
        if (st != FND_OK) return fail(st);
    }
    void* p = nullptr;
    st = dev_alloc(ctx, &p, kCounterWords * sizeof(uint32_t));
    if (st != FND_OK) return fail(st);
    ctx->d_counters = static_cast<uint32_t*>(p);
    st = dev_alloc(ctx, &p, 8 * sizeof(unsigned long long));
    if (st != FND_OK) return fail(st);
    ctx->d_stats = static_cast<unsigned long long*>(p);
    *out = ctx;
    return FND_OK;
}

void fnd_pt_destroy(fnd_pt_ctx* ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->scenes) free_scene(ctx, kv.second);
    ctx->scenes.clear();
    for (auto& kv : ctx->buffers) {
        dev_free(ctx, kv.second->ptr, kv.second->bytes);
        ctx_delete(ctx, kv.second);
    }
    ctx->buffers.clear();
    for (auto& kv : ctx->framebuffers) {
        Framebuffer* f = kv.second;
        dev_free(ctx, f->accum, 0);
        dev_free(ctx, f->accum64, 0);
        dev_free(ctx, f->rgba8, 0);
        ctx_delete(ctx, f);
    }
    ctx->framebuffers.clear();
    commands_free(ctx);
    render_state_free(ctx);
    cache_free(ctx);
    dev_free(ctx, ctx->d_counters, kCounterWords * sizeof(uint32_t));
    dev_free(ctx, ctx->d_present, 255 * sizeof(float));
    dev_free(ctx, ctx->d_relmse, (kRelMseParts + 1) * sizeof(double));
    dev_free(ctx, ctx->d_stats, 8 * sizeof(unsigned long long));
    dev_free(ctx, ctx->d_spill, ctx->spill_bytes);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    for (int k = 0; k < 2; ++k) {
        dev_free(ctx, ctx->d_order[k], ctx->order_bytes[k]);
        dev_free(ctx, ctx->d_order_scratch[k], ray_order_scratch_bytes());
        dev_free(ctx, ctx->d_order_counter[k], kCounterWords * sizeof(uint32_t));
        if (ctx->order_ready[k]) (void)hipEventDestroy(ctx->order_ready[k]);
        if (ctx->order_free[k]) (void)hipEventDestroy(ctx->order_free[k]);
    }
    dev_free(ctx, ctx->d_tmp, ctx->tmp_bytes);
    for (auto& ev : ctx->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : ctx->timing_events) (void)hipEventDestroy(ev);
    ctx->timing_events.clear();
    if (ctx->side) {
        (void)hipStreamSynchronize(ctx->side);
        (void)hipStreamDestroy(ctx->side);
    }
    if (ctx->train) {
        (void)hipStreamSynchronize(ctx->train);
        (void)hipStreamDestroy(ctx->train);
    }
    if (ctx->train_ev) (void)hipEventDestroy(ctx->train_ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    const fnd_host_allocator a = ctx->alloc;
    const bool has = ctx->has_alloc;
    ctx->~fnd_pt_ctx();
    host_free(has ? &a : nullptr, ctx);
}

const char* fnd_pt_last_error(const fnd_pt_ctx* ctx)
{
    if (!ctx) return "null context";
    return ctx->last_error;
}

fnd_status fnd_pt_wait(fnd_pt_ctx* ctx)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx) return FND_E_INVALID_ARG;
        fnd_status st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        if (st == FND_OK) resolve_timings(ctx);
        finish_training(ctx);
        return st;
    });
}

fnd_status fnd_pt_get_stats(fnd_pt_ctx* ctx, fnd_pt_stats* out)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || !out) return FND_E_INVALID_ARG;
        ctx->stats.device_bytes = ctx->used;
        *out = ctx->stats;
        return FND_OK;
    });
}

fnd_status fnd_pt_reset_stats(fnd_pt_ctx* ctx)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx) return FND_E_INVALID_ARG;
        const uint32_t cus = ctx->stats.num_cus;
        std::memset(&ctx->stats, 0, sizeof(ctx->stats));
        ctx->stats.num_cus = cus;
        return FND_OK;
    });
}

fnd_status fnd_pt_set_tuning(fnd_pt_ctx* ctx, const fnd_tuning* t)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || !t) return FND_E_INVALID_ARG;
        fnd_tuning n = ctx->tuning;
        fnd_tuning d;
        default_tuning(d);
        n.block_size = t->block_size ? t->block_size : d.block_size;
        n.stack_lds = t->stack_lds ? t->stack_lds : d.stack_lds;
        n.waves_per_cu = t->waves_per_cu ? t->waves_per_cu : d.waves_per_cu;
        n.refill = t->refill ? t->refill : d.refill;
        n.queue_capacity = t->queue_capacity ? t->queue_capacity : d.queue_capacity;
        n.tile = t->tile ? t->tile : d.tile;
        if (!trace_config_supported(n.block_size, n.stack_lds))
            return set_error(ctx, FND_E_UNSUPPORTED, "no traversal kernel for block %u / stack %u", n.block_size,
                             n.stack_lds);
        if (n.refill > 64 || n.waves_per_cu > 64)
            return set_error(ctx, FND_E_INVALID_ARG, "refill must be 1..64 and waves_per_cu 1..64");
        if (n.queue_capacity < 1024 || n.queue_capacity > (1u << 26))
            return set_error(ctx, FND_E_INVALID_ARG, "queue_capacity must be in [1024, 2^26]");
        if (n.tile > 256) return set_error(ctx, FND_E_INVALID_ARG, "tile must be <= 256");
        ctx->tuning = n;
        return FND_OK;
    });
}

fnd_status fnd_pt_get_tuning(fnd_pt_ctx* ctx, fnd_tuning* out)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || !out) return FND_E_INVALID_ARG;
        *out = ctx->tuning;
        return FND_OK;
    });
}

// ---------------------------------------------------------------- scenes

fnd_status fnd_pt_scene_create(fnd_pt_ctx* ctx, const fnd_scene_desc* d, fnd_handle* out)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || !d || !out) return ctx ? set_error(ctx, FND_E_INVALID_ARG, "null argument") : FND_E_INVALID_ARG;
        *out = FND_INVALID_HANDLE;
        const double t0 = now_ms();
        HostScene hs;
        std::string err;
        fnd_status st = encode_scene(d, hs, err);
        if (st != FND_OK) return set_error(ctx, st, "%s", err.c_str());
        Scene* s = ctx_new<Scene>(ctx);
        if (!s) return set_error(ctx, FND_E_OOM_HOST, "out of host memory");
        auto bail = [&](fnd_status e) {
            free_scene(ctx, s);
            return e;
        };
        s->sz_nodes = hs.nodes.size() * sizeof(WideNode);
        s->sz_tris = hs.tris.size() * sizeof(LeafTri);
        s->sz_insts = std::max<size_t>(1, hs.insts.size()) * sizeof(InstanceRec);
        s->sz_materials = hs.materials.size() * sizeof(DeviceMaterial);
        s->sz_lights = hs.lights.size() * sizeof(DeviceLight);
        void* p = nullptr;
        st = dev_alloc(ctx, &p, s->sz_nodes);
        if (st != FND_OK) return bail(st);
        s->nodes = static_cast<WideNode*>(p);
        st = dev_alloc(ctx, &p, s->sz_tris);
        if (st != FND_OK) return bail(st);
        s->tris = static_cast<float4*>(p);
        st = dev_alloc(ctx, &p, s->sz_insts);
        if (st != FND_OK) return bail(st);
        s->insts = static_cast<InstanceRec*>(p);
        st = dev_alloc(ctx, &p, s->sz_materials);
        if (st != FND_OK) return bail(st);
        s->materials = static_cast<DeviceMaterial*>(p);
        st = dev_alloc(ctx, &p, s->sz_lights);
        if (st != FND_OK) return bail(st);
        s->lights = static_cast<DeviceLight*>(p);
        st = copy_sync(ctx, s->nodes, hs.nodes.data(), s->sz_nodes, hipMemcpyHostToDevice, "upload nodes");
        if (st == FND_OK) st = copy_sync(ctx, s->tris, hs.tris.data(), s->sz_tris, hipMemcpyHostToDevice, "upload triangles");
        if (st == FND_OK && !hs.insts.empty())
            st = copy_sync(ctx, s->insts, hs.insts.data(), hs.insts.size() * sizeof(InstanceRec), hipMemcpyHostToDevice,
                           "upload instances");
        if (st == FND_OK)
            st = copy_sync(ctx, s->materials, hs.materials.data(), s->sz_materials, hipMemcpyHostToDevice, "upload materials");
        if (st == FND_OK) st = copy_sync(ctx, s->lights, hs.lights.data(), s->sz_lights, hipMemcpyHostToDevice, "upload lights");
        if (st != FND_OK) return bail(st);
        s->num_lights = hs.num_lights;
        s->num_materials = uint32_t(hs.materials.size());
        std::memcpy(s->background, hs.background, 12);
        s->two_level = hs.two_level;
        s->root = hs.root;
        s->info = hs.info;
        s->info.build_ms = now_ms() - t0;
        const fnd_handle h = make_handle(kHandleScene, ctx->next_id++);
        try {
            ctx->scenes[h] = s;
        } catch (...) {
            free_scene(ctx, s);
            throw;
        }
        *out = h;
        if (ctx->log) {
            char msg[256];
            snprintf(msg, sizeof msg, "scene: %llu tris, %llu nodes, depth %u, %.1f ms",
                     (unsigned long long)s->info.num_triangles, (unsigned long long)s->info.num_nodes, s->info.max_depth,
                     s->info.build_ms);
            ctx->log(ctx->log_user, 1, msg);
        }
        return FND_OK;
    });
}

fnd_status fnd_pt_scene_destroy(fnd_pt_ctx* ctx, fnd_handle h)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx) return FND_E_INVALID_ARG;
        Scene* s = find_scene(ctx, h);
        if (!s) return set_error(ctx, FND_E_NOT_FOUND, "unknown scene handle");
        (void)hipStreamSynchronize(ctx->stream);
        free_scene(ctx, s);
        ctx->scenes.erase(h);
        return FND_OK;
    });
}

fnd_status fnd_pt_scene_info(fnd_pt_ctx* ctx, fnd_handle h, fnd_scene_info* out)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || !out) return FND_E_INVALID_ARG;
        Scene* s = find_scene(ctx, h);
        if (!s) return set_error(ctx, FND_E_NOT_FOUND, "unknown scene handle");
        *out = s->info;
        return FND_OK;
    });
}

fnd_status fnd_pt_scene_export(fnd_pt_ctx* ctx, fnd_handle h, void* nodes, uint64_t* nodes_bytes, void* tris,
                               uint64_t* tris_bytes)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || !nodes_bytes || !tris_bytes) return FND_E_INVALID_ARG;
        Scene* s = find_scene(ctx, h);
        if (!s) return set_error(ctx, FND_E_NOT_FOUND, "unknown scene handle");
        const uint64_t nb = s->info.num_nodes * sizeof(WideNode), tb = s->info.num_triangles * sizeof(LeafTri);
        if (nodes) {
            if (*nodes_bytes < nb) return set_error(ctx, FND_E_INVALID_ARG, "node buffer too small");
            fnd_status st = copy_sync(ctx, nodes, s->nodes, nb, hipMemcpyDeviceToHost, "export nodes");
            if (st != FND_OK) return st;
        }
        if (tris) {
            if (*tris_bytes < tb) return set_error(ctx, FND_E_INVALID_ARG, "triangle buffer too small");
            fnd_status st = copy_sync(ctx, tris, s->tris, tb, hipMemcpyDeviceToHost, "export triangles");
            if (st != FND_OK) return st;
        }
        *nodes_bytes = nb;
        *tris_bytes = tb;
        return FND_OK;
    });
}

// ---------------------------------------------------------------- buffers

fnd_status fnd_pt_buffer_create(fnd_pt_ctx* ctx, uint64_t bytes, fnd_handle* out)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || !out) return FND_E_INVALID_ARG;
        *out = FND_INVALID_HANDLE;
        if (bytes == 0) return set_error(ctx, FND_E_INVALID_ARG, "zero-sized buffer");
        Buffer* b = ctx_new<Buffer>(ctx);
        if (!b) return set_error(ctx, FND_E_OOM_HOST, "out of host memory");
        fnd_status st = dev_alloc(ctx, &b->ptr, bytes);
        if (st != FND_OK) {
            ctx_delete(ctx, b);
            return st;
        }
        b->bytes = bytes;
        const fnd_handle h = make_handle(kHandleBuffer, ctx->next_id++);
        try {
            ctx->buffers[h] = b;
        } catch (...) {
            dev_free(ctx, b->ptr, b->bytes);
            ctx_delete(ctx, b);
            throw;
        }
        *out = h;
        return FND_OK;
    });
}

fnd_status fnd_pt_buffer_destroy(fnd_pt_ctx* ctx, fnd_handle h)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx) return FND_E_INVALID_ARG;
        auto it = ctx->buffers.find(h);
        if (handle_kind(h) != kHandleBuffer || it == ctx->buffers.end())
            return set_error(ctx, FND_E_NOT_FOUND, "unknown buffer handle");
        (void)hipStreamSynchronize(ctx->stream);
        dev_free(ctx, it->second->ptr, it->second->bytes);
        ctx_delete(ctx, it->second);
        ctx->buffers.erase(it);
        return FND_OK;
    });
}

}  // extern "C"

Buffer* fnd::find_buffer(fnd_pt_ctx* ctx, fnd_handle h)
{
    if (handle_kind(h) != kHandleBuffer) return nullptr;
    auto it = ctx->buffers.find(h);
    return it == ctx->buffers.end() ? nullptr : it->second;
}

extern "C" {

fnd_status fnd_pt_buffer_write(fnd_pt_ctx* ctx, fnd_handle h, uint64_t offset, const void* src, uint64_t bytes)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || (!src && bytes)) return FND_E_INVALID_ARG;
        Buffer* b = find_buffer(ctx, h);
        if (!b) return set_error(ctx, FND_E_NOT_FOUND, "unknown buffer handle");
        if (offset > b->bytes || bytes > b->bytes - offset) return set_error(ctx, FND_E_INVALID_ARG, "write out of range");
        if (!bytes) return FND_OK;
        return copy_sync(ctx, static_cast<char*>(b->ptr) + offset, src, bytes, hipMemcpyHostToDevice, "buffer write");
    });
}

fnd_status fnd_pt_buffer_read(fnd_pt_ctx* ctx, fnd_handle h, uint64_t offset, void* dst, uint64_t bytes)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || (!dst && bytes)) return FND_E_INVALID_ARG;
        Buffer* b = find_buffer(ctx, h);
        if (!b) return set_error(ctx, FND_E_NOT_FOUND, "unknown buffer handle");
        if (offset > b->bytes || bytes > b->bytes - offset) return set_error(ctx, FND_E_INVALID_ARG, "read out of range");
        if (!bytes) return FND_OK;
        return copy_sync(ctx, dst, static_cast<const char*>(b->ptr) + offset, bytes, hipMemcpyDeviceToHost, "buffer read");
    });
}

// ---------------------------------------------------------------- ray queries

}  // extern "C"

fnd_status fnd::trace_device(fnd_pt_ctx* ctx, Scene* s, const Ray* rays, float4* hits, uint32_t n, uint32_t flags,
                             bool staged)
{
    const bool any = (flags & FND_TRACE_ANYHIT) != 0;
    const bool count = (flags & FND_TRACE_COUNT) != 0;
    // ray indices are formed as grid threads + a 32-bit work counter
    if (n > kMaxRaysPerTrace) return set_error(ctx, FND_E_INVALID_ARG, "at most %u rays per call", kMaxRaysPerTrace);
    TraceLaunch l{};
    l.mode = any ? kModeAny : kModeClosest;
    l.two_level = s->two_level;
    l.multi = false;
    l.count = count;
    l.block = count ? 128u : ctx->tuning.block_size;
    l.stack_lds = count ? 16u : ctx->tuning.stack_lds;
    l.grid = trace_grid(ctx, l);
    fnd_status st = ensure_spill(ctx, s->info.stack_entries, l.stack_lds, l.grid * l.block);
    if (st != FND_OK) return st;
    TraceArgs a{};
    a.scene = SceneDev{s->nodes, s->tris, s->insts, s->root, s->two_level ? 1u : 0u};
    a.rays = rays;
    a.hits = hits;
    a.nrays = n;
    a.counter = ctx->d_counters;
    a.spill = ctx->d_spill;
    a.spill_stride = l.grid * l.block;
    a.refill = std::max(1u, std::min(64u, ctx->tuning.refill));
    a.stats = ctx->d_stats;
    // counting launches describe the walks in the caller's order
    const bool order = !count && n >= kOrderMinRays;
    const uint32_t slot = ctx->order_slot;
    if (order) {
        st = ensure_order_slot(ctx, slot, n);
        if (st != FND_OK) return st;
        a.ray_order = ctx->d_order[slot];
        a.ray_order_use = ray_order_flag(ctx->d_order_scratch[slot]);
        a.counter = ctx->d_order_counter[slot];
        ctx->order_slot ^= 1u;
    }
    if (!order) {  // the ordering launches clear the work counter themselves
        st = check_hip(ctx, hipMemsetAsync(ctx->d_counters, 0, kCounterWords * sizeof(uint32_t), ctx->stream),
                       "reset counter");
        if (st != FND_OK) return st;
    }
    if (count) {
        st = check_hip(ctx, hipMemsetAsync(ctx->d_stats, 0, 8 * sizeof(unsigned long long), ctx->stream), "reset stats");
        if (st != FND_OK) return st;
    }
    hipEvent_t e0, e1;
    st = timing_pair(ctx, &e0, &e1);
    if (st != FND_OK) return st;
    if (order) {
        // The order is made on the side stream, once the traversal that last
        // used this slot is done, so it overlaps the previous batch's traversal
        // (whose drain leaves most CUs idle).  Two blocks per CU: same-box runs
        // of the 10M-triangle incoherent set gave 0.622-0.626, 0.609 and 0.618
        // ms per call for 1, 2 and 4.
        st = check_hip(ctx, hipStreamWaitEvent(ctx->side, ctx->order_free[slot], 0), "ray order");
        if (st == FND_OK && staged) {
            // the rays were uploaded on the main stream just now: the side
            // stream reads them after the upload
            st = check_hip(ctx, hipEventRecord(ctx->order_ready[slot], ctx->stream), "ray order");
            if (st == FND_OK) st = check_hip(ctx, hipStreamWaitEvent(ctx->side, ctx->order_ready[slot], 0), "ray order");
        }
        if (st == FND_OK)
            st = check_hip(ctx, launch_ray_order(rays, n, s->info.bounds_lo, s->info.bounds_hi, ctx->d_order_scratch[slot],
                                                 ctx->d_order[slot], reinterpret_cast<uint16_t*>(ctx->d_order[slot] + n),
                                                 ctx->d_order_counter[slot], 2u * ctx->num_cus, ctx->side),
                           "ray order");
        if (st == FND_OK) st = check_hip(ctx, hipEventRecord(ctx->order_ready[slot], ctx->side), "ray order");
        if (st == FND_OK) st = check_hip(ctx, hipStreamWaitEvent(ctx->stream, ctx->order_ready[slot], 0), "ray order");
        if (st != FND_OK) return st;
    }
    (void)hipEventRecord(e0, ctx->stream);
    st = check_hip(ctx, launch_trace(l, a, ctx->stream), "trace launch");
    if (st != FND_OK) return st;
    (void)hipEventRecord(e1, ctx->stream);
    if (order) {
        st = check_hip(ctx, hipEventRecord(ctx->order_free[slot], ctx->stream), "ray order");
        if (st != FND_OK) return st;
    }
    ctx->stats.last_grid = l.grid;
    ctx->stats.rays_traced += n;
    if (flags & FND_ASYNC) return FND_OK;
    st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "trace");
    if (st != FND_OK) return st;
    resolve_timings(ctx);
    ctx->stats.last_call_ms = ctx->stats.last_kernel_ms;
    if (count) {
        unsigned long long c[8];
        st = copy_sync(ctx, c, ctx->d_stats, sizeof c, hipMemcpyDeviceToHost, "read counters");
        if (st != FND_OK) return st;
        ctx->stats.node_visits += c[0];
        ctx->stats.tri_tests += c[1];
        ctx->stats.instance_visits += c[2];
    }
    return FND_OK;
}

extern "C" {

fnd_status fnd_pt_trace_buffers(fnd_pt_ctx* ctx, fnd_handle scene, fnd_handle rays, fnd_handle hits, uint32_t n,
                                uint32_t flags)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx) return FND_E_INVALID_ARG;
        Scene* s = find_scene(ctx, scene);
        if (!s) return set_error(ctx, FND_E_NOT_FOUND, "unknown scene handle");
        Buffer* rb = find_buffer(ctx, rays);
        Buffer* hb = find_buffer(ctx, hits);
        if (!rb || !hb) return set_error(ctx, FND_E_NOT_FOUND, "unknown buffer handle");
        if (flags & ~(FND_TRACE_ANYHIT | FND_TRACE_COUNT | FND_ASYNC))
            return set_error(ctx, FND_E_INVALID_ARG, "unknown trace flags 0x%x", flags);
        if (uint64_t(n) * sizeof(Ray) > rb->bytes || uint64_t(n) * sizeof(Hit) > hb->bytes)
            return set_error(ctx, FND_E_INVALID_ARG, "buffers too small for %u rays", n);
        if (n == 0) return FND_OK;
        return trace_device(ctx, s, static_cast<const Ray*>(rb->ptr), static_cast<float4*>(hb->ptr), n, flags);
    });
}

fnd_status fnd_pt_trace_noalloc(fnd_pt_ctx* ctx, fnd_handle scene, const fnd_ray* rays, fnd_hit* hits, uint32_t n,
                                uint32_t flags)
{
    if (!ctx) return FND_E_INVALID_ARG;
    return guard(ctx, [&]() -> fnd_status {
        Scene* s = find_scene(ctx, scene);
        if (!s) return set_error(ctx, FND_E_NOT_FOUND, "unknown scene handle");
        if (flags & ~(FND_TRACE_ANYHIT | FND_ASYNC))
            return set_error(ctx, FND_E_INVALID_ARG, "unknown trace flags 0x%x (no FND_TRACE_COUNT here)", flags);
        if (n == 0) return FND_OK;
        if (!rays || !hits || reinterpret_cast<uintptr_t>(rays) % 16 || reinterpret_cast<uintptr_t>(hits) % 16)
            return set_error(ctx, FND_E_INVALID_ARG, "rays and hits must be 16-byte aligned device arrays");
        if (n > kMaxRaysPerTrace) return set_error(ctx, FND_E_INVALID_ARG, "at most %u rays per call", kMaxRaysPerTrace);
        if (!ctx->scratch) return set_error(ctx, FND_E_DEVICE, "the scratch array has no device address");
        TraceLaunch l{};
        l.mode = (flags & FND_TRACE_ANYHIT) ? kModeAny : kModeClosest;
        l.two_level = s->two_level;
        l.multi = false;
        l.count = false;
        l.block = ctx->tuning.block_size;
        l.stack_lds = ctx->tuning.stack_lds;
        l.grid = trace_grid(ctx, l);
        // the spill area holds (stack entries - LDS entries) 8-byte entries per
        // thread of the grid: fewer persistent blocks when the full grid's do
        // not fit (each thread still finishes every ray it takes)
        const uint64_t per_thread =
            uint64_t(s->info.stack_entries > l.stack_lds ? s->info.stack_entries - l.stack_lds : 0) * sizeof(uint2);
        if (per_thread) {
            const uint64_t fit = (kNoAllocScratchBytes - kNoAllocSpillOffset) / per_thread / l.block;
            if (fit == 0)
                return set_error(ctx, FND_E_INVALID_ARG,
                                 "the scene's traversal stack (%u entries) needs %llu spill bytes per block; the fixed "
                                 "scratch holds %llu",
                                 s->info.stack_entries, (unsigned long long)(per_thread * l.block),
                                 (unsigned long long)(kNoAllocScratchBytes - kNoAllocSpillOffset));
            l.grid = uint32_t(std::min<uint64_t>(l.grid, fit));
        }
        TraceArgs a{};
        a.scene = SceneDev{s->nodes, s->tris, s->insts, s->root, s->two_level ? 1u : 0u};
        a.rays = reinterpret_cast<const Ray*>(rays);
        a.hits = reinterpret_cast<float4*>(hits);
        a.nrays = n;
        a.counter = static_cast<uint32_t*>(ctx->scratch);
        a.spill = reinterpret_cast<uint2*>(static_cast<char*>(ctx->scratch) + kNoAllocSpillOffset);
        a.spill_stride = l.grid * l.block;
        a.refill = std::max(1u, std::min(64u, ctx->tuning.refill));
        launch_zero_counter(a.counter, ctx->stream);
        fnd_status st = check_hip(ctx, launch_trace(l, a, ctx->stream), "trace launch");
        if (st != FND_OK) return st;
        ctx->stats.last_grid = l.grid;
        ctx->stats.rays_traced += n;
        if (flags & FND_ASYNC) return FND_OK;
        return check_hip(ctx, hipStreamSynchronize(ctx->stream), "trace");
    });
}

fnd_status fnd_pt_trace(fnd_pt_ctx* ctx, fnd_handle scene, const fnd_ray* rays, uint32_t n, fnd_hit* hits, uint32_t flags)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx) return FND_E_INVALID_ARG;
        if (n && (!rays || !hits)) return set_error(ctx, FND_E_INVALID_ARG, "null ray or hit array");
        Scene* s = find_scene(ctx, scene);
        if (!s) return set_error(ctx, FND_E_NOT_FOUND, "unknown scene handle");
        if (flags & ~(FND_TRACE_ANYHIT | FND_TRACE_COUNT))
            return set_error(ctx, FND_E_INVALID_ARG, "unknown trace flags 0x%x (FND_ASYNC needs device buffers)", flags);
        if (n == 0) return FND_OK;
        const uint64_t need = uint64_t(n) * (sizeof(Ray) + sizeof(Hit));
        if (need > ctx->tmp_bytes) {
            dev_free(ctx, ctx->d_tmp, ctx->tmp_bytes);
            ctx->d_tmp = nullptr;
            ctx->tmp_bytes = 0;
            fnd_status st = dev_alloc(ctx, &ctx->d_tmp, need);
            if (st != FND_OK) return st;
            ctx->tmp_bytes = need;
        }
        Ray* dr = static_cast<Ray*>(ctx->d_tmp);
        float4* dh = reinterpret_cast<float4*>(dr + n);
        fnd_status st = check_hip(ctx, hipMemcpyAsync(dr, rays, uint64_t(n) * sizeof(Ray), hipMemcpyHostToDevice, ctx->stream),
                                  "upload rays");
        if (st != FND_OK) return st;
        st = trace_device(ctx, s, dr, dh, n, flags, true);
        if (st != FND_OK) return st;
        st = check_hip(ctx, hipMemcpyAsync(hits, dh, uint64_t(n) * sizeof(Hit), hipMemcpyDeviceToHost, ctx->stream),
                       "download hits");
        if (st != FND_OK) return st;
        return check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    });
}

}  // extern "C"
