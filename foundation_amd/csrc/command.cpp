// command.cpp — compute command lists, submission and fences.
//
// The compute half of the reference's RHI, which records a frame into an
// RHICommandList in the order Begin, SetPipeline, BindDescriptorSet, draw,
// End, and submits it to a queue with a fence that the host then waits on
// (src/Platform/RHI/Command.hpp:32-116, src/Platform/RHI/Device.hpp:15-37,
// src/Renderer/Renderer.cpp:312-394), but has no dispatch, no compute
// pipeline and no storage-image descriptor.  A list here records on the host
// (the checks of each call happen while recording, as a validation layer
// would report them); fnd_pt_submit runs the recorded commands in order on
// the context's HIP stream:
//  * consecutive path-trace dispatches into different images become one
//    batched fnd_pt_render call (one persistent launch sequence for all);
//  * trace dispatches are traversal launches over the bound buffers;
//  * present and clear dispatches are queued kernels and fills;
// and a fence is a HIP event recorded after the last of them.
#include <algorithm>
#include <cstring>

#include "ctx.hpp"

using namespace fnd;

namespace fnd {

constexpr uint32_t kMaxBindings = 4;
constexpr uint32_t kMaxConstants = 64;

struct Command {
    uint32_t pipeline;
    uint32_t x, y, z;
    uint32_t nset;
    fnd_descriptor set[kMaxBindings];
    uint32_t nconst;
    unsigned char constants[kMaxConstants];
};

// Lists and fences a context made, destroyed with it.
struct CommandRegistry {
    CtxVector<fnd_pt_cmd*> lists;
    CtxVector<fnd_pt_fence*> fences;
    explicit CommandRegistry(const CtxAlloc<char>& a) : lists(a), fences(a) {}
};

}  // namespace fnd

struct fnd_pt_cmd {
    fnd_pt_ctx* ctx;
    fnd::CtxVector<fnd::Command> commands;
    bool recording = false;
    bool ended = false;
    bool refused = false;   // a call since the last begin was refused
    uint32_t pipeline = 0;
    uint32_t nset = 0;
    fnd_descriptor set[fnd::kMaxBindings] = {};
    uint32_t nconst = 0;
    unsigned char constants[fnd::kMaxConstants] = {};
    explicit fnd_pt_cmd(fnd_pt_ctx* c) : ctx(c), commands(c->host_alloc_ref()) {}
};

struct fnd_pt_fence {
    fnd_pt_ctx* ctx;
    hipEvent_t event = nullptr;
    bool submitted = false;
};

namespace fnd {

namespace {

CommandRegistry* registry(fnd_pt_ctx* ctx)
{
    if (!ctx->commands) {
        void* p = host_alloc(ctx->host_alloc_ptr(), sizeof(CommandRegistry), alignof(CommandRegistry));
        if (!p) throw std::bad_alloc();
        ctx->commands = new (p) CommandRegistry(ctx->host_alloc_ref());
    }
    return static_cast<CommandRegistry*>(ctx->commands);
}

template <class T>
void forget(CtxVector<T*>& v, T* x)
{
    auto it = std::find(v.begin(), v.end(), x);
    if (it != v.end()) v.erase(it);
}

void release_list(fnd_pt_cmd* c)
{
    fnd_pt_ctx* ctx = c->ctx;
    c->~fnd_pt_cmd();
    host_free(ctx->host_alloc_ptr(), c);
}

void release_fence(fnd_pt_fence* f)
{
    fnd_pt_ctx* ctx = f->ctx;
    if (f->event) (void)hipEventDestroy(f->event);
    f->~fnd_pt_fence();
    host_free(ctx->host_alloc_ptr(), f);
}

const fnd_descriptor* binding(const fnd_pt_cmd* c, uint32_t b)
{
    for (uint32_t i = 0; i < c->nset; ++i)
        if (c->set[i].binding == b) return &c->set[i];
    return nullptr;
}

fnd_status refuse(fnd_pt_cmd* c, const char* fmt, const char* what, uint32_t a = 0, uint32_t b = 0)
{
    c->refused = true;
    return set_error(c->ctx, FND_E_INVALID_ARG, fmt, what, a, b);
}

// The scene, framebuffer or storage-buffer range a binding names, checked.
Scene* bound_scene(fnd_pt_cmd* c, uint32_t b)
{
    const fnd_descriptor* d = binding(c, b);
    if (!d || d->type != FND_DESCRIPTOR_ACCELERATION_STRUCTURE) return nullptr;
    return find_scene(c->ctx, d->resource);
}

Framebuffer* bound_image(fnd_pt_cmd* c, uint32_t b)
{
    const fnd_descriptor* d = binding(c, b);
    if (!d || d->type != FND_DESCRIPTOR_STORAGE_IMAGE) return nullptr;
    return find_framebuffer(c->ctx, d->resource);
}

// Device address of `need` bytes of a storage-buffer binding, 16-byte
// aligned, or null.
char* bound_buffer(fnd_pt_cmd* c, uint32_t b, uint64_t need)
{
    const fnd_descriptor* d = binding(c, b);
    if (!d || d->type != FND_DESCRIPTOR_STORAGE_BUFFER) return nullptr;
    Buffer* buf = find_buffer(c->ctx, d->resource);
    if (!buf || d->offset > buf->bytes || d->offset % 16) return nullptr;
    const uint64_t avail = d->range ? std::min(d->range, buf->bytes - d->offset) : buf->bytes - d->offset;
    if (d->range && d->range > buf->bytes - d->offset) return nullptr;
    if (need > avail) return nullptr;
    return static_cast<char*>(buf->ptr) + d->offset;
}

// Checks a dispatch against the bound pipeline, bindings and constants.
fnd_status check_dispatch(fnd_pt_cmd* c, uint32_t x, uint32_t y, uint32_t z)
{
    switch (c->pipeline) {
    case FND_PIPELINE_PATH_TRACE: {
        if (!bound_scene(c, 0)) return refuse(c, "%s: binding 0 must be a scene", "path trace");
        Framebuffer* f = bound_image(c, 1);
        if (!f) return refuse(c, "%s: binding 1 must be a framebuffer (storage image)", "path trace");
        if (c->nconst != sizeof(fnd_path_trace_constants))
            return refuse(c, "%s: constants must be fnd_path_trace_constants (%u bytes)", "path trace",
                          uint32_t(sizeof(fnd_path_trace_constants)));
        if (x != f->width || y != f->height || z == 0)
            return refuse(c, "%s: dispatch must be (image width, image height, spp >= 1); the image is %ux%u",
                          "path trace", f->width, f->height);
        fnd_path_trace_constants k;
        std::memcpy(&k, c->constants, sizeof k);
        if (camera_problem(k.camera)) return refuse(c, "%s: degenerate camera", "path trace");
        return FND_OK;
    }
    case FND_PIPELINE_TRACE_CLOSEST:
    case FND_PIPELINE_TRACE_ANYHIT:
        if (!bound_scene(c, 0)) return refuse(c, "%s: binding 0 must be a scene", "trace");
        if (x == 0 || y != 1 || z != 1 || x > kMaxRaysPerTrace)
            return refuse(c, "%s: dispatch must be (rays, 1, 1), 1 <= rays < 2^31", "trace");
        if (!bound_buffer(c, 1, uint64_t(x) * sizeof(Ray)) || !bound_buffer(c, 2, uint64_t(x) * sizeof(Hit)))
            return refuse(c, "%s: bindings 1 (rays) and 2 (hits) must be 16-byte aligned storage buffers holding %u records",
                          "trace", x);
        return FND_OK;
    case FND_PIPELINE_PRESENT: {
        Framebuffer* f = bound_image(c, 0);
        const fnd_descriptor* out = binding(c, 1);
        if (!f) return refuse(c, "%s: binding 0 must be a framebuffer (storage image)", "present");
        if (!out || out->type != FND_DESCRIPTOR_STORAGE_IMAGE || out->resource != FND_INVALID_HANDLE ||
            !out->device_ptr || reinterpret_cast<uintptr_t>(out->device_ptr) % 4 || out->range % 4 ||
            out->range < 4ull * f->width || out->range / 4 > 0xFFFFFFFFull)
            return refuse(c, "%s: binding 1 must be R8G8B8A8 device memory, 4-byte aligned, row pitch >= %u, a multiple of 4",
                          "present", 4 * f->width);
        if (c->nconst != sizeof(float)) return refuse(c, "%s: constants must be one float (exposure)", "present");
        float exposure;
        std::memcpy(&exposure, c->constants, sizeof exposure);
        if (!(exposure >= 0.0f)) return refuse(c, "%s: exposure must be >= 0", "present");
        if (x != f->width || y != f->height || z != 1)
            return refuse(c, "%s: dispatch must be (image width, image height, 1); the image is %ux%u", "present",
                          f->width, f->height);
        return FND_OK;
    }
    case FND_PIPELINE_CLEAR: {
        Framebuffer* f = bound_image(c, 0);
        if (!f) return refuse(c, "%s: binding 0 must be a framebuffer (storage image)", "clear");
        if (x != f->width || y != f->height || z != 1)
            return refuse(c, "%s: dispatch must be (image width, image height, 1); the image is %ux%u", "clear",
                          f->width, f->height);
        return FND_OK;
    }
    default:
        return refuse(c, "%s: no pipeline set", "dispatch");
    }
}

// Runs one list.  Path-trace dispatches are gathered into one batched render
// until another kind of command, or a second dispatch into an image already
// in the batch, ends the batch.
fnd_status run_list(fnd_pt_ctx* ctx, const fnd_pt_cmd* c)
{
    CtxVector<fnd_render_job> batch(ctx->host_alloc_ref());
    auto flush = [&]() -> fnd_status {
        if (batch.empty()) return FND_OK;
        const fnd_status st = fnd_pt_render(ctx, batch.data(), uint32_t(batch.size()));
        batch.clear();
        return st;
    };
    for (const Command& k : c->commands) {
        auto find = [&](uint32_t b) -> const fnd_descriptor* {
            for (uint32_t i = 0; i < k.nset; ++i)
                if (k.set[i].binding == b) return &k.set[i];
            return nullptr;
        };
        if (k.pipeline == FND_PIPELINE_PATH_TRACE) {
            fnd_path_trace_constants pc;
            std::memcpy(&pc, k.constants, sizeof pc);
            fnd_render_job j{};
            j.scene = find(0)->resource;
            j.framebuffer = find(1)->resource;
            j.camera = pc.camera;
            j.spp = k.z;
            j.max_depth = pc.max_depth;
            j.rr_depth = pc.rr_depth;
            j.flags = pc.flags;
            j.seed = pc.seed;
            j.sample_offset = pc.sample_offset;
            for (const fnd_render_job& b : batch)
                if (b.framebuffer == j.framebuffer) {
                    const fnd_status st = flush();
                    if (st != FND_OK) return st;
                    break;
                }
            batch.push_back(j);
            continue;
        }
        fnd_status st = flush();
        if (st != FND_OK) return st;
        if (k.pipeline == FND_PIPELINE_TRACE_CLOSEST || k.pipeline == FND_PIPELINE_TRACE_ANYHIT) {
            Scene* s = find_scene(ctx, find(0)->resource);
            Buffer* rb = find_buffer(ctx, find(1)->resource);
            Buffer* hb = find_buffer(ctx, find(2)->resource);
            if (!s || !rb || !hb) return set_error(ctx, FND_E_NOT_FOUND, "submit: a trace dispatch's resource was destroyed");
            const uint64_t need_r = uint64_t(k.x) * sizeof(Ray), need_h = uint64_t(k.x) * sizeof(Hit);
            if (find(1)->offset + need_r > rb->bytes || find(2)->offset + need_h > hb->bytes)
                return set_error(ctx, FND_E_INVALID_ARG, "submit: a trace dispatch's buffer shrank");
            st = trace_device(ctx, s, reinterpret_cast<const Ray*>(static_cast<char*>(rb->ptr) + find(1)->offset),
                              reinterpret_cast<float4*>(static_cast<char*>(hb->ptr) + find(2)->offset), k.x,
                              (k.pipeline == FND_PIPELINE_TRACE_ANYHIT ? FND_TRACE_ANYHIT : 0u) | FND_ASYNC);
        } else if (k.pipeline == FND_PIPELINE_PRESENT) {
            Framebuffer* f = find_framebuffer(ctx, find(0)->resource);
            if (!f) return set_error(ctx, FND_E_NOT_FOUND, "submit: a present dispatch's framebuffer was destroyed");
            float exposure;
            std::memcpy(&exposure, k.constants, sizeof exposure);
            st = present_async(ctx, f, exposure, static_cast<uint32_t*>(find(1)->device_ptr),
                               uint32_t(find(1)->range / 4));
        } else if (k.pipeline == FND_PIPELINE_CLEAR) {
            Framebuffer* f = find_framebuffer(ctx, find(0)->resource);
            if (!f) return set_error(ctx, FND_E_NOT_FOUND, "submit: a clear dispatch's framebuffer was destroyed");
            st = clear_async(ctx, f);
        }
        if (st != FND_OK) return st;
    }
    return flush();
}

}  // namespace

void commands_free(fnd_pt_ctx* ctx)
{
    auto* r = static_cast<CommandRegistry*>(ctx->commands);
    if (!r) return;
    for (fnd_pt_cmd* c : r->lists) release_list(c);
    for (fnd_pt_fence* f : r->fences) release_fence(f);
    r->~CommandRegistry();
    host_free(ctx->host_alloc_ptr(), r);
    ctx->commands = nullptr;
}

}  // namespace fnd

extern "C" {

fnd_status fnd_pt_cmd_create(fnd_pt_ctx* ctx, fnd_pt_cmd** out)
{
    if (!ctx || !out) return FND_E_INVALID_ARG;
    *out = nullptr;
    return guard(ctx, [&]() -> fnd_status {
        CommandRegistry* r = registry(ctx);
        void* p = host_alloc(ctx->host_alloc_ptr(), sizeof(fnd_pt_cmd), alignof(fnd_pt_cmd));
        if (!p) return set_error(ctx, FND_E_OOM_HOST, "out of host memory");
        auto* c = new (p) fnd_pt_cmd(ctx);
        try {
            r->lists.push_back(c);
        } catch (...) {
            release_list(c);
            throw;
        }
        *out = c;
        return FND_OK;
    });
}

void fnd_pt_cmd_destroy(fnd_pt_cmd* c)
{
    if (!c) return;
    if (auto* r = static_cast<CommandRegistry*>(c->ctx->commands)) forget(r->lists, c);
    release_list(c);
}

fnd_status fnd_pt_cmd_begin(fnd_pt_cmd* c)
{
    if (!c) return FND_E_INVALID_ARG;
    c->commands.clear();
    c->recording = true;
    c->ended = false;
    c->refused = false;
    c->pipeline = 0;
    c->nset = 0;
    c->nconst = 0;
    return FND_OK;
}

fnd_status fnd_pt_cmd_set_pipeline(fnd_pt_cmd* c, uint32_t pipeline)
{
    if (!c) return FND_E_INVALID_ARG;
    if (!c->recording) return refuse(c, "%s: the list is not recording (call fnd_pt_cmd_begin)", "set pipeline");
    if (pipeline < FND_PIPELINE_PATH_TRACE || pipeline > FND_PIPELINE_CLEAR)
        return refuse(c, "%s: unknown pipeline %u", "set pipeline", pipeline);
    c->pipeline = pipeline;
    return FND_OK;
}

fnd_status fnd_pt_cmd_bind_descriptors(fnd_pt_cmd* c, const fnd_descriptor* set, uint32_t count)
{
    if (!c || (!set && count)) return FND_E_INVALID_ARG;
    if (!c->recording) return refuse(c, "%s: the list is not recording (call fnd_pt_cmd_begin)", "bind descriptors");
    if (count > kMaxBindings) return refuse(c, "%s: at most %u bindings", "bind descriptors", kMaxBindings);
    for (uint32_t i = 0; i < count; ++i) {
        if (set[i].type < FND_DESCRIPTOR_ACCELERATION_STRUCTURE || set[i].type > FND_DESCRIPTOR_STORAGE_IMAGE)
            return refuse(c, "%s: binding %u has an unknown type %u", "bind descriptors", set[i].binding, set[i].type);
        for (uint32_t k = 0; k < i; ++k)
            if (set[k].binding == set[i].binding)
                return refuse(c, "%s: binding %u appears twice", "bind descriptors", set[i].binding);
    }
    std::copy(set, set + count, c->set);
    c->nset = count;
    return FND_OK;
}

fnd_status fnd_pt_cmd_push_constants(fnd_pt_cmd* c, const void* data, uint32_t bytes)
{
    if (!c || (!data && bytes)) return FND_E_INVALID_ARG;
    if (!c->recording) return refuse(c, "%s: the list is not recording (call fnd_pt_cmd_begin)", "push constants");
    if (bytes > kMaxConstants) return refuse(c, "%s: at most %u bytes", "push constants", kMaxConstants);
    if (bytes) std::memcpy(c->constants, data, bytes);
    c->nconst = bytes;
    return FND_OK;
}

fnd_status fnd_pt_cmd_dispatch(fnd_pt_cmd* c, uint32_t x, uint32_t y, uint32_t z)
{
    if (!c) return FND_E_INVALID_ARG;
    return guard(c->ctx, [&]() -> fnd_status {
        if (!c->recording) return refuse(c, "%s: the list is not recording (call fnd_pt_cmd_begin)", "dispatch");
        const fnd_status st = check_dispatch(c, x, y, z);
        if (st != FND_OK) return st;
        Command k{};
        k.pipeline = c->pipeline;
        k.x = x;
        k.y = y;
        k.z = z;
        k.nset = c->nset;
        std::copy(c->set, c->set + c->nset, k.set);
        k.nconst = c->nconst;
        std::memcpy(k.constants, c->constants, c->nconst);
        c->commands.push_back(k);
        return FND_OK;
    });
}

fnd_status fnd_pt_cmd_end(fnd_pt_cmd* c)
{
    if (!c) return FND_E_INVALID_ARG;
    if (!c->recording) return refuse(c, "%s: the list is not recording (call fnd_pt_cmd_begin)", "end");
    c->recording = false;
    c->ended = true;
    return c->refused ? set_error(c->ctx, FND_E_INVALID_ARG, "end: a call of this recording was refused") : FND_OK;
}

uint32_t fnd_pt_cmd_size(const fnd_pt_cmd* c) { return c ? uint32_t(c->commands.size()) : 0u; }

fnd_status fnd_pt_fence_create(fnd_pt_ctx* ctx, fnd_pt_fence** out)
{
    if (!ctx || !out) return FND_E_INVALID_ARG;
    *out = nullptr;
    return guard(ctx, [&]() -> fnd_status {
        CommandRegistry* r = registry(ctx);
        void* p = host_alloc(ctx->host_alloc_ptr(), sizeof(fnd_pt_fence), alignof(fnd_pt_fence));
        if (!p) return set_error(ctx, FND_E_OOM_HOST, "out of host memory");
        auto* f = new (p) fnd_pt_fence();
        f->ctx = ctx;
        fnd_status st = check_hip(ctx, hipEventCreateWithFlags(&f->event, hipEventDisableTiming), "hipEventCreate");
        if (st != FND_OK) {
            release_fence(f);
            return st;
        }
        try {
            r->fences.push_back(f);
        } catch (...) {
            release_fence(f);
            throw;
        }
        *out = f;
        return FND_OK;
    });
}

void fnd_pt_fence_destroy(fnd_pt_fence* f)
{
    if (!f) return;
    if (f->submitted) (void)hipEventSynchronize(f->event);
    if (auto* r = static_cast<CommandRegistry*>(f->ctx->commands)) forget(r->fences, f);
    release_fence(f);
}

fnd_status fnd_pt_submit(fnd_pt_ctx* ctx, fnd_pt_cmd* const* cmds, uint32_t count, fnd_pt_fence* fence)
{
    if (!ctx || (!cmds && count)) return FND_E_INVALID_ARG;
    return guard(ctx, [&]() -> fnd_status {
        if (fence && fence->ctx != ctx) return set_error(ctx, FND_E_INVALID_ARG, "submit: the fence belongs to another context");
        for (uint32_t i = 0; i < count; ++i) {
            if (!cmds[i] || cmds[i]->ctx != ctx) return set_error(ctx, FND_E_INVALID_ARG, "submit: list %u is not this context's", i);
            if (!cmds[i]->ended || cmds[i]->refused)
                return set_error(ctx, FND_E_INVALID_ARG, "submit: list %u is not ended or has a refused call", i);
        }
        for (uint32_t i = 0; i < count; ++i) {
            const fnd_status st = run_list(ctx, cmds[i]);
            if (st != FND_OK) return st;
        }
        if (fence) {
            fnd_status st = check_hip(ctx, hipEventRecord(fence->event, ctx->stream), "submit: fence");
            if (st != FND_OK) return st;
            fence->submitted = true;
        }
        return FND_OK;
    });
}

fnd_status fnd_pt_fence_wait(fnd_pt_ctx* ctx, fnd_pt_fence* fence)
{
    if (!ctx || !fence || fence->ctx != ctx) return FND_E_INVALID_ARG;
    return guard(ctx, [&]() -> fnd_status {
        if (!fence->submitted) return FND_OK;
        return check_hip(ctx, hipEventSynchronize(fence->event), "fence wait");
    });
}

}  // extern "C"
