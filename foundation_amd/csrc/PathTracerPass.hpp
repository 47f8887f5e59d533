// PathTracerPass.hpp — the C++ face of the path-tracing pass, over the
// extern "C" fnd_pt ABI (include/fnd_pt.h).  Header-only, no HIP types.
//
// Shaped after the reference's Renderer so it can stand in its place:
//  * constructed from a device choice and a borrowed host allocator, as
//    Renderer(RHIApplicationObjectHandle<RHIDevice>, Core::Allocator*)
//    (src/Renderer/Renderer.hpp:41, src/Renderer/Renderer.cpp:34-37); any
//    type with the Allocate(size, alignment) / Reallocate(ptr, size,
//    alignment) / Deallocate(ptr) members of Core::Allocator
//    (src/Core/Allocator/Allocator.hpp:27-31) plugs in;
//  * Record(cmd) records one frame's compute commands into a command list
//    (ComputeCommandList, the compute half of the reference's RHICommandList),
//    and Execute() records, submits and waits on the frame's fence, in place of
//    Record + Submit + the fence wait in Renderer::Draw
//    (src/Renderer/Renderer.cpp:312-394);
//  * the destructor waits for the device to go idle first, as
//    Renderer::~Renderer does (src/Renderer/Renderer.cpp:402-406);
//  * a failed call is fatal, as CHECK -> BugCheck -> std::terminate is
//    (src/Core/Core.hpp:17, src/Core/Core.cpp:7-10), unless the caller installs
//    a handler that throws instead.
//
// Single-threaded, like the reference's one render loop
// (src/Editor/Editor.cpp:20-23).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/fnd_pt.h"

namespace fnd {

// What happens on a non-OK status.  The default prints the status and the
// context's last error and terminates.
using FailureHandler = void (*)(fnd_status status, const char* call, const char* detail);

inline void terminate_on_failure(fnd_status status, const char* call, const char* detail)
{
    std::fprintf(stderr, "fnd_pt: %s failed: %s (%d): %s\n", call, fnd_pt_status_string(status), int(status),
                 detail ? detail : "");
    std::terminate();
}

inline void throw_on_failure(fnd_status status, const char* call, const char* detail)
{
    throw std::runtime_error(std::string(call) + ": " + fnd_pt_status_string(status) + ": " + (detail ? detail : ""));
}

inline FailureHandler& failure_handler()
{
    static FailureHandler handler = terminate_on_failure;
    return handler;
}

// C callbacks over a Core::Allocator-shaped object.
template <class Alloc>
fnd_host_allocator make_host_allocator(Alloc* a)
{
    fnd_host_allocator h{};
    h.user = a;
    h.alloc = [](void* u, size_t size, size_t align) -> void* { return static_cast<Alloc*>(u)->Allocate(size, align); };
    h.realloc = [](void* u, void* p, size_t size, size_t align) -> void* {
        return static_cast<Alloc*>(u)->Reallocate(p, size, align);
    };
    h.free = [](void* u, void* p) { static_cast<Alloc*>(u)->Deallocate(p); };
    return h;
}

// Core::StackAllocator's interface (src/Core/Allocator/StackAllocator.hpp:9-63)
// over fnd_arena: bump allocation from one block, no frees, Reset per frame.
// The addresses are computed only, so the block may be device memory.
class StackAllocator {
public:
    StackAllocator() { fnd_arena_reset(&arena_, nullptr, 0); }
    StackAllocator(void* memory, uint64_t size) { Reset(memory, size); }

    void Reset(void* memory, uint64_t size) { fnd_arena_reset(&arena_, memory, size); }
    void Reset() { fnd_arena_reset(&arena_, nullptr, 0); }
    // nullptr for size 0, without memory, or when the block would pass the end
    void* Allocate(size_t size) { return fnd_arena_alloc(&arena_, size); }
    void* Allocate(size_t size, size_t alignment) { return fnd_arena_alloc_aligned(&arena_, size, alignment); }
    void Deallocate(void*) {}
    void Deallocate(void*, size_t) {}
    void* Reallocate(void*, size_t, size_t) { throw std::runtime_error("StackAllocator does not support reallocation"); }
    uint64_t GetUsedMemory() const noexcept { return fnd_arena_used(&arena_); }
    explicit operator bool() const noexcept { return arena_.memory != 0; }

private:
    fnd_arena arena_{};
};

// A compute command list in the chaining style of the reference's
// RHICommandList (src/Platform/RHI/Command.hpp:32-116), over fnd_pt_cmd_*:
// Begin, SetPipeline, BindDescriptorSet, PushConstants, Dispatch, End.  The
// list belongs to the context that made it (the reference's command pool);
// failures go to failure_handler(), as every call of the pass does.
class ComputeCommandList {
public:
    explicit ComputeCommandList(fnd_pt_ctx* ctx) : ctx_(ctx) { check(fnd_pt_cmd_create(ctx, &cmd_), "fnd_pt_cmd_create"); }
    ComputeCommandList(const ComputeCommandList&) = delete;
    ComputeCommandList& operator=(const ComputeCommandList&) = delete;
    ~ComputeCommandList() { fnd_pt_cmd_destroy(cmd_); }

    ComputeCommandList& Begin() { return check(fnd_pt_cmd_begin(cmd_), "fnd_pt_cmd_begin"); }
    ComputeCommandList& SetPipeline(fnd_pipeline pipeline)
    {
        return check(fnd_pt_cmd_set_pipeline(cmd_, pipeline), "fnd_pt_cmd_set_pipeline");
    }
    ComputeCommandList& BindDescriptorSet(const fnd_descriptor* set, uint32_t count)
    {
        return check(fnd_pt_cmd_bind_descriptors(cmd_, set, count), "fnd_pt_cmd_bind_descriptors");
    }
    template <class T>
    ComputeCommandList& PushConstants(const T& constants)
    {
        return check(fnd_pt_cmd_push_constants(cmd_, &constants, uint32_t(sizeof(T))), "fnd_pt_cmd_push_constants");
    }
    ComputeCommandList& Dispatch(uint32_t x, uint32_t y = 1, uint32_t z = 1)
    {
        return check(fnd_pt_cmd_dispatch(cmd_, x, y, z), "fnd_pt_cmd_dispatch");
    }
    void End() { check(fnd_pt_cmd_end(cmd_), "fnd_pt_cmd_end"); }
    uint32_t Size() const { return fnd_pt_cmd_size(cmd_); }
    fnd_pt_cmd* Handle() const { return cmd_; }

    static fnd_descriptor Scene(uint32_t binding, fnd_handle scene)
    {
        fnd_descriptor d{};
        d.binding = binding;
        d.type = FND_DESCRIPTOR_ACCELERATION_STRUCTURE;
        d.resource = scene;
        return d;
    }
    static fnd_descriptor Image(uint32_t binding, fnd_handle framebuffer)
    {
        fnd_descriptor d{};
        d.binding = binding;
        d.type = FND_DESCRIPTOR_STORAGE_IMAGE;
        d.resource = framebuffer;
        return d;
    }
    static fnd_descriptor DeviceImage(uint32_t binding, void* device_rgba, uint64_t row_pitch)
    {
        fnd_descriptor d{};
        d.binding = binding;
        d.type = FND_DESCRIPTOR_STORAGE_IMAGE;
        d.resource = FND_INVALID_HANDLE;
        d.range = row_pitch;
        d.device_ptr = device_rgba;
        return d;
    }
    static fnd_descriptor Buffer(uint32_t binding, fnd_handle buffer, uint64_t offset = 0, uint64_t range = 0)
    {
        fnd_descriptor d{};
        d.binding = binding;
        d.type = FND_DESCRIPTOR_STORAGE_BUFFER;
        d.resource = buffer;
        d.offset = offset;
        d.range = range;
        return d;
    }

private:
    ComputeCommandList& check(fnd_status st, const char* call)
    {
        if (st != FND_OK) failure_handler()(st, call, fnd_pt_last_error(ctx_));
        return *this;
    }
    fnd_pt_ctx* ctx_;
    fnd_pt_cmd* cmd_ = nullptr;
};

class PathTracerPass {
public:
    struct Desc {
        int device = 0;
        uint64_t device_memory_budget = 0;  // 0 = unlimited
        uint32_t width = 1920;              // the reference's swapchain size (src/Renderer/Renderer.cpp:40-41)
        uint32_t height = 1080;
        uint32_t spp = 1;                   // samples added per Execute()
        uint32_t max_depth = 8;
        uint32_t rr_depth = 3;
        uint64_t seed = 1;
        bool use_cache = false;             // terminate paths into the radiance cache
        bool train_cache = false;           // train the cache on this pass's own paths every frame
        bool fp64 = false;                  // also accumulate in double (reference images)
    };

    explicit PathTracerPass(const Desc& desc) : desc_(desc) { init(nullptr); }

    template <class Alloc>
    PathTracerPass(const Desc& desc, Alloc* allocator) : desc_(desc), host_alloc_(make_host_allocator(allocator))
    {
        init(&host_alloc_);
    }

    PathTracerPass(const PathTracerPass&) = delete;
    PathTracerPass& operator=(const PathTracerPass&) = delete;

    ~PathTracerPass()
    {
        if (!ctx_) return;
        (void)fnd_pt_wait(ctx_);
        cmd_.reset();
        fnd_pt_destroy(ctx_);  // also releases scenes, buffers, framebuffers and the fence
    }

    // ---- scene and view
    fnd_handle CreateScene(const fnd_scene_desc& desc)
    {
        fnd_handle h = FND_INVALID_HANDLE;
        check(fnd_pt_scene_create(ctx_, &desc, &h), "fnd_pt_scene_create");
        return h;
    }
    void DestroyScene(fnd_handle scene)
    {
        if (scene == scene_) scene_ = FND_INVALID_HANDLE;
        check(fnd_pt_scene_destroy(ctx_, scene), "fnd_pt_scene_destroy");
    }
    fnd_scene_info SceneInfo(fnd_handle scene)
    {
        fnd_scene_info info{};
        check(fnd_pt_scene_info(ctx_, scene, &info), "fnd_pt_scene_info");
        return info;
    }
    // Select what Execute() renders.  A new scene or camera restarts the
    // accumulation.
    void SetView(fnd_handle scene, const fnd_camera& camera)
    {
        scene_ = scene;
        camera_ = camera;
        Reset();
    }

    void ConfigureCache(const fnd_cache_desc* desc)
    {
        fnd_cache_desc d{};
        if (!desc) {
            check(fnd_pt_cache_default_desc(&d), "fnd_pt_cache_default_desc");
            desc = &d;
        }
        check(fnd_pt_cache_configure(ctx_, desc), "fnd_pt_cache_configure");
    }

    // ---- one frame
    // The pass's commands of one frame: a path-trace dispatch that adds
    // desc.spp samples per pixel to the pass's image, then, when a
    // presentation target is set, the R8G8B8A8 conversion into it.  The
    // reference records its raster pass the same way (Renderer::Record,
    // src/Renderer/Renderer.cpp:312-366).
    void Record(ComputeCommandList& cmd)
    {
        if (scene_ == FND_INVALID_HANDLE) check(FND_E_INVALID_ARG, "PathTracerPass::Record (no scene set)");
        fnd_path_trace_constants k{};
        k.camera = camera_;
        k.max_depth = desc_.max_depth;
        k.rr_depth = desc_.rr_depth;
        k.flags = (desc_.use_cache ? FND_JOB_USE_CACHE : 0u) | (desc_.train_cache ? FND_JOB_TRAIN_CACHE : 0u);
        k.sample_offset = samples_ + recorded_;
        k.seed = desc_.seed;
        const fnd_descriptor set[2] = {ComputeCommandList::Scene(0, scene_), ComputeCommandList::Image(1, fb_)};
        cmd.SetPipeline(FND_PIPELINE_PATH_TRACE).BindDescriptorSet(set, 2).PushConstants(k).Dispatch(desc_.width,
                                                                                                      desc_.height,
                                                                                                      desc_.spp);
        recorded_ += desc_.spp;
        if (present_) {
            const fnd_descriptor out[2] = {ComputeCommandList::Image(0, fb_),
                                           ComputeCommandList::DeviceImage(1, present_, present_pitch_)};
            cmd.SetPipeline(FND_PIPELINE_PRESENT).BindDescriptorSet(out, 2).PushConstants(exposure_).Dispatch(
                desc_.width, desc_.height, 1);
        }
    }

    // Records, submits and waits for one frame, in place of Record + Submit +
    // the fence wait of Renderer::Draw (src/Renderer/Renderer.cpp:382-394).
    void Execute()
    {
        cmd_->Begin();
        recorded_ = 0;
        Record(*cmd_);
        cmd_->End();
        fnd_pt_cmd* list = cmd_->Handle();
        check(fnd_pt_submit(ctx_, &list, 1, fence_), "fnd_pt_submit");
        check(fnd_pt_fence_wait(ctx_, fence_), "fnd_pt_fence_wait");
        samples_ += recorded_;
        recorded_ = 0;
    }

    // Every Execute() also converts the frame to R8G8B8A8 into this device
    // memory (the swapchain image's memory imported into HIP, for instance),
    // rows row_pitch bytes apart; nullptr stops it.
    void SetPresentTarget(void* device_rgba, uint64_t row_pitch, float exposure = 1.0f)
    {
        present_ = device_rgba;
        present_pitch_ = row_pitch;
        exposure_ = exposure;
    }

    void Reset()
    {
        check(fnd_pt_framebuffer_clear(ctx_, fb_), "fnd_pt_framebuffer_clear");
        samples_ = 0;
    }

    // ---- output
    // Mean radiance, width * height * 3 floats, row 0 at the top.
    void ReadRGB(std::vector<float>& rgb)
    {
        rgb.resize(size_t(desc_.width) * desc_.height * 3);
        uint32_t n = 0;
        check(fnd_pt_framebuffer_read(ctx_, fb_, rgb.data(), &n), "fnd_pt_framebuffer_read");
    }
    void ReadRGB64(std::vector<double>& rgb)
    {
        rgb.resize(size_t(desc_.width) * desc_.height * 3);
        uint32_t n = 0;
        check(fnd_pt_framebuffer_read_f64(ctx_, fb_, rgb.data(), &n), "fnd_pt_framebuffer_read_f64");
    }
    // R8G8B8A8_UNORM, the reference's presentation format
    // (src/Renderer/Renderer.cpp:39-44), converted on the device:
    // round(255 * clamp(x, 0, 1)^(1/2.2)), alpha 255 (fnd_pt_framebuffer_read_rgba8).
    void ReadRGBA8(std::vector<uint8_t>& rgba, float exposure = 1.0f)
    {
        rgba.resize(size_t(desc_.width) * desc_.height * 4);
        uint32_t n = 0;
        check(fnd_pt_framebuffer_read_rgba8(ctx_, fb_, exposure, rgba.data(), &n), "fnd_pt_framebuffer_read_rgba8");
    }

    // Per-frame ray queues from the caller's device memory (256-byte aligned,
    // kept alive by the caller), or from the library's own block again with
    // nullptr (fnd_pt_set_frame_arena).  QueueBytes(n) is the size a batch of
    // n paths takes; a smaller block runs smaller batches.
    void SetFrameArena(void* device_memory, uint64_t bytes)
    {
        check(fnd_pt_set_frame_arena(ctx_, device_memory, bytes), "fnd_pt_set_frame_arena");
    }
    fnd_frame_arena_info FrameArena()
    {
        fnd_frame_arena_info i{};
        check(fnd_pt_frame_arena_info(ctx_, &i), "fnd_pt_frame_arena_info");
        return i;
    }
    static uint64_t QueueBytes(uint32_t paths)
    {
        uint64_t b = 0;
        return fnd_pt_queue_bytes(paths, &b) == FND_OK ? b : 0;
    }

    // The same bytes written into device memory of the caller's (for instance
    // the swapchain image's memory imported into HIP), rows row_pitch bytes
    // apart: the frame reaches the presentation engine without host staging
    // (fnd_pt_framebuffer_present).
    void PresentTo(void* device_rgba, uint64_t row_pitch, float exposure = 1.0f)
    {
        uint32_t n = 0;
        check(fnd_pt_framebuffer_present(ctx_, fb_, exposure, device_rgba, row_pitch, &n),
              "fnd_pt_framebuffer_present");
    }

    uint32_t Samples() const { return samples_; }
    const Desc& Description() const { return desc_; }
    fnd_pt_ctx* Handle() const { return ctx_; }
    fnd_pt_stats Stats()
    {
        fnd_pt_stats s{};
        check(fnd_pt_get_stats(ctx_, &s), "fnd_pt_get_stats");
        return s;
    }

private:
    void init(const fnd_host_allocator* alloc)
    {
        fnd_pt_create_desc cd{};
        cd.device = desc_.device;
        cd.device_memory_budget = desc_.device_memory_budget;
        check(fnd_pt_create(&cd, alloc, &ctx_), "fnd_pt_create");
        check(fnd_pt_framebuffer_create(ctx_, desc_.width, desc_.height, desc_.fp64 ? FND_FB_FP64 : 0u, &fb_),
              "fnd_pt_framebuffer_create");
        cmd_.reset(new ComputeCommandList(ctx_));
        check(fnd_pt_fence_create(ctx_, &fence_), "fnd_pt_fence_create");
    }

    void check(fnd_status st, const char* call)
    {
        if (st != FND_OK) failure_handler()(st, call, ctx_ ? fnd_pt_last_error(ctx_) : "");
    }

    Desc desc_;
    fnd_host_allocator host_alloc_{};
    fnd_pt_ctx* ctx_ = nullptr;
    fnd_handle fb_ = FND_INVALID_HANDLE;
    fnd_handle scene_ = FND_INVALID_HANDLE;
    fnd_camera camera_{};
    uint32_t samples_ = 0;
    uint32_t recorded_ = 0;  // samples recorded into the list being built
    std::unique_ptr<ComputeCommandList> cmd_;
    fnd_pt_fence* fence_ = nullptr;
    void* present_ = nullptr;
    uint64_t present_pitch_ = 0;
    float exposure_ = 1.0f;
};

}  // namespace fnd
