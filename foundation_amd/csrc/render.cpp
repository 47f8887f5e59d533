// render.cpp — framebuffers and the batched wavefront render entry point.
//
// fnd_pt_render takes the place of the reference's per-frame raster pass
// (Renderer::Record + Queue::Submit, src/Renderer/Renderer.cpp:312-366,386-393)
// and, like Renderer::Draw, returns after the device has finished
// (src/Renderer/Renderer.cpp:394).
//
// All jobs of one call share one sequence of launches: the global path space
// is the concatenation over jobs of spp x padded pixels (sample-major, pixels
// in T x T tile order); batches are consecutive ranges of it, so one batch may
// hold paths of several scenes and resolutions.  Per batch: generate, then per
// bounce extension trace -> shade -> shadow trace -> shadow resolve, then
// accumulate.  Stage kernels read their queue lengths from device memory, so
// the host never waits inside a render.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "cache_state.hpp"
#include "ctx.hpp"
#include "wavefront.hpp"

using namespace fnd;

namespace fnd {

// Drain polling: after iteration k the counters are copied to pinned host
// memory; the host reads iteration k - kPollLag's copy before it queues
// iteration k + 1, so the device always has kPollLag iterations queued and at
// most kPollLag empty ones run after the window drains.  A window has drained
// when every path was started and no extension ray is left to shade (a path
// that ended is retired within its iteration).
constexpr uint32_t kPollLag = 3;
constexpr uint32_t kPollRing = kPollLag + 1;

struct RenderState {
    WfBuffers b{};
    uint32_t* poll_host = nullptr;  // kPollRing copies of the counter slots (cnt()), pinned host memory
    hipEvent_t poll_ev[kPollRing] = {};
    hipEvent_t side_ev[2] = {};     // early cache queries ready / answered
    fnd_arena frame{};        // reset at every render call, the queues are carved from it
    void* owned = nullptr;    // the library's block (when the caller gives none)
    uint64_t owned_bytes = 0;
    uint32_t batch = 0;       // paths per batch of the last call
    JobDev* d_jobs = nullptr;
    uint32_t jobs_cap = 0;
    ShadeScene* d_shade = nullptr;
    SceneDev* d_scenes = nullptr;
    uint32_t scenes_cap = 0;
};

void render_state_free(fnd_pt_ctx* ctx)
{
    auto* rs = static_cast<RenderState*>(ctx->render_state);
    if (!rs) return;
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    if (rs->poll_host) (void)hipHostFree(rs->poll_host);
    for (auto& e : rs->poll_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : rs->side_ev)
        if (e) (void)hipEventDestroy(e);
    dev_free(ctx, rs->owned, rs->owned_bytes);
    dev_free(ctx, rs->d_jobs, rs->jobs_cap * sizeof(JobDev));
    dev_free(ctx, rs->d_shade, rs->scenes_cap * sizeof(ShadeScene));
    dev_free(ctx, rs->d_scenes, rs->scenes_cap * sizeof(SceneDev));
    ctx_delete(ctx, rs);
    ctx->render_state = nullptr;
}

// Alignment of every queue in the frame arena (hipMalloc's own is at least
// this, so a carved buffer starts on a cache line like an allocated one).
constexpr uint64_t kQueueAlign = 256;

// Bump-allocates the buffers of one wavefront batch of `cap` paths from `ar`.
// False when the arena runs out (the pointers are then partly null).
static bool carve_queues(fnd_arena* ar, uint64_t cap, WfBuffers& b)
{
    bool ok = true;
    auto A = [&](auto** p, uint64_t n) {
        using T = std::remove_reference_t<decltype(**p)>;
        void* q = fnd_arena_alloc_aligned(ar, std::max<uint64_t>(n, 1) * sizeof(T), kQueueAlign);
        ok = ok && q != nullptr;
        *p = static_cast<T*>(q);
    };
    A(&b.p_site, cap); A(&b.p_state, cap); A(&b.p_thr, cap); A(&b.p_L, cap);
    A(&b.w_L, cap * kWindowSlots);
    A(&b.free_list[0], cap); A(&b.free_list[1], cap);
    for (int q = 0; q < 2; ++q) {
        A(&b.ext_rays[q], cap); A(&b.ext_path[q], cap); A(&b.ext_scene[q], cap);
    }
    A(&b.ext_hits, cap);
    A(&b.sh_rays, cap); A(&b.sh_path, cap); A(&b.sh_scene, cap); A(&b.sh_contrib, cap); A(&b.sh_hits, cap);
    A(&b.counters, kCntCount * kCntStride);
    A(&b.trace_ctr, kCounterWords);
    A(&b.rays, 1);
    A(&b.q_query, cap); A(&b.q_path, cap); A(&b.q_weight, cap); A(&b.q_pred, cap);
    return ok;
}

// Bytes carve_queues takes for `cap` paths from a kQueueAlign-aligned block.
static uint64_t queue_bytes(uint64_t cap)
{
    fnd_arena m{};
    fnd_arena_reset(&m, reinterpret_cast<void*>(uintptr_t(kQueueAlign)), UINT64_MAX - 2 * kQueueAlign);
    WfBuffers b{};
    carve_queues(&m, cap, b);
    return fnd_arena_used(&m);
}

namespace {

RenderState* render_state(fnd_pt_ctx* ctx)
{
    if (!ctx->render_state) ctx->render_state = ctx_new<RenderState>(ctx);
    return static_cast<RenderState*>(ctx->render_state);
}

// The call's frame: reset the arena over the caller's block or the library's
// own and carve the queues of a batch of `want` paths, or of the largest
// batch of at least `least` paths that fits the caller's block.
fnd_status begin_frame(fnd_pt_ctx* ctx, uint32_t want, uint32_t least, uint32_t* batch)
{
    RenderState* rs = render_state(ctx);
    if (!rs) return set_error(ctx, FND_E_OOM_HOST, "out of host memory");
    fnd_arena_reset(&rs->frame, nullptr, 0);  // nothing of the previous frame survives a failed call
    rs->b = WfBuffers{};
    rs->batch = 0;
    uint32_t cap = want;
    void* block = ctx->frame_memory;
    uint64_t size = ctx->frame_bytes;
    if (block) {
        if (queue_bytes(cap) > size) {
            uint32_t lo = 0, hi = cap;  // queue_bytes(lo) fits (or lo = 0), queue_bytes(hi) does not
            while (hi - lo > 1) {
                const uint32_t mid = lo + (hi - lo) / 2;
                (queue_bytes(mid) <= size ? lo : hi) = mid;
            }
            cap = lo;
        }
        if (cap < least)
            return set_error(ctx, FND_E_OOM_DEVICE,
                             "frame arena of %llu bytes holds %u paths; the call needs at least %u (%llu bytes)",
                             (unsigned long long)size, cap, least, (unsigned long long)queue_bytes(least));
    } else {
        const uint64_t need = queue_bytes(cap);
        if (need > rs->owned_bytes) {  // grow: the previous call has finished (render is synchronous)
            (void)hipStreamSynchronize(ctx->stream);
            dev_free(ctx, rs->owned, rs->owned_bytes);
            rs->owned = nullptr;
            rs->owned_bytes = 0;
            fnd_status st = dev_alloc(ctx, &rs->owned, need);
            if (st != FND_OK) return st;
            rs->owned_bytes = need;
        }
        block = rs->owned;
        size = rs->owned_bytes;
    }
    fnd_arena_reset(&rs->frame, block, size);
    if (!carve_queues(&rs->frame, cap, rs->b))  // cannot happen: the sizes were checked
        return set_error(ctx, FND_E_OOM_DEVICE, "frame arena exhausted");
    rs->batch = cap;
    *batch = cap;
    return FND_OK;
}

fnd_status ensure_tables(fnd_pt_ctx* ctx, uint32_t njobs, uint32_t nscenes)
{
    auto* rs = static_cast<RenderState*>(ctx->render_state);
    if (njobs > rs->jobs_cap) {
        dev_free(ctx, rs->d_jobs, rs->jobs_cap * sizeof(JobDev));
        rs->d_jobs = nullptr;
        rs->jobs_cap = 0;
        void* p = nullptr;
        const uint32_t cap = std::max(njobs, 64u);
        fnd_status st = dev_alloc(ctx, &p, cap * sizeof(JobDev));
        if (st != FND_OK) return st;
        rs->d_jobs = static_cast<JobDev*>(p);
        rs->jobs_cap = cap;
    }
    if (nscenes > rs->scenes_cap) {
        dev_free(ctx, rs->d_shade, rs->scenes_cap * sizeof(ShadeScene));
        dev_free(ctx, rs->d_scenes, rs->scenes_cap * sizeof(SceneDev));
        rs->d_shade = nullptr;
        rs->d_scenes = nullptr;
        rs->scenes_cap = 0;
        const uint32_t cap = std::max(nscenes, 64u);
        void* p = nullptr;
        fnd_status st = dev_alloc(ctx, &p, cap * sizeof(ShadeScene));
        if (st != FND_OK) return st;
        rs->d_shade = static_cast<ShadeScene*>(p);
        st = dev_alloc(ctx, &p, cap * sizeof(SceneDev));
        if (st != FND_OK) return st;
        rs->d_scenes = static_cast<SceneDev*>(p);
        rs->scenes_cap = cap;
    }
    return FND_OK;
}

constexpr uint64_t kTrainShare = 4;

uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Training-path selection of one job (see JobDev): `count` (at most a quarter)
// of the job's first ts samples per pixel, ts = min(spp, (2^32 - 1) / (W * H)) so every product
// below stays under 2^64.  The multipliers are primes near 2^32 / golden
// ratio (Fibonacci hashing spreads consecutive t over the image); at most one
// of them can divide the modulus.
void select_training(JobDev& j, uint32_t spp, uint32_t base, uint32_t count, uint64_t frame)
{
    const uint64_t P = uint64_t(j.width) * j.height;
    const uint64_t ts = std::max<uint64_t>(1, std::min<uint64_t>(spp, 0xFFFFFFFFull / P));
    j.train_mod = P * ts;
    j.train_mul = (j.train_mod % 2654435789ull) != 0 ? 2654435789ull : 2654435803ull;
    j.train_add = mix64(j.seed ^ mix64(uint64_t(j.sample_offset) ^ (frame << 32))) % j.train_mod;
    j.train_barrett = ~0ull / j.train_mod;
    j.train_base = base;
    // at most one path in kTrainShare trains, so most pixels keep the cached estimate
    j.train_count = uint32_t(std::min<uint64_t>(count, std::max<uint64_t>(1, j.train_mod / kTrainShare)));
}

}  // namespace

// Why a camera cannot form a basis, or null: non-finite values, eye == at, a
// field of view outside (0, 180) degrees, or an up vector of no length or
// (nearly) parallel to the view direction (the right vector would be the
// normalisation of a zero cross product, NaN for every ray).
const char* camera_problem(const fnd_camera& c)
{
    double f[3], up[3];
    for (int i = 0; i < 3; ++i) {
        if (!std::isfinite(c.eye[i]) || !std::isfinite(c.at[i]) || !std::isfinite(c.up[i]))
            return "non-finite camera";
        f[i] = double(c.at[i]) - double(c.eye[i]);
        up[i] = c.up[i];
    }
    if (!(c.fov_y_deg > 0.0f && c.fov_y_deg < 180.0f)) return "field of view outside (0, 180) degrees";
    const double lf = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    const double lu = std::sqrt(up[0] * up[0] + up[1] * up[1] + up[2] * up[2]);
    if (!(lf > 0.0)) return "eye == at";
    if (!(lu > 0.0)) return "up vector of zero length";
    const double r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    const double lr = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (!(lr > 1e-6 * lf * lu)) return "up vector parallel to the view direction";
    return nullptr;
}

namespace {

void camera_basis(const fnd_camera& c, uint32_t w, uint32_t h, JobDev& j)
{
    double f[3], up[3] = {c.up[0], c.up[1], c.up[2]};
    for (int i = 0; i < 3; ++i) f[i] = double(c.at[i]) - double(c.eye[i]);
    auto norm = [](double* v) {
        const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int i = 0; i < 3; ++i) v[i] /= l;
    };
    norm(f);
    double r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    norm(r);
    const double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    for (int i = 0; i < 3; ++i) {
        j.eye[i] = c.eye[i];
        j.fwd[i] = float(f[i]);
        j.right[i] = float(r[i]);
        j.up[i] = float(u[i]);
    }
    j.tan_half = float(std::tan(double(c.fov_y_deg) * M_PI / 360.0));
    j.aspect = float(double(w) / double(h));
}

// The 255 byte thresholds ((k - 0.5) / 255)^2.2 of the R8G8B8A8 transfer,
// computed on the host in double precision and uploaded once per context.
fnd_status present_table(fnd_pt_ctx* ctx)
{
    if (ctx->d_present) return FND_OK;
    float t[255];
    for (int k = 1; k <= 255; ++k) t[k - 1] = float(std::pow((double(k) - 0.5) / 255.0, 2.2));
    void* p = nullptr;
    fnd_status st = dev_alloc(ctx, &p, sizeof t);
    if (st != FND_OK) return st;
    ctx->d_present = static_cast<float*>(p);
    return copy_sync(ctx, ctx->d_present, t, sizeof t, hipMemcpyHostToDevice, "present table");
}

}  // namespace

fnd_status present_async(fnd_pt_ctx* ctx, Framebuffer* f, float exposure, uint32_t* out, uint32_t pitch)
{
    fnd_status st = present_table(ctx);
    if (st != FND_OK) return st;
    const uint64_t pixels = uint64_t(f->width) * f->height;
    const float inv = f->samples ? 1.0f / float(f->samples) : 1.0f;
    launch_present(f->accum, pixels, inv, exposure, ctx->d_present, out, f->width, pitch, ctx->stream);
    return check_hip(ctx, hipGetLastError(), "present");
}

fnd_status clear_async(fnd_pt_ctx* ctx, Framebuffer* f)
{
    const uint64_t n = uint64_t(f->width) * f->height * 3;
    fnd_status st = check_hip(ctx, hipMemsetAsync(f->accum, 0, n * sizeof(float), ctx->stream), "clear");
    if (st == FND_OK && f->accum64)
        st = check_hip(ctx, hipMemsetAsync(f->accum64, 0, n * sizeof(double), ctx->stream), "clear");
    if (st == FND_OK) f->samples = 0;
    return st;
}

}  // namespace fnd

extern "C" {

fnd_status fnd_pt_framebuffer_create(fnd_pt_ctx* ctx, uint32_t width, uint32_t height, uint32_t flags, fnd_handle* out)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || !out) return FND_E_INVALID_ARG;
        *out = FND_INVALID_HANDLE;
        if (width == 0 || height == 0 || width > 16384 || height > 16384)
            return set_error(ctx, FND_E_INVALID_ARG, "framebuffer size %ux%u out of range", width, height);
        if (flags & ~FND_FB_FP64) return set_error(ctx, FND_E_INVALID_ARG, "unknown framebuffer flags");
        auto* f = ctx_new<Framebuffer>(ctx);
        if (!f) return set_error(ctx, FND_E_OOM_HOST, "out of host memory");
        f->width = width;
        f->height = height;
        f->flags = flags;
        const uint64_t n = uint64_t(width) * height * 3;
        void* p = nullptr;
        fnd_status st = dev_alloc(ctx, &p, n * sizeof(float));
        if (st != FND_OK) {
            ctx_delete(ctx, f);
            return st;
        }
        f->accum = static_cast<float*>(p);
        if (flags & FND_FB_FP64) {
            st = dev_alloc(ctx, &p, n * sizeof(double));
            if (st != FND_OK) {
                dev_free(ctx, f->accum, n * sizeof(float));
                ctx_delete(ctx, f);
                return st;
            }
            f->accum64 = static_cast<double*>(p);
        }
        const fnd_handle h = make_handle(kHandleFramebuffer, ctx->next_id++);
        try {
            ctx->framebuffers[h] = f;
        } catch (...) {
            dev_free(ctx, f->accum, n * sizeof(float));
            dev_free(ctx, f->accum64, n * sizeof(double));
            ctx_delete(ctx, f);
            throw;
        }
        *out = h;
        return fnd_pt_framebuffer_clear(ctx, h);
    });
}

fnd_status fnd_pt_framebuffer_destroy(fnd_pt_ctx* ctx, fnd_handle h)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx) return FND_E_INVALID_ARG;
        Framebuffer* f = find_framebuffer(ctx, h);
        if (!f) return set_error(ctx, FND_E_NOT_FOUND, "unknown framebuffer handle");
        (void)hipStreamSynchronize(ctx->stream);
        const uint64_t n = uint64_t(f->width) * f->height * 3;
        dev_free(ctx, f->accum, n * sizeof(float));
        dev_free(ctx, f->accum64, n * sizeof(double));
        dev_free(ctx, f->rgba8, n / 3 * sizeof(uint32_t));
        ctx_delete(ctx, f);
        ctx->framebuffers.erase(h);
        return FND_OK;
    });
}

fnd_status fnd_pt_framebuffer_clear(fnd_pt_ctx* ctx, fnd_handle h)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx) return FND_E_INVALID_ARG;
        Framebuffer* f = find_framebuffer(ctx, h);
        if (!f) return set_error(ctx, FND_E_NOT_FOUND, "unknown framebuffer handle");
        fnd_status st = clear_async(ctx, f);
        if (st != FND_OK) return st;
        return check_hip(ctx, hipStreamSynchronize(ctx->stream), "clear");
    });
}

fnd_status fnd_pt_framebuffer_read(fnd_pt_ctx* ctx, fnd_handle h, float* rgb, uint32_t* out_samples)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || !rgb) return FND_E_INVALID_ARG;
        Framebuffer* f = find_framebuffer(ctx, h);
        if (!f) return set_error(ctx, FND_E_NOT_FOUND, "unknown framebuffer handle");
        const uint64_t n = uint64_t(f->width) * f->height * 3;
        fnd_status st = check_hip(ctx, hipMemcpyAsync(rgb, f->accum, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
                                  "read framebuffer");
        if (st == FND_OK) st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "read framebuffer");
        if (st != FND_OK) return st;
        if (f->samples) {
            const float inv = 1.0f / float(f->samples);
            for (uint64_t i = 0; i < n; ++i) rgb[i] *= inv;
        }
        if (out_samples) *out_samples = f->samples;
        return FND_OK;
    });
}

// Present path: the reference presents R8G8B8A8_UNORM at 1920x1080
// (src/Renderer/Renderer.cpp:39-44).  Byte b of a channel is the number of
// thresholds ((k - 0.5) / 255)^2.2, k = 1..255, at or below x = clamp(mean *
// exposure), i.e. round(255 * x^(1/2.2)) with the boundaries computed once
// on the host in double precision, so any host can reproduce the bytes.
fnd_status fnd_pt_framebuffer_read_rgba8(fnd_pt_ctx* ctx, fnd_handle h, float exposure, uint8_t* rgba,
                                         uint32_t* out_samples)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || !rgba || !(exposure >= 0.0f)) return FND_E_INVALID_ARG;
        Framebuffer* f = find_framebuffer(ctx, h);
        if (!f) return set_error(ctx, FND_E_NOT_FOUND, "unknown framebuffer handle");
        const uint64_t pixels = uint64_t(f->width) * f->height;
        fnd_status st = present_table(ctx);
        if (st != FND_OK) return st;
        if (!f->rgba8) {
            void* p = nullptr;
            if ((st = dev_alloc(ctx, &p, pixels * sizeof(uint32_t))) != FND_OK) return st;
            f->rgba8 = static_cast<uint32_t*>(p);
        }
        const float inv = f->samples ? 1.0f / float(f->samples) : 1.0f;
        launch_present(f->accum, pixels, inv, exposure, ctx->d_present, f->rgba8, f->width, f->width, ctx->stream);
        st = check_hip(ctx, hipGetLastError(), "present");
        if (st == FND_OK)
            st = check_hip(ctx, hipMemcpyAsync(rgba, f->rgba8, pixels * 4, hipMemcpyDeviceToHost, ctx->stream),
                           "read framebuffer");
        if (st == FND_OK) st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "read framebuffer");
        if (st != FND_OK) return st;
        if (out_samples) *out_samples = f->samples;
        return FND_OK;
    });
}

// The same conversion written straight into device memory of the caller's
// (for instance an image imported from the presentation engine), rows
// `row_pitch` bytes apart: no copy through host memory.
fnd_status fnd_pt_framebuffer_present(fnd_pt_ctx* ctx, fnd_handle h, float exposure, void* device_rgba,
                                      uint64_t row_pitch, uint32_t* out_samples)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || !device_rgba || !(exposure >= 0.0f)) return FND_E_INVALID_ARG;
        Framebuffer* f = find_framebuffer(ctx, h);
        if (!f) return set_error(ctx, FND_E_NOT_FOUND, "unknown framebuffer handle");
        if (reinterpret_cast<uintptr_t>(device_rgba) % 4 || row_pitch % 4 || row_pitch < 4ull * f->width ||
            row_pitch / 4 > 0xFFFFFFFFull)
            return set_error(ctx, FND_E_INVALID_ARG,
                             "present target must be 4-byte aligned with a row pitch of at least %u bytes, a multiple of 4",
                             4 * f->width);
        fnd_status st = present_async(ctx, f, exposure, static_cast<uint32_t*>(device_rgba), uint32_t(row_pitch / 4));
        if (st == FND_OK) st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "present");
        if (st != FND_OK) return st;
        if (out_samples) *out_samples = f->samples;
        return FND_OK;
    });
}

fnd_status fnd_pt_framebuffer_read_f64(fnd_pt_ctx* ctx, fnd_handle h, double* rgb, uint32_t* out_samples)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || !rgb) return FND_E_INVALID_ARG;
        Framebuffer* f = find_framebuffer(ctx, h);
        if (!f) return set_error(ctx, FND_E_NOT_FOUND, "unknown framebuffer handle");
        if (!f->accum64) return set_error(ctx, FND_E_INVALID_ARG, "framebuffer was created without FND_FB_FP64");
        const uint64_t n = uint64_t(f->width) * f->height * 3;
        fnd_status st = check_hip(ctx, hipMemcpyAsync(rgb, f->accum64, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
                                  "read framebuffer");
        if (st == FND_OK) st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "read framebuffer");
        if (st != FND_OK) return st;
        if (f->samples) {
            const double inv = 1.0 / double(f->samples);
            for (uint64_t i = 0; i < n; ++i) rgb[i] *= inv;
        }
        if (out_samples) *out_samples = f->samples;
        return FND_OK;
    });
}

fnd_status fnd_pt_framebuffer_rel_mse(fnd_pt_ctx* ctx, fnd_handle h, fnd_handle r, double eps, double* out)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || !out) return FND_E_INVALID_ARG;
        Framebuffer* f = find_framebuffer(ctx, h);
        Framebuffer* g = find_framebuffer(ctx, r);
        if (!f || !g) return set_error(ctx, FND_E_NOT_FOUND, "unknown framebuffer handle");
        if (!g->accum64) return set_error(ctx, FND_E_INVALID_ARG, "reference framebuffer was created without FND_FB_FP64");
        if (f->width != g->width || f->height != g->height)
            return set_error(ctx, FND_E_INVALID_ARG, "framebuffers differ in size (%ux%u against %ux%u)", f->width,
                             f->height, g->width, g->height);
        if (!(eps > 0.0)) return set_error(ctx, FND_E_INVALID_ARG, "eps must be positive");
        if (!ctx->d_relmse) {
            void* p = nullptr;
            fnd_status st = dev_alloc(ctx, &p, (kRelMseParts + 1) * sizeof(double));
            if (st != FND_OK) return st;
            ctx->d_relmse = static_cast<double*>(p);
        }
        const uint64_t n = uint64_t(f->width) * f->height * 3;
        const float xinv = f->samples ? 1.0f / float(f->samples) : 1.0f;
        const double yinv = g->samples ? 1.0 / double(g->samples) : 1.0;
        launch_rel_mse(f->accum, xinv, g->accum64, yinv, n, eps, ctx->d_relmse, ctx->d_relmse + kRelMseParts,
                       ctx->stream);
        fnd_status st = check_hip(ctx, hipGetLastError(), "relMSE");
        if (st == FND_OK) st = copy_sync(ctx, out, ctx->d_relmse + kRelMseParts, sizeof(double), hipMemcpyDeviceToHost, "relMSE");
        return st;
    });
}

fnd_status fnd_pt_render(fnd_pt_ctx* ctx, const fnd_render_job* jobs, uint32_t num_jobs)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || (!jobs && num_jobs)) return FND_E_INVALID_ARG;
        if (num_jobs == 0) return FND_OK;
        // a call that failed after queueing cache queries on the side stream may
        // have left them running; they read the queues this call is about to reuse
        (void)hipStreamSynchronize(ctx->side);
        const uint32_t T = std::max(1u, ctx->tuning.tile);
        // the call's temporaries come from the caller's allocator, like the context's own objects
        const CtxAlloc<char> ha = ctx->host_alloc_ref();
        CtxMap<Scene*, uint32_t> slot_of(ha);
        CtxVector<Scene*> slots(ha);
        CtxVector<JobDev> hj(num_jobs, JobDev{}, ha);
        CtxVector<Framebuffer*> fbs(num_jobs, nullptr, ha);
        uint64_t total = 0;
        uint32_t max_depth = 0, max_stack = 0;
        bool any_use = false, any_train = false;
        Scene* cache_scene = nullptr;
        for (uint32_t i = 0; i < num_jobs; ++i) {
            const fnd_render_job& in = jobs[i];
            Scene* s = find_scene(ctx, in.scene);
            if (!s) return set_error(ctx, FND_E_NOT_FOUND, "job %u: unknown scene handle", i);
            Framebuffer* fb = find_framebuffer(ctx, in.framebuffer);
            if (!fb) return set_error(ctx, FND_E_NOT_FOUND, "job %u: unknown framebuffer handle", i);
            for (uint32_t k = 0; k < i; ++k)
                if (fbs[k] == fb)
                    return set_error(ctx, FND_E_INVALID_ARG, "jobs %u and %u accumulate into one framebuffer", k, i);
            if (in.spp == 0 || in.max_depth > 64)
                return set_error(ctx, FND_E_INVALID_ARG, "job %u: spp must be >= 1 and max_depth <= 64", i);
            if (in.flags & ~(FND_JOB_USE_CACHE | FND_JOB_TRAIN_CACHE))
                return set_error(ctx, FND_E_INVALID_ARG, "job %u: unknown flags 0x%x", i, in.flags);
            if (in.flags & (FND_JOB_USE_CACHE | FND_JOB_TRAIN_CACHE)) {
                if (!cache_get(ctx))
                    return set_error(ctx, FND_E_INVALID_ARG, "job %u: radiance cache flags need fnd_pt_cache_configure", i);
                if (!cache_scene) cache_scene = s;
                any_use |= (in.flags & FND_JOB_USE_CACHE) != 0;
                any_train |= (in.flags & FND_JOB_TRAIN_CACHE) != 0;
            }
            if (const char* why = camera_problem(in.camera))
                return set_error(ctx, FND_E_INVALID_ARG, "job %u: degenerate camera (%s)", i, why);
            if (uint64_t(in.sample_offset) + in.spp > 0xFFFFFFFFull)
                return set_error(ctx, FND_E_INVALID_ARG, "job %u: sample index overflows 32 bits", i);
            auto it = slot_of.find(s);
            if (it == slot_of.end()) {
                it = slot_of.emplace(s, uint32_t(slots.size())).first;
                slots.push_back(s);
            }
            JobDev& j = hj[i];
            std::memset(&j, 0, sizeof(j));
            j.scene_slot = it->second;
            j.width = fb->width;
            j.height = fb->height;
            j.tile = T;
            j.tiles_x = (fb->width + T - 1) / T;
            const uint64_t padded = uint64_t(j.tiles_x) * ((fb->height + T - 1) / T) * T * T;
            if (padded >= (1ull << 32)) return set_error(ctx, FND_E_INVALID_ARG, "job %u: image too large", i);
            j.padded = uint32_t(padded);
            j.max_depth = in.max_depth;
            j.rr_depth = in.rr_depth;
            j.sample_offset = in.sample_offset;
            j.flags = in.flags & FND_JOB_USE_CACHE;
            j.seed = in.seed;
            j.path_base = total;
            total += padded * in.spp;
            j.path_end = total;
            camera_basis(in.camera, fb->width, fb->height, j);
            j.accum = fb->accum;
            j.accum64 = fb->accum64;
            fbs[i] = fb;
            max_depth = std::max(max_depth, in.max_depth);
            max_stack = std::max(max_stack, s->info.stack_entries);
        }
        CacheState* cs = cache_get(ctx);
        if (cache_scene && !cs->bounds_set) cache_set_bounds(cs, cache_scene->info.bounds_lo, cache_scene->info.bounds_hi, true);
        if (any_train) {
            // the frame's training paths: desc.train_paths split over the training jobs
            uint32_t ntrain = 0;
            for (uint32_t i = 0; i < num_jobs; ++i) ntrain += (jobs[i].flags & FND_JOB_TRAIN_CACHE) ? 1u : 0u;
            const uint32_t per = std::max(1u, cs->desc.train_paths / ntrain);
            uint32_t base = 0;
            for (uint32_t i = 0; i < num_jobs; ++i) {
                if (!(jobs[i].flags & FND_JOB_TRAIN_CACHE) || base >= kCacheTrainMaxPaths) continue;
                select_training(hj[i], jobs[i].spp, base, std::min(per, kCacheTrainMaxPaths - base), cs->frames);
                base += hj[i].train_count;
            }
        }
        const uint32_t want = uint32_t(std::min<uint64_t>(std::max(1024u, ctx->tuning.queue_capacity), total));
        uint32_t cap = 0;
        fnd_status st = begin_frame(ctx, want, uint32_t(std::min<uint64_t>(1024u, total)), &cap);
        if (st != FND_OK) return st;
        st = ensure_tables(ctx, num_jobs, uint32_t(slots.size()));
        if (st != FND_OK) return st;
        auto* rs = static_cast<RenderState*>(ctx->render_state);
        CtxVector<ShadeScene> hs(slots.size(), ShadeScene{}, ha);
        CtxVector<SceneDev> hd(slots.size(), SceneDev{}, ha);
        for (size_t k = 0; k < slots.size(); ++k) {
            const Scene* s = slots[k];
            hd[k] = SceneDev{s->nodes, s->tris, s->insts, s->root, s->two_level ? 1u : 0u};
            ShadeScene& z = hs[k];
            z.tris = s->tris;
            z.insts = s->insts;
            z.materials = s->materials;
            z.lights = s->lights;
            z.num_lights = s->num_lights;
            z.two_level = s->two_level ? 1u : 0u;
            std::memcpy(z.background, s->background, 12);
            z.pad = 0;
        }
        st = check_hip(ctx, hipMemcpyAsync(rs->d_shade, hs.data(), hs.size() * sizeof(ShadeScene), hipMemcpyHostToDevice,
                                           ctx->stream), "upload scenes");
        if (st == FND_OK)
            st = check_hip(ctx, hipMemcpyAsync(rs->d_scenes, hd.data(), hd.size() * sizeof(SceneDev), hipMemcpyHostToDevice,
                                               ctx->stream), "upload scenes");
        if (st != FND_OK) return st;

        TraceLaunch tl{};
        // A call over one scene passes its traversal constants as kernel
        // arguments (no per-ray scene lookup at each refill, no per-lane scene
        // registers), and a single-level scene walks without the world-ray
        // registers of two-level scenes.
        tl.multi = slots.size() > 1;
        tl.two_level = tl.multi || slots[0]->two_level;
        tl.block = ctx->tuning.block_size;
        tl.stack_lds = ctx->tuning.stack_lds;
        tl.mode = kModeMixedInternal;  // the mode every launch below uses
        tl.grid = trace_grid(ctx, tl);
        st = ensure_spill(ctx, max_stack, tl.stack_lds, tl.grid * tl.block);
        if (st != FND_OK) return st;

        WfArgs wa{};
        wa.jobs = rs->d_jobs;
        wa.scenes = rs->d_shade;
        wa.b = rs->b;
        if (cs) {
            wa.b.rec = cs->rec;
            wa.b.rec_count = cs->rec_count;
            wa.b.t_list = cs->t_list;
            wa.b.rec_max = std::min(cs->desc.train_depth, kCacheMaxRecords);
            wa.b.glossy_min = cs->desc.glossy_min_roughness;
            wa.b.glossy_any_bounce = cs->desc.glossy_any_bounce;
            wa.b.t_query = cs->s_query;
            wa.b.t_target = cs->s_target;
            wa.b.t_cap = cs->sample_cap;
        }
        const uint32_t stage_grid = std::max(1u, ctx->num_cus * 8u);  // 256-thread blocks; the other stages
                                                                       // launch as many threads
        // Launches with fewer rays than coop_rays trace with a group of lanes
        // per ray (trace_coop.hpp): kCoopRaysPerWave per wave of the
        // cooperative grid.  Measured on 1080p frames of the 1.5M-triangle
        // scene: 6, 12, 24 and 48 rays per wave gave 4.92, 5.01, 5.06 and 6.52
        // ms per plain frame (5.42 without); one wave per ray with 2, 4 and 6
        // gave 5.06, 5.01 and 5.00.
        constexpr uint32_t kCoopRaysPerWave = 6;
        uint32_t coop_grid = 0, coop_rays = 0;
        {
            TraceLaunch lc = tl;
            lc.coop = true;
            coop_grid = std::min(tl.grid, trace_blocks_per_cu(lc) * ctx->num_cus);
            coop_rays = kCoopRaysPerWave * (coop_grid * tl.block / 64u);
        }

        // Drain polling resources, made once per render state.
        if (!rs->poll_host) {
            void* h = nullptr;
            st = check_hip(ctx, hipHostMalloc(&h, kPollRing * kCntCount * kCntStride * sizeof(uint32_t)), "hipHostMalloc");
            if (st != FND_OK) return st;
            rs->poll_host = static_cast<uint32_t*>(h);
            for (auto& e : rs->poll_ev) {
                st = check_hip(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
                if (st != FND_OK) return st;
            }
            for (auto& e : rs->side_ev) {
                st = check_hip(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
                if (st != FND_OK) return st;
            }
        }

        // One pass of the wavefront over a job table already on the device, window
        // by window.  A window's first paths fill the slots; then every iteration
        // shades the traced extension rays, gives the slots freed in the previous
        // iteration the window's next paths (their camera rays join the same
        // queue), traces extension and shadow rays in one launch, adds the shadow
        // results and retires the paths that ended.  Paths ending into the
        // radiance cache are queried once the window has drained, so a window
        // that uses the cache holds at most one path per slot.
        auto run = [&](uint32_t njobs, uint64_t paths, uint32_t slots, uint32_t depth, bool apply_cache,
                       bool collect) -> fnd_status {
            wa.num_jobs = njobs;
            const uint64_t wmax = (apply_cache || collect) ? uint64_t(slots) : uint64_t(slots) * kWindowSlots;
            TraceArgs ta{};
            ta.scene = hd[0];
            ta.scenes = rs->d_scenes;
            ta.hits = rs->b.ext_hits;
            ta.counter = rs->b.trace_ctr;
            ta.spill = ctx->d_spill;
            ta.spill_stride = tl.grid * tl.block;
            ta.refill = std::max(1u, std::min(64u, ctx->tuning.refill));
            ta.rays2 = rs->b.sh_rays;
            ta.hits2 = rs->b.sh_hits;
            ta.ray_scene2 = rs->b.sh_scene;
            ta.ray_total = rs->b.rays;
            TraceLaunch l = tl;
            l.mode = kModeMixedInternal;
            // One traversal launch per iteration: the extension rays (closest hit)
            // and the shadow rays of the vertices just shaded (any hit) share it,
            // so their tails overlap.
            // A launch that may hold few rays is queued twice: the single-lane
            // kernel returns at once below coop_rays rays, the cooperative one
            // (one wave per ray) at or above it.  The ray count is only known
            // on the device; either way the records are the same.
            auto trace = [&](uint32_t q, uint32_t sh, const char* what, bool maybe_few = true) -> fnd_status {
                ta.rays = rs->b.ext_rays[q];
                ta.ray_scene = rs->b.ext_scene[q];
                ta.nrays_dev = &rs->b.counters[cnt(q)];
                ta.nrays2_dev = &rs->b.counters[cnt(sh)];
                const bool dual = maybe_few && coop_rays != 0;
                ta.coop_below = dual ? coop_rays : 0u;
                fnd_status e = check_hip(ctx, launch_trace(l, ta, ctx->stream), what);
                if (e != FND_OK || !dual) return e;
                TraceLaunch lc = l;
                lc.coop = true;
                lc.grid = coop_grid;
                return check_hip(ctx, launch_trace(lc, ta, ctx->stream), what);
            };
            for (uint64_t g0 = 0; g0 < paths; g0 += wmax) {
                const uint32_t W = uint32_t(std::min<uint64_t>(wmax, paths - g0));
                const uint32_t n0 = std::min(W, slots);
                const uint32_t wgrid = std::min(stage_grid, (W + 255u) / 256u);
                // a window uses n0 slots
                const uint32_t sgrid = std::min(stage_grid, (n0 + 255u) / 256u);
                const uint32_t ggrid = std::min(stage_grid * 256u / kGenerateBlock, (n0 + kGenerateBlock - 1u) / kGenerateBlock);
                const uint32_t hgrid = std::min(stage_grid * 256u / kShadeBlock, (n0 + kShadeBlock - 1u) / kShadeBlock);
                // A window that fits in the slots has no regeneration: slot i holds
                // window path i to its end, so its radiance stays in place and the
                // bounce count bounds the iterations (no polling).
                const bool regen = W > slots;
                wa.b.window = W;
                wa.b.w_L = regen ? rs->b.w_L : rs->b.p_L;
                wa.b.retire = regen ? 1u : 0u;
                uint32_t cur = kCntExtA;
                launch_zero_slots(wa, (1u << kCntExtA) | (1u << kCntQuery) | (1u << kCntFree0), ctx->stream);
                launch_prep(wa, cur ^ 1u, 1u, kCntShadow, ctx->stream);  // empty shadow queue and lists, zero work counter
                launch_generate(wa, g0, n0, ggrid, ctx->stream);
                // the window's first paths: n0 camera rays at most (fewer for tile padding)
                fnd_status e = trace(cur, kCntShadow, "primary trace", n0 < coop_rays + 1024u);
                if (e != FND_OK) return e;
                const uint32_t cg = std::max(1u, std::min<uint32_t>(ctx->num_cus * 4u, (W + 127u) / 128u));
                bool early = false;  // queries answered on the side stream
                if (!regen) {
                    // After the shade of bounce 1 nearly every path that ends in the
                    // radiance cache has queried it.  Those queries are answered on
                    // the side stream while the main stream traces the last bounces,
                    // launches that hold few rays and leave most CUs idle; the rest
                    // are answered after the loop, on the main stream.
                    const uint32_t ksplit = std::min(1u, depth);
                    for (uint32_t k = 0; k <= depth; ++k) {
                        const uint32_t nxt = cur ^ 1u, fw = (k & 1u) ^ 1u;  // (the free lists go unused)
                        // the bounces' shadow queues alternate between two counters, so the
                        // shadow resolve of bounce k clears what bounce k + 1 appends to
                        // (its extension queue is the one bounce k shaded) and no counter
                        // launch runs between bounces
                        const uint32_t sh = kCntShadow + (k & 1u);
                        if (k == 0) launch_prep(wa, nxt, fw, sh, ctx->stream);
                        launch_shade(wa, cur, fw, sh, hgrid, ctx->stream);
                        if (apply_cache && k == ksplit && k < depth) {
                            launch_snapshot(wa, kCntQuery, kCntQuerySplit, ctx->stream);
                            e = check_hip(ctx, hipEventRecord(rs->side_ev[0], ctx->stream), "side stream");
                            if (e == FND_OK) e = check_hip(ctx, hipStreamWaitEvent(ctx->side, rs->side_ev[0], 0), "side stream");
                            if (e == FND_OK) e = join_training(ctx, ctx->side);
                            if (e != FND_OK) return e;
                            launch_cache_forward(cs->view, rs->b.q_query, 1, 0, &rs->b.counters[cnt(kCntQuerySplit)], rs->b.q_pred,
                                                 nullptr, 0, cg, ctx->side);
                            e = check_hip(ctx, hipEventRecord(rs->side_ev[1], ctx->side), "side stream");
                            if (e != FND_OK) return e;
                            early = true;
                        }
                        if (k < depth) {  // the last shade appends no rays
                            e = trace(nxt, sh, "extension + shadow trace");
                            if (e != FND_OK) return e;
                            launch_shadow_resolve(wa, fw, sh, cur, kCntShadow + ((k + 1u) & 1u), kCntCount, sgrid,
                                                  ctx->stream);
                        }
                        cur = nxt;
                    }
                } else {
                    // every path is shaded at most depth + 1 times, and while paths are
                    // left no slot stays idle, so the window drains within this many
                    // iterations (plus the polling lag)
                    const uint64_t max_iters = (uint64_t(W) / slots + 2u) * (uint64_t(depth) + 3u) + kPollLag + 2u;
                    bool drained = false;
                    for (uint64_t k = 0; !drained; ++k) {
                        if (k >= max_iters)
                            return set_error(ctx, FND_E_DEVICE,
                                             "wavefront window of %u paths did not drain in %llu iterations", W,
                                             (unsigned long long)max_iters);
                        const uint32_t nxt = cur ^ 1u;
                        const uint32_t fr = uint32_t(k & 1u), fw = fr ^ 1u;  // free lists read / written
                        // as above, the shadow resolve of iteration k clears what iteration
                        // k + 1 appends to: the extension queue iteration k shaded, the other
                        // shadow counter and the free list iteration k's regeneration read
                        const uint32_t sh = kCntShadow + uint32_t(k & 1u);
                        if (k == 0) launch_prep(wa, nxt, fw, sh, ctx->stream);
                        launch_shade(wa, cur, fw, sh, hgrid, ctx->stream);
                        launch_regen(wa, g0, fr, nxt, sgrid, ctx->stream);
                        e = trace(nxt, sh, "extension + shadow trace");
                        if (e != FND_OK) return e;
                        launch_shadow_resolve(wa, fw, sh, cur, kCntShadow + uint32_t((k + 1u) & 1u), kCntFree0 + fr, sgrid,
                                              ctx->stream);
                        const uint32_t ring = uint32_t(k % kPollRing);
                        e = check_hip(ctx, hipMemcpyAsync(rs->poll_host + ring * kCntCount * kCntStride, rs->b.counters,
                                                          kCntCount * kCntStride * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                                          ctx->stream),
                                      "drain poll");
                        if (e == FND_OK) e = check_hip(ctx, hipEventRecord(rs->poll_ev[ring], ctx->stream), "drain poll");
                        if (e != FND_OK) return e;
                        cur = nxt;
                        if (k >= kPollLag) {
                            const uint32_t old = uint32_t((k - kPollLag) % kPollRing);
                            e = check_hip(ctx, hipEventSynchronize(rs->poll_ev[old]), "drain poll");
                            if (e != FND_OK) return e;
                            // queue `old_next` is what iteration k - kPollLag left to shade
                            const uint32_t* c = rs->poll_host + old * kCntCount * kCntStride;
                            const uint32_t old_next = uint32_t(((k - kPollLag) & 1u) ? kCntExtA : kCntExtB);
                            drained = c[cnt(old_next)] == 0u && c[cnt(kCntNext)] >= W;
                        }
                    }
                }
                // the previous frame's training must have updated the cache before
                // this window reads it or gathers the next training samples
                if (apply_cache || collect) {
                    e = join_training(ctx, ctx->stream);
                    if (e != FND_OK) return e;
                }
                if (apply_cache) {
                    // the queries not answered early: [split, count), then wait for the early ones
                    launch_cache_forward(cs->view, rs->b.q_query, 1, 0, &rs->b.counters[cnt(kCntQuery)], rs->b.q_pred, nullptr, 0,
                                         cg, ctx->stream, early ? &rs->b.counters[cnt(kCntQuerySplit)] : nullptr);
                    if (early) {
                        e = check_hip(ctx, hipStreamWaitEvent(ctx->stream, rs->side_ev[1], 0), "side stream");
                        if (e != FND_OK) return e;
                    }
                    launch_cache_apply(wa, wgrid, ctx->stream);
                }
                if (collect)
                    launch_cache_collect(wa, g0, W, std::max(1u, std::min(wgrid, (wa.b.t_slots + 255u) / 256u)), ctx->stream);
                launch_accumulate(wa, g0, W, wgrid, ctx->stream);
                e = check_hip(ctx, hipGetLastError(), "render launch");
                if (e != FND_OK) return e;
            }
            return FND_OK;
        };

        st = check_hip(ctx, hipMemsetAsync(rs->b.rays, 0, sizeof(unsigned long long), ctx->stream), "clear");
        if (st != FND_OK) return st;
        (void)hipEventRecord(ctx->ev[2], ctx->stream);

        // ---- the frame; training paths among its own paths record their cache vertices
        if (any_train) {
            wa.b.t_slots = 0;
            for (uint32_t i = 0; i < num_jobs; ++i) wa.b.t_slots += hj[i].train_count;
            launch_zero_slots(wa, 1u << kCntTrain, ctx->stream);
            launch_fill_u64(wa.b.t_list, wa.b.t_slots, ~0ull, ctx->stream);
        }
        st = check_hip(ctx, hipMemcpyAsync(rs->d_jobs, hj.data(), hj.size() * sizeof(JobDev), hipMemcpyHostToDevice,
                                           ctx->stream), "upload jobs");
        if (st != FND_OK) return st;
        st = run(num_jobs, total, cap, max_depth, any_use, any_train);
        if (st != FND_OK) return st;

        // ---- optimise on the samples the frame gathered.  The steps run on the
        // training stream, overlapped with the next call until it first reads the
        // cache (join_training); only the next frame needs them, so this call's
        // time ends with its image.
        if (any_train) {
            uint32_t nsamp = 0;
            st = check_hip(ctx, hipMemcpyAsync(&nsamp, &rs->b.counters[cnt(kCntTrain)], 4, hipMemcpyDeviceToHost, ctx->stream),
                           "read sample count");
            if (st == FND_OK) st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "training paths");
            if (st != FND_OK) return st;
            nsamp = std::min(nsamp, cs->sample_cap);
            if (nsamp) {
                // step k takes samples k, k + S, k + 2S, ... (per of them): every
                // step spans the whole frame, and the steps are disjoint
                const uint32_t steps = std::min(cs->desc.train_steps, nsamp);
                const uint32_t per = std::max(1u, std::min(cs->desc.train_batch, nsamp / steps));
                const uint32_t S = nsamp / per;
                // the main stream is idle here (the sample count was read after it)
                for (uint32_t k = 0; k < steps; ++k) {
                    st = cache_step(ctx, cs, cs->s_query + k, cs->s_target + k, per, S, true, nullptr, nullptr, nullptr,
                                    ctx->train);
                    if (st != FND_OK) return st;
                }
                st = check_hip(ctx, hipEventRecord(ctx->train_ev, ctx->train), "cache training");
                if (st != FND_OK) return st;
                ctx->train_pending = true;
                cs->last_samples = nsamp;
            }
            cs->frames += 1;
        }
        (void)hipEventRecord(ctx->ev[3], ctx->stream);
        st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "render");
        if (st != FND_OK) return st;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]);
        ctx->stats.last_call_ms = ms;
        unsigned long long rays = 0;
        st = check_hip(ctx, hipMemcpyAsync(&rays, rs->b.rays, sizeof rays, hipMemcpyDeviceToHost, ctx->stream),
                       "read ray count");
        if (st == FND_OK) st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "read ray count");
        if (st != FND_OK) return st;
        ctx->stats.render_rays += rays;
        ctx->stats.paths += total;
        for (uint32_t i = 0; i < num_jobs; ++i) fbs[i]->samples += jobs[i].spp;
        return FND_OK;
    });
}

fnd_status fnd_camera_basis(const fnd_camera* camera, uint32_t width, uint32_t height, float* out14)
{
    if (!camera || !out14 || width == 0 || height == 0) return FND_E_INVALID_ARG;
    if (camera_problem(*camera)) return FND_E_INVALID_ARG;
    JobDev j{};
    camera_basis(*camera, width, height, j);
    for (int i = 0; i < 3; ++i) {
        out14[i] = j.eye[i];
        out14[3 + i] = j.fwd[i];
        out14[6 + i] = j.right[i];
        out14[9 + i] = j.up[i];
    }
    out14[12] = j.tan_half;
    out14[13] = j.aspect;
    return FND_OK;
}

fnd_status fnd_pt_camera_rays(fnd_pt_ctx* ctx, const fnd_camera* camera, uint32_t width, uint32_t height,
                              uint64_t seed, uint32_t sample, fnd_ray* out)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || !camera || !out) return FND_E_INVALID_ARG;
        if (width == 0 || height == 0 || width > 16384 || height > 16384)
            return set_error(ctx, FND_E_INVALID_ARG, "image size %ux%u out of range", width, height);
        if (const char* why = camera_problem(*camera))
            return set_error(ctx, FND_E_INVALID_ARG, "degenerate camera (%s)", why);
        JobDev j{};
        camera_basis(*camera, width, height, j);
        j.width = width;
        j.height = height;
        j.seed = seed;
        const uint64_t bytes = uint64_t(width) * height * sizeof(Ray);
        void* p = nullptr;
        fnd_status st = dev_alloc(ctx, &p, bytes);
        if (st != FND_OK) return st;
        launch_camera_rays(j, sample, static_cast<Ray*>(p), ctx->stream);
        st = check_hip(ctx, hipGetLastError(), "camera rays");
        if (st == FND_OK) st = copy_sync(ctx, out, p, bytes, hipMemcpyDeviceToHost, "camera rays");
        dev_free(ctx, p, bytes);
        return st;
    });
}

fnd_status fnd_pt_rng_draws(fnd_pt_ctx* ctx, uint64_t seed, uint32_t sample, uint32_t pixel0, uint32_t count,
                            uint32_t bounce, uint32_t draws, uint32_t* out)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || (!out && count && draws)) return FND_E_INVALID_ARG;
        if (uint64_t(count) * draws > (1ull << 28))
            return set_error(ctx, FND_E_INVALID_ARG, "at most 2^28 draws per call");
        if (uint64_t(pixel0) + count > (1ull << 32))
            return set_error(ctx, FND_E_INVALID_ARG, "pixel index overflows 32 bits");
        if (count == 0 || draws == 0) return FND_OK;
        const uint64_t bytes = uint64_t(count) * draws * sizeof(uint32_t);
        void* p = nullptr;
        fnd_status st = dev_alloc(ctx, &p, bytes);
        if (st != FND_OK) return st;
        launch_rng_draws(seed, sample, pixel0, count, bounce, draws, static_cast<uint32_t*>(p), ctx->stream);
        st = check_hip(ctx, hipGetLastError(), "generator draws");
        if (st == FND_OK) st = copy_sync(ctx, out, p, bytes, hipMemcpyDeviceToHost, "generator draws");
        dev_free(ctx, p, bytes);
        return st;
    });
}

fnd_status fnd_pt_queue_bytes(uint32_t capacity, uint64_t* bytes)
{
    if (!bytes || capacity == 0) return FND_E_INVALID_ARG;
    *bytes = queue_bytes(capacity);
    return FND_OK;
}

fnd_status fnd_pt_set_frame_arena(fnd_pt_ctx* ctx, void* memory, uint64_t bytes)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx) return FND_E_INVALID_ARG;
        if (memory && reinterpret_cast<uintptr_t>(memory) % kQueueAlign)
            return set_error(ctx, FND_E_INVALID_ARG, "frame arena must be %llu-byte aligned",
                             (unsigned long long)kQueueAlign);
        if (memory && bytes == 0) return set_error(ctx, FND_E_INVALID_ARG, "frame arena of 0 bytes");
        fnd_status st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "set frame arena");
        if (st == FND_OK) st = check_hip(ctx, hipStreamSynchronize(ctx->side), "set frame arena");
        if (st != FND_OK) return st;
        if (memory) {  // the library's block is no longer used: give it back
            if (auto* rs = static_cast<RenderState*>(ctx->render_state)) {
                dev_free(ctx, rs->owned, rs->owned_bytes);
                rs->owned = nullptr;
                rs->owned_bytes = 0;
            }
        }
        if (auto* rs = static_cast<RenderState*>(ctx->render_state)) {
            fnd_arena_reset(&rs->frame, nullptr, 0);
            rs->b = WfBuffers{};
            rs->batch = 0;
        }
        ctx->frame_memory = memory;
        ctx->frame_bytes = memory ? bytes : 0;
        return FND_OK;
    });
}

fnd_status fnd_pt_frame_arena_info(fnd_pt_ctx* ctx, fnd_frame_arena_info* out)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || !out) return FND_E_INVALID_ARG;
        std::memset(out, 0, sizeof(*out));
        const auto* rs = static_cast<const RenderState*>(ctx->render_state);
        out->caller_owned = ctx->frame_memory ? 1u : 0u;
        out->size = ctx->frame_memory ? ctx->frame_bytes : (rs ? rs->owned_bytes : 0);
        if (rs) {
            out->used = fnd_arena_used(&rs->frame);
            out->batch_paths = rs->batch;
        }
        return FND_OK;
    });
}

}  // extern "C"
