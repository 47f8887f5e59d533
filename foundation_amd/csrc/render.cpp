// render.cpp — framebuffers and the batched wavefront render entry point.
#include <cstring>
#include <vector>

#include "ctx.hpp"

using namespace fnd;

namespace fnd {
void render_state_free(fnd_pt_ctx* ctx) { (void)ctx; }
}  // namespace fnd

extern "C" {

fnd_status fnd_pt_framebuffer_create(fnd_pt_ctx* ctx, uint32_t width, uint32_t height, uint32_t flags, fnd_handle* out)
{
    if (!ctx || !out) return FND_E_INVALID_ARG;
    *out = FND_INVALID_HANDLE;
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return set_error(ctx, FND_E_INVALID_ARG, "framebuffer size %ux%u out of range", width, height);
    if (flags & ~FND_FB_FP64) return set_error(ctx, FND_E_INVALID_ARG, "unknown framebuffer flags");
    auto* f = new (std::nothrow) Framebuffer();
    if (!f) return set_error(ctx, FND_E_OOM_HOST, "out of host memory");
    f->width = width;
    f->height = height;
    f->flags = flags;
    const uint64_t n = uint64_t(width) * height * 3;
    void* p = nullptr;
    fnd_status st = dev_alloc(ctx, &p, n * sizeof(float));
    if (st != FND_OK) {
        delete f;
        return st;
    }
    f->accum = static_cast<float*>(p);
    if (flags & FND_FB_FP64) {
        st = dev_alloc(ctx, &p, n * sizeof(double));
        if (st != FND_OK) {
            dev_free(ctx, f->accum, n * sizeof(float));
            delete f;
            return st;
        }
        f->accum64 = static_cast<double*>(p);
    }
    const fnd_handle h = make_handle(kHandleFramebuffer, ctx->next_id++);
    ctx->framebuffers[h] = f;
    *out = h;
    return fnd_pt_framebuffer_clear(ctx, h);
}

fnd_status fnd_pt_framebuffer_destroy(fnd_pt_ctx* ctx, fnd_handle h)
{
    if (!ctx) return FND_E_INVALID_ARG;
    Framebuffer* f = find_framebuffer(ctx, h);
    if (!f) return set_error(ctx, FND_E_NOT_FOUND, "unknown framebuffer handle");
    (void)hipStreamSynchronize(ctx->stream);
    const uint64_t n = uint64_t(f->width) * f->height * 3;
    dev_free(ctx, f->accum, n * sizeof(float));
    dev_free(ctx, f->accum64, n * sizeof(double));
    delete f;
    ctx->framebuffers.erase(h);
    return FND_OK;
}

fnd_status fnd_pt_framebuffer_clear(fnd_pt_ctx* ctx, fnd_handle h)
{
    if (!ctx) return FND_E_INVALID_ARG;
    Framebuffer* f = find_framebuffer(ctx, h);
    if (!f) return set_error(ctx, FND_E_NOT_FOUND, "unknown framebuffer handle");
    const uint64_t n = uint64_t(f->width) * f->height * 3;
    fnd_status st = check_hip(ctx, hipMemsetAsync(f->accum, 0, n * sizeof(float), ctx->stream), "clear");
    if (st == FND_OK && f->accum64)
        st = check_hip(ctx, hipMemsetAsync(f->accum64, 0, n * sizeof(double), ctx->stream), "clear");
    if (st != FND_OK) return st;
    f->samples = 0;
    return check_hip(ctx, hipStreamSynchronize(ctx->stream), "clear");
}

fnd_status fnd_pt_framebuffer_read(fnd_pt_ctx* ctx, fnd_handle h, float* rgb, uint32_t* out_samples)
{
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
}

fnd_status fnd_pt_framebuffer_read_f64(fnd_pt_ctx* ctx, fnd_handle h, double* rgb, uint32_t* out_samples)
{
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
}

fnd_status fnd_pt_render(fnd_pt_ctx* ctx, const fnd_render_job* jobs, uint32_t num_jobs)
{
    if (!ctx || (!jobs && num_jobs)) return FND_E_INVALID_ARG;
    return set_error(ctx, FND_E_UNSUPPORTED, "render is not built yet");
}

}  // extern "C"
