// cache_api.cpp — radiance-cache state, optimiser step and the fnd_pt_cache_* ABI.
//
// The reference has no learned component; this is the MI355X build's own.
// Parameters live on the device as FP32 masters with Adam moments; the
// kernels read an FP16 hash table and FP16/BF16 MLP weight blocks that are
// refreshed after every step.  A training step is:
//   forward (FP16 MFMA, stores bf16 activations) -> relative-L2 loss ->
//   backward (BF16 MFMA, scatters hash-grid gradients) -> weight gradients
//   (BF16 MFMA over samples) -> Adam on MLP and grid -> FP16/BF16 refresh.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "cache_state.hpp"
#include "ctx.hpp"

using namespace fnd;

namespace fnd {

namespace {

template <class T>
fnd_status calloc_dev(fnd_pt_ctx* ctx, CacheState* cs, T** p, uint64_t count)
{
    void* q = nullptr;
    const uint64_t bytes = std::max<uint64_t>(count, 1) * sizeof(T);
    fnd_status st = dev_alloc(ctx, &q, bytes);
    if (st != FND_OK) return st;
    cs->allocs.emplace_back(q, bytes);
    *p = static_cast<T*>(q);
    return check_hip(ctx, hipMemsetAsync(q, 0, bytes, ctx->stream), "clear cache buffer");
}

fnd_status validate(fnd_pt_ctx* ctx, const fnd_cache_desc& d)
{
    if (d.levels != kCacheLevels || d.features != kCacheFeatures || d.log2_table != kCacheLog2T ||
        d.base_resolution != 16 || d.max_resolution != 2048 || d.hidden != kCacheHid ||
        d.hidden_layers != kCacheHidLayers)
        return set_error(ctx, FND_E_UNSUPPORTED,
                         "cache kernels are built for 16 levels x 2 features, 2^19 entries, resolutions 16..2048, "
                         "5 x 64 hidden");
    if (!(d.learning_rate > 0.0f) || d.train_paths == 0 || d.train_paths > kCacheTrainMaxPaths || d.train_steps == 0 ||
        d.train_steps > 64 || d.train_depth == 0 || d.train_depth > kCacheMaxRecords || d.train_batch < 256 ||
        d.train_batch > kCacheBatch || !(d.glossy_min_roughness >= 0.0f) || d.grid_gradient > FND_CACHE_GRAD_FP32)
        return set_error(ctx, FND_E_INVALID_ARG, "cache: learning_rate > 0, train_paths in 1..65536, "
                                                 "train_steps in 1..64, train_depth in 1..16, train_batch in 256..65536, "
                                                 "grid_gradient FND_CACHE_GRAD_BF16 or FND_CACHE_GRAD_FP32");
    return FND_OK;
}

}  // namespace

void cache_free(fnd_pt_ctx* ctx)
{
    auto* cs = static_cast<CacheState*>(ctx->cache);
    if (!cs) return;
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);  // inference may read the weights there
    finish_training(ctx);
    for (auto& a : cs->allocs) dev_free(ctx, a.first, a.second);
    ctx_delete(ctx, cs);
    ctx->cache = nullptr;
}

CacheState* cache_get(fnd_pt_ctx* ctx) { return static_cast<CacheState*>(ctx->cache); }

// The grid covers the cube [lo, lo + extent], extent = largest side of the
// box.  Bounds taken from a scene are first widened by 2% about the centre.
void cache_set_bounds(CacheState* cs, const float* lo, const float* hi, bool pad)
{
    float ext = 0.0f;
    for (int a = 0; a < 3; ++a) ext = std::max(ext, hi[a] - lo[a]);
    ext = std::max(ext, 1e-6f);
    if (pad) ext *= 1.02f;
    for (int a = 0; a < 3; ++a) cs->view.lo[a] = pad ? 0.5f * (lo[a] + hi[a]) - 0.5f * ext : lo[a];
    cs->view.inv_extent = 1.0f / ext;
    cs->bounds_set = true;
}

// One optimiser step on n device-resident samples.
fnd_status join_training(fnd_pt_ctx* ctx, hipStream_t s)
{
    if (!ctx->train_pending) return FND_OK;
    return check_hip(ctx, hipStreamWaitEvent(s, ctx->train_ev, 0), "wait for cache training");
}

void finish_training(fnd_pt_ctx* ctx)
{
    if (ctx->train) (void)hipStreamSynchronize(ctx->train);
    ctx->train_pending = false;
}

fnd_status cache_step(fnd_pt_ctx* ctx, CacheState* cs, const CacheQuery* q, const float4* target, uint32_t n,
                      uint32_t sample_stride, bool apply, float* dmlp_host, float* dgrid_host, float* loss_host,
                      hipStream_t s)
{
    if (n == 0) return FND_OK;
    if (n > cs->tb.cap) return set_error(ctx, FND_E_INVALID_ARG, "cache batch of %u exceeds %u", n, cs->tb.cap);
    launch_zero(cs->tb.grad_mlp, kMlpParams, s);
    // the lazy grid update clears what it reads; only a step without update
    // (gradients for the caller) leaves the grid gradient to be cleared here
    const bool fp32 = cs->desc.grid_gradient == FND_CACHE_GRAD_FP32;
    if (cs->grid_grad_dirty) launch_zero(reinterpret_cast<float*>(cs->tb.grad_grid), fp32 ? kGridParams : kGridParams / 2, s);
    cs->grid_grad_dirty = !apply;
    launch_zero(cs->tb.loss, 1, s);
    const uint32_t grid = std::max(1u, std::min<uint32_t>(ctx->num_cus * 4u, (n + 127u) / 128u));
    launch_cache_forward(cs->view, q, sample_stride, n, nullptr, cs->tb.pred, cs->tb.act, cs->tb.cap, grid, s);
    launch_cache_loss(cs->tb.pred, target, sample_stride, n, cs->tb.dz, cs->tb.cap, cs->tb.loss, s);
    launch_cache_backward(cs->view, q, sample_stride, n, cs->tb.act, cs->tb.dz, cs->tb.cap, cs->tb.grad_grid, fp32, s);
    launch_cache_dw(cs->tb.act, cs->tb.dz, n, cs->tb.cap, cs->tb.grad_mlp, s);
    fnd_status st = check_hip(ctx, hipGetLastError(), "cache step");
    if (st != FND_OK) return st;
    if (apply) {
        cs->step += 1;
        const float b1 = 0.9f, b2 = 0.99f, eps = 1e-15f;
        const float bc1 = 1.0f - std::pow(b1, float(cs->step)), bc2 = 1.0f - std::pow(b2, float(cs->step));
        launch_adam(cs->w_master, cs->tb.grad_mlp, cs->w_m, cs->w_v, kMlpParams, cs->desc.learning_rate, b1, b2, eps,
                    bc1, bc2, s);
        launch_adam_grid(cs->g_master, cs->tb.grad_grid, fp32, cs->g_m, cs->g_v, kGridParams, cs->desc.learning_rate,
                         b1, b2, eps, bc1, bc2, cs->grid16, s);
        launch_arrange(cs->w_master, cs->wfwd, cs->wbwd, s);
        st = check_hip(ctx, hipGetLastError(), "cache update");
        if (st != FND_OK) return st;
    }
    if (loss_host || dmlp_host || dgrid_host) {
        if (loss_host) st = check_hip(ctx, hipMemcpyAsync(loss_host, cs->tb.loss, 4, hipMemcpyDeviceToHost, s), "read loss");
        if (st == FND_OK && dmlp_host)
            st = check_hip(ctx, hipMemcpyAsync(dmlp_host, cs->tb.grad_mlp, kMlpParams * 4, hipMemcpyDeviceToHost, s),
                           "read gradients");
        // an FP32 grid gradient is copied as it is; a bf16 one into the front
        // half of the caller's array, then widened in place from the back
        // (element i's source bytes [2i, 2i+2) lie at or after its
        // destination's start and are read first)
        if (st == FND_OK && dgrid_host)
            st = check_hip(ctx, hipMemcpyAsync(dgrid_host, cs->tb.grad_grid, kGridParams * (fp32 ? 4 : 2),
                                               hipMemcpyDeviceToHost, s),
                           "read gradients");
        if (st == FND_OK) st = check_hip(ctx, hipStreamSynchronize(s), "cache step");
        if (st == FND_OK && dgrid_host && !fp32) {
            auto* bytes = reinterpret_cast<unsigned char*>(dgrid_host);
            for (uint64_t i = kGridParams; i-- > 0;) {
                uint16_t h;
                std::memcpy(&h, bytes + 2 * i, 2);
                const uint32_t w = uint32_t(h) << 16;
                std::memcpy(bytes + 4 * i, &w, 4);
            }
        }
    }
    return st;
}

}  // namespace fnd

namespace {

fnd_status upload_queries(fnd_pt_ctx* ctx, CacheState* cs, const float* queries, const float* targets, uint32_t n)
{
    if (n > cs->tb.cap) return set_error(ctx, FND_E_INVALID_ARG, "at most %u queries per call", cs->tb.cap);
    std::vector<CacheQuery> hq(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = queries + 10ull * i;
        hq[i].pos = make_float4(p[0], p[1], p[2], p[9]);
        hq[i].nrm = make_float4(p[3], p[4], p[5], 0.0f);
        hq[i].dir = make_float4(p[6], p[7], p[8], 0.0f);
    }
    fnd_status st = check_hip(ctx, hipMemcpyAsync(cs->io_query, hq.data(), n * sizeof(CacheQuery), hipMemcpyHostToDevice,
                                                  ctx->stream), "upload queries");
    if (st != FND_OK) return st;
    if (targets) {
        std::vector<float4> ht(n);
        for (uint32_t i = 0; i < n; ++i)
            ht[i] = make_float4(targets[3ull * i], targets[3ull * i + 1], targets[3ull * i + 2], 0.0f);
        st = check_hip(ctx, hipMemcpyAsync(cs->io_target, ht.data(), n * sizeof(float4), hipMemcpyHostToDevice,
                                           ctx->stream), "upload targets");
        if (st != FND_OK) return st;
    }
    return check_hip(ctx, hipStreamSynchronize(ctx->stream), "upload");
}

fnd_status need_cache(fnd_pt_ctx* ctx, CacheState** out)
{
    if (!ctx) return FND_E_INVALID_ARG;
    finish_training(ctx);
    *out = cache_get(ctx);
    if (!*out) return set_error(ctx, FND_E_INVALID_ARG, "radiance cache is not configured");
    if (!(*out)->bounds_set) return set_error(ctx, FND_E_INVALID_ARG, "radiance cache has no bounds yet");
    return FND_OK;
}

}  // namespace

extern "C" {

fnd_status fnd_pt_cache_default_desc(fnd_cache_desc* d)
{
    if (!d) return FND_E_INVALID_ARG;
    std::memset(d, 0, sizeof(*d));
    d->enable = 1;
    d->levels = kCacheLevels;
    d->features = kCacheFeatures;
    d->log2_table = kCacheLog2T;
    d->base_resolution = 16;
    d->max_resolution = 2048;
    d->hidden = kCacheHid;
    d->hidden_layers = kCacheHidLayers;
    d->learning_rate = 1e-2f;
    d->train_paths = 1u << 16;
    d->train_steps = 2;
    d->train_depth = 1;
    d->train_batch = 1u << 14;
    d->glossy_min_roughness = 0.5f;
    d->glossy_any_bounce = 1;
    d->seed = 1;
    return FND_OK;
}

fnd_status fnd_pt_cache_configure(fnd_pt_ctx* ctx, const fnd_cache_desc* d)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx || !d) return FND_E_INVALID_ARG;
        if (!d->enable) {
            cache_free(ctx);
            return FND_OK;
        }
        fnd_status st = validate(ctx, *d);
        if (st != FND_OK) return st;
        cache_free(ctx);
        auto* cs = ctx_new<CacheState>(ctx);
        if (!cs) return set_error(ctx, FND_E_OOM_HOST, "out of host memory");
        ctx->cache = cs;
        cs->desc = *d;
        auto fail = [&](fnd_status e) {
            cache_free(ctx);
            return e;
        };
        if ((st = calloc_dev(ctx, cs, &cs->w_master, kMlpParams)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &cs->w_m, kMlpParams)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &cs->w_v, kMlpParams)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &cs->wfwd, kMlpFwdElems)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &cs->wbwd, kMlpBwdElems)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &cs->g_master, kGridParams)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &cs->g_m, kGridParams)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &cs->g_v, kGridParams)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &cs->grid16, kGridParams / 2)) != FND_OK) return fail(st);
        CacheTrainBufs& tb = cs->tb;
        tb.cap = kCacheBatch;
        if ((st = calloc_dev(ctx, cs, &tb.act, uint64_t(kActRows) * tb.cap)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &tb.dz, uint64_t(kDzRows) * tb.cap)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &tb.pred, tb.cap)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &tb.grad_mlp, kMlpParams)) != FND_OK) return fail(st);
        // room for the FP32 form (two floats per entry); the bf16 form uses the first half
        if ((st = calloc_dev(ctx, cs, &tb.grad_grid, kGridParams)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &tb.loss, 1)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &cs->io_query, tb.cap)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &cs->io_target, tb.cap)) != FND_OK) return fail(st);
        cs->sample_cap = kCacheTrainMaxPaths * d->train_depth;
        if ((st = calloc_dev(ctx, cs, &cs->s_query, cs->sample_cap)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &cs->s_target, cs->sample_cap)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &cs->rec, uint64_t(kCacheTrainMaxPaths) * kCacheMaxRecords * kCacheRecordF4)) != FND_OK)
            return fail(st);
        if ((st = calloc_dev(ctx, cs, &cs->rec_count, kCacheTrainMaxPaths)) != FND_OK) return fail(st);
        if ((st = calloc_dev(ctx, cs, &cs->t_list, kCacheTrainMaxPaths)) != FND_OK) return fail(st);
        // level resolutions: N_l = floor(N_min * b^l), b = (N_max / N_min)^(1 / (L - 1))
        const double b = std::exp((std::log(2048.0) - std::log(16.0)) / double(kCacheLevels - 1));
        cs->view.dense_mask = 0;
        for (int l = 0; l < kCacheLevels; ++l) {
            const int r = int(std::floor(16.0 * std::pow(b, double(l)) + 1e-9));
            cs->view.res[l] = r;
            const double cells = double(r + 1) * double(r + 1) * double(r + 1);
            if (cells <= double(kCacheT)) cs->view.dense_mask |= 1u << l;
        }
        cs->view.grid16 = cs->grid16;
        cs->view.wfwd = cs->wfwd;
        cs->view.wbwd = cs->wbwd;
        if (d->bounds_hi[0] > d->bounds_lo[0] || d->bounds_hi[1] > d->bounds_lo[1] || d->bounds_hi[2] > d->bounds_lo[2])
            cache_set_bounds(cs, d->bounds_lo, d->bounds_hi, false);
        // initial parameters: MLP uniform +-sqrt(6 / (fan_in + fan_out)) (Glorot), grid uniform +-1e-4
        std::mt19937_64 rng(d->seed);
        std::vector<float> w(kMlpParams);
        auto fill = [&](int off, int fin, int fout) {
            const float lim = std::sqrt(6.0f / float(fin + fout));
            std::uniform_real_distribution<float> U(-lim, lim);
            for (int i = 0; i < fin * fout; ++i) w[off + i] = U(rng);
        };
        fill(kW0, kCacheIn, kCacheHid);
        for (int l = 0; l < 4; ++l) fill(kW1 + l * 4096, kCacheHid, kCacheHid);
        fill(kW5, kCacheHid, 16);
        std::vector<float> gridv(kGridParams);
        std::uniform_real_distribution<float> G(-1e-4f, 1e-4f);
        for (auto& x : gridv) x = G(rng);
        return fnd_pt_cache_set_params(ctx, w.data(), gridv.data());
    });
}

fnd_status fnd_pt_cache_param_count(fnd_pt_ctx* ctx, uint64_t* mlp, uint64_t* grid)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx) return FND_E_INVALID_ARG;
        if (mlp) *mlp = kMlpParams;
        if (grid) *grid = kGridParams;
        return FND_OK;
    });
}

fnd_status fnd_pt_cache_set_params(fnd_pt_ctx* ctx, const float* mlp, const float* grid)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx) return FND_E_INVALID_ARG;
        finish_training(ctx);
        CacheState* cs = cache_get(ctx);
        if (!cs) return set_error(ctx, FND_E_INVALID_ARG, "radiance cache is not configured");
        fnd_status st = FND_OK;
        if (mlp) {
            st = check_hip(ctx, hipMemcpyAsync(cs->w_master, mlp, kMlpParams * 4, hipMemcpyHostToDevice, ctx->stream),
                           "upload weights");
            if (st != FND_OK) return st;
            launch_arrange(cs->w_master, cs->wfwd, cs->wbwd, ctx->stream);
        }
        if (grid) {
            std::vector<uint32_t> h16(kGridParams / 2);
            for (uint64_t i = 0; i < kGridParams / 2; ++i) {
                const _Float16 a = _Float16(grid[2 * i]), b = _Float16(grid[2 * i + 1]);
                uint16_t ua, ub;
                std::memcpy(&ua, &a, 2);
                std::memcpy(&ub, &b, 2);
                h16[i] = uint32_t(ua) | (uint32_t(ub) << 16);
            }
            st = check_hip(ctx, hipMemcpyAsync(cs->g_master, grid, kGridParams * 4, hipMemcpyHostToDevice, ctx->stream),
                           "upload grid");
            if (st == FND_OK)
                st = check_hip(ctx, hipMemcpyAsync(cs->grid16, h16.data(), kGridParams * 2, hipMemcpyHostToDevice, ctx->stream),
                               "upload grid");
            if (st == FND_OK) st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "upload grid");
            if (st != FND_OK) return st;
        }
        // a new parameter set restarts the optimiser
        launch_zero(cs->w_m, kMlpParams, ctx->stream);
        launch_zero(cs->w_v, kMlpParams, ctx->stream);
        launch_zero(cs->g_m, kGridParams, ctx->stream);
        launch_zero(cs->g_v, kGridParams, ctx->stream);
        cs->step = 0;
        return check_hip(ctx, hipStreamSynchronize(ctx->stream), "set params");
    });
}

fnd_status fnd_pt_cache_get_params(fnd_pt_ctx* ctx, float* mlp, float* grid)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx) return FND_E_INVALID_ARG;
        finish_training(ctx);
        CacheState* cs = cache_get(ctx);
        if (!cs) return set_error(ctx, FND_E_INVALID_ARG, "radiance cache is not configured");
        fnd_status st = FND_OK;
        if (mlp)
            st = check_hip(ctx, hipMemcpyAsync(mlp, cs->w_master, kMlpParams * 4, hipMemcpyDeviceToHost, ctx->stream),
                           "read weights");
        if (st == FND_OK && grid)
            st = check_hip(ctx, hipMemcpyAsync(grid, cs->g_master, kGridParams * 4, hipMemcpyDeviceToHost, ctx->stream),
                           "read grid");
        if (st != FND_OK) return st;
        return check_hip(ctx, hipStreamSynchronize(ctx->stream), "read params");
    });
}

fnd_status fnd_pt_cache_infer(fnd_pt_ctx* ctx, const float* queries, uint32_t n, float* out)
{
    return guard(ctx, [&]() -> fnd_status {
        CacheState* cs;
        fnd_status st = need_cache(ctx, &cs);
        if (st != FND_OK) return st;
        if (n && (!queries || !out)) return set_error(ctx, FND_E_INVALID_ARG, "null array");
        for (uint32_t b0 = 0; b0 < n; b0 += cs->tb.cap) {
            const uint32_t m = std::min(cs->tb.cap, n - b0);
            st = upload_queries(ctx, cs, queries + 10ull * b0, nullptr, m);
            if (st != FND_OK) return st;
            const uint32_t grid = std::max(1u, std::min<uint32_t>(ctx->num_cus * 4u, (m + 127u) / 128u));
            launch_cache_forward(cs->view, cs->io_query, 1, m, nullptr, cs->tb.pred, nullptr, 0, grid, ctx->stream);
            std::vector<float4> hp(m);
            st = check_hip(ctx, hipMemcpyAsync(hp.data(), cs->tb.pred, m * sizeof(float4), hipMemcpyDeviceToHost, ctx->stream),
                           "read predictions");
            if (st == FND_OK) st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "infer");
            if (st != FND_OK) return st;
            for (uint32_t i = 0; i < m; ++i) {
                out[3ull * (b0 + i)] = hp[i].x;
                out[3ull * (b0 + i) + 1] = hp[i].y;
                out[3ull * (b0 + i) + 2] = hp[i].z;
            }
        }
        return FND_OK;
    });
}

fnd_status fnd_pt_cache_gradients(fnd_pt_ctx* ctx, const float* queries, const float* targets, uint32_t n, float* loss,
                                  float* dmlp, float* dgrid)
{
    return guard(ctx, [&]() -> fnd_status {
        CacheState* cs;
        fnd_status st = need_cache(ctx, &cs);
        if (st != FND_OK) return st;
        if (!queries || !targets || n == 0) return set_error(ctx, FND_E_INVALID_ARG, "empty batch");
        st = upload_queries(ctx, cs, queries, targets, n);
        if (st != FND_OK) return st;
        float l = 0;
        st = cache_step(ctx, cs, cs->io_query, cs->io_target, n, 1, false, dmlp, dgrid, &l, ctx->stream);
        if (loss) *loss = l;
        return st;
    });
}

fnd_status fnd_pt_cache_train(fnd_pt_ctx* ctx, const float* queries, const float* targets, uint32_t n, uint32_t steps,
                              float* loss)
{
    return guard(ctx, [&]() -> fnd_status {
        CacheState* cs;
        fnd_status st = need_cache(ctx, &cs);
        if (st != FND_OK) return st;
        if (!queries || !targets || n == 0) return set_error(ctx, FND_E_INVALID_ARG, "empty batch");
        st = upload_queries(ctx, cs, queries, targets, n);
        if (st != FND_OK) return st;
        float l = 0;
        for (uint32_t k = 0; k < steps && st == FND_OK; ++k)
            st = cache_step(ctx, cs, cs->io_query, cs->io_target, n, 1, true, nullptr, nullptr, k + 1 == steps ? &l : nullptr,
                            ctx->stream);
        if (loss) *loss = l;
        return st;
    });
}

}  // extern "C"
