// cache_state.hpp — host-side state of the radiance cache (one per context).
#pragma once
#include <utility>
#include <vector>

#include "../../include/fnd_pt.h"
#include "cache.hpp"

struct fnd_pt_ctx;

namespace fnd {

constexpr uint32_t kCacheBatch = 1u << 16;          // samples per optimiser step (buffer size)
constexpr uint32_t kCacheTrainMaxPaths = 1u << 16;  // training paths per frame
constexpr uint32_t kCacheMaxRecords = 16;           // diffuse vertices recorded per training path

struct CacheState {
    fnd_cache_desc desc{};
    std::vector<std::pair<void*, uint64_t>> allocs;
    float* w_master = nullptr;
    float* w_m = nullptr;
    float* w_v = nullptr;
    _Float16* wfwd = nullptr;
    uint16_t* wbwd = nullptr;
    float* g_master = nullptr;
    float* g_m = nullptr;
    float* g_v = nullptr;
    uint32_t* grid16 = nullptr;
    CacheTrainBufs tb{};
    CacheQuery* io_query = nullptr;   // host API staging
    float4* io_target = nullptr;
    CacheQuery* s_query = nullptr;    // training samples gathered from traced paths
    float4* s_target = nullptr;
    uint32_t sample_cap = 0;
    float4* rec = nullptr;            // per training path: kCacheMaxRecords records of kCacheRecordF4 float4
    uint32_t* rec_count = nullptr;
    unsigned long long* t_list = nullptr;  // per training slot: global path index (kCacheTrainMaxPaths)
    CacheView view{};
    bool bounds_set = false;
    bool grid_grad_dirty = false;     // grid gradient holds values no update has consumed
    uint32_t step = 0;
    uint32_t last_samples = 0;        // training samples gathered in the last training frame
    uint64_t frames = 0;
};

CacheState* cache_get(fnd_pt_ctx* ctx);
void cache_free(fnd_pt_ctx* ctx);
void cache_set_bounds(CacheState* cs, const float* lo, const float* hi, bool pad);
// One optimiser step on n samples q[s * sample_stride], target[s * sample_stride],
// queued on stream s (the host waits on s only to return gradients or the loss).
fnd_status cache_step(fnd_pt_ctx* ctx, CacheState* cs, const CacheQuery* q, const float4* target, uint32_t n,
                      uint32_t sample_stride, bool apply, float* dmlp_host, float* dgrid_host, float* loss_host,
                      hipStream_t s);
// Stream s waits for the render-queued training (a no-op when none is queued).
fnd_status join_training(fnd_pt_ctx* ctx, hipStream_t s);

}  // namespace fnd
