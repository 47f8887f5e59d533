// wavefront.hpp — device-side tables and buffers of the wavefront path tracer.
#pragma once
#include <hip/hip_runtime.h>

#include "cache.hpp"
#include "pt_layout.hpp"
#include "scene_encode.hpp"

namespace fnd {

// JobDev.flags: FND_JOB_USE_CACHE (1) from the caller.
//
// Radiance-cache training paths are a subset of a job's own paths, chosen by
// an affine bijection of the path's index q = sample_in_call * W * H + pixel:
// t = (q * train_mul + train_add) mod train_mod with gcd(train_mul,
// train_mod) = 1, and the path trains iff t < train_count.  The set has
// exactly train_count members, and t is the path's record slot (plus
// train_base), so selection needs no atomics and the image does not depend on
// scheduling.
constexpr uint32_t kNoTrainSlot = 0xFFFFFFFFu;

struct JobDev {
    uint32_t scene_slot;
    uint32_t width, height;
    uint32_t tile;            // pixel order: T x T tiles, row-major tiles
    uint32_t tiles_x;
    uint32_t padded;          // tiles_x * tiles_y * T * T paths per sample
    uint32_t max_depth;
    uint32_t rr_depth;
    uint32_t sample_offset;
    uint32_t flags;
    uint64_t seed;
    uint32_t train_base;      // first record slot of this job's training paths
    uint32_t train_count;     // training paths of this job in this call (0: none)
    uint64_t train_mod;       // W * H * spp
    uint64_t train_mul;       // coprime with train_mod
    uint64_t train_add;       // < train_mod, changes every call
    uint64_t train_barrett;   // floor((2^64 - 1) / train_mod), for the reduction mod train_mod
    uint64_t path_base;       // first global path index of this job in the render call
    uint64_t path_end;
    float eye[3];
    float tan_half;
    float fwd[3];
    float aspect;
    float right[3];
    float pad0;
    float up[3];
    float pad1;
    float* accum;             // W*H*3
    double* accum64;          // W*H*3 or null
};

struct ShadeScene {
    const float4* tris;
    const InstanceRec* insts;
    const DeviceMaterial* materials;
    const DeviceLight* lights;
    uint32_t num_lights;
    uint32_t two_level;
    float background[3];
    float pad;
};

// Counter slots (u32) of one render call.
enum WfCounter : uint32_t {
    kCntExtA = 0,
    kCntExtB = 1,
    kCntShadow = 2,  // shadow queue; 2 and 3 used alternately by the iterations of a window
    // 4: unused (the traversal's work counter is WfBuffers::trace_ctr)
    kCntQuery = 5,
    kCntTrain = 6,
    kCntNext = 7,    // next window path to start (may pass the window's end)
    kCntFree0 = 8,   // free slots, two lists used alternately
    kCntFree1 = 9,
    kCntQuerySplit = 10,  // radiance-cache queries answered early, on the side stream
    kCntCount = 16
};
// Each counter slot has its own 256-byte line: atomics to one line serialise
// (measured on MI355X: 88 M/s for one line, whether the atomics hit one word
// or several words of it), and the shade and regeneration launches append to
// several queues at once.
constexpr uint32_t kCntStride = 64;
__host__ __device__ constexpr uint32_t cnt(uint32_t slot) { return slot * kCntStride; }

// Threads per block of the stages that append to the queues: each block
// takes one atomic per queue per pass over its share, so larger blocks mean
// fewer atomics on the hot counter lines, but a block's waves also meet at
// the append's barriers.  Generation (54 VGPRs, two 1024-thread blocks per
// CU) took 103 -> 60 us per 1080p frame with 1024 threads.  The shade is
// not bound by its appends: per 1080p frame it took 565 us in 256-thread
// blocks, 620 in 512 and 731 in 1024 (one block per CU, its 16 waves in step
// at the barriers).
constexpr uint32_t kGenerateBlock = 1024;
constexpr uint32_t kShadeBlock = 256;
constexpr uint32_t kStageBlock = 256;  // regeneration, shadow resolve

// Paths are rendered in windows of consecutive global path indices; a window
// holds at most kWindowSlots times as many paths as there are slots, and a
// slot whose path ends takes the window's next path (regeneration), so the
// traversal launches stay full until the window's last paths drain.
constexpr uint32_t kWindowSlots = 8;

struct WfBuffers {
    // per path slot
    // The small fields are packed per slot: slots are read and written in
    // queue order, i.e. scattered, so one record touches one line where
    // separate arrays touched one line each.
    uint4* p_site;            // (job, pixel: y * W + x or 0xFFFFFFFF for padding, sample,
                              //  gid: the slot's path, its index in the window)
    uint2* p_state;           // (training record slot or kNoTrainSlot, vertices shaded so far)
    float4* p_thr;            // throughput, w = solid-angle pdf of the last BSDF sample (0: camera)
    float4* p_L;              // radiance
    // per window path: final radiance (cache term added after the window drains)
    float4* w_L;
    uint32_t window;          // paths in the current window
    uint32_t retire;          // 1: the window regenerates, ended paths are retired and their slots freed
    // slots whose path has ended and been recorded, free for the next
    // iteration (kCntFree0 + i; the two lists alternate)
    uint32_t* free_list[2];
    // extension queues (double-buffered) and hits
    Ray* ext_rays[2];
    uint32_t* ext_path[2];
    uint32_t* ext_scene[2];
    float4* ext_hits;
    // shadow queue
    Ray* sh_rays;
    uint32_t* sh_path;
    uint32_t* sh_scene;
    float4* sh_contrib;
    float4* sh_hits;
    uint32_t* counters;       // kCntCount slots, kCntStride words apart (index with cnt())
    uint32_t* trace_ctr;      // the traversal launches' work counter (kCounterWords)
    unsigned long long* rays; // rays traced (extension + shadow), cumulative: added by the trace launches
    // radiance-cache queries (at most one per path per batch)
    CacheQuery* q_query;
    uint32_t* q_path;         // window index of the querying path
    float4* q_weight;         // throughput * albedo at the query vertex
    float4* q_pred;
    // training records: per training slot, rec_max records of kCacheRecordF4 float4
    float4* rec;
    uint32_t* rec_count;
    unsigned long long* t_list; // per training slot: the global path index (~0 until generated)
    uint32_t t_slots;         // training slots of this call
    uint32_t rec_max;
    float glossy_min;         // GGX at least this rough queries and trains the cache like Lambert
    uint32_t glossy_any_bounce; // from this bounce on every GGX vertex does
    CacheQuery* t_query;      // gathered training samples
    float4* t_target;
    uint32_t t_cap;
};

struct WfArgs {
    const JobDev* jobs;
    uint32_t num_jobs;
    const ShadeScene* scenes;
    WfBuffers b;
};

void launch_generate(const WfArgs& a, uint64_t g0, uint32_t n, uint32_t grid, hipStream_t s);
void launch_regen(const WfArgs& a, uint64_t g0, uint32_t fidx, uint32_t next, uint32_t grid, hipStream_t s);
// out[y * pitch + x] for pixel y * width + x (pitch in pixels, >= width)
// relMSE of x (FP32 sums, times xinv) against y (FP64 sums, times yinv) over
// n values into out[0], with kRelMseParts partial sums in part.
constexpr uint32_t kRelMseParts = 1024;
void launch_rel_mse(const float* x, float xinv, const double* y, double yinv, uint64_t n, double eps, double* part,
                    double* out, hipStream_t s);
void launch_present(const float* accum, uint64_t pixels, float inv, float exposure, const float* thresholds,
                    uint32_t* out, uint32_t width, uint32_t pitch, hipStream_t s);
void launch_prep(const WfArgs& a, uint32_t next, uint32_t free_zero, uint32_t shadow, hipStream_t s);
void launch_shade(const WfArgs& a, uint32_t cur, uint32_t fw, uint32_t shadow, uint32_t grid, hipStream_t s);
// zero_ext, zero_shadow, zero_free: counter slots the launch clears for the
// next iteration (kCntCount: none), and with them the traversal's work
// counter (when zero_shadow is a slot).
void launch_shadow_resolve(const WfArgs& a, uint32_t fw, uint32_t shadow, uint32_t zero_ext, uint32_t zero_shadow,
                           uint32_t zero_free, uint32_t grid, hipStream_t s);
void launch_accumulate(const WfArgs& a, uint64_t g0, uint32_t n, uint32_t grid, hipStream_t s);
void launch_cache_apply(const WfArgs& a, uint32_t grid, hipStream_t s);
void launch_zero_slots(const WfArgs& a, uint32_t mask, hipStream_t s);
void launch_cache_collect(const WfArgs& a, uint64_t g0, uint32_t n, uint32_t grid, hipStream_t s);
void launch_fill_u64(unsigned long long* p, uint32_t n, unsigned long long v, hipStream_t s);
// counters[dst] = counters[src]
void launch_snapshot(const WfArgs& a, uint32_t src, uint32_t dst, hipStream_t s);
// Diagnostics: every pixel's camera ray of one sample; keyed generator outputs.
void launch_camera_rays(const JobDev& job, uint32_t sample, Ray* out, hipStream_t s);
void launch_rng_draws(uint64_t seed, uint32_t sample, uint32_t pixel0, uint32_t npix, uint32_t bounce, uint32_t ndraws,
                      uint32_t* out, hipStream_t s);

}  // namespace fnd
