// cache.hpp — the learned radiance cache: sizes, device views, launchers.
//
// A 16-level hash-grid encoding (Mueller, Evans, Schied, Keller, "Instant
// Neural Graphics Primitives with a Multiresolution Hash Encoding", 2022)
// feeds a 5 x 64 ReLU MLP without biases (Mueller, Rousselle, Novak, Keller,
// "Real-time Neural Radiance Caching for Path Tracing", 2021) that predicts
// the albedo-demodulated indirect radiance leaving a diffuse surface point.
//
// Shapes are fixed at compile time so every layer maps onto whole MFMA tiles:
//   input  48 = 16 levels x 2 features (hash) + 16 extras
//               (oriented normal, 1, normalised position, 0, outgoing
//               direction, roughness, glossy flag, zeros)
//   hidden 64 x 5 (ReLU), output 16 rows of which 3 are RGB.
// Forward uses v_mfma_f32_16x16x32_f16 with FP16 weights and activations,
// FP32 accumulation; backward uses v_mfma_f32_16x16x32_bf16.  Only these two
// opcodes are used: chaining the accumulator of one MFMA opcode into SrcC of
// a different one (16x16x32 into 16x16x16) read stale accumulators on gfx950.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fnd {

constexpr int kCacheLevels = 16;
constexpr int kCacheFeatures = 2;
constexpr int kCacheLog2T = 19;
constexpr uint32_t kCacheT = 1u << kCacheLog2T;
constexpr int kCacheIn = 48;
constexpr int kCacheHid = 64;
constexpr int kCacheOut = 16;
constexpr int kCacheHidLayers = 5;
// plain fp32 weight layout: W0 [48][64], W1..W4 [64][64], W5 [64][16]
constexpr int kW0 = 0, kW1 = 3072, kW5 = 19456;
constexpr int kMlpParams = 20480;
// MFMA operand blocks, every layer on the K=32 opcodes (the 16 extras of
// layer 0 and the 16 outputs of layer 5 are zero-padded to K = 32):
//   forward  : layer 0 4 x (512 hash + 512 extras), hidden 4 x 4096, out 1024
//   backward : W5 4 x 512, hidden 4 x 4096, W0 rows 0..31 4 x 512
constexpr int kMlpFwdElems = 21504;
constexpr int kMlpBwdElems = 20480;
constexpr uint64_t kGridParams = uint64_t(kCacheLevels) * kCacheT * kCacheFeatures;
// training activation rows ([row][sample], bf16): X0 48 rows, H0..H4 64 rows
constexpr int kActRows = kCacheIn + kCacheHidLayers * kCacheHid;  // 368
// gradient rows: dZ0..dZ4 64 rows, dZ5 4 rows (RGB + zero)
constexpr int kDzRows = kCacheHidLayers * kCacheHid + 4;         // 324
constexpr int kDz5Row = kCacheHidLayers * kCacheHid;

struct CacheView {
    const uint32_t* grid16;     // [L][T] packed half2 features
    const _Float16* wfwd;       // forward weight blocks, kMlpFwdElems
    const uint16_t* wbwd;       // backward weight blocks (bf16 bits), kMlpBwdElems
    float lo[3];
    float inv_extent;           // position -> [0,1]^3
    int res[kCacheLevels];      // grid resolution per level
    uint32_t dense_mask;        // bit l: level l indexes densely
};

// A query: world position (w = GGX roughness, 0 for a Lambert vertex),
// oriented normal, outgoing direction (0 for a Lambert vertex).
struct CacheQuery {
    float4 pos;
    float4 nrm;
    float4 dir;
};

// A training record: position (w = roughness), oriented normal, outgoing
// direction, weight (throughput * albedo), radiance mark (w = 1 once set).
constexpr uint32_t kCacheRecordF4 = 5;

struct CacheTrainBufs {
    uint16_t* act;              // [kActRows][cap] bf16
    uint16_t* dz;               // [kDzRows][cap] bf16
    float4* pred;               // [cap]
    float* grad_mlp;            // kMlpParams
    uint32_t* grad_grid;        // kGridParams words: per entry two FP32 features, or (bf16 form, first
                                // half) the two features as packed bf16
    float* loss;                // 1
    uint32_t cap;               // samples the buffers hold (multiple of 32)
};

// qstride / tstride: sample s reads q[s * qstride] (and target[s * tstride]).
// The samples are [*lo_dev, *n_dev) when those are given, else [0, n).
void launch_cache_forward(const CacheView& v, const CacheQuery* q, uint32_t qstride, uint32_t n, const uint32_t* n_dev,
                          float4* pred, uint16_t* act, uint32_t act_stride, uint32_t grid, hipStream_t s,
                          const uint32_t* lo_dev = nullptr);
void launch_cache_loss(const float4* pred, const float4* target, uint32_t tstride, uint32_t n, uint16_t* dz,
                       uint32_t stride, float* loss, hipStream_t s);
void launch_cache_backward(const CacheView& v, const CacheQuery* q, uint32_t qstride, uint32_t n, const uint16_t* act,
                           uint16_t* dz, uint32_t stride, uint32_t* grad_grid, bool fp32, hipStream_t s);
void launch_cache_dw(const uint16_t* act, const uint16_t* dz, uint32_t n, uint32_t stride, float* grad_mlp,
                     hipStream_t s);
// lazy: skip pairs whose gradients are both 0 and clear the gradients read
// Adam over n FP32 parameters with FP32 gradients (the MLP).
void launch_adam(float* p, const float* g, float* m, float* v, uint64_t n, float lr, float b1, float b2, float eps,
                 float bc1, float bc2, hipStream_t s);
// Lazy Adam over the hash grid: gradients are packed bf16 pairs; pairs with a
// zero gradient keep their moments, the others are updated, their gradient
// cleared and their FP16 copy refreshed into out_half2.
void launch_adam_grid(float* p, uint32_t* g, bool fp32, float* m, float* v, uint64_t n, float lr, float b1, float b2,
                      float eps, float bc1, float bc2, uint32_t* out_half2, hipStream_t s);
void launch_arrange(const float* w, _Float16* wfwd, uint16_t* wbwd, hipStream_t s);
void launch_zero(float* p, uint64_t n, hipStream_t s);

}  // namespace fnd
