// cache.hip — radiance-cache kernels: hash-grid encoding, fused MLP forward
// and backward on MFMA, loss, weight gradients, Adam.
//
// Layout (lanes of a wave64, checked on gfx950): for v_mfma_f32_16x16xK
// with E = K/4 elements per lane,
//   A[m][k] sits in lane m + 16*(k/E), element k%E;
//   B[k][n] in lane n + 16*(k/E), element k%E;
//   D[m][n] in lane n + 16*(m/4), element m%4.
// The MLP runs "transposed": Z = W^T X with the weights as A and 16 samples as
// the 16 columns of B.  A D tile (16 neurons x 16 samples) then already sits
// where the next layer's B operand wants it, once the K index of the next
// layer is permuted: K=32 slot 8g+e holds neuron 32s+4g+e for e<4 and neuron
// 32s+16+4g+(e-4) for e>=4 (perm32 below).  The permutation is folded into the
// weight blocks, so activations never leave registers between layers.
#include "cache.hpp"

#include <algorithm>

namespace fnd {

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

namespace {

constexpr uint32_t kPrime1 = 2654435761u;
constexpr uint32_t kPrime2 = 805459861u;

__host__ __device__ __forceinline__ int perm32(int q)
{
    const int g = q >> 3, e = q & 7;
    return e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4);
}

__device__ __forceinline__ uint16_t bf16_bits(float x)
{
    return __builtin_bit_cast(uint16_t, static_cast<__bf16>(x));
}

__device__ __forceinline__ float bf16_float(uint16_t b)
{
    return __uint_as_float(uint32_t(b) << 16);
}

// Corner indices and trilinear weights of one level (Instant NGP, eq. 4):
// scaled position x*N, corners floor(.) + {0,1}^3, dense index while
// (N+1)^3 <= T, else the xor-of-primes spatial hash modulo T.
__device__ __forceinline__ void level_corners(float px, float py, float pz, int res, bool dense, uint32_t idx[8],
                                              float w[8])
{
    const float sx = px * float(res), sy = py * float(res), sz = pz * float(res);
    int ix = int(floorf(sx)), iy = int(floorf(sy)), iz = int(floorf(sz));
    ix = min(max(ix, 0), res - 1);
    iy = min(max(iy, 0), res - 1);
    iz = min(max(iz, 0), res - 1);
    const float fx = sx - float(ix), fy = sy - float(iy), fz = sz - float(iz);
    const uint32_t stride = uint32_t(res + 1);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const uint32_t cx = uint32_t(ix + (c & 1)), cy = uint32_t(iy + ((c >> 1) & 1)), cz = uint32_t(iz + (c >> 2));
        w[c] = ((c & 1) ? fx : 1.0f - fx) * (((c >> 1) & 1) ? fy : 1.0f - fy) * ((c >> 2) ? fz : 1.0f - fz);
        idx[c] = dense ? cx + cy * stride + cz * stride * stride : ((cx) ^ (cy * kPrime1) ^ (cz * kPrime2)) & (kCacheT - 1u);
    }
}

__device__ __forceinline__ void normalise(const CacheView& v, float4 p, float& x, float& y, float& z)
{
    x = fminf(fmaxf((p.x - v.lo[0]) * v.inv_extent, 0.0f), 1.0f);
    y = fminf(fmaxf((p.y - v.lo[1]) * v.inv_extent, 0.0f), 1.0f);
    z = fminf(fmaxf((p.z - v.lo[2]) * v.inv_extent, 0.0f), 1.0f);
}

// Hash features of levels 4g..4g+3 (B slots 8g..8g+7) for one sample.  The
// level loop is fully unrolled so every vector element index is a constant.
__device__ __forceinline__ h8 encode4(const CacheView& v, float x, float y, float z, int g)
{
    float f[8];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int l = 4 * g + q;
        uint32_t idx[8];
        float w[8];
        level_corners(x, y, z, v.res[l], (v.dense_mask >> l) & 1u, idx, w);
        float f0 = 0.0f, f1 = 0.0f;
        const uint32_t* tab = v.grid16 + size_t(l) * kCacheT;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const uint32_t hv = tab[idx[c]];
            const _Float16 a = __builtin_bit_cast(_Float16, uint16_t(hv & 0xffffu));
            const _Float16 b = __builtin_bit_cast(_Float16, uint16_t(hv >> 16));
            f0 = fmaf(w[c], float(a), f0);
            f1 = fmaf(w[c], float(b), f1);
        }
        f[2 * q] = f0;
        f[2 * q + 1] = f1;
    }
    h8 out;
#pragma unroll
    for (int e = 0; e < 8; ++e) out[e] = _Float16(f[e]);
    return out;
}

// Extra inputs 32..47 as a K=32 B operand (slots 16..31 are zero padding):
// lanes of group 0 hold oriented normal, 1, normalised position, 0; lanes of
// group 1 the outgoing direction, roughness, glossy flag (roughness > 0), 0.
__device__ __forceinline__ h8 extras8(float x, float y, float z, const CacheQuery& q, int g)
{
    h8 e = {0, 0, 0, 0, 0, 0, 0, 0};
    if (g == 0) {
        e[0] = _Float16(q.nrm.x);
        e[1] = _Float16(q.nrm.y);
        e[2] = _Float16(q.nrm.z);
        e[3] = _Float16(1.0f);
        e[4] = _Float16(x);
        e[5] = _Float16(y);
        e[6] = _Float16(z);
    } else if (g == 1) {
        e[0] = _Float16(q.dir.x);
        e[1] = _Float16(q.dir.y);
        e[2] = _Float16(q.dir.z);
        e[3] = _Float16(q.pos.w);
        e[4] = _Float16(q.pos.w > 0.0f ? 1.0f : 0.0f);
    }
    return e;
}

// forward weight block offsets (halfs)
__device__ __forceinline__ int fwd_l0(int r, int s) { return (r * 2 + s) * 512; }
__device__ __forceinline__ int fwd_hid(int l, int r, int s) { return 4096 + (l - 1) * 4096 + (r * 2 + s) * 512; }
__device__ __forceinline__ int fwd_out(int s) { return 20480 + s * 512; }
// backward weight block offsets (bf16)
__device__ __forceinline__ int bwd_out(int r) { return r * 512; }
__device__ __forceinline__ int bwd_hid(int l, int r, int s) { return 2048 + (l - 1) * 4096 + (r * 2 + s) * 512; }
__device__ __forceinline__ int bwd_l0(int r, int s) { return 18432 + (r * 2 + s) * 512; }

constexpr int kP = 1;  // 16-sample tiles per wave

template <bool TRAIN>
__global__ __launch_bounds__(256) void k_cache_forward(CacheView v, const CacheQuery* q, uint32_t qstride,
                                                       uint32_t n_arg, const uint32_t* n_dev, float4* pred,
                                                       uint16_t* act, uint32_t stride, const uint32_t* lo_dev)
{
    __shared__ h8 wl[kMlpFwdElems / 8];
    for (int i = threadIdx.x; i < kMlpFwdElems / 8; i += 256) wl[i] = reinterpret_cast<const h8*>(v.wfwd)[i];
    __syncthreads();
    const _Float16* w = reinterpret_cast<const _Float16*>(wl);
    const uint32_t n = n_dev ? *n_dev : n_arg;
    const uint32_t lo = lo_dev ? *lo_dev : 0u;
    const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
    const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6), nwaves = gridDim.x * 4;
    for (uint32_t base = lo + wave * 16u * kP; base < n; base += nwaves * 16u * kP) {
        h8 bh[kP];
        h8 bx[kP];
#pragma unroll
        for (int t = 0; t < kP; ++t) {
            const uint32_t s = base + 16u * t + col;
            float x = 0, y = 0, z = 0;
            if (s < n) {
                const CacheQuery cq = q[size_t(s) * qstride];
                normalise(v, cq.pos, x, y, z);
                bh[t] = encode4(v, x, y, z, g);
                bx[t] = extras8(x, y, z, cq, g);
            } else {
                bh[t] = h8{0, 0, 0, 0, 0, 0, 0, 0};
                bx[t] = h8{0, 0, 0, 0, 0, 0, 0, 0};
            }
            if (TRAIN && s < stride) {
#pragma unroll
                for (int e = 0; e < 8; ++e) act[size_t(8 * g + e) * stride + s] = bf16_bits(float(bh[t][e]));
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    if (g < 2) act[size_t(32 + 8 * g + e) * stride + s] = bf16_bits(float(bx[t][e]));
            }
        }
        f4 d[kP][4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const h8 a0 = *reinterpret_cast<const h8*>(w + fwd_l0(r, 0) + lane * 8);
            const h8 a1 = *reinterpret_cast<const h8*>(w + fwd_l0(r, 1) + lane * 8);
#pragma unroll
            for (int t = 0; t < kP; ++t) {
                f4 acc = {0, 0, 0, 0};
                acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, bh[t], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, bx[t], acc, 0, 0, 0);
                d[t][r] = acc;
            }
        }
        h8 b[kP][2];
        auto relu_pack = [&](int layer) {
#pragma unroll
            for (int t = 0; t < kP; ++t) {
                const uint32_t s = base + 16u * t + col;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const float h = fmaxf(d[t][r][j], 0.0f);
                        b[t][r >> 1][(r & 1) * 4 + j] = _Float16(h);
                        if (TRAIN && s < stride)
                            act[size_t(kCacheIn + layer * kCacheHid + 16 * r + 4 * g + j) * stride + s] =
                                bf16_bits(s < n ? h : 0.0f);
                    }
                }
            }
        };
        relu_pack(0);
#pragma unroll 1
        for (int l = 1; l <= 4; ++l) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const h8 a0 = *reinterpret_cast<const h8*>(w + fwd_hid(l, r, 0) + lane * 8);
                const h8 a1 = *reinterpret_cast<const h8*>(w + fwd_hid(l, r, 1) + lane * 8);
#pragma unroll
                for (int t = 0; t < kP; ++t) {
                    f4 acc = {0, 0, 0, 0};
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, b[t][0], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, b[t][1], acc, 0, 0, 0);
                    d[t][r] = acc;
                }
            }
            relu_pack(l);
        }
        const h8 o0 = *reinterpret_cast<const h8*>(w + fwd_out(0) + lane * 8);
        const h8 o1 = *reinterpret_cast<const h8*>(w + fwd_out(1) + lane * 8);
#pragma unroll
        for (int t = 0; t < kP; ++t) {
            f4 acc = {0, 0, 0, 0};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(o0, b[t][0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(o1, b[t][1], acc, 0, 0, 0);
            const uint32_t s = base + 16u * t + col;
            if (g == 0 && s < n) pred[s] = make_float4(acc[0], acc[1], acc[2], 0.0f);
        }
    }
}

// Relative L2 (NRC): sum over RGB of (p - y)^2 / (sg(p)^2 + 0.01); the
// gradient row dZ5 = 2 (p - y) / (p^2 + 0.01) in bf16, row 3 = 0.
__global__ __launch_bounds__(256) void k_cache_loss(const float4* pred, const float4* target, uint32_t tstride,
                                                    uint32_t n, uint16_t* dz, uint32_t stride, float* loss)
{
    __shared__ float red[256];
    float acc = 0.0f;
    for (uint32_t s = blockIdx.x * 256u + threadIdx.x; s < stride; s += gridDim.x * 256u) {
        float gr[3] = {0, 0, 0};
        if (s < n) {
            const float4 p = pred[s], y = target[size_t(s) * tstride];
            const float pv[3] = {p.x, p.y, p.z}, yv[3] = {y.x, y.y, y.z};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float den = pv[c] * pv[c] + 0.01f;
                const float e = pv[c] - yv[c];
                acc += e * e / den;
                gr[c] = 2.0f * e / den;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) dz[size_t(kDz5Row + c) * stride + s] = bf16_bits(gr[c]);
        dz[size_t(kDz5Row + 3) * stride + s] = 0;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// F32: the grid gradient is summed with FP32 atomics (two per corner),
// otherwise with one packed bf16 atomic per corner.
template <bool F32>
__global__ __launch_bounds__(256) void k_cache_backward(CacheView v, const CacheQuery* q, uint32_t qstride, uint32_t n,
                                                        const uint16_t* act, uint16_t* dz, uint32_t stride,
                                                        uint32_t* grad_grid)
{
    __shared__ s4 wl[kMlpBwdElems / 4];
    for (int i = threadIdx.x; i < kMlpBwdElems / 4; i += 256) wl[i] = reinterpret_cast<const s4*>(v.wbwd)[i];
    __syncthreads();
    const uint16_t* w = reinterpret_cast<const uint16_t*>(wl);
    const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
    const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6), nwaves = gridDim.x * 4;
    const uint32_t n32 = (n + 31u) & ~31u;  // the forward pass wrote rows up to here
    for (uint32_t base = wave * 16u * kP; base < n32; base += nwaves * 16u * kP) {
        f4 d[kP][4];
        // dH4 = W5 dZ5 (K = 32: the 4 dZ5 rows sit in slots 0..3 of lane
        // group 0, the other slots are zero)
        b8 bz[kP];
#pragma unroll
        for (int t = 0; t < kP; ++t) {
            const uint32_t s = base + 16u * t + col;
            b8 z = {};
            if (g == 0) {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    z[e] = __builtin_bit_cast(__bf16, dz[size_t(kDz5Row + e) * stride + s]);
            }
            bz[t] = z;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const b8 a = *reinterpret_cast<const b8*>(w + bwd_out(r) + lane * 8);
#pragma unroll
            for (int t = 0; t < kP; ++t) {
                f4 acc = {0, 0, 0, 0};
                d[t][r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bz[t], acc, 0, 0, 0);
            }
        }
        b8 b[kP][2];
        // layers 4 .. 0: mask by H_l > 0, store dZ_l, propagate
#pragma unroll 1
        for (int l = 4; l >= 0; --l) {
#pragma unroll
            for (int t = 0; t < kP; ++t) {
                const uint32_t s = base + 16u * t + col;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int neuron = 16 * r + 4 * g + j;
                        const uint16_t hb = act[size_t(kCacheIn + l * kCacheHid + neuron) * stride + s];
                        const float dzv = bf16_float(hb) > 0.0f ? d[t][r][j] : 0.0f;
                        const uint16_t zb = bf16_bits(dzv);
                        dz[size_t(l * kCacheHid + neuron) * stride + s] = zb;
                        b[t][r >> 1][(r & 1) * 4 + j] = __builtin_bit_cast(__bf16, zb);
                    }
                }
            }
            if (l > 0) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const b8 a0 = *reinterpret_cast<const b8*>(w + bwd_hid(l, r, 0) + lane * 8);
                    const b8 a1 = *reinterpret_cast<const b8*>(w + bwd_hid(l, r, 1) + lane * 8);
#pragma unroll
                    for (int t = 0; t < kP; ++t) {
                        f4 acc = {0, 0, 0, 0};
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b[t][0], acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b[t][1], acc, 0, 0, 0);
                        d[t][r] = acc;
                    }
                }
            }
        }
        // dX0 (hash features 0..31) = W0 dZ0, then scatter into the table
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const b8 a0 = *reinterpret_cast<const b8*>(w + bwd_l0(r, 0) + lane * 8);
            const b8 a1 = *reinterpret_cast<const b8*>(w + bwd_l0(r, 1) + lane * 8);
#pragma unroll
            for (int t = 0; t < kP; ++t) {
                f4 acc = {0, 0, 0, 0};
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b[t][0], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b[t][1], acc, 0, 0, 0);
                d[t][r] = acc;
            }
        }
#pragma unroll
        for (int t = 0; t < kP; ++t) {
            const uint32_t s = base + 16u * t + col;
            if (s >= n) continue;
            float x, y, z;
            normalise(v, q[size_t(s) * qstride].pos, x, y, z);
#pragma unroll
            for (int r = 0; r < 2; ++r) {
#pragma unroll 1
                for (int h = 0; h < 2; ++h) {
                    const int l = 8 * r + 2 * g + h;  // features 16r + 4g + 2h, +1
                    const float g0 = d[t][r][2 * h], g1 = d[t][r][2 * h + 1];
                    if (g0 == 0.0f && g1 == 0.0f) continue;
                    uint32_t idx[8];
                    float wc[8];
                    level_corners(x, y, z, v.res[l], (v.dense_mask >> l) & 1u, idx, wc);
                    if (F32) {
                        // FP32 sums: two global_atomic_add_f32 per corner
                        auto* gf = (__attribute__((address_space(1))) float*)(grad_grid) + size_t(l) * kCacheT * 2;
#pragma unroll
                        for (int c = 0; c < 8; ++c) {
                            (void)__builtin_amdgcn_global_atomic_fadd_f32(gf + 2 * size_t(idx[c]), wc[c] * g0);
                            (void)__builtin_amdgcn_global_atomic_fadd_f32(gf + 2 * size_t(idx[c]) + 1, wc[c] * g1);
                        }
                    } else {
                        // one packed bf16 atomic per corner carries both features
                        // (global_atomic_pk_add_bf16): half the atomics and half the
                        // table bytes of two FP32 adds
                        auto* gl = (__attribute__((address_space(1))) s2*)(grad_grid + size_t(l) * kCacheT);
#pragma unroll
                        for (int c = 0; c < 8; ++c) {
                            s2 val;
                            val[0] = short(bf16_bits(wc[c] * g0));
                            val[1] = short(bf16_bits(wc[c] * g1));
                            (void)__builtin_amdgcn_global_atomic_fadd_v2bf16(gl + idx[c], val);
                        }
                    }
                }
            }
        }
    }
}

// dW_l[in][out] = sum over samples of X_l[in][s] dZ_l[out][s] (K = samples).
__global__ __launch_bounds__(256) void k_cache_dw(const uint16_t* act, const uint16_t* dz, uint32_t n,
                                                  uint32_t stride, uint32_t chunk, float* grad)
{
    const int layer = blockIdx.y;
    const int in_rows = layer == 0 ? kCacheIn : kCacheHid;
    const int out_cols = layer == 5 ? 16 : kCacheHid;
    const int x_row0 = layer == 0 ? 0 : kCacheIn + (layer - 1) * kCacheHid;
    const int z_row0 = layer == 5 ? kDz5Row : layer * kCacheHid;
    const int z_rows = layer == 5 ? 4 : kCacheHid;
    const int w_off = layer == 0 ? kW0 : (layer == 5 ? kW5 : kW1 + (layer - 1) * 4096);
    const int mt = in_rows / 16, nt = out_cols / 16;
    const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15, wv = threadIdx.x >> 6;
    const uint32_t n32 = (n + 31u) & ~31u;
    const uint32_t s0 = blockIdx.x * chunk, s1 = min(n32, s0 + chunk);
    for (int tile = wv; tile < mt * nt; tile += 4) {
        const int mi = tile / nt, ni = tile % nt;
        const uint16_t* xa = act + size_t(x_row0 + 16 * mi + col) * stride;
        const int zr = 16 * ni + col;
        const uint16_t* zb = zr < z_rows ? dz + size_t(z_row0 + zr) * stride : nullptr;
        f4 acc = {0, 0, 0, 0};
        for (uint32_t k0 = s0; k0 < s1; k0 += 32u) {
            const b8 a = *reinterpret_cast<const b8*>(xa + k0 + 8 * g);
            b8 bb = {};
            if (zb) bb = *reinterpret_cast<const b8*>(zb + k0 + 8 * g);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int in = 16 * mi + 4 * g + j, out = 16 * ni + col;
            atomicAdd(grad + w_off + in * out_cols + out, acc[j]);
        }
    }
}

// Adam (Kingma & Ba 2015) with bias-correction factors bc1 = 1 - b1^t,
// bc2 = 1 - b2^t.  MODE 0: plain FP32 gradients (the MLP).  MODE 1 and 2
// (the hash grid): gradients arrive per table entry as one packed bf16 pair
// (1) or two FP32 values (2); an entry whose two gradients are both 0 is left
// alone, moments included (Mueller et al. 2022, section 4: updates of
// zero-gradient table entries are skipped); every gradient read is cleared,
// so the next step starts from zero without a separate pass; the FP16 copy
// the forward pass reads is refreshed.
template <int MODE>
__global__ __launch_bounds__(256) void k_adam(float* p, void* gr, float* m, float* v, uint64_t n, float lr,
                                              float b1, float b2, float eps, float bc1, float bc2, uint32_t* out_half2)
{
    constexpr bool GRID = MODE != 0;
    const uint64_t pairs = n / 2;
    for (uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x; i < pairs; i += uint64_t(gridDim.x) * 256u) {
        float2 gg;
        if (MODE == 1) {
            uint32_t* gb = static_cast<uint32_t*>(gr);
            const uint32_t w = gb[i];
            if ((w & 0x7FFF7FFFu) == 0u) continue;  // both +-0: untouched this step
            gb[i] = 0u;
            gg = make_float2(__builtin_bit_cast(float, w << 16), __builtin_bit_cast(float, w & 0xFFFF0000u));
        } else if (MODE == 2) {
            float2* gf = static_cast<float2*>(gr);
            gg = gf[i];
            if (gg.x == 0.0f && gg.y == 0.0f) continue;  // untouched this step
            gf[i] = make_float2(0.0f, 0.0f);
        } else {
            gg = static_cast<const float2*>(gr)[i];
        }
        float2 pp = reinterpret_cast<float2*>(p)[i];
        float2 mm = reinterpret_cast<float2*>(m)[i];
        float2 vv = reinterpret_cast<float2*>(v)[i];
        mm.x = b1 * mm.x + (1.0f - b1) * gg.x;
        mm.y = b1 * mm.y + (1.0f - b1) * gg.y;
        vv.x = b2 * vv.x + (1.0f - b2) * gg.x * gg.x;
        vv.y = b2 * vv.y + (1.0f - b2) * gg.y * gg.y;
        pp.x -= lr * (mm.x / bc1) / (sqrtf(vv.x / bc2) + eps);
        pp.y -= lr * (mm.y / bc1) / (sqrtf(vv.y / bc2) + eps);
        reinterpret_cast<float2*>(p)[i] = pp;
        reinterpret_cast<float2*>(m)[i] = mm;
        reinterpret_cast<float2*>(v)[i] = vv;
        if (GRID) {
            const uint32_t lo = __builtin_bit_cast(uint16_t, _Float16(pp.x));
            const uint32_t hi = __builtin_bit_cast(uint16_t, _Float16(pp.y));
            out_half2[i] = lo | (hi << 16);
        }
    }
}

// Plain fp32 weights -> forward (fp16) and backward (bf16) operand blocks.
// A block is one 16 x 32 A operand: element e of lane L is A[m][k] with
// m = L & 15, k = 8 (L >> 4) + e.
__global__ __launch_bounds__(256) void k_arrange(const float* wt, _Float16* wfwd, uint16_t* wbwd)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int o = i % 512, lane = o / 8, e = o % 8, m = lane & 15, k = 8 * (lane >> 4) + e;
    if (i < kMlpFwdElems) {
        float val;
        if (i < 4096) {  // layer 0: per r, hash inputs 0..31 then extras 32..47 (+16 zero slots)
            const int r = i / 1024, s = (i / 512) % 2;
            if (s == 0)
                val = wt[kW0 + k * 64 + 16 * r + m];
            else
                val = k < 16 ? wt[kW0 + (32 + k) * 64 + 16 * r + m] : 0.0f;
        } else if (i < 20480) {
            const int j = i - 4096, l = 1 + j / 4096, rs = (j % 4096) / 512, r = rs / 2, s = rs % 2;
            val = wt[kW1 + (l - 1) * 4096 + (32 * s + perm32(k)) * 64 + 16 * r + m];
        } else {
            const int s = (i - 20480) / 512;
            val = wt[kW5 + (32 * s + perm32(k)) * 16 + m];
        }
        wfwd[i] = _Float16(val);
    }
    if (i < kMlpBwdElems) {
        float val;
        if (i < 2048) {  // W5 as A (64 x 16 outputs, +16 zero slots), per r
            const int r = i / 512;
            val = k < 16 ? wt[kW5 + (16 * r + m) * 16 + k] : 0.0f;
        } else if (i < 18432) {
            const int j = i - 2048, l = 1 + j / 4096, rs = (j % 4096) / 512, r = rs / 2, s = rs % 2;
            val = wt[kW1 + (l - 1) * 4096 + (16 * r + m) * 64 + 32 * s + perm32(k)];
        } else {
            const int rs = (i - 18432) / 512, r = rs / 2, s = rs % 2;
            val = wt[kW0 + (16 * r + m) * 64 + 32 * s + perm32(k)];
        }
        wbwd[i] = bf16_bits(val);
    }
}

__global__ __launch_bounds__(256) void k_zero(float* p, uint64_t n)
{
    for (uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256u) p[i] = 0.0f;
}

}  // namespace

void launch_cache_forward(const CacheView& v, const CacheQuery* q, uint32_t qstride, uint32_t n, const uint32_t* n_dev,
                          float4* pred, uint16_t* act, uint32_t act_stride, uint32_t grid, hipStream_t s,
                          const uint32_t* lo_dev)
{
    if (act)
        hipLaunchKernelGGL(k_cache_forward<true>, dim3(grid), dim3(256), 0, s, v, q, qstride, n, n_dev, pred, act,
                           act_stride, lo_dev);
    else
        hipLaunchKernelGGL(k_cache_forward<false>, dim3(grid), dim3(256), 0, s, v, q, qstride, n, n_dev, pred, act, 0u,
                           lo_dev);
}

void launch_cache_loss(const float4* pred, const float4* target, uint32_t tstride, uint32_t n, uint16_t* dz,
                       uint32_t stride, float* loss, hipStream_t s)
{
    const uint32_t grid = std::min<uint32_t>(1024u, (stride + 255u) / 256u);
    hipLaunchKernelGGL(k_cache_loss, dim3(grid), dim3(256), 0, s, pred, target, tstride, n, dz, stride, loss);
}

void launch_cache_backward(const CacheView& v, const CacheQuery* q, uint32_t qstride, uint32_t n, const uint16_t* act,
                           uint16_t* dz, uint32_t stride, uint32_t* grad_grid, bool fp32, hipStream_t s)
{
    // one 16-sample tile per wave: the scatter's atomics want every CU
    const uint32_t grid = std::max(1u, std::min<uint32_t>(4096u, (n + 63u) / 64u));
    if (fp32)
        hipLaunchKernelGGL(k_cache_backward<true>, dim3(grid), dim3(256), 0, s, v, q, qstride, n, act, dz, stride,
                           grad_grid);
    else
        hipLaunchKernelGGL(k_cache_backward<false>, dim3(grid), dim3(256), 0, s, v, q, qstride, n, act, dz, stride,
                           grad_grid);
}

void launch_cache_dw(const uint16_t* act, const uint16_t* dz, uint32_t n, uint32_t stride, float* grad_mlp,
                     hipStream_t s)
{
    const uint32_t chunk = 1024;
    const uint32_t chunks = std::max(1u, (n + chunk - 1) / chunk);
    hipLaunchKernelGGL(k_cache_dw, dim3(chunks, 6), dim3(256), 0, s, act, dz, n, stride, chunk, grad_mlp);
}

static dim3 adam_grid(uint64_t n)
{
    const uint64_t blocks = std::min<uint64_t>(8192u, (n / 2 + 255u) / 256u);
    return dim3(uint32_t(std::max<uint64_t>(1, blocks)));
}

void launch_adam(float* p, const float* g, float* m, float* v, uint64_t n, float lr, float b1, float b2, float eps,
                 float bc1, float bc2, hipStream_t s)
{
    hipLaunchKernelGGL(k_adam<0>, adam_grid(n), dim3(256), 0, s, p, const_cast<float*>(g), m, v, n, lr, b1, b2,
                       eps, bc1, bc2, nullptr);
}

void launch_adam_grid(float* p, uint32_t* g, bool fp32, float* m, float* v, uint64_t n, float lr, float b1, float b2,
                      float eps, float bc1, float bc2, uint32_t* out_half2, hipStream_t s)
{
    if (fp32)
        hipLaunchKernelGGL(k_adam<2>, adam_grid(n), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps, bc1, bc2,
                           out_half2);
    else
        hipLaunchKernelGGL(k_adam<1>, adam_grid(n), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps, bc1, bc2,
                           out_half2);
}

void launch_arrange(const float* w, _Float16* wfwd, uint16_t* wbwd, hipStream_t s)
{
    hipLaunchKernelGGL(k_arrange, dim3((std::max(kMlpFwdElems, kMlpBwdElems) + 255) / 256), dim3(256), 0, s, w, wfwd,
                       wbwd);
}

void launch_zero(float* p, uint64_t n, hipStream_t s)
{
    const uint64_t blocks = std::min<uint64_t>(8192u, (n + 255u) / 256u);
    hipLaunchKernelGGL(k_zero, dim3(uint32_t(std::max<uint64_t>(1, blocks))), dim3(256), 0, s, p, n);
}

}  // namespace fnd
