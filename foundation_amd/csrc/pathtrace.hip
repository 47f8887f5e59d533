// pathtrace.hip — wavefront path tracer stages (generate, shade, shadow
// resolve, accumulate).  Extension and shadow rays are traced by the
// persistent traversal kernel (trace_kernel.hpp) between these stages.
//
// Estimator (unidirectional path tracing, pbrt-v3 conventions): at the vertex
// reached after `bounce` segments, emission is added with the power-heuristic
// weight against light sampling; the path stops there if bounce == max_depth;
// otherwise one light sample (next-event estimation, shadow ray) and one BSDF
// sample extend it, with Russian roulette from bounce rr_depth on
// (survival q = min(max throughput, 0.95)).
//
// Appends to the next extension queue and to the shadow queue are compacted
// per wave: a ballot of the lanes that append, one atomicAdd per wave, and
// each lane's slot from the popcount of the ballot below it.
#include <algorithm>

#include "pt_device.hpp"
#include "pt_shading.hpp"
#include "wavefront.hpp"

namespace fnd {

namespace {

__device__ __forceinline__ uint32_t lanes_below_mask(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

// Wave-compacted append: returns this lane's slot (valid only where pred).
__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t* counter)
{
    const uint64_t mask = __ballot(pred);
    if (mask == 0) return 0;
    const int leader = __builtin_ctzll(mask);
    uint32_t base = 0;
    if (int(__lane_id()) == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
    base = uint32_t(__builtin_amdgcn_readlane(int(base), leader));
    return base + lanes_below_mask(mask);
}

// Block-compacted append to Q queues at once, for blocks of BLOCK threads:
// one global atomic per queue per block instead of one per wave.  The queue
// counters are single hot addresses that every block of a launch hits, and
// device-scope atomics on one address serialise (about 88 M/s on MI355X): a
// stage over 2M paths in 256-thread blocks takes 8k atomics on one counter,
// about 93 us.  Every thread of the block must call it; returns each lane's
// slots (valid only where its predicate holds).
template <int Q, uint32_t BLOCK>
__device__ __forceinline__ void block_append(const bool (&pred)[Q], uint32_t* const (&counter)[Q], uint32_t (&slot)[Q])
{
    constexpr uint32_t kStageWaves = BLOCK / 64;
    __shared__ uint32_t s_cnt[Q][kStageWaves];
    __shared__ uint32_t s_base[Q];
    const uint32_t w = threadIdx.x >> 6;
    uint64_t m[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) m[q] = __ballot(pred[q]);
    if (__lane_id() == 0) {
#pragma unroll
        for (int q = 0; q < Q; ++q) s_cnt[q][w] = uint32_t(__popcll(m[q]));
    }
    __syncthreads();
    if (threadIdx.x < uint32_t(Q)) {
        const uint32_t q = threadIdx.x;
        uint32_t total = 0;
        for (uint32_t v = 0; v < kStageWaves; ++v) total += s_cnt[q][v];
        s_base[q] = total ? atomicAdd(counter[q], total) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        uint32_t before = 0;
        for (uint32_t v = 0; v < w; ++v) before += s_cnt[q][v];
        slot[q] = s_base[q] + before + lanes_below_mask(m[q]);
    }
    __syncthreads();  // s_cnt and s_base are reused by the next call
}

__device__ __forceinline__ uint32_t find_job(const JobDev* jobs, uint32_t num_jobs, uint64_t g)
{
    uint32_t lo = 0, hi = num_jobs;  // last job with path_base <= g
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].path_base <= g) lo = mid; else hi = mid;
    }
    return lo;
}

// ---- light selection adapted to the shading point
//
// With at most kAdaptiveLights lights, NEE picks light l with probability
// 0.9 * w_l / sum(w) + 0.1 * (its power share), where w_l is the light's
// power times a geometric estimate of how much of it reaches the point:
// the largest cosine, over the light's four corners, at the receiver's
// normal (0 when the whole quad lies behind the tangent plane: the cosine
// numerator is linear over the quad), times the point's distance from the
// light plane (0 behind a one-sided light) over d^3 to the light's centre
// (d^2 at least a quarter of the area).  The power floor keeps every light
// possible, so the estimator stays unbiased.  MIS needs the same probability
// again when a BSDF-sampled ray hits the light: both evaluations take the
// offset ray origin and the normal quantised to 16+16-bit octahedral
// (Cigolle et al., JCGT 3(2), 2014), whose code the path keeps in p_L.w.
constexpr uint32_t kAdaptiveLights = 8;

__device__ __forceinline__ uint32_t oct_encode(V3 n)
{
    const float s = fabsf(n.x) + fabsf(n.y) + fabsf(n.z);
    float u = n.x / s, v = n.y / s;
    if (n.z < 0.0f) {
        const float ou = u;
        u = (1.0f - fabsf(v)) * (ou >= 0.0f ? 1.0f : -1.0f);
        v = (1.0f - fabsf(ou)) * (v >= 0.0f ? 1.0f : -1.0f);
    }
    const uint32_t qu = uint32_t(fminf(fmaxf(u * 0.5f + 0.5f, 0.0f), 1.0f) * 65535.0f + 0.5f);
    const uint32_t qv = uint32_t(fminf(fmaxf(v * 0.5f + 0.5f, 0.0f), 1.0f) * 65535.0f + 0.5f);
    return qu | (qv << 16);
}

__device__ __forceinline__ V3 oct_decode(uint32_t c)
{
    float u = float(c & 0xFFFFu) * (2.0f / 65535.0f) - 1.0f;
    float v = float(c >> 16) * (2.0f / 65535.0f) - 1.0f;
    const float z = 1.0f - fabsf(u) - fabsf(v);
    if (z < 0.0f) {
        const float ou = u;
        u = (1.0f - fabsf(v)) * (ou >= 0.0f ? 1.0f : -1.0f);
        v = (1.0f - fabsf(ou)) * (v >= 0.0f ? 1.0f : -1.0f);
    }
    return normalize(v3(u, v, z));
}

__device__ __forceinline__ float light_weight(const DeviceLight& lt, V3 p, V3 n)
{
    const V3 c0 = v3(lt.corner[0], lt.corner[1], lt.corner[2]);
    const V3 e0 = v3(lt.edge0[0], lt.edge0[1], lt.edge0[2]), e1 = v3(lt.edge1[0], lt.edge1[1], lt.edge1[2]);
    const V3 nl = normalize(cross(e0, e1));
    const float side = dot(nl, p - c0);
    if (!(side > 0.0f)) return 0.0f;
    float cx = 0.0f;
    const V3 cs[4] = {c0, c0 + e0, c0 + e1, c0 + e0 + e1};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const V3 v = cs[k] - p;
        const float l2 = dot(v, v);
        if (l2 > 0.0f) cx = fmaxf(cx, dot(n, v) / sqrtf(l2));
    }
    if (!(cx > 0.0f)) return 0.0f;
    const V3 v = c0 + (e0 + e1) * 0.5f - p;
    const float d2 = fmaxf(dot(v, v), 0.25f * lt.area);
    const float power = (lt.radiance[0] + lt.radiance[1] + lt.radiance[2]) * lt.area;
    return power * cx * side / (d2 * sqrtf(d2));
}

// The light NEE takes for uniform u at (p, n), with its selection probability.
__device__ __forceinline__ uint32_t pick_light(const ShadeScene& sc, V3 p, V3 n, float u, float& pdf)
{
    if (sc.num_lights > kAdaptiveLights) {
        uint32_t li = 0;
        while (li + 1 < sc.num_lights && u >= sc.lights[li].cdf) ++li;
        pdf = sc.lights[li].pdf_select;
        return li;
    }
    float w[kAdaptiveLights];
    float total = 0.0f;
#pragma unroll
    for (uint32_t l = 0; l < kAdaptiveLights; ++l) {
        w[l] = l < sc.num_lights ? light_weight(sc.lights[l], p, n) : 0.0f;
        total += w[l];
    }
    const float inv = total > 0.0f ? 0.9f / total : 0.0f;
    const float floor_share = total > 0.0f ? 0.1f : 1.0f;
    float acc = 0.0f;
    uint32_t pick = sc.num_lights - 1;
    float pick_pdf = w[0] * inv + floor_share * sc.lights[0].pdf_select;
    bool found = false;
#pragma unroll
    for (uint32_t l = 0; l < kAdaptiveLights; ++l) {
        if (l < sc.num_lights) {
            const float pl = w[l] * inv + floor_share * sc.lights[l].pdf_select;
            acc += pl;
            if (!found && u < acc) {
                pick = l;
                pick_pdf = pl;
                found = true;
            }
            if (!found && l == sc.num_lights - 1) pick_pdf = pl;
        }
    }
    pdf = pick_pdf;
    return pick;
}

// Probability that NEE at (p, n) picks light li.
__device__ __forceinline__ float light_pick_pdf(const ShadeScene& sc, uint32_t li, V3 p, V3 n)
{
    if (sc.num_lights > kAdaptiveLights) return sc.lights[li].pdf_select;
    float total = 0.0f, wl = 0.0f;
#pragma unroll
    for (uint32_t l = 0; l < kAdaptiveLights; ++l) {
        const float w = l < sc.num_lights ? light_weight(sc.lights[l], p, n) : 0.0f;
        total += w;
        if (l == li) wl = w;
    }
    // the same products pick_light forms, so the two agree to the bit
    const float inv = total > 0.0f ? 0.9f / total : 0.0f;
    const float floor_share = total > 0.0f ? 0.1f : 1.0f;
    return wl * inv + floor_share * sc.lights[li].pdf_select;
}

__device__ __forceinline__ float ray_eps(V3 p)
{
    return 1e-4f * fmaxf(1.0f, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
}

// Job, pixel (0xFFFFFFFF for tile padding) and sample of global path g: the
// path space is, per job, spp x padded pixels, sample-major, pixels in T x T
// tile order.
struct PathSite {
    uint32_t job, pixel, sample, s, x, y;
};

__device__ __forceinline__ PathSite locate(const WfArgs& a, uint64_t g)
{
    PathSite q;
    q.job = find_job(a.jobs, a.num_jobs, g);
    const JobDev& job = a.jobs[q.job];
    const uint64_t local = g - job.path_base;
    q.s = uint32_t(local / job.padded);
    const uint32_t k = uint32_t(local % job.padded);
    const uint32_t T = job.tile;
    const uint32_t t = k / (T * T), r = k % (T * T);
    q.x = (t % job.tiles_x) * T + r % T;
    q.y = (t / job.tiles_x) * T + r / T;
    q.sample = job.sample_offset + q.s;
    q.pixel = (q.x < job.width && q.y < job.height) ? q.y * job.width + q.x : 0xFFFFFFFFu;
    return q;
}

// Camera ray of pixel (x, y) = pixel index `pixel` for `sample`: glm::lookAt
// basis of the job (formed on the host in double), vertical field of view,
// row 0 at the top; the jitter is the first two draws of the path's bounce-0
// generator.
__device__ __forceinline__ void camera_ray(const JobDev& job, uint32_t x, uint32_t y, uint32_t sample, uint32_t pixel,
                                           Ray& ray)
{
    Rng rng = rng_for(job.seed, sample, pixel, 0u);
    const float jx = rng_float(rng), jy = rng_float(rng);
    const float sx = (2.0f * (float(x) + jx) / float(job.width) - 1.0f) * job.tan_half * job.aspect;
    const float sy = (1.0f - 2.0f * (float(y) + jy) / float(job.height)) * job.tan_half;
    const V3 f = v3(job.fwd[0], job.fwd[1], job.fwd[2]);
    const V3 r = v3(job.right[0], job.right[1], job.right[2]);
    const V3 u = v3(job.up[0], job.up[1], job.up[2]);
    const V3 d = normalize(f + r * sx + u * sy);
    ray.org[0] = job.eye[0]; ray.org[1] = job.eye[1]; ray.org[2] = job.eye[2];
    ray.tmin = 0.0f;
    ray.dir[0] = d.x; ray.dir[1] = d.y; ray.dir[2] = d.z;
    ray.tmax = 3.0e38f;
}

// Starts window path `gid` (global path g0 + gid) in `slot`: the slot's
// state, and the camera ray.  False for tile padding (no ray; radiance 0).
__device__ __forceinline__ bool start_path(const WfArgs& a, uint64_t g0, uint32_t gid, uint32_t slot, Ray& ray)
{
    const uint64_t g = g0 + gid;
    const PathSite q = locate(a, g);
    const JobDev& job = a.jobs[q.job];
    uint32_t train = kNoTrainSlot;
    bool live = false;
    if (q.pixel != 0xFFFFFFFFu) {
        if (job.train_count) {
            const uint64_t qi = uint64_t(q.s) * (uint64_t(job.width) * job.height) + q.pixel;
            if (qi < job.train_mod) {
                // (qi * mul + add) mod train_mod, x < 2^64 (qi < train_mod < 2^32,
                // mul < 2^32).  Barrett with floor((2^64 - 1) / train_mod): the
                // quotient estimate is low by at most 2.
                const uint64_t x = qi * job.train_mul + job.train_add;
                uint64_t ti = x - __umul64hi(x, job.train_barrett) * job.train_mod;
                if (ti >= job.train_mod) ti -= job.train_mod;
                if (ti >= job.train_mod) ti -= job.train_mod;
                if (ti < job.train_count) {
                    train = job.train_base + uint32_t(ti);
                    a.b.rec_count[train] = 0u;
                    a.b.t_list[train] = g;  // slots are unique: no atomic
                }
            }
        }
        camera_ray(job, q.x, q.y, q.sample, q.pixel, ray);
        live = true;
    }
    a.b.p_site[slot] = make_uint4(q.job, q.pixel, q.sample, gid);
    a.b.p_state[slot] = make_uint2(train, 0u);
    a.b.p_thr[slot] = make_float4(1.f, 1.f, 1.f, 0.f);
    a.b.p_L[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
    return live;
}

// Records the radiance of the path in `slot`, which has ended, and frees the
// slot for the next iteration's regeneration (the caller appends it to a free
// list).
__device__ __forceinline__ void retire(const WfArgs& a, uint32_t slot, float4 L)
{
    a.b.w_L[a.b.p_site[slot].w] = make_float4(L.x, L.y, L.z, 0.0f);
}

// The window's first n paths into slots 0..n-1; camera rays to extension
// queue A, padding slots (radiance 0) straight to free list 0.
__global__ __launch_bounds__(kGenerateBlock) void k_generate(WfArgs a, uint64_t g0, uint32_t n)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) a.b.counters[cnt(kCntNext)] = n;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        Ray ray{};
        const bool live = i < n && start_path(a, g0, i, i, ray);
        const bool dead = i < n && !live;
        if (dead) retire(a, i, make_float4(0.f, 0.f, 0.f, 0.f));
        uint32_t slot[2];
        block_append<2, kGenerateBlock>({live, dead}, {&a.b.counters[cnt(kCntExtA)], &a.b.counters[cnt(kCntFree0)]}, slot);
        if (live) {
            a.b.ext_rays[0][slot[0]] = ray;
            a.b.ext_path[0][slot[0]] = i;
            a.b.ext_scene[0][slot[0]] = a.jobs[a.b.p_site[i].x].scene_slot;
        }
        if (dead) a.b.free_list[0][slot[1]] = i;
    }
}

// Regeneration: each free slot of list fidx takes the window's next path
// (while any is left); camera rays to extension queue `next`, padding slots
// (radiance 0) straight to the other free list.
__global__ __launch_bounds__(kStageBlock) void k_regen(WfArgs a, uint64_t g0, uint32_t fidx, uint32_t next)
{
    const uint32_t nf = a.b.counters[cnt(kCntFree0 + fidx)];
    // No path left: return, so the counter stops growing.  One read per block,
    // shared, so the whole block returns or none of it does (block_append
    // synchronises the block).
    __shared__ uint32_t s_next;
    if (threadIdx.x == 0) s_next = a.b.counters[cnt(kCntNext)];
    __syncthreads();
    if (s_next >= a.b.window) return;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < nf; base += stride) {
        const uint32_t i = base + threadIdx.x;
        const bool want = i < nf;
        uint32_t gs[1];
        block_append<1, kStageBlock>({want}, {&a.b.counters[cnt(kCntNext)]}, gs);
        const bool has = want && gs[0] < a.b.window;
        const uint32_t slot = want ? a.b.free_list[fidx][i] : 0u;
        Ray ray{};
        const bool live = has && start_path(a, g0, gs[0], slot, ray);
        const bool dead = has && !live;
        if (dead) retire(a, slot, make_float4(0.f, 0.f, 0.f, 0.f));
        uint32_t q[2];
        block_append<2, kStageBlock>({live, dead}, {&a.b.counters[cnt(next)], &a.b.counters[cnt(kCntFree0 + (fidx ^ 1u))]}, q);
        if (live) {
            a.b.ext_rays[next][q[0]] = ray;
            a.b.ext_path[next][q[0]] = slot;
            a.b.ext_scene[next][q[0]] = a.jobs[a.b.p_site[slot].x].scene_slot;
        }
        if (dead) a.b.free_list[fidx ^ 1u][q[1]] = slot;
    }
}

__global__ void k_fill_u64(unsigned long long* p, uint32_t n, unsigned long long v)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_snapshot(uint32_t* counters, uint32_t src, uint32_t dst)
{
    if (threadIdx.x == 0) counters[cnt(dst)] = counters[cnt(src)];
}

// Zeroes the counter slots whose bit is set in mask.
__global__ void k_zero_slots(uint32_t* counters, uint32_t mask)
{
    if (threadIdx.x < kCntCount && ((mask >> threadIdx.x) & 1u)) counters[cnt(threadIdx.x)] = 0u;
}

// Zeroes the counters one iteration appends to: the next extension queue,
// the shadow queue, the traversal work counter's shards and the free list
// this iteration's retired slots go to.
__global__ void k_prep(uint32_t* counters, uint32_t* trace_ctr, uint32_t next, uint32_t free_zero, uint32_t shadow)
{
    if (threadIdx.x == 0) {
        counters[cnt(next)] = 0u;
        counters[cnt(shadow)] = 0u;
        counters[cnt(kCntFree0 + free_zero)] = 0u;
    }
    zero_trace_counter(trace_ctr, threadIdx.x);
}

__global__ __launch_bounds__(kShadeBlock) void k_shade(WfArgs a, uint32_t cur, uint32_t fw, uint32_t shq)
{
    const uint32_t n = a.b.counters[cnt(cur)];
    const uint32_t nxt = cur ^ 1u;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        bool cont = false, shadow = false, query = false;
        Ray next_ray{}, sh_ray{};
        V3 sh_contrib = v3(0, 0, 0);
        CacheQuery cq{};
        V3 q_w = v3(0, 0, 0);
        uint32_t path = 0, scene_slot = 0, gid = 0;
        if (i < n) {
            path = a.b.ext_path[cur][i];
            const uint4 site = a.b.p_site[path];
            const uint2 state = a.b.p_state[path];
            gid = site.w;
            const uint32_t bounce = state.y;
            const JobDev& job = a.jobs[site.x];
            scene_slot = job.scene_slot;
            const ShadeScene& sc = a.scenes[scene_slot];
            const float4 hit = a.b.ext_hits[i];
            const Ray ray = a.b.ext_rays[cur][i];
            float4 thr4 = a.b.p_thr[path];
            float4 L4 = a.b.p_L[path];
            V3 thr = v3(thr4.x, thr4.y, thr4.z);
            V3 L = v3(L4.x, L4.y, L4.z);
            const float pdf_prev = thr4.w;
            const V3 dir = v3(ray.dir[0], ray.dir[1], ray.dir[2]);
            const uint32_t tslot = state.x;
            const bool record = tslot != kNoTrainSlot;
            uint32_t rc = 0;
            if (record) {
                // the previous diffuse vertex's target starts after its own light sample
                rc = a.b.rec_count[tslot];
                if (rc > 0) {
                    float4* last = a.b.rec + (size_t(tslot) * a.b.rec_max + rc - 1) * kCacheRecordF4;
                    if (last[4].w == 0.0f) last[4] = make_float4(L.x, L.y, L.z, 1.0f);
                }
            }
            uint32_t ncode = 0;  // this vertex's quantised normal, for the next vertex's MIS
            const uint32_t tri = __float_as_uint(hit.y);
            if (tri == 0xFFFFFFFFu) {
                L = L + thr * v3(sc.background[0], sc.background[1], sc.background[2]);
            } else {
                const float t = hit.x;
                const uint32_t inst = __float_as_uint(hit.z);
                const float4* tp = sc.tris + 3ull * tri;
                const float4 A = tp[0], B = tp[1], C = tp[2];
                const uint32_t mat_id = __float_as_uint(B.w);
                const int32_t light_id = __float_as_int(C.w);
                V3 ng = cross(v3(B.x - A.x, B.y - A.y, B.z - A.z), v3(C.x - A.x, C.y - A.y, C.z - A.z));
                if (inst != 0xFFFFFFFFu) {
                    const float* m = sc.insts[inst].w2o;
                    ng = v3(m[0] * ng.x + m[4] * ng.y + m[8] * ng.z, m[1] * ng.x + m[5] * ng.y + m[9] * ng.z,
                            m[2] * ng.x + m[6] * ng.y + m[10] * ng.z);
                }
                ng = normalize(ng);
                const bool front = dot(ng, dir) < 0.0f;
                const V3 nf = front ? ng : -ng;  // faces the incoming ray
                const V3 p = v3(ray.org[0], ray.org[1], ray.org[2]) + dir * t;
                const DeviceMaterial mat = sc.materials[mat_id];
                const V3 Le = v3(mat.emission[0], mat.emission[1], mat.emission[2]);
                const bool emissive = mat.type == 2u;
                if ((Le.x > 0.f || Le.y > 0.f || Le.z > 0.f) && (!emissive || front)) {
                    float w = 1.0f;
                    if (bounce > 0 && pdf_prev > 0.0f && light_id >= 0) {
                        const DeviceLight& lt = sc.lights[light_id];
                        const V3 nl = normalize(cross(v3(lt.edge0[0], lt.edge0[1], lt.edge0[2]),
                                                      v3(lt.edge1[0], lt.edge1[1], lt.edge1[2])));
                        const float cl = fabsf(dot(nl, dir));
                        // the previous vertex's NEE would have picked this light with:
                        const float psel = light_pick_pdf(sc, uint32_t(light_id), v3(ray.org[0], ray.org[1], ray.org[2]),
                                                          oct_decode(__float_as_uint(L4.w)));
                        const float pdf_l = cl > 0.0f ? psel * t * t / (lt.area * cl) : 0.0f;
                        w = power_heuristic(pdf_prev, pdf_l);
                    }
                    L = L + thr * Le * w;
                }
                if (!emissive && bounce < job.max_depth) {
                    Rng rng = rng_for(job.seed, site.z, site.y, bounce + 1u);
                    const Onb frame = onb_from(nf);
                    const V3 wo = to_local(frame, -dir);
                    const V3 albedo = v3(mat.albedo[0], mat.albedo[1], mat.albedo[2]);
                    const float eps = ray_eps(p);
                    const V3 org = p + nf * eps;
                    ncode = oct_encode(nf);
                    // next-event estimation
                    const float ul = rng_float(rng), ua = rng_float(rng), ub = rng_float(rng);
                    if (sc.num_lights > 0) {
                        float psel;
                        const uint32_t li = pick_light(sc, org, oct_decode(ncode), ul, psel);
                        const DeviceLight& lt = sc.lights[li];
                        const V3 e0 = v3(lt.edge0[0], lt.edge0[1], lt.edge0[2]);
                        const V3 e1 = v3(lt.edge1[0], lt.edge1[1], lt.edge1[2]);
                        const V3 x = v3(lt.corner[0], lt.corner[1], lt.corner[2]) + e0 * ua + e1 * ub;
                        const V3 nl = normalize(cross(e0, e1));
                        V3 wi = x - p;
                        const float d2 = dot(wi, wi);
                        wi = wi * (1.0f / sqrtf(d2));
                        const float cl = -dot(nl, wi);
                        if (cl > 0.0f && d2 > 0.0f) {
                            const float pdf_l = psel * d2 / (lt.area * cl);
                            const BsdfEval ev = bsdf_eval(mat.type, albedo, mat.roughness, wo, to_local(frame, wi));
                            if (ev.pdf > 0.0f) {
                                const float w = power_heuristic(pdf_l, ev.pdf);
                                sh_contrib = thr * ev.f_cos * v3(lt.radiance[0], lt.radiance[1], lt.radiance[2]) * (w / pdf_l);
                                if (sh_contrib.x > 0.f || sh_contrib.y > 0.f || sh_contrib.z > 0.f) {
                                    const V3 to = x - org;
                                    sh_ray.org[0] = org.x; sh_ray.org[1] = org.y; sh_ray.org[2] = org.z;
                                    sh_ray.tmin = 0.0f;
                                    sh_ray.dir[0] = to.x; sh_ray.dir[1] = to.y; sh_ray.dir[2] = to.z;
                                    sh_ray.tmax = 1.0f - 1e-3f;
                                    shadow = true;
                                }
                            }
                        }
                    }
                    // cache vertices: Lambert, and GGX at least glossy_min rough (with the
                    // outgoing direction and roughness as inputs); from bounce
                    // glossy_any_bounce on any GGX vertex, whose throughput is small by then
                    const bool diffuse = mat.type == 0u;
                    const bool cachable =
                        diffuse || (mat.type == 1u && (mat.roughness >= a.b.glossy_min || bounce >= a.b.glossy_any_bounce));
                    const float q_rough = diffuse ? 0.0f : mat.roughness;
                    const V3 q_dir = diffuse ? v3(0, 0, 0) : -dir;
                    // a training path records its first rec_max cache vertices and ends in
                    // the cache at the next one: earlier targets then include the cache's
                    // estimate of the remaining tail (self-training, as in NRC)
                    const bool train_end = record && cachable && rc >= a.b.rec_max;
                    if (record && cachable && rc < a.b.rec_max) {
                        float4* r = a.b.rec + (size_t(tslot) * a.b.rec_max + rc) * kCacheRecordF4;
                        r[0] = make_float4(p.x, p.y, p.z, q_rough);
                        r[1] = make_float4(nf.x, nf.y, nf.z, 0.0f);
                        r[2] = make_float4(q_dir.x, q_dir.y, q_dir.z, 0.0f);
                        r[3] = make_float4(thr.x * albedo.x, thr.y * albedo.y, thr.z * albedo.z, 0.0f);
                        r[4] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        a.b.rec_count[tslot] = rc + 1;
                    }
                    if (cachable && (job.flags & FND_JOB_USE_CACHE) && (!record || train_end)) {
                        // terminate into the cache: it stands for the indirect light leaving p.
                        // Training paths go on past their first cache vertices.
                        query = true;
                        cq.pos = make_float4(p.x, p.y, p.z, q_rough);
                        cq.nrm = make_float4(nf.x, nf.y, nf.z, 0.0f);
                        cq.dir = make_float4(q_dir.x, q_dir.y, q_dir.z, 0.0f);
                        q_w = thr * albedo;
                    }
                    // BSDF sampling
                    const float u1 = rng_float(rng), u2 = rng_float(rng), urr = rng_float(rng);
                    const BsdfSample bs = query ? BsdfSample{v3(0, 0, 1), v3(0, 0, 0), 0.0f, false}
                                                : bsdf_sample(mat.type, albedo, mat.roughness, wo, u1, u2);
                    if (bs.ok) {
                        V3 nthr = thr * bs.weight;
                        bool alive = max3(nthr) > 0.0f;
                        if (alive && bounce >= job.rr_depth) {
                            const float q = fminf(max3(nthr), 0.95f);
                            if (urr >= q) alive = false; else nthr = nthr * (1.0f / q);
                        }
                        if (alive) {
                            const V3 wi = normalize(to_world(frame, bs.wi));
                            next_ray.org[0] = org.x; next_ray.org[1] = org.y; next_ray.org[2] = org.z;
                            next_ray.tmin = 0.0f;
                            next_ray.dir[0] = wi.x; next_ray.dir[1] = wi.y; next_ray.dir[2] = wi.z;
                            next_ray.tmax = 3.0e38f;
                            thr = nthr;
                            cont = true;
                            a.b.p_thr[path] = make_float4(thr.x, thr.y, thr.z, bs.pdf);
                            a.b.p_state[path] = make_uint2(tslot, bounce + 1u);
                        }
                    }
                }
            }
            a.b.p_L[path] = make_float4(L.x, L.y, L.z, __uint_as_float(ncode));
            // a path that does not continue has ended: retired here, or by the
            // shadow-ray resolve when its last light sample is still in flight
            if (a.b.retire && !cont && !shadow) retire(a, path, make_float4(L.x, L.y, L.z, 0.0f));
        }
        // one block-aggregated append for the four lists
        const bool freed = a.b.retire && i < n && !cont && !shadow;
        uint32_t slots[4];
        block_append<4, kShadeBlock>({cont, query, shadow, freed},
                        {&a.b.counters[cnt(nxt)], &a.b.counters[cnt(kCntQuery)], &a.b.counters[cnt(shq)],
                         &a.b.counters[cnt(kCntFree0 + fw)]},
                        slots);
        const uint32_t es = slots[0], qs = slots[1], ss = slots[2];
        if (freed) a.b.free_list[fw][slots[3]] = path;
        if (cont) {
            a.b.ext_rays[nxt][es] = next_ray;
            a.b.ext_path[nxt][es] = path;
            a.b.ext_scene[nxt][es] = scene_slot;
        }
        if (query) {
            a.b.q_query[qs] = cq;
            a.b.q_path[qs] = gid;
            a.b.q_weight[qs] = make_float4(q_w.x, q_w.y, q_w.z, 0.f);
        }
        if (shadow) {
            a.b.sh_rays[ss] = sh_ray;
            a.b.sh_path[ss] = path;
            a.b.sh_scene[ss] = scene_slot;
            // w = 1: the path ended at this vertex, the resolve retires it
            a.b.sh_contrib[ss] = make_float4(sh_contrib.x, sh_contrib.y, sh_contrib.z, (a.b.retire && !cont) ? 1.f : 0.f);
        }
    }
}

// One shadow ray per path per bounce, so the adds below never collide.  A
// path whose last vertex cast this ray is retired after the add, and its slot
// goes to free list fw.  Block 0 also clears the counters the next iteration
// appends to (zero_ext, zero_sh, zero_free; kCntCount: none) and then the
// traversal work counter: the iteration's shade, regeneration and traces,
// which read them last, have finished, and this launch reads none of them (it
// reads shadow queue shq and appends to free list fw).  No counter launch runs
// between iterations this way.
__global__ __launch_bounds__(kStageBlock) void k_shadow_resolve(WfArgs a, uint32_t fw, uint32_t shq, uint32_t zero_ext,
                                                                uint32_t zero_sh, uint32_t zero_free)
{
    const uint32_t n = a.b.counters[cnt(shq)];
    if (blockIdx.x == 0 && zero_sh < kCntCount) {
        if (threadIdx.x == 0) {
            if (zero_ext < kCntCount) a.b.counters[cnt(zero_ext)] = 0u;
            if (zero_free < kCntCount) a.b.counters[cnt(zero_free)] = 0u;
            a.b.counters[cnt(zero_sh)] = 0u;
        }
        zero_trace_counter(a.b.trace_ctr, threadIdx.x);
    }
    if (!a.b.retire) {  // a window without regeneration: paths stay in their slots
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
            if (__float_as_uint(a.b.sh_hits[i].y) != 0xFFFFFFFFu) continue;
            const uint32_t path = a.b.sh_path[i];
            const float4 c = a.b.sh_contrib[i];
            float4 L = a.b.p_L[path];
            L.x += c.x; L.y += c.y; L.z += c.z;
            a.b.p_L[path] = L;
        }
        return;
    }
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        bool freed = false;
        uint32_t path = 0;
        if (i < n) {
            path = a.b.sh_path[i];
            const float4 c = a.b.sh_contrib[i];
            const bool lit = __float_as_uint(a.b.sh_hits[i].y) == 0xFFFFFFFFu;
            freed = c.w != 0.0f;
            if (lit || freed) {
                float4 L = a.b.p_L[path];
                if (lit) {
                    L.x += c.x; L.y += c.y; L.z += c.z;
                    a.b.p_L[path] = L;
                }
                if (freed) retire(a, path, L);
            }
        }
        uint32_t q[1];
        block_append<1, kStageBlock>({freed}, {&a.b.counters[cnt(kCntFree0 + fw)]}, q);
        if (freed) a.b.free_list[fw][q[0]] = path;
    }
}

// Per pixel, in increasing sample order: the thread of a pixel's first path in
// the window adds the window's samples of that pixel one after another, so the
// framebuffer sums are the same for any window size, slot count or tile order.
__global__ __launch_bounds__(256) void k_accumulate(WfArgs a, uint64_t g0, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint64_t g = g0 + i;
        const PathSite site = locate(a, g);
        const uint32_t pixel = site.pixel;
        if (pixel == 0xFFFFFFFFu) continue;
        const JobDev& job = a.jobs[site.job];
        if (!job.accum) continue;
        const uint64_t lo = job.path_base > g0 ? job.path_base : g0;
        if (g >= lo + job.padded) continue;  // not the first sample of this pixel in the batch
        const uint64_t end = job.path_end < g0 + n ? job.path_end : g0 + n;
        float* acc = job.accum + 3ull * pixel;
        float r = acc[0], gr = acc[1], b = acc[2];
        double r64 = 0, g64 = 0, b64 = 0;
        double* acc64 = job.accum64 ? job.accum64 + 3ull * pixel : nullptr;
        if (acc64) { r64 = acc64[0]; g64 = acc64[1]; b64 = acc64[2]; }
        for (uint64_t h = g; h < end; h += job.padded) {
            const float4 L = a.b.w_L[h - g0];
            r += L.x; gr += L.y; b += L.z;
            if (acc64) { r64 += double(L.x); g64 += double(L.y); b64 += double(L.z); }
        }
        acc[0] = r; acc[1] = gr; acc[2] = b;
        if (acc64) { acc64[0] = r64; acc64[1] = g64; acc64[2] = b64; }
    }
}

// Cache predictions replace the rest of the path: L += throughput * albedo * max(pred, 0).
__global__ __launch_bounds__(256) void k_cache_apply(WfArgs a)
{
    const uint32_t n = a.b.counters[cnt(kCntQuery)];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t gid = a.b.q_path[i];
        const float4 w = a.b.q_weight[i], p = a.b.q_pred[i];
        float4 L = a.b.w_L[gid];
        L.x += w.x * fmaxf(p.x, 0.0f);
        L.y += w.y * fmaxf(p.y, 0.0f);
        L.z += w.z * fmaxf(p.z, 0.0f);
        a.b.w_L[gid] = L;
    }
}

// Training targets from the finished training paths of a window: at a
// recorded diffuse vertex with weight = throughput * albedo, target =
// (L_path - L_mark) / weight, the albedo-demodulated radiance the path
// gathered after that vertex's own light sample (per channel; vertices with a
// vanishing weight are skipped).
__global__ __launch_bounds__(256) void k_cache_collect(WfArgs a, uint64_t g0, uint32_t n)
{
    // the training slots whose path lies in this window [g0, g0 + n)
    const uint32_t ns = a.b.t_slots;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < ns; base += stride) {
        const uint32_t j = base + threadIdx.x;
        const unsigned long long g = j < ns ? a.b.t_list[j] : ~0ull;
        const bool mine = g >= g0 && g - g0 < n;
        const uint32_t tslot = mine ? j : kNoTrainSlot;  // t_list is indexed by training slot
        const uint32_t rc = tslot != kNoTrainSlot ? a.b.rec_count[tslot] : 0u;
        const float4 L = rc ? a.b.w_L[uint32_t(g - g0)] : make_float4(0.f, 0.f, 0.f, 0.f);
        // one wave-aggregated append per record index (every lane takes part)
        for (uint32_t k = 0; k < a.b.rec_max; ++k) {
            bool take = false;
            CacheQuery q{};
            float4 target{};
            if (k < rc) {
                const float4* r = a.b.rec + (size_t(tslot) * a.b.rec_max + k) * kCacheRecordF4;
                const float4 w = r[3];
                if (w.x > 1e-4f && w.y > 1e-4f && w.z > 1e-4f) {
                    const float4 m = r[4].w != 0.0f ? r[4] : L;
                    q.pos = r[0];
                    q.nrm = r[1];
                    q.dir = r[2];
                    target = make_float4(fmaxf(L.x - m.x, 0.0f) / w.x, fmaxf(L.y - m.y, 0.0f) / w.y,
                                         fmaxf(L.z - m.z, 0.0f) / w.z, 0.0f);
                    take = true;
                }
            }
            const uint32_t slot = wave_append(take, &a.b.counters[cnt(kCntTrain)]);
            if (take && slot < a.b.t_cap) {
                a.b.t_query[slot] = q;
                a.b.t_target[slot] = target;
            }
        }
    }
}

// Every pixel's camera ray for one sample, as the renderer starts its paths.
__global__ __launch_bounds__(256) void k_camera_rays(JobDev job, uint32_t sample, Ray* out)
{
    const uint32_t n = job.width * job.height;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        Ray ray;
        camera_ray(job, i % job.width, i / job.width, sample, i, ray);
        out[i] = ray;
    }
}

// The first `ndraws` outputs of the generator of (seed, sample, pixel, bounce)
// for pixels pixel0 .. pixel0 + npix - 1.
__global__ __launch_bounds__(256) void k_rng_draws(uint64_t seed, uint32_t sample, uint32_t pixel0, uint32_t npix,
                                                   uint32_t bounce, uint32_t ndraws, uint32_t* out)
{
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
        Rng r = rng_for(seed, sample, pixel0 + p, bounce);
        for (uint32_t k = 0; k < ndraws; ++k) out[size_t(p) * ndraws + k] = rng_next(r);
    }
}

}  // namespace

void launch_zero_slots(const WfArgs& a, uint32_t mask, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_slots, dim3(1), dim3(64), 0, s, a.b.counters, mask);
}

void launch_cache_apply(const WfArgs& a, uint32_t grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_cache_apply, dim3(grid), dim3(256), 0, s, a);
}

void launch_cache_collect(const WfArgs& a, uint64_t g0, uint32_t n, uint32_t grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_cache_collect, dim3(grid), dim3(256), 0, s, a, g0, n);
}

void launch_snapshot(const WfArgs& a, uint32_t src, uint32_t dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_snapshot, dim3(1), dim3(64), 0, s, a.b.counters, src, dst);
}

void launch_fill_u64(unsigned long long* p, uint32_t n, unsigned long long v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_u64, dim3(std::max(1u, std::min(1024u, (n + 255u) / 256u))), dim3(256), 0, s, p, n, v);
}

void launch_generate(const WfArgs& a, uint64_t g0, uint32_t n, uint32_t grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_generate, dim3(grid), dim3(kGenerateBlock), 0, s, a, g0, n);
}

void launch_regen(const WfArgs& a, uint64_t g0, uint32_t fidx, uint32_t next, uint32_t grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_regen, dim3(grid), dim3(kStageBlock), 0, s, a, g0, fidx, next);
}



// One pixel per thread: mean = sum * inv (the same product as the FP32
// readback), times exposure; each channel's byte is the number of the 255
// ascending thresholds at or below it (binary search; NaN and x <= 0 give 0).
__global__ __launch_bounds__(256) void k_present(const float* accum, uint64_t pixels, float inv, float exposure,
                                                 const float* thresholds, uint32_t* out, uint32_t width,
                                                 uint32_t pitch)
{
    __shared__ float t[256];
    for (uint32_t k = threadIdx.x; k < 256; k += blockDim.x) t[k] = k < 255 ? thresholds[k] : __builtin_inff();
    __syncthreads();
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < pixels; i += uint64_t(gridDim.x) * blockDim.x) {
        uint32_t word = 0xFF000000u;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float x = accum[3 * i + c] * inv;
            if (exposure != 1.0f) x *= exposure;
            uint32_t b = 0;  // thresholds t[0..254] ascending; count those <= x
#pragma unroll
            for (uint32_t step = 128; step > 0; step >>= 1)
                if (t[b + step - 1] <= x) b += step;
            word |= b << (8 * c);
        }
        out[(i / width) * pitch + i % width] = word;  // row pitch in pixels
    }
}

// relMSE of a framebuffer against an FP64 reference: each block sums its
// strided share in double, the shares are added in block order by one block
// (a tree over 256 threads), so the result does not depend on timing.
__global__ __launch_bounds__(256) void k_rel_mse_part(const float* x, float xinv, const double* y, double yinv,
                                                      uint64_t n, double eps, double* part)
{
    __shared__ double red[256];
    double s = 0.0;
    for (uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256u) {
        const double xv = double(x[i] * xinv);  // the FP32 mean, as the FP32 readback forms it
        const double yv = y[i] * yinv;
        const double d = xv - yv;
        s += d * d / (yv * yv + eps);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_rel_mse_sum(const double* part, uint32_t parts, double inv_n, double* out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < parts; i += 256u) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] * inv_n;
}

void launch_rel_mse(const float* x, float xinv, const double* y, double yinv, uint64_t n, double eps, double* part,
                    double* out, hipStream_t s)
{
    const uint32_t grid = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(kRelMseParts, (n + 255) / 256)));
    hipLaunchKernelGGL(k_rel_mse_part, dim3(grid), dim3(256), 0, s, x, xinv, y, yinv, n, eps, part);
    hipLaunchKernelGGL(k_rel_mse_sum, dim3(1), dim3(256), 0, s, part, grid, n ? 1.0 / double(n) : 0.0, out);
}

void launch_present(const float* accum, uint64_t pixels, float inv, float exposure, const float* thresholds,
                    uint32_t* out, uint32_t width, uint32_t pitch, hipStream_t s)
{
    const uint32_t grid = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(4096, (pixels + 255) / 256)));
    hipLaunchKernelGGL(k_present, dim3(grid), dim3(256), 0, s, accum, pixels, inv, exposure, thresholds, out, width,
                       pitch);
}

void launch_prep(const WfArgs& a, uint32_t next, uint32_t free_zero, uint32_t shadow, hipStream_t s)
{
    hipLaunchKernelGGL(k_prep, dim3(1), dim3(64), 0, s, a.b.counters, a.b.trace_ctr, next, free_zero, shadow);
}

void launch_shade(const WfArgs& a, uint32_t cur, uint32_t fw, uint32_t shadow, uint32_t grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_shade, dim3(grid), dim3(kShadeBlock), 0, s, a, cur, fw, shadow);
}

void launch_shadow_resolve(const WfArgs& a, uint32_t fw, uint32_t shadow, uint32_t zero_ext, uint32_t zero_shadow,
                           uint32_t zero_free, uint32_t grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_shadow_resolve, dim3(grid), dim3(kStageBlock), 0, s, a, fw, shadow, zero_ext, zero_shadow,
                       zero_free);
}

void launch_accumulate(const WfArgs& a, uint64_t g0, uint32_t n, uint32_t grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_accumulate, dim3(grid), dim3(256), 0, s, a, g0, n);
}

void launch_camera_rays(const JobDev& job, uint32_t sample, Ray* out, hipStream_t s)
{
    const uint64_t n = uint64_t(job.width) * job.height;
    const uint32_t grid = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(4096, (n + 255) / 256)));
    hipLaunchKernelGGL(k_camera_rays, dim3(grid), dim3(256), 0, s, job, sample, out);
}

void launch_rng_draws(uint64_t seed, uint32_t sample, uint32_t pixel0, uint32_t npix, uint32_t bounce, uint32_t ndraws,
                      uint32_t* out, hipStream_t s)
{
    const uint32_t grid = std::max(1u, std::min(4096u, (npix + 255u) / 256u));
    hipLaunchKernelGGL(k_rng_draws, dim3(grid), dim3(256), 0, s, seed, sample, pixel0, npix, bounce, ndraws, out);
}

}  // namespace fnd
