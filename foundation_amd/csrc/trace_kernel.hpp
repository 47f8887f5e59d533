// trace_kernel.hpp — the persistent traversal kernel (closest-hit / any-hit).
//
// Design for CDNA4 (wave64, 160 KB LDS per CU, no ray-tracing units):
//  * One ray per lane.  A wave keeps its 64 lanes busy by refetching rays:
//    once `refill` lanes are idle, one lane takes a block of ray indices with a
//    single atomicAdd on the wave's shard of the work counter and the idle
//    lanes take theirs by popcount of the ballot below them (Aila & Laine,
//    "Understanding the Efficiency of Ray Traversal on GPUs", HPG 2009,
//    dynamic fetch; here done with wave ballots).
//  * The traversal stack is per lane in LDS, laid out [entry][thread] so any
//    mix of stack depths across a wave hits distinct banks; entries past the
//    LDS part go to a per-thread spill area in HBM sized from the tree depth.
//    An entry is (word, entry distance); on pop, entries farther than the best
//    hit so far are dropped.
//  * Each loop iteration does at most one unit of work per lane: one 64-byte
//    node (4 quantised slab tests, sort, push), one triangle, or an instance
//    entry; the kinds that run in an iteration are picked by lanes per cost.
//    A node stores each slot's finished stack entry (child index, or tagged
//    triangle range or instance), so a node step pushes what it loaded.
//  * Two-level scenes: an instance entry moves the ray into object space and
//    pushes a return marker; popping the marker restores the world ray and
//    the pop continues.
#pragma once
#include "pt_device.hpp"

namespace fnd {

// Rays per trace launch: ray indices are formed from grid threads and a
// 32-bit shard counter and must not wrap.
constexpr uint32_t kMaxRaysPerTrace = 0x7FFFFFFFu;
constexpr uint32_t kWave = 64;

struct SceneDev {
    const WideNode* nodes;
    const float4* tris;
    const InstanceRec* insts;
    uint32_t root;
    uint32_t two_level;
};

enum TraceMode : int {
    kModeClosest = 0,          // fnd_hit records (global primitive ids)
    kModeAny = 1,              // fnd_hit records, first accepted hit
    kModeClosestInternal = 2,  // (t, leaf triangle index, instance slot) for shading
    kModeAnyInternal = 3,      // visibility for shadow rays
    kModeMixedInternal = 4     // one launch: set 1 closest (as 2), then set 2 any-hit (as 3)
};

struct TraceArgs {
    SceneDev scene;            // used when the kernel is not per-ray scene
    const SceneDev* scenes;    // per-ray scene table
    const uint32_t* ray_scene; // per-ray scene index
    const Ray* rays;
    float4* hits;
    uint32_t nrays;
    const uint32_t* nrays_dev; // if set, the ray count is read here (wavefront queues)
    uint32_t* counter;         // work counter: kCounterWords words, shards zero at launch
    uint2* spill;
    uint32_t spill_stride;     // threads in the grid
    uint32_t refill;
    unsigned long long* stats; // node visits, triangle tests, instance entries (COUNT)
    // kModeMixedInternal: the second ray set (any-hit), indices after the first
    const Ray* rays2;
    float4* hits2;
    const uint32_t* ray_scene2;
    const uint32_t* nrays2_dev;
    unsigned long long* ray_total;  // if set, this launch's ray count is added here (statistics)
    // kModeMixedInternal: a launch whose two sets hold fewer rays than this
    // returns at once, and the cooperative launch (trace_coop.hpp) queued
    // beside it traces them instead, and the other way round (0: off)
    uint32_t coop_below;
    // kModeClosest / kModeAny: if set and *ray_order_use is nonzero, the
    // launch's i-th ray is ray_order[i] (records still go to hits[ray]); a
    // spatial order of the batch.  (The renderer's extension rays, in tile
    // order, measured slower ordered: 5.10 against 4.92 ms per plain 1080p
    // frame, 663 against 687 Msamples/s for the 64-spp render.)
    const uint32_t* ray_order;
    const uint32_t* ray_order_use;
};

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

// Lanes-per-cost weights of the three kinds of work, about 1000 / (VALU
// instructions of the kind's branch): node 190, triangle 85, instance 122.
constexpr uint32_t kCostWeightNode = 5, kCostWeightTri = 12, kCostWeightInst = 8;

template <int MODE, bool TWO, bool MULTI, bool COUNT, int SLDS, int BLOCK>
__device__ __forceinline__ void trace_body(const TraceArgs& args)
{
    constexpr bool MIXED = MODE == kModeMixedInternal;
    constexpr bool ANY_CONST = !MIXED && (MODE & 1) != 0;
    constexpr bool INTERNAL = MODE >= kModeClosestInternal;
    __shared__ uint2 stk[SLDS * BLOCK];

    const uint32_t tid = threadIdx.x;
    const uint32_t gtid = blockIdx.x * BLOCK + tid;
    const uint32_t nrays1 = args.nrays_dev ? *args.nrays_dev : args.nrays;
    const uint32_t nrays = MIXED ? nrays1 + *args.nrays2_dev : nrays1;
    if constexpr (MIXED) {
        if (nrays < args.coop_below) return;
    }
    if (args.ray_total && gtid == 0) atomicAdd(args.ray_total, (unsigned long long)nrays);
    const bool use_order = (MODE == kModeClosest || MODE == kModeAny) && args.ray_order && *args.ray_order_use != 0u;
    bool lane_any = ANY_CONST;  // per lane in the mixed mode, a constant otherwise

    SceneDev sc = args.scene;
    bool active = false, retired = false;
    uint32_t ray_idx = 0;
    RayK r;   // the ray in the current space
    // The world ray, kept for instance entries and returns.  Every node of a
    // two-level scene's top level is flagged top and has only instance
    // leaves, so world space needs origin and 1/d for slab tests and the
    // direction for the transform; the shear constants of r are rebuilt by
    // each instance entry.
    float wox = 0.f, woy = 0.f, woz = 0.f, wdx = 0.f, wdy = 0.f, wdz = 0.f, wix = 0.f, wiy = 0.f, wiz = 0.f;
    float tmin = 0.f, tmax = 0.f, tbest = 0.f, bV = 0.f, bW = 0.f, bdet = 1.f;
    uint32_t best = 0xFFFFFFFFu, best_aux = 0u;
    uint32_t cur = 0, sp = 0, id_base = 0, inst_slot = 0xFFFFFFFFu;
    uint32_t c_nodes = 0, c_tris = 0, c_inst = 0;

    auto push = [&](uint32_t e, float t) {
        const uint2 v = make_uint2(e, __float_as_uint(t));
        if (sp < SLDS)
            stk[sp * BLOCK + tid] = v;
        else
            args.spill[size_t(sp - SLDS) * args.spill_stride + gtid] = v;
        ++sp;
    };

    // Ray fetch.  The first fetch is static (lane gtid takes ray gtid), so a
    // launch whose queue is short or empty retires its surplus waves without
    // an atomic; later fetches take exactly the idle lanes' worth of rays,
    // past the grid's first gtid range, from the wave's shard of the work
    // counter.  Shard s hands out the blocks of 64 rays s, s + kCounterShards,
    // ... (shard-local index l is ray grid_threads + (l / 64 * kCounterShards
    // + s) * 64 + l % 64), so the shards advance through the rays together and
    // the rays in flight stay a sliding window over the ray array, whose
    // neighbouring rays share nodes in L2 (measured faster than fixed
    // per-wave shares of the array).  One address would serialise every
    // wave's refill atomic: at a refill threshold of 48 the single counter was
    // busy for about 40% of the incoherent 10M-triangle launch, and lower
    // thresholds were slower still; with the shards, refills can be frequent.
    const uint32_t grid_threads = gridDim.x * BLOCK;
    const uint32_t shard = (gtid / kWave) & (kCounterShards - 1u);
    uint32_t* const shard_ctr = args.counter + shard * kCounterStride;
    bool first_fetch = true;
    for (;;) {
        const uint64_t idle = __ballot(!active && !retired);
        const uint64_t act = __ballot(active);
        if (idle != 0 && (act == 0 || uint32_t(__popcll(idle)) >= args.refill)) {
            const uint32_t nidle = uint32_t(__popcll(idle));
            const int leader = __builtin_ctzll(idle);
            uint32_t idx;
            if (first_fetch) {
                idx = gtid;  // every lane is idle
                first_fetch = false;
            } else {
                uint32_t base = 0;
                if (int(__lane_id()) == leader) base = atomicAdd(shard_ctr, nidle);
                const uint32_t l = uint32_t(__builtin_amdgcn_readlane(int(base), leader)) + lanes_below(idle);
                idx = grid_threads + ((l / kWave) * kCounterShards + shard) * kWave + l % kWave;
            }
            if (!active && !retired) {
                if (idx < nrays) {
                    const bool second = MIXED && idx >= nrays1;
                    ray_idx = second ? idx - nrays1 : idx;
                    if (MIXED) lane_any = second;
                    if constexpr (MODE == kModeClosest || MODE == kModeAny) {
                        if (use_order) ray_idx = args.ray_order[ray_idx];
                    }
                    const float4* rp = reinterpret_cast<const float4*>((second ? args.rays2 : args.rays) + ray_idx);
                    const float4 ra = rp[0];
                    const float4 rb = rp[1];
                    // A non-finite origin or direction, or a NaN range bound, is
                    // refused: its range is made empty (tmin = +inf, best bound
                    // -inf), so the root's slab tests fail (fmaxf/fminf drop the
                    // NaN terms) and the ray finishes as a miss (t = tmax as
                    // given) after one node step, without a walk (NaN slab
                    // reciprocals alone would pass every box test).
                    const float fin = ((ra.x - ra.x) + (ra.y - ra.y)) + ((ra.z - ra.z) + (rb.x - rb.x)) +
                                      ((rb.y - rb.y) + (rb.z - rb.z));
                    const bool refused = fin != fin || ra.w != ra.w || rb.w != rb.w;
                    if (MULTI) sc = args.scenes[(second ? args.ray_scene2 : args.ray_scene)[ray_idx]];
                    // canonical once here, so the slab tests' fmaxf/fminf need not
                    // quiet them on every node
                    tmin = refused ? __builtin_inff() : __builtin_canonicalizef(ra.w);
                    tmax = refused ? rb.w : __builtin_canonicalizef(rb.w);
                    ray_setup(r, ra.x, ra.y, ra.z, rb.x, rb.y, rb.z);
                    if (TWO) {
                        wox = ra.x; woy = ra.y; woz = ra.z;
                        wdx = rb.x; wdy = rb.y; wdz = rb.z;
                        wix = r.ix; wiy = r.iy; wiz = r.iz;
                    }
                    tbest = refused ? -__builtin_inff() : tmax;
                    best = 0xFFFFFFFFu;
                    best_aux = 0xFFFFFFFFu;
                    bV = bW = 0.f;
                    bdet = 1.f;
                    sp = 0;
                    cur = sc.root;
                    id_base = 0;
                    inst_slot = 0xFFFFFFFFu;
                    active = true;
                } else {
                    retired = true;
                }
            }
        }
        if (__ballot(active) == 0) break;
        if (!active) continue;

        // Work selection: a lane's next unit is a node, a triangle or an
        // instance entry, and each kind is a branch the whole wave pays for
        // once any lane takes it.  Every kind whose lanes-per-cost is at least
        // half the best kind's runs this iteration; the others wait a turn.
        // Per-lane order is unchanged, so results are too.
        const bool is_node = cur < kEntryInstance;
        const bool is_tri = (cur & kEntryTri) != 0u;
        const uint32_t s_node = uint32_t(__popcll(__ballot(is_node))) * kCostWeightNode;
        const uint32_t s_tri = uint32_t(__popcll(__ballot(is_tri))) * kCostWeightTri;
        const uint32_t s_inst = TWO ? uint32_t(__popcll(__ballot(!is_node && !is_tri))) * kCostWeightInst : 0u;
        const uint32_t s_best = max(s_node, max(s_tri, s_inst));
        const bool run_node = 2u * s_node >= s_best, run_tri = 2u * s_tri >= s_best, run_inst = 2u * s_inst >= s_best;

        bool pop = false, finish = false;
        if (is_node && run_node) {
            // ---- one 64-byte wide node: four quantised slab tests
            if (COUNT) ++c_nodes;
            const uint4* np = reinterpret_cast<const uint4*>(sc.nodes + cur);
            const uint4 n0 = np[0];
            const uint4 n1 = np[1];
            const uint4 n2 = np[2];
            const uint4 n3 = np[3];
            const int ex = int(int8_t(n0.w & 0xffu));
            const int ey = int(int8_t((n0.w >> 8) & 0xffu));
            const int ez = int(int8_t((n0.w >> 16) & 0xffu));
            const float sx = scale_exp2(r.ix, ex);
            const float sy = scale_exp2(r.iy, ey);
            const float sz = scale_exp2(r.iz, ez);
            const float cx = (__uint_as_float(n0.x) - r.ox) * r.ix;
            const float cy = (__uint_as_float(n0.y) - r.oy) * r.iy;
            const float cz = (__uint_as_float(n0.z) - r.oz) * r.iz;
            // entries = n1; qlo = (n2.x, n2.y, n2.z), qhi = (n2.w, n3.x, n3.y)
            // (selecting the words with v_perm_b32 and per-ray selectors instead of
            // these compares measured no faster: fewer node VALU, more per ray set-up)
            const bool px_ = r.ix >= 0.f, py_ = r.iy >= 0.f, pz_ = r.iz >= 0.f;
            const uint32_t nx = px_ ? n2.x : n2.w, fx = px_ ? n2.w : n2.x;
            const uint32_t ny = py_ ? n2.y : n3.x, fy = py_ ? n3.x : n2.y;
            const uint32_t nz = pz_ ? n2.z : n3.y, fz = pz_ ? n3.y : n2.z;
            float tk[4];
            uint32_t ek[4] = {n1.x, n1.y, n1.z, n1.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float tn = fmaxf(fmaxf(fmaf(ubyte_f(nx, k), sx, cx), fmaf(ubyte_f(ny, k), sy, cy)),
                                       fmaxf(fmaf(ubyte_f(nz, k), sz, cz), tmin));
                const float tf = fminf(fminf(fmaf(ubyte_f(fx, k), sx, cx), fmaf(ubyte_f(fy, k), sy, cy)),
                                       fminf(fmaf(ubyte_f(fz, k), sz, cz), tbest));
                // an empty slot's entry is 0 (node 0 is the root, never a child)
                const bool hit = ek[k] != 0u && tn <= tf * kSlack;
                tk[k] = hit ? tn : __builtin_inff();
            }
            const bool any_l = MIXED ? lane_any : ANY_CONST;
            if (!any_l) {
                // sorting network (0,1)(2,3)(0,2)(1,3)(1,2): nearest child first
                auto cs = [&](int a, int b) {
                    const bool s = tk[b] < tk[a];
                    const float ta = s ? tk[b] : tk[a], tb = s ? tk[a] : tk[b];
                    const uint32_t ea = s ? ek[b] : ek[a], eb = s ? ek[a] : ek[b];
                    tk[a] = ta; tk[b] = tb; ek[a] = ea; ek[b] = eb;
                };
                cs(0, 1); cs(2, 3); cs(0, 2); cs(1, 3); cs(1, 2);
                // the hit children are a prefix of the sorted four: push 3, 2, 1
                // (farthest first) where they are hits
                const uint32_t v1 = tk[1] != __builtin_inff(), v2 = tk[2] != __builtin_inff(),
                               v3 = tk[3] != __builtin_inff();
                if (sp + 3u <= uint32_t(SLDS)) {
                    // all three slots are in LDS: write them unconditionally at
                    // prefix positions (a miss lands on a slot a later write or
                    // the next push reuses), no branches
                    const uint32_t p3 = sp, p2 = sp + v3, p1 = p2 + v2;
                    stk[p3 * BLOCK + tid] = make_uint2(ek[3], __float_as_uint(tk[3]));
                    stk[p2 * BLOCK + tid] = make_uint2(ek[2], __float_as_uint(tk[2]));
                    stk[p1 * BLOCK + tid] = make_uint2(ek[1], __float_as_uint(tk[1]));
                    sp = p1 + v1;
                } else {
                    if (v3) push(ek[3], tk[3]);
                    if (v2) push(ek[2], tk[2]);
                    if (v1) push(ek[1], tk[1]);
                }
                if (tk[0] != __builtin_inff()) cur = ek[0]; else pop = true;
            } else {
                bool have = false;
#pragma unroll
                for (int k = 3; k >= 0; --k) {
                    if (tk[k] != __builtin_inff()) {
                        if (have) push(cur, 0.f);
                        cur = ek[k];
                        have = true;
                    }
                }
                if (!have) pop = true;
            }
        } else if (is_tri && run_tri) {
            // ---- one triangle of a leaf range
            if (COUNT) ++c_tris;
            const uint32_t ti = (cur >> 3) & 0x0FFFFFFFu;
            const float4* tp = sc.tris + 3ull * ti;
            const float4 a = tp[0];
            const float4 b = tp[1];
            const float4 c = tp[2];
            float t, Vn, Wn, detn;
            if (isect_tri(r, a, b, c, t, Vn, Wn, detn) && t > tmin && t < tmax) {
                const uint32_t id = INTERNAL ? ti : id_base + __float_as_uint(a.w);
                if (MIXED ? lane_any : ANY_CONST) {
                    tbest = t; best = id; bV = Vn; bW = Wn; bdet = detn; best_aux = inst_slot;
                    finish = true;
                } else if (t < tbest || (t == tbest && id < best)) {
                    tbest = t; best = id; bV = Vn; bW = Wn; bdet = detn; best_aux = inst_slot;
                }
            }
            if (!finish) {
                if ((cur & 7u) > 1u)
                    cur += 7u;  // first + 1, count - 1
                else
                    pop = true;
            }
        } else if (TWO && !is_node && !is_tri && run_inst) {
            // ---- enter an instance: object-space ray, return marker
            if (COUNT) ++c_inst;
            const uint32_t ii = (cur >> 3) & 0x07FFFFFFu;
            const float4* ip = reinterpret_cast<const float4*>(sc.insts + ii);
            const float4 m0 = ip[0], m1 = ip[1], m2 = ip[2];
            const uint4 m3 = reinterpret_cast<const uint4*>(ip)[3];
            const float mm[12] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w, m2.x, m2.y, m2.z, m2.w};
            float rox, roy, roz, rdx, rdy, rdz;
            xform_ray(mm, wox, woy, woz, wdx, wdy, wdz, rox, roy, roz, rdx, rdy, rdz);
            ray_setup(r, rox, roy, roz, rdx, rdy, rdz);
            push(kEntryReturn, -__builtin_inff());
            cur = m3.x;
            id_base = m3.y;
            inst_slot = ii;
        }
        if (pop) {
            bool left_instance = false;
            for (;;) {
                if (sp == 0) {
                    finish = true;
                    break;
                }
                --sp;
                // one 8-byte LDS read; the spill read is volatile so the two are
                // not merged into one generic (flat) load
                uint2 e;
                if (sp < SLDS) {
                    e = stk[sp * BLOCK + tid];
                } else {
                    const volatile unsigned long long* q = reinterpret_cast<const volatile unsigned long long*>(
                        args.spill + size_t(sp - SLDS) * args.spill_stride + gtid);
                    const unsigned long long w = *q;
                    e = make_uint2(uint32_t(w), uint32_t(w >> 32));
                }
                if (TWO && e.x == kEntryReturn) {
                    left_instance = true;  // back to the world ray; keep popping
                    continue;
                }
                if ((MIXED ? lane_any : ANY_CONST) || __uint_as_float(e.y) <= tbest * kSlack) {
                    cur = e.x;
                    break;
                }
            }
            if (TWO && left_instance) {
                // back in the top level: slab-test fields only (no triangles there)
                r.ox = wox; r.oy = woy; r.oz = woz;
                r.ix = wix; r.iy = wiy; r.iz = wiz;
                id_base = 0;
                inst_slot = 0xFFFFFFFFu;
            }
        }
        if (finish) {
            const bool miss = best == 0xFFFFFFFFu;
            float4 out;
            if (INTERNAL) {
                out = make_float4(miss ? tmax : tbest, __uint_as_float(best), __uint_as_float(miss ? 0xFFFFFFFFu : best_aux),
                                  0.f);
            } else {
                out = make_float4(miss ? tmax : tbest, __uint_as_float(best), miss ? 0.f : bV / bdet,
                                  miss ? 0.f : bW / bdet);
            }
            ((MIXED && lane_any) ? args.hits2 : args.hits)[ray_idx] = out;
            active = false;
        }
    }
    if (COUNT) {
        atomicAdd(args.stats + 0, (unsigned long long)c_nodes);
        atomicAdd(args.stats + 1, (unsigned long long)c_tris);
        atomicAdd(args.stats + 2, (unsigned long long)c_inst);
    }
}

template <int MODE, bool TWO, bool MULTI, bool COUNT, int SLDS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace(TraceArgs args)
{
    trace_body<MODE, TWO, MULTI, COUNT, SLDS, BLOCK>(args);
}

// The renderer's launches (per-ray scenes).  amdgpu_waves_per_eu(6) holds
// them to 80 VGPRs, six waves per SIMD: without it the mixed-mode body took
// 85 (five waves); held to 80 it keeps the same instructions with no spills
// and 64-spp renders run 3% faster.  The ray-query kernels stay without the
// hint, which cost them 2% (same registers, another schedule).
template <int MODE, bool TWO, bool MULTI, bool COUNT, int SLDS, int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6))) void k_trace_render(TraceArgs args)
{
    trace_body<MODE, TWO, MULTI, COUNT, SLDS, BLOCK>(args);
}

// Host-side launcher: picks the instantiation matching (block, stack) and mode.
struct TraceLaunch {
    int mode;
    bool two_level;
    bool multi;
    bool count;
    bool coop;     // kModeMixedInternal: one wave per ray (trace_coop.hpp), for launches that hold few rays
    uint32_t block;
    uint32_t stack_lds;
    uint32_t grid;
};

hipError_t launch_trace(const TraceLaunch& l, const TraceArgs& args, hipStream_t stream);
// Blocks of the launch's kernel that fit on one CU at once (0 if unknown).
uint32_t trace_blocks_per_cu(const TraceLaunch& l);
bool trace_config_supported(uint32_t block, uint32_t stack_lds);

}  // namespace fnd
