// trace_coop.hpp — cooperative traversal for launches that hold few rays.
//
// The last bounces of a frame trace few rays (tens of thousands on a GPU
// that runs 390k lanes), and each such launch lasts as long as its slowest
// walk: a lane's node steps are sequential, one dependent 64-byte fetch
// after another.  Here a group of G lanes walks one ray, G = 64 / 2^k chosen
// at launch so that the grid's groups about cover the rays (one wave per ray
// for the fewest rays, eight lanes per ray for the most).  Each lane keeps
// its own stack, as in trace_kernel.hpp, and the group's idle lanes take the
// bottom entries (the subtrees nearest the root) of its busy lanes' stacks,
// so the walk's frontier spreads over the group within a few steps and the
// ray's sequential chain shrinks towards the depth of the tree.
//
// The record is the one the single-lane walk returns: a closest-hit record
// is the lexicographic minimum of (t, triangle) over the triangles the walk
// tests in (tmin, tmax), whichever lane tests them, and the group reduces it
// at the end; culling with the group's best distance so far only skips boxes
// and stack entries that cannot hold a hit at or before it (the same
// conservative slack).  An any-hit ray ends when any lane accepts a hit.
// The triangle test, the instance transform and the ray set-up are the same
// functions as the single-lane walk, so the same triangles give the same bits.
//
// Two-level scenes: an entry handed over carries its context, the world or
// an instance, and the lane that takes it forms that space's ray itself.  A
// lane's stack holds entries of its base context at the bottom, then, while
// it is inside an instance entered from the world, a return marker and that
// instance's entries; a marker that reaches the bottom is dropped and the
// lane's base context becomes the instance.
#pragma once
#include "trace_kernel.hpp"

namespace fnd {

template <bool TWO, bool MULTI, int SLDS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace_coop(TraceArgs args)
{
    static_assert(BLOCK % kWave == 0, "whole waves");
    __shared__ uint2 stk[SLDS * BLOCK];
    __shared__ uint2 x_entry[BLOCK];  // entries handed over this step, by the giver's rank in its group
    __shared__ uint32_t x_ctx[BLOCK];  // their contexts: instance slot, 0xFFFFFFFF for the world

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t gtid = blockIdx.x * BLOCK + tid;
    const uint32_t nrays1 = *args.nrays_dev;
    const uint32_t nrays = nrays1 + *args.nrays2_dev;
    if (nrays >= args.coop_below) return;  // the single-lane launch beside this one traces them
    if (args.ray_total && gtid == 0) atomicAdd(args.ray_total, (unsigned long long)nrays);

    // group size: the smallest G in {64, 32, 16, 8} whose groups cover the rays
    const uint32_t waves = gridDim.x * (BLOCK / kWave);
    uint32_t G = 64;
    while (G > 8 && uint64_t(waves) * (kWave / G) < nrays) G >>= 1;
    const uint32_t glane = lane & (G - 1);
    const uint32_t gbase = lane - glane;
    const uint32_t xslot = (tid - lane) + gbase;
    const uint64_t gmask = G == 64 ? ~0ull : (((1ull << G) - 1ull) << gbase);
    const uint32_t groups = gridDim.x * (BLOCK / G);

    SceneDev sc = args.scene;
    RayK r;
    float wox = 0.f, woy = 0.f, woz = 0.f, wdx = 0.f, wdy = 0.f, wdz = 0.f, wix = 0.f, wiy = 0.f, wiz = 0.f;
    uint32_t sp = 0, bot = 0;

    auto push = [&](uint32_t e, float t) {
        const uint2 v = make_uint2(e, __float_as_uint(t));
        if (sp < SLDS)
            stk[sp * BLOCK + tid] = v;
        else
            args.spill[size_t(sp - SLDS) * args.spill_stride + gtid] = v;
        ++sp;
    };
    auto entry_at = [&](uint32_t pos) -> uint2 {
        if (pos < SLDS) return stk[pos * BLOCK + tid];
        const volatile unsigned long long* q =
            reinterpret_cast<const volatile unsigned long long*>(args.spill + size_t(pos - SLDS) * args.spill_stride + gtid);
        const unsigned long long w = *q;
        return make_uint2(uint32_t(w), uint32_t(w >> 32));
    };

    // the group's ray (the same in all its lanes): the group's index first,
    // then from the work counter
    bool live = false, retired = false, any = false;
    uint32_t idx = gtid / G, ray_idx = 0;
    float tmin = 0.f, tmax = 0.f;
    // the lane's part of the walk
    float tbest = 0.f;
    uint32_t best = 0xFFFFFFFFu, best_aux = 0xFFFFFFFFu;
    bool has = false, found = false;
    uint32_t cur = 0, id_base = 0, inst_slot = 0xFFFFFFFFu, base_ctx = 0xFFFFFFFFu;

    for (;;) {
        if (!live && !retired) {
            if (idx >= nrays) {
                retired = true;
            } else {
                any = idx >= nrays1;
                ray_idx = any ? idx - nrays1 : idx;
                const float4* rp = reinterpret_cast<const float4*>((any ? args.rays2 : args.rays) + ray_idx);
                const float4 ra = rp[0];
                const float4 rb = rp[1];
                // refused rays as in the single-lane walk: empty range, a miss after the root
                const float fin =
                    ((ra.x - ra.x) + (ra.y - ra.y)) + ((ra.z - ra.z) + (rb.x - rb.x)) + ((rb.y - rb.y) + (rb.z - rb.z));
                const bool refused = fin != fin || ra.w != ra.w || rb.w != rb.w;
                if (MULTI) sc = args.scenes[(any ? args.ray_scene2 : args.ray_scene)[ray_idx]];
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
                has = glane == 0;
                found = false;
                cur = sc.root;
                id_base = 0;
                inst_slot = 0xFFFFFFFFu;
                base_ctx = 0xFFFFFFFFu;
                sp = 0;
                bot = 0;
                live = true;
            }
        }
        if (__ballot(live) == 0) break;
        if (!live) continue;

        if ((__ballot(has) & gmask) == 0 || (any && (__ballot(found) & gmask) != 0)) {
            // the group's ray is done: its record is the first lane's accepted
            // any hit, or the lexicographic minimum of (t, triangle) over the lanes
            if (any) {
                const uint64_t f = __ballot(found) & gmask;
                if (f != 0) {
                    const int src = __builtin_ctzll(f);
                    tbest = __shfl(tbest, src);
                    best = uint32_t(__shfl(int(best), src));
                    best_aux = uint32_t(__shfl(int(best_aux), src));
                }
            } else {
                for (uint32_t o = G >> 1; o > 0; o >>= 1) {
                    const float t2 = __shfl_xor(tbest, int(o));
                    const uint32_t b2 = uint32_t(__shfl_xor(int(best), int(o)));
                    const uint32_t a2 = uint32_t(__shfl_xor(int(best_aux), int(o)));
                    if (t2 < tbest || (t2 == tbest && b2 < best)) {
                        tbest = t2;
                        best = b2;
                        best_aux = a2;
                    }
                }
            }
            if (glane == 0) {
                const bool miss = best == 0xFFFFFFFFu;
                (any ? args.hits2 : args.hits)[ray_idx] = make_float4(
                    miss ? tmax : tbest, __uint_as_float(best), __uint_as_float(miss ? 0xFFFFFFFFu : best_aux), 0.f);
            }
            // the next ray: past the grid's first one per group, from the counter
            uint32_t nxt = 0;
            if (glane == 0) nxt = atomicAdd(args.counter, 1u);
            idx = groups + uint32_t(__shfl(int(nxt), int(gbase)));
            live = false;
            has = false;
            continue;
        }

        if (TWO && bot < sp && entry_at(bot).x == kEntryReturn) {
            // everything left on this lane's stack lies in its current instance
            ++bot;
            base_ctx = inst_slot;
        }
        // idle lanes take the bottom entries of the group's lanes with entries to spare
        const uint64_t idle = __ballot(!has) & gmask;
        const uint64_t giver = __ballot(bot < sp) & gmask;
        if (idle != 0 && giver != 0) {
            const uint32_t k = min(uint32_t(__popcll(idle)), uint32_t(__popcll(giver)));
            if (bot < sp) {
                const uint32_t g = lanes_below(giver);
                if (g < k) {
                    x_entry[xslot + g] = entry_at(bot);
                    x_ctx[xslot + g] = base_ctx;
                    ++bot;
                }
            }
            // the exchange is between lanes of this wave: a wave-scope fence
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (!has) {
                const uint32_t i = lanes_below(idle);
                if (i < k) {
                    const uint2 e = x_entry[xslot + i];
                    const uint32_t c = x_ctx[xslot + i];
                    cur = e.x;
                    sp = 0;
                    bot = 0;
                    base_ctx = c;
                    has = true;
                    if (TWO) {
                        if (c == 0xFFFFFFFFu) {
                            ray_setup(r, wox, woy, woz, wdx, wdy, wdz);
                            id_base = 0;
                            inst_slot = 0xFFFFFFFFu;
                        } else {
                            const float4* ip = reinterpret_cast<const float4*>(sc.insts + c);
                            const float4 m0 = ip[0], m1 = ip[1], m2 = ip[2];
                            const uint4 m3 = reinterpret_cast<const uint4*>(ip)[3];
                            const float mm[12] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w, m2.x, m2.y, m2.z, m2.w};
                            float rox, roy, roz, rdx, rdy, rdz;
                            xform_ray(mm, wox, woy, woz, wdx, wdy, wdz, rox, roy, roz, rdx, rdy, rdz);
                            ray_setup(r, rox, roy, roz, rdx, rdy, rdz);
                            id_base = m3.y;
                            inst_slot = c;
                        }
                    }
                }
            }
            // the exchange is between lanes of this wave: a wave-scope fence
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        // the group's best distance so far bounds every lane's culling
        float tcull = tbest;
        if (!any) {
            for (uint32_t o = G >> 1; o > 0; o >>= 1) tcull = fminf(tcull, __shfl_xor(tcull, int(o)));
        }
        if (!has) continue;

        // work selection as in the single-lane walk
        const bool is_node = cur < kEntryInstance;
        const bool is_tri = (cur & kEntryTri) != 0u;
        const uint32_t s_node = uint32_t(__popcll(__ballot(is_node))) * kCostWeightNode;
        const uint32_t s_tri = uint32_t(__popcll(__ballot(is_tri))) * kCostWeightTri;
        const uint32_t s_inst = TWO ? uint32_t(__popcll(__ballot(!is_node && !is_tri))) * kCostWeightInst : 0u;
        const uint32_t s_best = max(s_node, max(s_tri, s_inst));
        const bool run_node = 2u * s_node >= s_best, run_tri = 2u * s_tri >= s_best, run_inst = 2u * s_inst >= s_best;

        bool pop = false;
        if (is_node && run_node) {
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
                                       fminf(fmaf(ubyte_f(fz, k), sz, cz), tcull));
                const bool hit = ek[k] != 0u && tn <= tf * kSlack;
                tk[k] = hit ? tn : __builtin_inff();
            }
            if (!any) {
                // nearest child first; the others pushed farthest first
                auto cs = [&](int a, int b) {
                    const bool s = tk[b] < tk[a];
                    const float ta = s ? tk[b] : tk[a], tb = s ? tk[a] : tk[b];
                    const uint32_t ea = s ? ek[b] : ek[a], eb = s ? ek[a] : ek[b];
                    tk[a] = ta; tk[b] = tb; ek[a] = ea; ek[b] = eb;
                };
                cs(0, 1); cs(2, 3); cs(0, 2); cs(1, 3); cs(1, 2);
                if (tk[3] != __builtin_inff()) push(ek[3], tk[3]);
                if (tk[2] != __builtin_inff()) push(ek[2], tk[2]);
                if (tk[1] != __builtin_inff()) push(ek[1], tk[1]);
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
            const uint32_t ti = (cur >> 3) & 0x0FFFFFFFu;
            const float4* tp = sc.tris + 3ull * ti;
            const float4 a = tp[0];
            const float4 b = tp[1];
            const float4 c = tp[2];
            float t, Vn, Wn, detn;
            if (isect_tri(r, a, b, c, t, Vn, Wn, detn) && t > tmin && t < tmax) {
                (void)Vn; (void)Wn; (void)detn; (void)id_base;
                if (any) {
                    tbest = t; best = ti; best_aux = inst_slot;
                    found = true;
                } else if (t < tbest || (t == tbest && ti < best)) {
                    tbest = t; best = ti; best_aux = inst_slot;
                }
            }
            if (!found) {
                if ((cur & 7u) > 1u)
                    cur += 7u;  // first + 1, count - 1
                else
                    pop = true;
            }
        } else if (TWO && !is_node && !is_tri && run_inst) {
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
                if (sp == bot) {
                    has = false;
                    break;
                }
                --sp;
                const uint2 e = entry_at(sp);
                if (TWO && e.x == kEntryReturn) {
                    left_instance = true;
                    continue;
                }
                if (any || __uint_as_float(e.y) <= tcull * kSlack) {
                    cur = e.x;
                    break;
                }
            }
            if (TWO && left_instance) {
                // back in the top level (instances only): slab-test fields only
                r.ox = wox; r.oy = woy; r.oz = woz;
                r.ix = wix; r.iy = wiy; r.iz = wiz;
                id_base = 0;
                inst_slot = 0xFFFFFFFFu;
                base_ctx = 0xFFFFFFFFu;
            }
        }
        if (any && found) has = false;
    }
}

}  // namespace fnd
