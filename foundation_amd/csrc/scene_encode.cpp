// scene_encode.cpp — host half of fnd_pt_scene_create: validates the caller's
// arrays, builds one wide tree per mesh and a top-level tree over instances,
// and lays everything out exactly as the kernels read it (pt_layout.hpp).
//
// This replaces the vertex/index staging of the reference's constructor
// (src/Renderer/Renderer.cpp:133-198), which copies a 4-vertex quad.  It is
// also exported on its own (fnd_scene_encode) so the encoded layout can be
// checked on a host without a GPU.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <thread>

#include "scene_encode.hpp"
#include "host_pool.hpp"

namespace fnd {

namespace {
double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

fnd_status fail(std::string& err, fnd_status st, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return st;
}
}  // namespace

fnd_status encode_scene(const fnd_scene_desc* d, HostScene& out, std::string& err)
{
    const double t0 = now_ms();
    out = HostScene{};
    if (!d) return fail(err, FND_E_INVALID_ARG, "null scene description");
    if (!d->meshes || d->num_meshes == 0) return fail(err, FND_E_INVALID_ARG, "scene has no meshes");
    if (d->num_instances == 0 && d->num_meshes != 1)
        return fail(err, FND_E_INVALID_ARG, "a scene without instances takes exactly one mesh");
    if (d->num_instances && !d->instances) return fail(err, FND_E_INVALID_ARG, "instances is null");
    if (d->num_instances > kMaxInstances)
        return fail(err, FND_E_INVALID_ARG, "at most %u instances", kMaxInstances);
    if (d->num_lights && !d->lights) return fail(err, FND_E_INVALID_ARG, "lights is null");
    if (d->num_materials && !d->materials) return fail(err, FND_E_INVALID_ARG, "materials is null");
    uint64_t total_tris = 0;
    for (uint32_t m = 0; m < d->num_meshes; ++m) {
        const fnd_mesh_desc& md = d->meshes[m];
        if (md.num_triangles == 0 || !md.indices || !md.positions || md.num_vertices == 0)
            return fail(err, FND_E_INVALID_ARG, "mesh %u is empty or has null arrays", m);
        for (uint64_t i = 0; i < 3ull * md.num_triangles; ++i)
            if (md.indices[i] >= md.num_vertices)
                return fail(err, FND_E_INVALID_ARG, "mesh %u: index %u out of range", m, md.indices[i]);
        for (uint64_t i = 0; i < 3ull * md.num_vertices; ++i)
            if (!std::isfinite(md.positions[i]))
                return fail(err, FND_E_INVALID_ARG, "mesh %u: non-finite vertex coordinate", m);
        if (md.material_ids)
            for (uint32_t t = 0; t < md.num_triangles; ++t)
                if (md.material_ids[t] >= std::max(1u, d->num_materials))
                    return fail(err, FND_E_INVALID_ARG, "mesh %u: material id out of range", m);
        if (md.light_ids)
            for (uint32_t t = 0; t < md.num_triangles; ++t)
                if (md.light_ids[t] >= int32_t(d->num_lights) || md.light_ids[t] < -1)
                    return fail(err, FND_E_INVALID_ARG, "mesh %u: light id out of range", m);
        total_tris += md.num_triangles;
    }
    if (total_tris >= kMaxTriangles)
        return fail(err, FND_E_UNSUPPORTED, "%llu triangles exceed the 2^28 limit of a scene",
                    (unsigned long long)total_tris);
    uint64_t total_prims = 0;
    for (uint32_t i = 0; i < d->num_instances; ++i) {
        const fnd_instance_desc& in = d->instances[i];
        if (in.mesh >= d->num_meshes)
            return fail(err, FND_E_INVALID_ARG, "instance %u names mesh %u of %u", i, in.mesh, d->num_meshes);
        for (int k = 0; k < 12; ++k)
            if (!std::isfinite(in.object_to_world[k]) || !std::isfinite(in.world_to_object[k]))
                return fail(err, FND_E_INVALID_ARG, "instance %u: non-finite transform", i);
        total_prims += d->meshes[in.mesh].num_triangles;
    }
    if (total_prims >= 0xFFFFFFFFull) return fail(err, FND_E_UNSUPPORTED, "more than 2^32-1 primitive ids");
    if (d->num_instances == 0) total_prims = total_tris;

    // ---- one tree per mesh: parallel over meshes, or over subtrees of one mesh
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<WideTree> trees(d->num_meshes);
    std::vector<std::string> errs(d->num_meshes);
    std::vector<int> ok(d->num_meshes, 0);
    auto build_one = [&](uint32_t m, unsigned threads) {
        const fnd_mesh_desc& md = d->meshes[m];
        std::vector<Box> boxes(md.num_triangles);
        for (uint32_t t = 0; t < md.num_triangles; ++t) {
            Box& b = boxes[t];
            for (int a = 0; a < 3; ++a) {
                const float x = md.positions[3ull * md.indices[3ull * t] + a];
                const float y = md.positions[3ull * md.indices[3ull * t + 1] + a];
                const float z = md.positions[3ull * md.indices[3ull * t + 2] + a];
                b.lo[a] = std::min(std::min(x, y), z);
                b.hi[a] = std::max(std::max(x, y), z);
            }
        }
        ok[m] = build_wide_tree(boxes.data(), md.num_triangles, kLeafMax, threads, trees[m], errs[m]) ? 1 : 0;
    };
    try {
        if (d->num_meshes == 1) {
            build_one(0, hw);
        } else {
            std::atomic<uint32_t> next{0};
            run_pool(std::min<unsigned>(hw, d->num_meshes), [&]() {
                for (;;) {
                    const uint32_t m = next.fetch_add(1);
                    if (m >= d->num_meshes) break;
                    build_one(m, 1);
                }
            });
        }
    } catch (const std::bad_alloc&) {
        return fail(err, FND_E_OOM_HOST, "out of host memory building the BVH");
    }
    for (uint32_t m = 0; m < d->num_meshes; ++m)
        if (!ok[m]) return fail(err, FND_E_INVALID_ARG, "mesh %u: %s", m, errs[m].c_str());

    // ---- concatenate trees and leaf triangles
    std::vector<uint32_t> node_off(d->num_meshes), tri_off(d->num_meshes);
    uint64_t nnodes = 0, ntris = 0;
    uint32_t max_blas_depth = 0;
    for (uint32_t m = 0; m < d->num_meshes; ++m) {
        node_off[m] = uint32_t(nnodes);
        tri_off[m] = uint32_t(ntris);
        nnodes += trees[m].nodes.size();
        ntris += trees[m].order.size();
        max_blas_depth = std::max(max_blas_depth, trees[m].max_depth);
    }
    WideTree top;
    Box scene_bounds;
    for (int a = 0; a < 3; ++a) {
        scene_bounds.lo[a] = INFINITY;
        scene_bounds.hi[a] = -INFINITY;
    }
    if (d->num_instances) {
        std::vector<Box> ib(d->num_instances);
        std::vector<uint32_t> id_base(d->num_instances);
        // Each instance's world box is the mapped union of its mesh's subtree
        // boxes `depth` levels below the root, not its mapped root box: a
        // rotated instance's box then hugs its geometry, and fewer rays enter
        // instances they miss.  The depth is the largest in [2, 7] that keeps
        // the work within 2^22 box transforms, and 2 where none does (16,384
        // boxes per mesh for 100 instances; 16 for a million).
        int depth = 2;
        while (depth < 7 && (uint64_t(d->num_instances) << (2 * (depth + 1))) <= (1ull << 22)) ++depth;
        std::vector<std::vector<DBox>> cover(d->num_meshes);
        for (uint32_t m = 0; m < d->num_meshes; ++m) cover[m] = cover_boxes(trees[m], depth);
        uint32_t base = 0;
        for (uint32_t i = 0; i < d->num_instances; ++i) {
            const fnd_instance_desc& in = d->instances[i];
            id_base[i] = base;
            base += d->meshes[in.mesh].num_triangles;
            ib[i] = transform_cover_conservative(cover[in.mesh], in.object_to_world);
            for (int a = 0; a < 3; ++a) {
                scene_bounds.lo[a] = std::min(scene_bounds.lo[a], ib[i].lo[a]);
                scene_bounds.hi[a] = std::max(scene_bounds.hi[a], ib[i].hi[a]);
            }
        }
        if (!build_wide_tree(ib.data(), d->num_instances, 1, hw, top, err)) {
            err = "top level: " + err;
            return FND_E_INVALID_ARG;
        }
        out.insts.resize(d->num_instances);
        for (uint32_t j = 0; j < d->num_instances; ++j) {
            const uint32_t i = top.order[j];
            const fnd_instance_desc& in = d->instances[i];
            InstanceRec& r = out.insts[j];
            std::memcpy(r.w2o, in.world_to_object, sizeof r.w2o);
            r.blas_root = node_off[in.mesh];
            r.id_base = id_base[i];
            r.mesh = in.mesh;
            r.orig_index = i;
        }
    } else {
        scene_bounds = trees[0].bounds;
    }
    const uint64_t top_off = nnodes;
    const uint64_t total_nodes = nnodes + top.nodes.size();
    if (total_nodes >= kMaxNodes) return fail(err, FND_E_UNSUPPORTED, "too many BVH nodes");
    try {
        out.nodes.resize(total_nodes);
        out.tris.resize(ntris);
    } catch (const std::bad_alloc&) {
        return fail(err, FND_E_OOM_HOST, "out of host memory staging the scene");
    }
    for (uint32_t m = 0; m < d->num_meshes; ++m) {
        const WideTree& tr = trees[m];
        for (size_t k = 0; k < tr.nodes.size(); ++k) {
            WideNode w = tr.nodes[k];
            w.child_base += node_off[m];
            w.prim_base += tri_off[m];
            for (int c = 0; c < 4; ++c) {
                if (!((w.imask >> (4 + c)) & 1u)) continue;
                // node indices move by the mesh's node offset, triangle ranges
                // by its triangle offset (first << 3 | count)
                w.entry[c] += ((w.imask >> c) & 1u) ? node_off[m] : tri_off[m] << 3;
            }
            out.nodes[node_off[m] + k] = w;
        }
        const fnd_mesh_desc& md = d->meshes[m];
        for (size_t j = 0; j < tr.order.size(); ++j) {
            const uint32_t t = tr.order[j];
            LeafTri& lt = out.tris[tri_off[m] + j];
            std::memcpy(lt.v0, md.positions + 3ull * md.indices[3ull * t], 12);
            std::memcpy(lt.v1, md.positions + 3ull * md.indices[3ull * t + 1], 12);
            std::memcpy(lt.v2, md.positions + 3ull * md.indices[3ull * t + 2], 12);
            lt.local_id = t;
            lt.material = md.material_ids ? md.material_ids[t] : 0u;
            lt.light = md.light_ids ? md.light_ids[t] : -1;
        }
    }
    for (size_t k = 0; k < top.nodes.size(); ++k) {
        WideNode w = top.nodes[k];
        w.child_base += uint32_t(top_off);
        for (int c = 0; c < 4; ++c) {
            if (!((w.imask >> (4 + c)) & 1u)) continue;
            // the top tree's leaves are instances, one per leaf: retag them
            w.entry[c] = ((w.imask >> c) & 1u) ? w.entry[c] + uint32_t(top_off)
                                               : kEntryInstance | (w.entry[c] & ~kEntryTri);
        }
        out.nodes[top_off + k] = w;
    }

    // ---- materials and lights
    out.materials.resize(std::max(1u, d->num_materials));
    if (d->num_materials == 0) out.materials[0] = DeviceMaterial{{0.5f, 0.5f, 0.5f}, FND_MAT_LAMBERT, {0, 0, 0}, 1.0f};
    for (uint32_t i = 0; i < d->num_materials; ++i) {
        const fnd_material& s = d->materials[i];
        if (s.type > FND_MAT_EMISSIVE) return fail(err, FND_E_INVALID_ARG, "material %u: unknown type", i);
        DeviceMaterial& m = out.materials[i];
        std::memcpy(m.albedo, s.albedo, 12);
        m.type = s.type;
        std::memcpy(m.emission, s.emission, 12);
        m.roughness = std::min(1.0f, std::max(0.02f, s.roughness));
    }
    out.lights.resize(std::max(1u, d->num_lights));
    {
        double total = 0;
        std::vector<double> power(d->num_lights);
        for (uint32_t i = 0; i < d->num_lights; ++i) {
            const fnd_quad_light& L = d->lights[i];
            const double cx = double(L.edge0[1]) * L.edge1[2] - double(L.edge0[2]) * L.edge1[1];
            const double cy = double(L.edge0[2]) * L.edge1[0] - double(L.edge0[0]) * L.edge1[2];
            const double cz = double(L.edge0[0]) * L.edge1[1] - double(L.edge0[1]) * L.edge1[0];
            const double area = std::sqrt(cx * cx + cy * cy + cz * cz);
            if (!(area > 0)) return fail(err, FND_E_INVALID_ARG, "light %u has zero area", i);
            const double lum = (double(L.radiance[0]) + L.radiance[1] + L.radiance[2]) / 3.0;
            power[i] = std::max(1e-12, lum * area);
            total += power[i];
            DeviceLight& D = out.lights[i];
            std::memcpy(D.corner, L.corner, 12);
            std::memcpy(D.edge0, L.edge0, 12);
            std::memcpy(D.edge1, L.edge1, 12);
            std::memcpy(D.radiance, L.radiance, 12);
            D.area = float(area);
        }
        double acc = 0;
        for (uint32_t i = 0; i < d->num_lights; ++i) {
            acc += power[i];
            out.lights[i].cdf = (i + 1 == d->num_lights) ? 1.0f : float(acc / total);
            out.lights[i].pdf_select = float(power[i] / total);
        }
    }
    out.num_lights = d->num_lights;
    std::memcpy(out.background, d->background, 12);
    out.two_level = d->num_instances != 0;
    out.root = out.two_level ? uint32_t(top_off) : 0u;
    fnd_scene_info& info = out.info;
    info.num_nodes = total_nodes;
    info.num_triangles = ntris;
    info.num_prims = total_prims;
    info.max_depth = max_blas_depth + top.max_depth;
    info.stack_entries = 3u * (max_blas_depth + top.max_depth) + 2u;
    info.device_bytes = total_nodes * sizeof(WideNode) + ntris * sizeof(LeafTri) +
                        out.insts.size() * sizeof(InstanceRec) + out.materials.size() * sizeof(DeviceMaterial) +
                        out.lights.size() * sizeof(DeviceLight);
    std::memcpy(info.bounds_lo, scene_bounds.lo, 12);
    std::memcpy(info.bounds_hi, scene_bounds.hi, 12);
    info.build_ms = now_ms() - t0;
    return FND_OK;
}

}  // namespace fnd

struct fnd_host_scene {
    fnd::HostScene s;
};

extern "C" {

fnd_status fnd_scene_encode(const fnd_scene_desc* desc, fnd_host_scene** out, char* err, size_t err_len)
{
    if (!out) return FND_E_INVALID_ARG;
    *out = nullptr;
    return fnd::guard_host(err, err_len, [&]() -> fnd_status {
        std::unique_ptr<fnd_host_scene> h(new fnd_host_scene());
        std::string e;
        const fnd_status st = fnd::encode_scene(desc, h->s, e);
        if (st != FND_OK) {
            if (err && err_len) snprintf(err, err_len, "%s", e.c_str());
            return st;
        }
        *out = h.release();
        return FND_OK;
    });
}

fnd_status fnd_host_scene_get(const fnd_host_scene* h, fnd_host_scene_view* v)
{
    if (!h || !v) return FND_E_INVALID_ARG;
    std::memset(v, 0, sizeof(*v));
    v->nodes = h->s.nodes.data();
    v->num_nodes = h->s.nodes.size();
    v->tris = h->s.tris.data();
    v->num_tris = h->s.tris.size();
    v->instances = h->s.insts.empty() ? nullptr : h->s.insts.data();
    v->num_instances = uint32_t(h->s.insts.size());
    v->root = h->s.root;
    v->two_level = h->s.two_level ? 1u : 0u;
    v->info = h->s.info;
    return FND_OK;
}

void fnd_host_scene_free(fnd_host_scene* h) { delete h; }

}  // extern "C"
