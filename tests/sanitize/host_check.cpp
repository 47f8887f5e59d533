// host_check.cpp — the host code of libfnd_pt (scene generators, BVH build
// and encoding, mesh ingestion, the frame arena) and the CPU oracle, built
// with AddressSanitizer and UndefinedBehaviorSanitizer and driven through
// their C interfaces on small inputs.  Test infrastructure: run by
// tests/test_sanitized_host.py, exits non-zero (or the sanitizers abort) on
// the first problem.
//
//  1. every generator kind at small sizes -> fnd_scene_encode; the oracle's
//     binary walk and its walk over the encoded bytes must give the same
//     closest-hit records on camera and random rays; the oracle renders a
//     small image in its three estimator modes; refused rays;
//  2. scene descriptions that must be refused;
//  3. mutated OBJ files (fnd_obj_load) and random stream descriptions
//     (fnd_mesh_from_streams) over buffers of exactly the stated size;
//  4. random frame-arena programs;
//  5. the radiance-cache oracle's loss and MLP gradients on a small batch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fnd_pt.h"
#include "../../oracle/pt_oracle.h"

namespace {

int g_failures = 0;

#define CHECK(cond, ...)                                               \
    do {                                                               \
        if (!(cond)) {                                                 \
            std::fprintf(stderr, "host_check: %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);                         \
            std::fprintf(stderr, "\n");                                \
            ++g_failures;                                              \
        }                                                              \
    } while (0)

struct Rng {
    orc_pcg32 r;
    explicit Rng(uint64_t seed) { orc_pcg32_srandom(&r, seed, 7); }
    uint32_t u32() { return orc_pcg32_next(&r); }
    float uniform() { return orc_pcg32_float(&r); }
    uint32_t below(uint32_t n) { return n ? u32() % n : 0; }
};

// The oracle's copy of a scene description (unindexed triangles).
orc_scene* oracle_scene(const fnd_scene_desc& d, bool shading)
{
    orc_scene* s = orc_scene_new();
    for (uint32_t m = 0; m < d.num_meshes; ++m) {
        const fnd_mesh_desc& md = d.meshes[m];
        std::vector<float> tv(size_t(md.num_triangles) * 9);
        for (uint32_t t = 0; t < md.num_triangles; ++t)
            for (int k = 0; k < 3; ++k)
                std::memcpy(&tv[9 * size_t(t) + 3 * k], md.positions + 3 * size_t(md.indices[3 * size_t(t) + k]), 12);
        CHECK(orc_scene_add_mesh(s, tv.data(), md.num_triangles) == int(m), "add mesh %u", m);
        if (shading) {
            std::vector<uint32_t> mat(md.num_triangles, 0);
            std::vector<int32_t> light(md.num_triangles, -1);
            if (md.material_ids) std::memcpy(mat.data(), md.material_ids, 4 * size_t(md.num_triangles));
            if (md.light_ids) std::memcpy(light.data(), md.light_ids, 4 * size_t(md.num_triangles));
            CHECK(orc_scene_set_mesh_attrs(s, m, mat.data(), light.data()) == 0, "mesh attrs");
        }
    }
    for (uint32_t i = 0; i < d.num_instances; ++i)
        CHECK(orc_scene_add_instance(s, d.instances[i].mesh, d.instances[i].object_to_world,
                                     d.instances[i].world_to_object) == int(i),
              "add instance %u", i);
    if (shading) {
        std::vector<float> mats(8 * size_t(d.num_materials ? d.num_materials : 1), 0.0f);
        for (uint32_t k = 0; k < d.num_materials; ++k) {
            std::memcpy(&mats[8 * k], &d.materials[k].type, 4);
            std::memcpy(&mats[8 * k + 1], d.materials[k].albedo, 12);
            mats[8 * k + 4] = d.materials[k].roughness;
            std::memcpy(&mats[8 * k + 5], d.materials[k].emission, 12);
        }
        CHECK(orc_scene_set_materials(s, mats.data(), d.num_materials ? d.num_materials : 1) == 0, "materials");
        std::vector<float> lights(12 * size_t(d.num_lights) + 1);
        for (uint32_t k = 0; k < d.num_lights; ++k) std::memcpy(&lights[12 * k], &d.lights[k], 48);
        CHECK(orc_scene_set_lights(s, d.num_lights ? lights.data() : nullptr, d.num_lights) == 0, "lights");
        orc_scene_set_background(s, d.background);
    }
    CHECK(orc_scene_build(s, 4, 16, 4) == 0, "oracle build");
    CHECK(orc_scene_check(s) == 0, "oracle tree invariants");
    return s;
}

void check_generated(uint32_t kind, uint32_t shrink_soup, uint32_t subdivision, uint32_t instances)
{
    fnd_gen_params p{};
    CHECK(fnd_gen_default_params(kind, &p) == FND_OK, "default params %u", kind);
    p.soup_triangles = shrink_soup;
    if (subdivision) p.subdivision = subdivision;
    if (instances) p.instances = instances;
    fnd_generated_scene* g = nullptr;
    CHECK(fnd_gen_scene(&p, &g) == FND_OK && g, "generate kind %u", kind);
    if (!g) return;
    fnd_scene_desc d{};
    fnd_camera cam{};
    CHECK(fnd_gen_scene_get(g, &d, &cam) == FND_OK, "scene get");
    char err[256] = {};
    fnd_host_scene* hs = nullptr;
    CHECK(fnd_scene_encode(&d, &hs, err, sizeof err) == FND_OK && hs, "encode kind %u: %s", kind, err);
    fnd_host_scene_view v{};
    if (hs) CHECK(fnd_host_scene_get(hs, &v) == FND_OK && v.num_nodes > 0, "host scene view");

    orc_scene* os = oracle_scene(d, true);
    const uint32_t w = 24, h = 16;
    const float cam10[10] = {cam.eye[0], cam.eye[1], cam.eye[2], cam.at[0], cam.at[1], cam.at[2],
                             cam.up[0], cam.up[1], cam.up[2], cam.fov_y_deg};
    std::vector<float> rays(size_t(w) * h * 8 + 8 * 64);
    orc_camera_rays(cam10, w, h, nullptr, nullptr, nullptr, 0, 1, 0, rays.data());
    Rng rng(kind * 131 + subdivision);
    for (size_t i = size_t(w) * h; i < rays.size() / 8; ++i) {  // random rays through the scene's box
        float* r = &rays[8 * i];
        for (int a = 0; a < 3; ++a) {
            r[a] = v.info.bounds_lo[a] + (v.info.bounds_hi[a] - v.info.bounds_lo[a]) * rng.uniform();
            r[4 + a] = rng.uniform() - 0.5f;
        }
        r[3] = 0.0f;
        r[7] = 3e38f;
    }
    // refused rays: non-finite components and NaN bounds
    rays[8 * (w * h) + 4] = NAN;
    rays[8 * (w * h + 1) + 0] = INFINITY;
    rays[8 * (w * h + 2) + 7] = NAN;
    const uint32_t n = uint32_t(rays.size() / 8);
    std::vector<float> a(4 * size_t(n)), b(4 * size_t(n));
    std::vector<uint32_t> cnt(3 * size_t(n));
    for (int anyhit = 0; anyhit < 2; ++anyhit) {
        orc_trace(os, rays.data(), n, anyhit, 0, 2, a.data(), nullptr);
        if (hs)
            orc_trace_wide(v.nodes, v.tris, v.instances, v.root, int(v.two_level), v.info.stack_entries, rays.data(), n,
                           anyhit, 2, b.data(), cnt.data());
        for (uint32_t i = 0; i < n && hs; ++i) {
            uint32_t pa, pb;
            std::memcpy(&pa, &a[4 * i + 1], 4);
            std::memcpy(&pb, &b[4 * i + 1], 4);
            if (anyhit)
                CHECK((pa == 0xFFFFFFFFu) == (pb == 0xFFFFFFFFu), "kind %u ray %u: any-hit differs", kind, i);
            else
                CHECK(std::memcmp(&a[4 * i], &b[4 * i], 16) == 0, "kind %u ray %u: wide walk differs", kind, i);
        }
    }
    CHECK(orc_wide_max_stack(1) != 0xFFFFFFFFu, "wide walk stack overflow");
    std::vector<float> acc(size_t(w) * h * 3);
    std::vector<double> acc64(size_t(w) * h * 3);
    for (uint32_t mode = 0; mode < 3; ++mode) {
        std::fill(acc.begin(), acc.end(), 0.0f);
        std::fill(acc64.begin(), acc64.end(), 0.0);
        orc_render(os, cam10, w, h, 2, 0, 5, 2, 3, mode, 2, acc.data(), acc64.data());
        for (float x : acc) CHECK(std::isfinite(x) && x >= 0.0f, "render mode %u: pixel %g", mode, double(x));
    }
    orc_scene_free(os);
    fnd_host_scene_free(hs);
    fnd_gen_scene_free(g);
}

void check_refused_scenes()
{
    const float pos[12] = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, NAN};
    const uint32_t good[3] = {0, 1, 2}, far[3] = {0, 1, 7}, nan_tri[3] = {0, 1, 3};
    const uint32_t mat_bad[1] = {5};
    char err[128];
    fnd_host_scene* hs = nullptr;
    fnd_mesh_desc m{};
    m.positions = pos;
    m.num_vertices = 4;
    m.num_triangles = 1;
    fnd_scene_desc d{};
    d.meshes = &m;
    d.num_meshes = 1;
    m.indices = far;
    CHECK(fnd_scene_encode(&d, &hs, err, sizeof err) == FND_E_INVALID_ARG && !hs, "index out of range accepted");
    m.indices = nan_tri;
    CHECK(fnd_scene_encode(&d, &hs, err, sizeof err) == FND_E_INVALID_ARG && !hs, "NaN vertex accepted");
    m.indices = good;
    m.material_ids = mat_bad;
    CHECK(fnd_scene_encode(&d, &hs, err, sizeof err) == FND_E_INVALID_ARG && !hs, "bad material accepted");
    m.material_ids = nullptr;
    fnd_instance_desc inst{};
    inst.mesh = 3;
    d.instances = &inst;
    d.num_instances = 1;
    CHECK(fnd_scene_encode(&d, &hs, err, sizeof err) == FND_E_INVALID_ARG && !hs, "instance of a missing mesh accepted");
    d.num_meshes = 0;
    d.num_instances = 0;
    CHECK(fnd_scene_encode(&d, &hs, err, sizeof err) != FND_OK && !hs, "empty scene accepted");
}

void check_obj_fuzz()
{
    const std::string base =
        "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\nf 1 2 3 4\nf -1 -2 -3\nf 1/1/1 2/1/1 3//1\n"
        "o x\nv 0.5 0.5 1\nf 5 1 2\n";
    const char* inserts[] = {"-", "/", "//", " 0", " -99999999999", " 4294967296", "nan", "inf", "\n", "#",
                             "f", "v", " 1e39", "\t", "f 1 2", "f 9 9 9\n", "v 1 2\n"};
    char path[] = "/tmp/fnd_host_check_XXXXXX";
    const int fd = mkstemp(path);
    CHECK(fd >= 0, "temp file");
    if (fd < 0) return;
    std::FILE* keep = fdopen(fd, "w");
    std::fclose(keep);
    Rng rng(42);
    int ok = 0;
    for (int c = 0; c < 600; ++c) {
        std::string s = base;
        const uint32_t edits = 1 + rng.below(6);
        for (uint32_t e = 0; e < edits; ++e) {
            const size_t at = rng.below(uint32_t(s.size() + 1));
            switch (rng.below(5)) {
            case 0: if (!s.empty()) s[at % s.size()] = char(rng.below(256)); break;
            case 1: s.insert(at, inserts[rng.below(sizeof inserts / sizeof *inserts)]); break;
            case 2: s.erase(at, rng.below(30)); break;
            case 3: s.resize(at); break;
            default: s.insert(at, std::string(1 + rng.below(70000), 'x')); break;
            }
        }
        std::FILE* f = std::fopen(path, "wb");
        std::fwrite(s.data(), 1, s.size(), f);
        std::fclose(f);
        fnd_obj_mesh* m = nullptr;
        char err[128];
        const fnd_status st = fnd_obj_load(path, &m, err, sizeof err);
        if (st == FND_OK) {
            ++ok;
            fnd_mesh_desc md{};
            CHECK(fnd_obj_mesh_get(m, &md) == FND_OK, "obj mesh get");
            for (uint64_t i = 0; i < 3ull * md.num_triangles; ++i)
                CHECK(md.indices[i] < md.num_vertices, "obj index out of range");
            fnd_obj_free(m);
        } else {
            CHECK(m == nullptr, "failed load returned a mesh");
        }
    }
    std::remove(path);
    CHECK(ok > 20, "obj fuzz accepted only %d inputs", ok);
}

void check_stream_fuzz()
{
    Rng rng(9);
    for (int c = 0; c < 4000; ++c) {
        const uint32_t nv = rng.below(30), stride = rng.below(40), comps = 1 + rng.below(4), off = rng.below(24);
        // a buffer a little shorter or longer than nv whole vertices
        const int64_t want = int64_t(nv) * stride - 8 + int64_t(rng.below(24));
        const uint64_t bytes = want > 0 ? uint64_t(want) : 0u;
        const uint32_t bits = rng.below(3) == 0 ? 8 : (rng.below(2) ? 16 : 32);
        const uint32_t ni = 3 * rng.below(8) + (rng.below(6) == 0 ? 1 : 0);
        // buffers of exactly the stated sizes, so any overread is reported
        unsigned char* vb = static_cast<unsigned char*>(std::malloc(bytes ? bytes : 1));
        for (uint64_t i = 0; i < bytes; ++i) vb[i] = (unsigned char)rng.u32();
        const size_t ib_bytes = size_t(ni) * (bits == 16 ? 2 : 4);
        unsigned char* ib = static_cast<unsigned char*>(std::malloc(ib_bytes ? ib_bytes : 1));
        for (uint32_t i = 0; i < ni; ++i) {
            const uint32_t x = rng.below(nv + 2);
            if (bits == 16) {
                const uint16_t y = uint16_t(x);
                std::memcpy(ib + 2 * i, &y, 2);
            } else {
                std::memcpy(ib + 4 * size_t(i), &x, 4);
            }
        }
        fnd_vertex_stream vs{vb, bytes, stride, off, comps, nv};
        fnd_index_stream is{ni ? ib : nullptr, bits, ni};
        std::vector<float> posv(3 * size_t(nv));
        std::vector<uint32_t> idx(ni);
        char err[128];
        (void)fnd_mesh_from_streams(&vs, &is, nv ? posv.data() : nullptr, ni ? idx.data() : nullptr, err, sizeof err);
        std::free(vb);
        std::free(ib);
    }
}

void check_arena()
{
    Rng rng(5);
    alignas(256) static unsigned char block[1 << 16];
    for (int prog = 0; prog < 200; ++prog) {
        fnd_arena a{};
        const uint64_t size = rng.below(sizeof block);
        fnd_arena_reset(&a, block + rng.below(64), size);
        uint64_t prev_end = 0;
        for (int k = 0; k < 100; ++k) {
            const uint64_t sz = rng.below(3000);
            const uint64_t al = 1ull << rng.below(12);
            void* p = rng.below(2) ? fnd_arena_alloc(&a, sz) : fnd_arena_alloc_aligned(&a, sz, al);
            if (p) {
                const uint64_t at = uint64_t(reinterpret_cast<uintptr_t>(p));
                CHECK(at >= a.memory && at + sz <= a.end, "arena block outside the memory");
                CHECK(at >= prev_end, "arena blocks overlap");
                prev_end = at + sz;
                std::memset(p, 0xA5, sz);  // the block really is usable memory
            }
            CHECK(fnd_arena_used(&a) <= size, "used past the end");
        }
    }
    fnd_arena a{};
    fnd_arena_reset(&a, nullptr, 0);
    CHECK(fnd_arena_alloc(&a, 8) == nullptr, "empty arena gave memory");
    CHECK(fnd_arena_alloc_aligned(&a, 8, 3) == nullptr, "non-power-of-two alignment accepted");
}

void check_cache_oracle()
{
    const uint32_t n = 16;
    std::vector<float> mlp(20480), q(10 * n), t(3 * n), pred(3 * n);
    std::vector<float> grid(size_t(16) << 20);
    Rng rng(3);
    for (float& x : mlp) x = (rng.uniform() - 0.5f) * 0.2f;
    for (float& x : grid) x = (rng.uniform() - 0.5f) * 1e-2f;
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) q[10 * i + k] = rng.uniform();
        q[10 * i + 5] = 1.0f;  // normal +z
        for (int k = 0; k < 3; ++k) t[3 * i + k] = rng.uniform();
    }
    std::vector<double> dmlp(20480);
    const float lo[3] = {0, 0, 0};
    const double loss = orc_cache_loss_grad(mlp.data(), grid.data(), q.data(), t.data(), n, lo, 1.0f, 1, 0,
                                            pred.data(), dmlp.data(), nullptr);
    CHECK(std::isfinite(loss) && loss >= 0.0, "cache loss %g", loss);
    std::vector<double> p(64), m(64), v(64), g(64, 0.5);
    orc_adam_step(p.data(), m.data(), v.data(), g.data(), 64, 1e-2, 0.9, 0.99, 1e-15, 1);
    CHECK(std::fabs(p[0] + 1e-2) < 1e-12, "adam first step %g", p[0]);
}

}  // namespace

int main()
{
    check_generated(FND_GEN_CORNELL_SOUP, 1500, 0, 0);
    check_generated(FND_GEN_DISPLACED_SPHERE, 800, 12, 0);
    check_generated(FND_GEN_INSTANCED_SPHERES, 0, 6, 6);
    check_generated(FND_GEN_FURNACE, 0, 4, 0);
    check_refused_scenes();
    check_obj_fuzz();
    check_stream_fuzz();
    check_arena();
    check_cache_oracle();
    if (g_failures) {
        std::fprintf(stderr, "host_check: %d failure(s)\n", g_failures);
        return 1;
    }
    std::printf("host_check: ok\n");
    return 0;
}
