// scene_gen.cpp — procedural scenes of the benchmark configurations.
//
// The reference has no scene content beyond a textured quad
// (src/Renderer/Renderer.cpp:133-198), so these generators define the inputs:
//  * Cornell box (5 quads + 2 boxes = 34 triangles) + a random triangle soup,
//    one quad light of radiance 15, Lambert albedo uniform in [0.2, 0.8]
//    per triangle and channel.
//  * Displaced icospheres: an icosahedron whose faces are split into n x n
//    triangles (20 n^2 triangles; n = 2^k gives the k-times subdivided
//    icosphere's count), projected to the unit sphere, radius displaced by
//    1 + amplitude * fBm value noise.  Vertices on shared edges are computed
//    from a canonical (sorted-corner) barycentric form, so the mesh is closed
//    and indexed.
//  * Random soup: centre uniform in a box, size log-uniform in [0.005, 0.1],
//    three vertices at centre + 0.57735 * size * (uniform unit vector).
// All randomness is PCG32 (O'Neill 2014) keyed by the generator seed.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/fnd_pt.h"
#include "host_pool.hpp"

namespace {

struct Pcg {
    uint64_t state = 0, inc = 1;
    Pcg(uint64_t initstate, uint64_t initseq)
    {
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    uint32_t next()
    {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    double uniform() { return double(next() >> 8) * 0x1p-24; }
};

struct MeshData {
    std::vector<float> pos;
    std::vector<uint32_t> idx;
    std::vector<uint32_t> mat;
    std::vector<int32_t> light;
};

uint32_t hash3(int32_t x, int32_t y, int32_t z, uint32_t seed)
{
    uint64_t h = uint64_t(uint32_t(x)) * 0x9E3779B97F4A7C15ULL;
    h ^= uint64_t(uint32_t(y)) * 0xC2B2AE3D27D4EB4FULL + (h << 6) + (h >> 2);
    h ^= uint64_t(uint32_t(z)) * 0x165667B19E3779F9ULL + (h << 6) + (h >> 2);
    h ^= uint64_t(seed) * 0x27D4EB2F165667C5ULL;
    h ^= h >> 33;
    h *= 0xFF51AFD7ED558CCDULL;
    h ^= h >> 33;
    h *= 0xC4CEB9FE1A85EC53ULL;
    h ^= h >> 33;
    return uint32_t(h);
}

// Value noise in [-1, 1]: hashed lattice values, trilinear with smoothstep.
double value_noise(double x, double y, double z, uint32_t seed)
{
    const double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    const int32_t ix = int32_t(fx), iy = int32_t(fy), iz = int32_t(fz);
    const double tx = x - fx, ty = y - fy, tz = z - fz;
    const double sx = tx * tx * (3 - 2 * tx), sy = ty * ty * (3 - 2 * ty), sz = tz * tz * (3 - 2 * tz);
    double c[8];
    for (int k = 0; k < 8; ++k)
        c[k] = double(hash3(ix + (k & 1), iy + ((k >> 1) & 1), iz + ((k >> 2) & 1), seed)) * (2.0 / 4294967295.0) - 1.0;
    const double x00 = c[0] + (c[1] - c[0]) * sx, x10 = c[2] + (c[3] - c[2]) * sx;
    const double x01 = c[4] + (c[5] - c[4]) * sx, x11 = c[6] + (c[7] - c[6]) * sx;
    const double y0 = x00 + (x10 - x00) * sy, y1 = x01 + (x11 - x01) * sy;
    return y0 + (y1 - y0) * sz;
}

double fbm(double x, double y, double z, uint32_t octaves, uint32_t seed)
{
    double sum = 0, norm = 0, amp = 1, f = 2.0;
    for (uint32_t o = 0; o < octaves; ++o) {
        sum += amp * value_noise(x * f, y * f, z * f, seed + 7919u * o);
        norm += amp;
        amp *= 0.5;
        f *= 2.0;
    }
    return norm > 0 ? sum / norm : 0.0;
}

const double kIcoT = (1.0 + std::sqrt(5.0)) / 2.0;
const double kIcoV[12][3] = {{-1, kIcoT, 0}, {1, kIcoT, 0},  {-1, -kIcoT, 0}, {1, -kIcoT, 0},
                             {0, -1, kIcoT}, {0, 1, kIcoT},  {0, -1, -kIcoT}, {0, 1, -kIcoT},
                             {kIcoT, 0, -1}, {kIcoT, 0, 1},  {-kIcoT, 0, -1}, {-kIcoT, 0, 1}};
const int kIcoF[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                          {11, 10, 2}, {10, 7, 6}, {7, 1, 8},  {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
                          {3, 8, 9},  {4, 9, 5},  {2, 4, 11}, {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};

// Closed displaced icosphere, n x n subdivision per face, outward winding.
void gen_sphere(MeshData& m, uint32_t n, double amplitude, uint32_t octaves, uint32_t seed, bool inward)
{
    std::unordered_map<uint64_t, uint32_t> map;
    map.reserve(size_t(10) * n * n + 2);
    auto vertex = [&](int c0, int w0, int c1, int w1, int c2, int w2) -> uint32_t {
        // canonical: corners sorted by index, zero weights dropped
        int cs[3] = {c0, c1, c2}, ws[3] = {w0, w1, w2};
        for (int a = 0; a < 3; ++a)
            for (int b = a + 1; b < 3; ++b)
                if (cs[b] < cs[a]) {
                    std::swap(cs[a], cs[b]);
                    std::swap(ws[a], ws[b]);
                }
        int kc[3], kw[3], nk = 0;
        for (int a = 0; a < 3; ++a)
            if (ws[a]) {
                kc[nk] = cs[a];
                kw[nk] = ws[a];
                ++nk;
            }
        uint64_t key = uint64_t(nk);
        for (int a = 0; a < nk; ++a) key = (key << 4 | uint64_t(kc[a])) << 13 | uint64_t(kw[a]);
        auto it = map.find(key);
        if (it != map.end()) return it->second;
        double p[3] = {0, 0, 0};
        for (int a = 0; a < nk; ++a)
            for (int i = 0; i < 3; ++i) p[i] += kIcoV[kc[a]][i] * (double(kw[a]) / double(n));
        const double len = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        for (int i = 0; i < 3; ++i) p[i] /= len;
        const double r = 1.0 + amplitude * fbm(p[0], p[1], p[2], octaves, seed);
        const uint32_t id = uint32_t(m.pos.size() / 3);
        for (int i = 0; i < 3; ++i) m.pos.push_back(float(p[i] * r));
        map.emplace(key, id);
        return id;
    };
    for (int f = 0; f < 20; ++f) {
        const int A = kIcoF[f][0], B = kIcoF[f][1], C = kIcoF[f][2];
        // grid point (i, j): weights A = n - i - j, B = i, C = j
        auto P = [&](int i, int j) { return vertex(A, int(n) - i - j, B, i, C, j); };
        for (int j = 0; j < int(n); ++j) {
            for (int i = 0; i + j < int(n); ++i) {
                const uint32_t a = P(i, j), b = P(i + 1, j), c = P(i, j + 1);
                if (inward) { m.idx.insert(m.idx.end(), {a, c, b}); } else { m.idx.insert(m.idx.end(), {a, b, c}); }
                if (i + j + 1 < int(n)) {
                    const uint32_t d = P(i + 1, j + 1);
                    if (inward) { m.idx.insert(m.idx.end(), {b, c, d}); } else { m.idx.insert(m.idx.end(), {b, d, c}); }
                }
            }
        }
    }
}

void gen_soup(MeshData& m, uint32_t count, const double lo[3], const double hi[3], uint64_t seed)
{
    Pcg rng(seed, 0x50f0u);
    const double l0 = std::log(0.005), l1 = std::log(0.1);
    for (uint32_t t = 0; t < count; ++t) {
        double c[3];
        for (int i = 0; i < 3; ++i) c[i] = lo[i] + (hi[i] - lo[i]) * rng.uniform();
        const double size = std::exp(l0 + (l1 - l0) * rng.uniform());
        const uint32_t base = uint32_t(m.pos.size() / 3);
        for (int v = 0; v < 3; ++v) {
            const double z = 2.0 * rng.uniform() - 1.0;
            const double phi = 2.0 * M_PI * rng.uniform();
            const double s = std::sqrt(std::max(0.0, 1.0 - z * z));
            const double u[3] = {s * std::cos(phi), s * std::sin(phi), z};
            for (int i = 0; i < 3; ++i) m.pos.push_back(float(c[i] + 0.57735 * size * u[i]));
        }
        m.idx.insert(m.idx.end(), {base, base + 1, base + 2});
    }
}

void add_quad(MeshData& m, const double c[3], const double e0[3], const double e1[3])
{
    const uint32_t b = uint32_t(m.pos.size() / 3);
    const double p[4][3] = {{c[0], c[1], c[2]},
                            {c[0] + e0[0], c[1] + e0[1], c[2] + e0[2]},
                            {c[0] + e0[0] + e1[0], c[1] + e0[1] + e1[1], c[2] + e0[2] + e1[2]},
                            {c[0] + e1[0], c[1] + e1[1], c[2] + e1[2]}};
    for (auto& v : p)
        for (int i = 0; i < 3; ++i) m.pos.push_back(float(v[i]));
    m.idx.insert(m.idx.end(), {b, b + 1, b + 2, b, b + 2, b + 3});
}

// Axis-aligned box rotated by `angle` about z, sitting on z = zmin.
void add_box(MeshData& m, double cx, double cy, double hx, double hy, double zmin, double h, double angle)
{
    const double ca = std::cos(angle), sa = std::sin(angle);
    auto P = [&](double x, double y, double z, double out[3]) {
        out[0] = cx + ca * x - sa * y;
        out[1] = cy + sa * x + ca * y;
        out[2] = z;
    };
    double v[8][3];
    for (int k = 0; k < 8; ++k) P((k & 1) ? hx : -hx, (k & 2) ? hy : -hy, (k & 4) ? zmin + h : zmin, v[k]);
    const uint32_t b = uint32_t(m.pos.size() / 3);
    for (auto& q : v)
        for (int i = 0; i < 3; ++i) m.pos.push_back(float(q[i]));
    const uint32_t f[12][3] = {{0, 2, 1}, {1, 2, 3}, {4, 5, 6}, {5, 7, 6}, {0, 1, 4}, {1, 5, 4},
                               {2, 6, 3}, {3, 6, 7}, {0, 4, 2}, {2, 4, 6}, {1, 3, 5}, {3, 7, 5}};
    for (auto& t : f) m.idx.insert(m.idx.end(), {b + t[0], b + t[1], b + t[2]});
}

// Rotation from a uniform random unit quaternion (Shoemake, Graphics Gems III).
void random_rotation(Pcg& rng, double R[3][3])
{
    const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
    const double a = std::sqrt(1 - u1), b = std::sqrt(u1);
    const double w = a * std::sin(2 * M_PI * u2), x = a * std::cos(2 * M_PI * u2);
    const double y = b * std::sin(2 * M_PI * u3), z = b * std::cos(2 * M_PI * u3);
    R[0][0] = 1 - 2 * (y * y + z * z); R[0][1] = 2 * (x * y - z * w); R[0][2] = 2 * (x * z + y * w);
    R[1][0] = 2 * (x * y + z * w); R[1][1] = 1 - 2 * (x * x + z * z); R[1][2] = 2 * (y * z - x * w);
    R[2][0] = 2 * (x * z - y * w); R[2][1] = 2 * (y * z + x * w); R[2][2] = 1 - 2 * (x * x + y * y);
}

}  // namespace

struct fnd_generated_scene {
    std::vector<MeshData> meshes;
    std::vector<fnd_mesh_desc> mesh_descs;
    std::vector<fnd_instance_desc> instances;
    std::vector<fnd_material> materials;
    std::vector<fnd_quad_light> lights;
    float background[3] = {0, 0, 0};
    fnd_camera camera{};
};

namespace {

fnd_material lambert(float r, float g, float b)
{
    fnd_material m{};
    m.type = FND_MAT_LAMBERT;
    m.albedo[0] = r; m.albedo[1] = g; m.albedo[2] = b;
    return m;
}

void add_light(fnd_generated_scene& s, MeshData& m, const double c[3], const double e0[3], const double e1[3],
               float radiance, uint32_t emissive_material)
{
    fnd_quad_light L{};
    for (int i = 0; i < 3; ++i) {
        L.corner[i] = float(c[i]);
        L.edge0[i] = float(e0[i]);
        L.edge1[i] = float(e1[i]);
        L.radiance[i] = radiance;
    }
    const int32_t li = int32_t(s.lights.size());
    s.lights.push_back(L);
    add_quad(m, c, e0, e1);
    for (int k = 0; k < 2; ++k) {
        m.mat.push_back(emissive_material);
        m.light.push_back(li);
    }
}

void set_camera(fnd_camera& c, float ex, float ey, float ez, float ax, float ay, float az)
{
    c.eye[0] = ex; c.eye[1] = ey; c.eye[2] = ez;
    c.at[0] = ax; c.at[1] = ay; c.at[2] = az;
    c.up[0] = 0; c.up[1] = 0; c.up[2] = 1;
    c.fov_y_deg = 45.0f;
}

// Four lights of `radiance` over a square of half-width `half` at height z,
// each `size` wide, facing down.
void add_four_lights(fnd_generated_scene& s, MeshData& m, double half, double z, double size, float radiance,
                     uint32_t emissive_material, uint32_t count)
{
    const double offs[4][2] = {{-half, -half}, {half, -half}, {-half, half}, {half, half}};
    for (uint32_t k = 0; k < count && k < 4; ++k) {
        const double c[3] = {offs[k][0] - size / 2, offs[k][1] - size / 2, z};
        const double e0[3] = {0, size, 0}, e1[3] = {size, 0, 0};  // cross(e0, e1) = -z: faces down
        add_light(s, m, c, e0, e1, radiance, emissive_material);
    }
}

void gen_cornell(fnd_generated_scene& s, const fnd_gen_params& p)
{
    s.meshes.resize(1);
    MeshData& m = s.meshes[0];
    Pcg rng(p.seed, 0xC0u);
    auto per_tri_material = [&](size_t ntri_before) {
        for (size_t t = ntri_before; t < m.idx.size() / 3; ++t) {
            const float r = float(0.2 + 0.6 * rng.uniform()), g = float(0.2 + 0.6 * rng.uniform()),
                        b = float(0.2 + 0.6 * rng.uniform());
            m.mat.push_back(uint32_t(s.materials.size()));
            m.light.push_back(-1);
            s.materials.push_back(lambert(r, g, b));
        }
    };
    // walls of [-1,1]^3, open toward -y, normals inward
    const double walls[5][3][3] = {
        {{-1, -1, -1}, {2, 0, 0}, {0, 2, 0}},   // floor z = -1 (normal +z)
        {{-1, -1, 1}, {0, 2, 0}, {2, 0, 0}},    // ceiling z = +1 (normal -z)
        {{-1, 1, -1}, {2, 0, 0}, {0, 0, 2}},    // back y = +1 (normal -y)
        {{-1, -1, -1}, {0, 2, 0}, {0, 0, 2}},   // left x = -1 (normal +x)
        {{1, -1, -1}, {0, 0, 2}, {0, 2, 0}},    // right x = +1 (normal -x)
    };
    for (auto& w : walls) add_quad(m, w[0], w[1], w[2]);
    add_box(m, -0.35, 0.35, 0.3, 0.3, -1.0, 1.2, 0.3);    // tall box
    add_box(m, 0.4, -0.3, 0.3, 0.3, -1.0, 0.6, -0.3);     // short box
    per_tri_material(0);
    const double lo[3] = {-0.9, -0.9, -0.9}, hi[3] = {0.9, 0.9, 0.9};
    const size_t before = m.idx.size() / 3;
    gen_soup(m, p.soup_triangles, lo, hi, p.seed + 1);
    per_tri_material(before);
    fnd_material em{};
    em.type = FND_MAT_EMISSIVE;
    em.emission[0] = em.emission[1] = em.emission[2] = 15.0f;
    const uint32_t emi = uint32_t(s.materials.size());
    s.materials.push_back(em);
    const double c[3] = {-0.25, -0.25, 0.999}, e0[3] = {0, 0.5, 0}, e1[3] = {0.5, 0, 0};
    add_light(s, m, c, e0, e1, 15.0f, emi);
    set_camera(s.camera, 0.0f, -3.9f, 0.0f, 0.0f, 0.0f, 0.0f);
}

void gen_material_table(fnd_generated_scene& s, uint64_t seed, uint32_t count)
{
    Pcg rng(seed, 0x3A7u);
    for (uint32_t k = 0; k < count; ++k) {
        fnd_material mt{};
        const bool ggx = (k & 1u) != 0u;
        mt.type = ggx ? FND_MAT_GGX : FND_MAT_LAMBERT;
        for (int i = 0; i < 3; ++i) mt.albedo[i] = float(ggx ? 0.5 + 0.5 * rng.uniform() : 0.2 + 0.6 * rng.uniform());
        mt.roughness = float(0.05 + 0.95 * rng.uniform());
        s.materials.push_back(mt);
    }
}

void gen_displaced_sphere(fnd_generated_scene& s, const fnd_gen_params& p)
{
    s.meshes.resize(1);
    MeshData& m = s.meshes[0];
    gen_sphere(m, p.subdivision, p.amplitude, p.octaves, uint32_t(p.seed), false);
    const double lo[3] = {-1.5, -1.5, -1.5}, hi[3] = {1.5, 1.5, 1.5};
    gen_soup(m, p.soup_triangles, lo, hi, p.seed + 1);
    // 16 materials, even = Lambert, odd = GGX; each triangle draws one
    gen_material_table(s, p.seed, 16);
    Pcg rng(p.seed, 0x77u);
    const size_t ntri = m.idx.size() / 3;
    for (size_t t = 0; t < ntri; ++t) {
        const bool ggx = rng.uniform() * 100.0 < double(p.ggx_fraction_pct);
        const uint32_t k = (rng.next() % 8u) * 2u + (ggx ? 1u : 0u);
        m.mat.push_back(k);
        m.light.push_back(-1);
    }
    fnd_material em{};
    em.type = FND_MAT_EMISSIVE;
    em.emission[0] = em.emission[1] = em.emission[2] = 10.0f;
    const uint32_t emi = uint32_t(s.materials.size());
    s.materials.push_back(em);
    add_four_lights(s, m, 1.0, 2.5, 1.0, 10.0f, emi, p.lights);
    set_camera(s.camera, 2.0f, 2.0f, 2.0f, 0.0f, 0.0f, 0.0f);
}

void gen_instanced(fnd_generated_scene& s, const fnd_gen_params& p)
{
    const uint32_t ninst = p.instances;
    s.meshes.resize(ninst + 1);
    // distinct meshes, one seed each, built in parallel
    {
        const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::atomic<uint32_t> next{0};
        fnd::run_pool(nt, [&]() {
            for (;;) {
                const uint32_t i = next.fetch_add(1);
                if (i >= ninst) break;
                gen_sphere(s.meshes[i], p.subdivision, p.amplitude, p.octaves, uint32_t(p.seed + i), false);
            }
        });
    }
    Pcg rng(p.seed, 0x1257u);
    const double R = 4.5;
    for (uint32_t i = 0; i < ninst; ++i) {
        MeshData& m = s.meshes[i];
        const bool ggx = rng.uniform() * 100.0 < double(p.ggx_fraction_pct);
        fnd_material mt{};
        mt.type = ggx ? FND_MAT_GGX : FND_MAT_LAMBERT;
        for (int c = 0; c < 3; ++c) mt.albedo[c] = float(ggx ? 0.5 + 0.5 * rng.uniform() : 0.2 + 0.6 * rng.uniform());
        mt.roughness = float(0.05 + 0.95 * rng.uniform());
        const uint32_t mi = uint32_t(s.materials.size());
        s.materials.push_back(mt);
        m.mat.assign(m.idx.size() / 3, mi);
        m.light.assign(m.idx.size() / 3, -1);
        double rot[3][3];
        random_rotation(rng, rot);
        const double sc = 0.5 + 1.5 * rng.uniform();
        const double t[3] = {-R + 2 * R * rng.uniform(), -R + 2 * R * rng.uniform(), -R + 2 * R * rng.uniform()};
        fnd_instance_desc in{};
        in.mesh = i;
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                in.object_to_world[4 * r + c] = float(rot[r][c] * sc);
                in.world_to_object[4 * r + c] = float(rot[c][r] / sc);
            }
            in.object_to_world[4 * r + 3] = float(t[r]);
            double acc = 0;
            for (int c = 0; c < 3; ++c) acc += rot[c][r] * t[c];
            in.world_to_object[4 * r + 3] = float(-acc / sc);
        }
        s.instances.push_back(in);
    }
    // light mesh, identity instance
    fnd_material em{};
    em.type = FND_MAT_EMISSIVE;
    em.emission[0] = em.emission[1] = em.emission[2] = 20.0f;
    const uint32_t emi = uint32_t(s.materials.size());
    s.materials.push_back(em);
    MeshData& lm = s.meshes[ninst];
    add_four_lights(s, lm, 4.0, R + 3.0, 3.0, 20.0f, emi, p.lights ? p.lights : 4);
    fnd_instance_desc li{};
    li.mesh = ninst;
    li.object_to_world[0] = li.object_to_world[5] = li.object_to_world[10] = 1.0f;
    li.world_to_object[0] = li.world_to_object[5] = li.world_to_object[10] = 1.0f;
    s.instances.push_back(li);
    set_camera(s.camera, 8.0f, 8.0f, 6.0f, 0.0f, 0.0f, 0.0f);
}

void gen_furnace(fnd_generated_scene& s, const fnd_gen_params& p)
{
    s.meshes.resize(1);
    MeshData& m = s.meshes[0];
    gen_sphere(m, p.subdivision, 0.0, 0, 0, true);
    fnd_material mt = lambert(p.albedo, p.albedo, p.albedo);
    mt.emission[0] = mt.emission[1] = mt.emission[2] = p.emission;
    s.materials.push_back(mt);
    m.mat.assign(m.idx.size() / 3, 0);
    m.light.assign(m.idx.size() / 3, -1);
    set_camera(s.camera, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f);
}

}  // namespace

extern "C" {

fnd_status fnd_gen_default_params(uint32_t kind, fnd_gen_params* out)
{
    if (!out) return FND_E_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    out->kind = kind;
    switch (kind) {
    case FND_GEN_CORNELL_SOUP:
        out->seed = 1;
        out->soup_triangles = 10000;
        out->lights = 1;
        return FND_OK;
    case FND_GEN_DISPLACED_SPHERE:
        out->seed = 1;
        out->soup_triangles = 200000;
        out->subdivision = 256;
        out->octaves = 3;
        out->amplitude = 0.1f;
        out->ggx_fraction_pct = 50;
        out->lights = 4;
        return FND_OK;
    case FND_GEN_INSTANCED_SPHERES:
        out->seed = 0;
        out->instances = 100;
        out->subdivision = 71;
        out->octaves = 3;
        out->amplitude = 0.1f;
        out->ggx_fraction_pct = 50;
        out->lights = 4;
        return FND_OK;
    case FND_GEN_FURNACE:
        out->subdivision = 8;
        out->albedo = 0.5f;
        out->emission = 1.0f;
        return FND_OK;
    default:
        return FND_E_INVALID_ARG;
    }
}

fnd_status fnd_gen_batch_params(uint32_t index, fnd_gen_params* out, uint32_t* image_side)
{
    if (!out || !image_side) return FND_E_INVALID_ARG;
    fnd_status st = fnd_gen_default_params(FND_GEN_DISPLACED_SPHERE, out);
    if (st != FND_OK) return st;
    Pcg rng(0xBA7C4ull, index);
    const double lo = std::log(1.0e4), hi = std::log(2.0e6);
    const double n = std::exp(lo + (hi - lo) * rng.uniform());
    const uint32_t total = uint32_t(std::lround(n));
    out->seed = index;
    out->soup_triangles = total / 8u;
    out->subdivision = std::max(1u, uint32_t(std::lround(std::sqrt(double(total - out->soup_triangles) / 20.0))));
    static const uint32_t sides[3] = {256, 512, 1024};
    *image_side = sides[rng.next() % 3u];
    return FND_OK;
}

fnd_status fnd_gen_scene(const fnd_gen_params* p, fnd_generated_scene** out)
{
    if (!p || !out) return FND_E_INVALID_ARG;
    *out = nullptr;
    if ((p->kind == FND_GEN_DISPLACED_SPHERE || p->kind == FND_GEN_INSTANCED_SPHERES || p->kind == FND_GEN_FURNACE) &&
        (p->subdivision == 0 || p->subdivision > 4096))
        return FND_E_INVALID_ARG;
    if (p->kind == FND_GEN_INSTANCED_SPHERES && (p->instances == 0 || p->instances > 100000)) return FND_E_INVALID_ARG;
    return fnd::guard_host(nullptr, 0, [&]() -> fnd_status {
        std::unique_ptr<fnd_generated_scene> s(new fnd_generated_scene());
        switch (p->kind) {
        case FND_GEN_CORNELL_SOUP: gen_cornell(*s, *p); break;
        case FND_GEN_DISPLACED_SPHERE: gen_displaced_sphere(*s, *p); break;
        case FND_GEN_INSTANCED_SPHERES: gen_instanced(*s, *p); break;
        case FND_GEN_FURNACE: gen_furnace(*s, *p); break;
        default: return FND_E_INVALID_ARG;
        }
        s->mesh_descs.resize(s->meshes.size());
        for (size_t i = 0; i < s->meshes.size(); ++i) {
            MeshData& m = s->meshes[i];
            fnd_mesh_desc& d = s->mesh_descs[i];
            d.positions = m.pos.data();
            d.num_vertices = uint32_t(m.pos.size() / 3);
            d.indices = m.idx.data();
            d.num_triangles = uint32_t(m.idx.size() / 3);
            d.material_ids = m.mat.empty() ? nullptr : m.mat.data();
            d.light_ids = m.light.empty() ? nullptr : m.light.data();
        }
        *out = s.release();
        return FND_OK;
    });
}

fnd_status fnd_gen_scene_get(const fnd_generated_scene* s, fnd_scene_desc* d, fnd_camera* cam)
{
    if (!s) return FND_E_INVALID_ARG;
    if (d) {
        std::memset(d, 0, sizeof(*d));
        d->meshes = s->mesh_descs.data();
        d->num_meshes = uint32_t(s->mesh_descs.size());
        d->instances = s->instances.empty() ? nullptr : s->instances.data();
        d->num_instances = uint32_t(s->instances.size());
        d->materials = s->materials.data();
        d->num_materials = uint32_t(s->materials.size());
        d->lights = s->lights.empty() ? nullptr : s->lights.data();
        d->num_lights = uint32_t(s->lights.size());
        for (int i = 0; i < 3; ++i) d->background[i] = s->background[i];
    }
    if (cam) *cam = s->camera;
    return FND_OK;
}

void fnd_gen_scene_free(fnd_generated_scene* s) { delete s; }

}  // extern "C"
