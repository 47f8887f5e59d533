// mesh_ingest.cpp — scene ingestion from the reference's buffer conventions
// and from Wavefront OBJ files (host only, no device).
//
// The reference describes geometry the Vulkan way: an interleaved vertex
// buffer whose position attribute has a stride, an offset and a float format
// (its only mesh uses {vec2 pos, vec3 color, vec2 texCoord},
// src/Renderer/Renderer.cpp:23-27), and an index buffer of R16_UINT or
// R32_UINT (src/Platform/RHI/Vulkan/Command.cpp:289-305).
// fnd_mesh_from_streams turns such a pair into the float3 positions and u32
// indices of fnd_mesh_desc; fnd_obj_load reads positions and faces of an OBJ
// file (polygons split into fans) so that assets can replace the generators.
#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/fnd_pt.h"
#include "host_pool.hpp"

namespace {

fnd_status fail(char* err, size_t err_len, fnd_status st, const char* fmt, ...)
{
    if (err && err_len) {
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(err, err_len, fmt, ap);
        va_end(ap);
    }
    return st;
}

}  // namespace

extern "C" fnd_status fnd_mesh_from_streams(const fnd_vertex_stream* vs, const fnd_index_stream* is,
                                            float* positions, uint32_t* indices, char* err, size_t err_len)
{
    if (err && err_len) err[0] = '\0';
    if (!vs || !is || (!positions && vs->num_vertices) || (!indices && is->num_indices))
        return fail(err, err_len, FND_E_INVALID_ARG, "null stream or output");
    if (vs->position_components != 2 && vs->position_components != 3)
        return fail(err, err_len, FND_E_INVALID_ARG, "position_components must be 2 or 3");
    const uint64_t attr = uint64_t(vs->position_components) * sizeof(float);
    if (vs->num_vertices) {
        if (!vs->data) return fail(err, err_len, FND_E_INVALID_ARG, "vertex data is null");
        if (vs->stride < attr || vs->position_offset > vs->stride - attr)
            return fail(err, err_len, FND_E_INVALID_ARG, "position attribute (offset %u, %llu bytes) exceeds stride %u",
                        vs->position_offset, (unsigned long long)attr, vs->stride);
        const uint64_t need = uint64_t(vs->num_vertices - 1) * vs->stride + vs->position_offset + attr;
        if (need > vs->bytes)
            return fail(err, err_len, FND_E_INVALID_ARG, "vertex buffer of %llu bytes holds fewer than %u vertices",
                        (unsigned long long)vs->bytes, vs->num_vertices);
    }
    if (is->index_bits != 16 && is->index_bits != 32)
        return fail(err, err_len, FND_E_INVALID_ARG, "index_bits must be 16 (R16_UINT) or 32 (R32_UINT)");
    if (is->num_indices % 3) return fail(err, err_len, FND_E_INVALID_ARG, "index count %u is not a multiple of 3",
                                         is->num_indices);
    if (is->num_indices && !is->data) return fail(err, err_len, FND_E_INVALID_ARG, "index data is null");
    const unsigned char* base = static_cast<const unsigned char*>(vs->data);
    for (uint32_t v = 0; v < vs->num_vertices; ++v) {
        float p[3] = {0.0f, 0.0f, 0.0f};
        std::memcpy(p, base + uint64_t(v) * vs->stride + vs->position_offset, attr);  // unaligned-safe
        positions[3ull * v] = p[0];
        positions[3ull * v + 1] = p[1];
        positions[3ull * v + 2] = p[2];
    }
    for (uint32_t i = 0; i < is->num_indices; ++i) {
        uint32_t x;
        if (is->index_bits == 16) {
            uint16_t h;
            std::memcpy(&h, static_cast<const unsigned char*>(is->data) + 2ull * i, 2);
            x = h;
        } else {
            std::memcpy(&x, static_cast<const unsigned char*>(is->data) + 4ull * i, 4);
        }
        if (x >= vs->num_vertices)
            return fail(err, err_len, FND_E_INVALID_ARG, "index %u at position %u is not below %u vertices", x, i,
                        vs->num_vertices);
        indices[i] = x;
    }
    return FND_OK;
}

struct fnd_obj_mesh {
    std::vector<float> positions;
    std::vector<uint32_t> indices;
};

namespace {

// Parses a 1-based (or negative, relative) OBJ vertex reference at *s:
// "7", "7/2", "7//3", "-1/...".  Returns false on a malformed token.
bool parse_ref(const char*& s, long long nverts, uint32_t& out)
{
    char* end = nullptr;
    errno = 0;
    const long long v = std::strtoll(s, &end, 10);
    if (end == s || errno) return false;
    s = end;
    while (*s && *s != ' ' && *s != '\t' && *s != '\r' && *s != '\n') ++s;  // skip /vt/vn
    const long long idx = v > 0 ? v - 1 : nverts + v;
    if (v == 0 || idx < 0 || idx >= nverts || idx > 0xFFFFFFFELL) return false;
    out = uint32_t(idx);
    return true;
}

}  // namespace

extern "C" fnd_status fnd_obj_load(const char* path, fnd_obj_mesh** out, char* err, size_t err_len)
{
    if (err && err_len) err[0] = '\0';
    if (!path || !out) return fail(err, err_len, FND_E_INVALID_ARG, "null path or output");
    *out = nullptr;
    std::unique_ptr<std::FILE, int (*)(std::FILE*)> file(std::fopen(path, "rb"), &std::fclose);
    if (!file) return fail(err, err_len, FND_E_NOT_FOUND, "cannot open %s", path);
    std::FILE* fp = file.get();
    return fnd::guard_host(err, err_len, [&]() -> fnd_status {
        std::unique_ptr<fnd_obj_mesh> m(new fnd_obj_mesh());
        std::vector<char> line(1 << 16);
        uint64_t lineno = 0;
        std::vector<uint32_t> face;
        fnd_status st = FND_OK;
        while (std::fgets(line.data(), int(line.size()), fp)) {
            ++lineno;
            if (!std::strchr(line.data(), '\n') && !std::feof(fp)) {
                st = fail(err, err_len, FND_E_INVALID_ARG, "%s:%llu: line longer than %zu bytes", path,
                          (unsigned long long)lineno, line.size() - 1);
                break;
            }
            const char* s = line.data();
            while (*s == ' ' || *s == '\t') ++s;
            if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
                float p[3];
                if (std::sscanf(s + 2, "%f %f %f", &p[0], &p[1], &p[2]) != 3) {
                    st = fail(err, err_len, FND_E_INVALID_ARG, "%s:%llu: vertex needs three coordinates", path,
                              (unsigned long long)lineno);
                    break;
                }
                m->positions.insert(m->positions.end(), p, p + 3);
            } else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
                s += 2;
                face.clear();
                const long long nverts = (long long)(m->positions.size() / 3);
                bool ok = true;
                for (;;) {
                    while (*s == ' ' || *s == '\t') ++s;
                    if (!*s || *s == '\r' || *s == '\n' || *s == '#') break;
                    uint32_t v;
                    if (!parse_ref(s, nverts, v)) {
                        ok = false;
                        break;
                    }
                    face.push_back(v);
                }
                if (!ok || face.size() < 3) {
                    st = fail(err, err_len, FND_E_INVALID_ARG, "%s:%llu: bad face", path, (unsigned long long)lineno);
                    break;
                }
                for (size_t k = 1; k + 1 < face.size(); ++k) {  // fan
                    m->indices.push_back(face[0]);
                    m->indices.push_back(face[k]);
                    m->indices.push_back(face[k + 1]);
                }
            }
            // other statements (vt, vn, o, g, usemtl, s, comments) carry nothing this path uses
        }
        if (st == FND_OK && m->positions.size() / 3 > 0xFFFFFFFFull)
            st = fail(err, err_len, FND_E_INVALID_ARG, "%s: too many vertices", path);
        if (st != FND_OK) return st;
        *out = m.release();
        return FND_OK;
    });
}

extern "C" fnd_status fnd_obj_mesh_get(const fnd_obj_mesh* m, fnd_mesh_desc* out)
{
    if (!m || !out) return FND_E_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    out->positions = m->positions.data();
    out->num_vertices = uint32_t(m->positions.size() / 3);
    out->indices = m->indices.data();
    out->num_triangles = uint32_t(m->indices.size() / 3);
    return FND_OK;
}

extern "C" void fnd_obj_free(fnd_obj_mesh* m) { delete m; }
