// pt_layout.hpp — memory layout of a scene in HBM, shared by the host encoder
// (bvh_build.cpp) and the traversal kernels (trace.hip).
//
// The reference has no acceleration structure (its only geometry is a
// 4-vertex quad uploaded as a vertex buffer, src/Renderer/Renderer.cpp:133-198),
// so this layout is designed for CDNA4 from scratch:
//
//  * WideNode: a 4-wide BVH node in exactly 64 bytes, one aligned 64-byte fetch
//    (four 16-byte loads by one lane).  Child boxes are quantised to 8 bits per
//    plane against a per-node origin and power-of-two scale (the idea of Ylitie,
//    Karras, Laine, "Efficient Incoherent Ray Traversal on GPUs Through
//    Compressed Wide BVHs", HPG 2017), rounded outward so a dequantised box
//    always contains the exact child box.
//  * Each slot holds its finished traversal stack entry (a node index, or a
//    tagged triangle range or instance), so the node step pushes what it
//    loads and derives nothing per child.  Internal children of a node are
//    still contiguous (child_base + rank among the internal slots) and the
//    triangles of its leaf children too (prim_base + offset).
//  * LeafTri: 48 bytes, three float4 whose .w of the first carries the
//    triangle's index within its mesh — a hit never needs a second lookup.
//  * InstanceRec: 64 bytes per instance in top-level leaf order: the
//    world-to-object map the rays are moved with, the root of its mesh's tree
//    and the first primitive id of the instance.
#pragma once
#include <cstdint>

namespace fnd {

constexpr uint32_t kWideArity = 4;
// Triangles per leaf child: one.  Against leaves of up to 4, the traversal
// tests 2.5x (10M instanced spheres) to 6x (displaced sphere + soup) fewer
// triangles for 5-15% more node visits and runs 8-31% faster; the tree has
// 3x the nodes (340 MB at 10M triangles).
constexpr uint32_t kLeafMax = 1;
constexpr uint32_t kMaxTriangles = 1u << 28;
constexpr uint32_t kMaxNodes = 1u << 30;

// Traversal stack entry tags (32-bit words).
constexpr uint32_t kEntryTri = 0x80000000u;      // | first << 3 | count
constexpr uint32_t kEntryInstance = 0x40000000u; // | index << 3 | 1 (index < 2^27 - 1)
constexpr uint32_t kEntryReturn = 0x7FFFFFFFu;   // back to the top level (low bits 7: no instance entry)
constexpr uint32_t kMaxInstances = (1u << 27) - 1u;

// The traversal's work counter is split into kCounterShards words
// kCounterStride words apart (their own 256-byte lines): one address
// serialises device-scope atomics (measured on MI355X: 88 M/s on one address,
// 670 M/s over 8, 3.3 G/s over 64), and a refill's reservation must not queue
// behind other waves'.  Shard s hands out the blocks of 64 rays s,
// s + kCounterShards, ... past the launch's static first pools, so the shards
// advance through the rays together.
constexpr uint32_t kCounterShards = 16;
constexpr uint32_t kCounterStride = 64;
constexpr uint32_t kCounterWords = kCounterShards * kCounterStride;
// Anything below kEntryInstance is a node index.

struct alignas(64) WideNode {
    float px, py, pz;        // quantisation origin
    int8_t ex, ey, ez;       // child box scale is 2^e per axis
    uint8_t imask;           // bit k: slot k is an internal node; bit 4 + k: slot k holds a child
    uint32_t entry[4];       // slot k's stack entry: child node index, kEntryTri | first << 3 | count,
                             // or kEntryInstance | index << 3 | 1; 0 for an empty slot
    uint32_t qlo[3];         // byte k of qlo[a]: slot k lower plane on axis a
    uint32_t qhi[3];         // byte k of qhi[a]: slot k upper plane on axis a
    uint32_t child_base;     // node index of the first internal child (not read by the traversal)
    uint32_t prim_base;      // first triangle (or instance) of the leaf children (not read by the traversal)
};
static_assert(sizeof(WideNode) == 64, "WideNode must be one 64-byte fetch");

struct alignas(16) LeafTri {
    float v0[3];
    uint32_t local_id;       // triangle index within its mesh
    float v1[3];
    uint32_t material;       // material index
    float v2[3];
    int32_t light;           // quad light index or -1
};
static_assert(sizeof(LeafTri) == 48, "LeafTri is three float4");

struct alignas(64) InstanceRec {
    float w2o[12];           // row-major 3x4 world -> object
    uint32_t blas_root;      // node index of the mesh's root
    uint32_t id_base;        // primitive id of the instance's triangle 0
    uint32_t mesh;
    uint32_t orig_index;     // index in the caller's instance array
};
static_assert(sizeof(InstanceRec) == 64, "InstanceRec is 64 bytes");

struct alignas(16) Ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
};

struct alignas(16) Hit {
    float t;
    uint32_t prim;
    float u, v;
};

}  // namespace fnd
