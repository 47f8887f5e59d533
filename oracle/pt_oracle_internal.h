/*
 * pt_oracle_internal.h — shared state of the oracle's translation units.
 * TEST INFRASTRUCTURE ONLY (see pt_oracle.h).
 */
#ifndef PT_ORACLE_INTERNAL_H
#define PT_ORACLE_INTERNAL_H

#include <stdint.h>

#define ORC_MAT_LAMBERT 0u
#define ORC_MAT_GGX 1u
#define ORC_MAT_EMISSIVE 2u

typedef struct onode {
    float lo[3], hi[3];
    uint32_t first; /* internal: left child (right = first + 1); leaf: first entry in prim_idx */
    uint32_t count; /* 0 = internal */
} onode;

typedef struct obvh {
    onode* nodes;
    uint32_t num_nodes;
    uint32_t* prim_idx;
    uint32_t num_prims;
    uint32_t max_depth;
} obvh;

typedef struct omesh {
    float* tris; /* 9 floats per triangle */
    uint32_t num_tris;
    obvh bvh;
    uint32_t* mat;   /* per triangle, NULL = material 0 */
    int32_t* light;  /* per triangle, NULL = none */
} omesh;

typedef struct oinst {
    uint32_t mesh;
    float o2w[12];
    float w2o[12];
    uint32_t id_base;
    float lo[3], hi[3];
} oinst;

typedef struct orc_material {
    uint32_t type;
    float albedo[3];
    float roughness;
    float emission[3];
} orc_material;

typedef struct orc_light {
    float corner[3], e0[3], e1[3], radiance[3];
    float n[3];
    float area, cdf, pdf_select;
} orc_light;

struct orc_scene {
    omesh* meshes;
    uint32_t num_meshes, cap_meshes;
    oinst* insts;
    uint32_t num_insts, cap_insts;
    obvh tlas;
    int built;
    orc_material* materials;
    uint32_t num_materials;
    orc_light* lights;
    uint32_t num_lights;
    float background[3];
};


typedef struct orc_shade_hit {
    float t;
    float n[3];      /* world-space geometric normal, not normalised */
    uint32_t material;
    int32_t light;
} orc_shade_hit;

typedef struct v3f_in { float x, y, z; } v3f_in;

/* A ray the traversal refuses: a non-finite origin or direction component,
 * or a NaN range bound.  It is answered as a miss (t = tmax as given, no
 * primitive, u = v = 0) without a walk, as the device does (the device's one
 * root-node test with an empty range is counted by the wide walk). */
static inline int orc_ray_rejected(const float* ray8)
{
    for (int k = 0; k < 8; ++k) {
        const float x = ray8[k];
        if (k == 3 || k == 7) {
            if (x != x) return 1;
        } else if (!(x - x == 0.0f)) {
            return 1;
        }
    }
    return 0;
}

static inline void orc_miss_record(const float* ray8, float* hit4)
{
    const uint32_t none = 0xFFFFFFFFu;
    hit4[0] = ray8[7];
    __builtin_memcpy(&hit4[1], &none, 4);
    hit4[2] = 0.0f;
    hit4[3] = 0.0f;
}

/* Closest hit along o + t d, 0 < t < 3e38 (same walk as orc_trace). */
int orc_closest_shade(const orc_scene* s, v3f_in o, v3f_in d, orc_shade_hit* out);
/* Any hit with 0 < t < tmax. */
int orc_occluded(const orc_scene* s, v3f_in o, v3f_in d, float tmax);

#endif
