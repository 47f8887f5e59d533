/*
 * pt_oracle.h — CPU oracle for the path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or as the timed host baseline.  The product (libfnd_pt.so) never links it.
 *
 * The reference (mos9527/Foundation) contains no path tracer: no BVH, ray
 * intersection, RNG, BSDF or compute dispatch (its only GPU code is the raster
 * shader src/Renderer/Triangle.slang:23-37).  This oracle therefore restates
 * the published algorithms the path is specified by, and each function names
 * its source.  The conventions it keeps from the reference are the camera
 * (glm::lookAt, Z up, 45 degree vertical field of view,
 * src/Renderer/Renderer.cpp:373-380) and 32-bit triangle indices
 * (src/Platform/RHI/Vulkan/Command.cpp:289-305).
 *
 * How each part is pinned (what a plausible mistake would fail):
 *   orc_pcg32_*         PCG reference demo vector (O'Neill 2014) and an
 *                       independent Python-int restatement in the tests.
 *   orc_intersect       exact rational orientation tests (Python fractions) on
 *                       random and constructed cases; watertight fan test;
 *                       axis-aligned and degenerate special cases.
 *   orc_scene_build     structural invariants (orc_scene_check).
 *   orc_trace (BVH)     brute force over every triangle (orc_trace with
 *                       brute = 1) on scenes up to 50k triangles.
 *   orc_trace (2-level) brute force over every instance and triangle.
 *   orc_render          closed forms (furnace, with and without Russian
 *                       roulette), unbiasedness of NEE vs BSDF sampling, pdf
 *                       normalisation (see the header of pt_oracle_render.c).
 *   orc_camera_rays     closed forms of the glm::lookAt view: the centre ray
 *                       is normalize(at - eye), the top and bottom edges lie
 *                       fov/2 above and below it, the side edges at
 *                       atan(tan(fov/2) * aspect), row 0 at the top.
 *   orc_sincos_2pi      libm sin/cos within a few ulp, exact at quarter turns.
 */
#ifndef PT_ORACLE_H
#define PT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NO_HIT 0xFFFFFFFFu

uint32_t orc_version(void);

/* PCG32 XSH-RR, O'Neill 2014 (pcg-random.org, pcg32_srandom_r / pcg32_random_r). */
typedef struct orc_pcg32 {
    uint64_t state;
    uint64_t inc;
} orc_pcg32;
void orc_pcg32_srandom(orc_pcg32* rng, uint64_t initstate, uint64_t initseq);
uint32_t orc_pcg32_next(orc_pcg32* rng);
float orc_pcg32_float(orc_pcg32* rng); /* (u >> 8) * 2^-24, exact, in [0, 1) */
void orc_pcg32_seq(uint64_t initstate, uint64_t initseq, uint32_t* out, uint32_t n);

/* Woop, Benthin, Wald: "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013.
 * FP32 only: the double-precision re-evaluation of exactly-zero edge functions
 * is left out; a zero counts as inside and the (t, id) tie rule decides.
 * ray8 = org[3], tmin, dir[3], tmax; tri9 = v0, v1, v2.  Returns 1 on an
 * intersection (t not range-checked), out3 = t, u (weight of v1), v (of v2). */
int orc_intersect(const float* ray8, const float* tri9, float* out3);

typedef struct orc_scene orc_scene;

orc_scene* orc_scene_new(void);
void orc_scene_free(orc_scene* s);
/* Unindexed triangles, 9 floats each.  Returns the mesh index or -1. */
int orc_scene_add_mesh(orc_scene* s, const float* tri_verts, uint32_t num_tris);
/* Row-major 3x4 maps.  Returns the instance index or -1. */
int orc_scene_add_instance(orc_scene* s, uint32_t mesh, const float* object_to_world, const float* world_to_object);
/* Binned SAH (Wald 2007) per mesh, then a top level over instance boxes.
 * Returns 0 on success. */
int orc_scene_build(orc_scene* s, int leaf_max, int bins, int nthreads);
/* Structural invariants of every tree; 0 = valid, else a code naming the
 * first broken invariant. */
int orc_scene_check(const orc_scene* s);
uint64_t orc_scene_num_prims(const orc_scene* s);

/* Closest-hit (anyhit = 0) or any-hit queries.  rays: n * 8 floats (fnd_ray),
 * hits: n * 4 words (t, prim, u, v as fnd_hit).  brute = 1 tests every
 * triangle of every instance instead of walking the trees.  counts, if not
 * NULL, receives per ray (node visits, triangle tests).  nthreads <= 0 uses
 * one thread per online CPU. */
void orc_trace(const orc_scene* s, const float* rays, uint32_t n, int anyhit, int brute, int nthreads,
               float* hits, uint32_t* counts);

/* Shading data for orc_render.  Materials: 8 floats each laid out as
 * fnd_material (type as uint32 bits, albedo[3], roughness, emission[3]);
 * lights: 12 floats each as fnd_quad_light. */
int orc_scene_set_mesh_attrs(orc_scene* s, uint32_t mesh, const uint32_t* mat_ids, const int32_t* light_ids);
int orc_scene_set_materials(orc_scene* s, const float* mats8, uint32_t n);
int orc_scene_set_lights(orc_scene* s, const float* lights12, uint32_t n);
void orc_scene_set_background(orc_scene* s, const float* rgb);

/* Adds spp samples per pixel to accum (W*H*3 running sums, row 0 at the top,
 * samples added in order) and to accum64 if not NULL.  cam10 = eye[3], at[3],
 * up[3], fov_y_deg.  mode 0 = MIS (next-event + BSDF), 1 = BSDF sampling
 * only, 2 = next-event only for light sources.  Returns rays traced. */
uint64_t orc_render(const orc_scene* s, const float* cam10, uint32_t width, uint32_t height, uint32_t spp,
                    uint32_t sample_offset, uint32_t max_depth, uint32_t rr_depth, uint64_t seed, uint32_t mode,
                    int nthreads, float* accum, double* accum64);

/* Camera rays as orc_render starts its paths (out8 per ray: origin, 0,
 * direction, 3e38).  xs/ys/jxy given: n rays through (xs[i] + jxy[2i],
 * ys[i] + jxy[2i+1]); xs NULL: every pixel in order with the jitter of
 * `sample` (PCG32 stream of (seed, sample, pixel, bounce 0)). */
void orc_camera_rays(const float* cam10, uint32_t width, uint32_t height, const uint32_t* xs, const uint32_t* ys,
                     const float* jxy, uint32_t n, uint64_t seed, uint32_t sample, float* out8);
/* The first ndraws outputs of the keyed PCG32 stream of (seed, sample,
 * pixel, bounce) for npix consecutive pixels from pixel0. */
void orc_rng_draws(uint64_t seed, uint32_t sample, uint32_t pixel0, uint32_t npix, uint32_t bounce, uint32_t ndraws,
                   uint32_t* out);
/* sin and cos of 2 pi u (out2 = sin, cos per input), the sampler's own
 * evaluation (exact quarter-turn reduction + Taylor polynomials). */
void orc_sincos_2pi(const float* u, uint32_t n, float* out2);

/* Local-frame BSDF sampling / evaluation (type 0 Lambert, 1 GGX), for the
 * sampling tests.  out7 = wi[3], weight[3], pdf; out4 = f*cos[3], pdf. */
int orc_bsdf_sample(int type, const float* albedo3, float rough, const float* wo3, float u1, float u2, float* out7);
void orc_bsdf_eval(int type, const float* albedo3, float rough, const float* wo3, const float* wi3, float* out4);

/* Host transcription of the GPU walk over the encoded wide-BVH bytes
 * (fnd_scene_encode output: 64-byte nodes, 48-byte leaf triangles, 64-byte
 * instance records).  counts, if not NULL, receives per ray (node visits,
 * triangle tests, instance entries) — the same numbers the kernel counts. */
void orc_trace_wide(const void* nodes, const void* tris, const void* instances, uint32_t root, int two_level,
                    uint32_t stack_entries, const float* rays, uint32_t n, int anyhit, int nthreads, float* hits,
                    uint32_t* counts);

/* Radiance cache in FP64 (see pt_oracle_cache.c).  mlp: 20480 floats in the
 * plain layout W0 [48][64], W1..W4 [64][64], W5 [64][16] (row = input);
 * grid: 16 * 2^19 * 2 floats.  Returns the summed relative-L2 loss. */
void orc_cache_levels(int* res16, uint32_t* dense_mask);
double orc_cache_loss_grad(const float* mlp, const float* grid, const float* queries, const float* targets, uint32_t n,
                           const float* lo3, float inv_extent, int round_fp16, int plain_l2, float* pred, double* dmlp,
                           double* dgrid);
void orc_adam_step(double* p, double* m, double* v, const double* g, uint64_t n, double lr, double b1, double b2,
                   double eps, uint32_t t);

/* Deepest traversal stack any orc_trace_wide walk reached (0xFFFFFFFF if a
 * walk hit its capacity); reset = 1 clears it. */
uint32_t orc_wide_max_stack(int reset);

#ifdef __cplusplus
}
#endif

#endif
