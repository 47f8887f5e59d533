/*
 * pt_oracle.c — CPU oracle: RNG, intersection, BVH and traversal.
 *
 * TEST INFRASTRUCTURE ONLY (see pt_oracle.h).  Built with -ffp-contract=off so
 * that every fused multiply-add is the explicit fmaf written here; the GPU
 * kernels evaluate the same expressions in the same order, which is what makes
 * closest-hit records comparable bit for bit.
 *
 * The reference has no implementation of any function below (its hot-path
 * grep finds nothing; the only GPU code is src/Renderer/Triangle.slang:23-37),
 * so each function cites the published algorithm it restates instead.
 */
#define _GNU_SOURCE
#include "pt_oracle.h"
#include "pt_oracle_internal.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

uint32_t orc_version(void) { return 1u; }

/* ------------------------------------------------------------------ PCG32 */
/* O'Neill, "PCG: A Family of Simple Fast Space-Efficient Statistically Good
 * Algorithms for Random Number Generation", 2014; pcg32_srandom_r and
 * pcg32_random_r of the minimal C implementation (XSH-RR 64/32). */
#define ORC_PCG_MULT 6364136223846793005ULL

uint32_t orc_pcg32_next(orc_pcg32* rng)
{
    uint64_t old = rng->state;
    rng->state = old * ORC_PCG_MULT + rng->inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
}

void orc_pcg32_srandom(orc_pcg32* rng, uint64_t initstate, uint64_t initseq)
{
    rng->state = 0u;
    rng->inc = (initseq << 1u) | 1u;
    orc_pcg32_next(rng);
    rng->state += initstate;
    orc_pcg32_next(rng);
}

float orc_pcg32_float(orc_pcg32* rng) { return (float)(orc_pcg32_next(rng) >> 8) * 0x1p-24f; }

void orc_pcg32_seq(uint64_t initstate, uint64_t initseq, uint32_t* out, uint32_t n)
{
    orc_pcg32 r;
    orc_pcg32_srandom(&r, initstate, initseq);
    for (uint32_t i = 0; i < n; ++i) out[i] = orc_pcg32_next(&r);
}

/* ------------------------------------------------------------------ rays */

#define ORC_SLACK (1.0f + 0x1p-16f) /* relative slack of every box cull */
#define ORC_TINY_DIR 0x1p-80f       /* |d| below this is replaced before 1/d */

typedef struct oray {
    float o[3], d[3];
    float tmin, tmax;
    float invd[3];
    int kx, ky, kz;
    float Sx, Sy, Sz;
} oray;

/* Woop et al. 2013, section 3 (ray-constant precomputation): kz is the axis of
 * largest |d| (first such axis on ties), kx and ky follow cyclically and are
 * swapped when d[kz] < 0 to keep the winding. */
static void oray_setup(oray* r, const float* o, const float* d, float tmin, float tmax)
{
    for (int i = 0; i < 3; ++i) {
        r->o[i] = o[i];
        r->d[i] = d[i];
        float di = d[i];
        if (fabsf(di) < ORC_TINY_DIR) di = copysignf(ORC_TINY_DIR, di);
        r->invd[i] = 1.0f / di;
    }
    r->tmin = tmin;
    r->tmax = tmax;
    int kz = 0;
    float m = fabsf(d[0]);
    if (fabsf(d[1]) > m) { kz = 1; m = fabsf(d[1]); }
    if (fabsf(d[2]) > m) { kz = 2; }
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (d[kz] < 0.0f) { int tmp = kx; kx = ky; ky = tmp; }
    r->kx = kx; r->ky = ky; r->kz = kz;
    /* one division fewer than Woop's d[kx] / d[kz]: Sz is the reciprocal
     * already formed for the slab tests and Sx, Sy are products with it; a
     * per-ray constant either way, so shared edges stay watertight */
    r->Sz = r->invd[kz];
    r->Sx = d[kx] * r->Sz;
    r->Sy = d[ky] * r->Sz;
}

/* Woop et al. 2013, section 3, FP32 only (see pt_oracle.h). */
static int oray_isect(const oray* r, const float* v0, const float* v1, const float* v2, float* t_out,
                      float* u_out, float* v_out)
{
    float A[3], B[3], C[3];
    for (int i = 0; i < 3; ++i) {
        A[i] = v0[i] - r->o[i];
        B[i] = v1[i] - r->o[i];
        C[i] = v2[i] - r->o[i];
    }
    const int kx = r->kx, ky = r->ky, kz = r->kz;
    const float Ax = fmaf(-r->Sx, A[kz], A[kx]);
    const float Ay = fmaf(-r->Sy, A[kz], A[ky]);
    const float Bx = fmaf(-r->Sx, B[kz], B[kx]);
    const float By = fmaf(-r->Sy, B[kz], B[ky]);
    const float Cx = fmaf(-r->Sx, C[kz], C[kx]);
    const float Cy = fmaf(-r->Sy, C[kz], C[ky]);
    /* Plain products, no fma (-ffp-contract=off): the edge function of an
     * edge seen from either of its triangles is the exact negation, which keeps
     * shared edges watertight and gives det == 0 for two equal vertices. */
    const float U = Cx * By - Cy * Bx;
    const float V = Ax * Cy - Ay * Cx;
    const float W = Bx * Ay - By * Ax;
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
    const float det = (U + V) + W;
    if (det == 0.0f) return 0;
    const float Az = r->Sz * A[kz];
    const float Bz = r->Sz * B[kz];
    const float Cz = r->Sz * C[kz];
    const float T = fmaf(U, Az, fmaf(V, Bz, W * Cz));
    *t_out = T / det;
    *u_out = V / det;
    *v_out = W / det;
    return 1;
}

int orc_intersect(const float* ray8, const float* tri9, float* out3)
{
    oray r;
    oray_setup(&r, ray8, ray8 + 4, ray8[3], ray8[7]);
    return oray_isect(&r, tri9, tri9 + 3, tri9 + 6, &out3[0], &out3[1], &out3[2]);
}

/* Slab test (Kay-Kajiya) with a relative slack on the far distance, the
 * conservative form of Ize, "Robust BVH Ray Traversal", JCGT 2(2), 2013.
 * The box survives iff max(near, tmin) <= min(far, tbest) * (1 + 2^-16). */
static int oray_box(const oray* r, const float* lo, const float* hi, float tbest, float* tnear)
{
    float tn = r->tmin, tf = tbest;
    for (int i = 0; i < 3; ++i) {
        const float a = (lo[i] - r->o[i]) * r->invd[i];
        const float b = (hi[i] - r->o[i]) * r->invd[i];
        tn = fmaxf(tn, fminf(a, b));
        tf = fminf(tf, fmaxf(a, b));
    }
    *tnear = tn;
    return tn <= tf * ORC_SLACK;
}

/* Instance transform: o' = M o + m, d' = M d with the fma nesting below
 * (row-major 3x4).  The GPU evaluates exactly these expressions. */
static void xform_ray(const float* m, const float* o, const float* d, float* oo, float* od)
{
    for (int i = 0; i < 3; ++i) {
        const float* row = m + 4 * i;
        oo[i] = fmaf(row[0], o[0], fmaf(row[1], o[1], fmaf(row[2], o[2], row[3])));
        od[i] = fmaf(row[0], d[0], fmaf(row[1], d[1], row[2] * d[2]));
    }
}

/* ------------------------------------------------------------------ BVH */

typedef struct obox { float lo[3], hi[3]; } obox;

static void obox_empty(obox* b)
{
    for (int i = 0; i < 3; ++i) { b->lo[i] = INFINITY; b->hi[i] = -INFINITY; }
}
static void obox_grow(obox* b, const obox* o)
{
    for (int i = 0; i < 3; ++i) {
        b->lo[i] = fminf(b->lo[i], o->lo[i]);
        b->hi[i] = fmaxf(b->hi[i], o->hi[i]);
    }
}
static double obox_area(const obox* b)
{
    double dx = (double)b->hi[0] - b->lo[0], dy = (double)b->hi[1] - b->lo[1], dz = (double)b->hi[2] - b->lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

typedef struct obuild_item { uint32_t node, begin, end, depth; } obuild_item;

/* Top-down binned SAH, Wald, "On fast Construction of SAH-based Bounding
 * Volume Hierarchies", IEEE RT 2007: centroids binned per axis, split cost
 * SA(L)*N(L) + SA(R)*N(R), leaves of at most leaf_max primitives, an index
 * median split when all centroids coincide. */
static int obvh_build(obvh* out, const obox* boxes, uint32_t n, int leaf_max, int bins)
{
    memset(out, 0, sizeof(*out));
    if (n == 0) return 0;
    if (bins < 2) bins = 2;
    if (bins > 64) bins = 64;
    if (leaf_max < 1) leaf_max = 1;
    float* cent = (float*)malloc(sizeof(float) * 3 * (size_t)n);
    out->prim_idx = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    const uint32_t cap = 2 * n;
    out->nodes = (onode*)malloc(sizeof(onode) * (size_t)cap);
    obuild_item* stack = (obuild_item*)malloc(sizeof(obuild_item) * (size_t)(n + 64));
    if (!cent || !out->prim_idx || !out->nodes || !stack) {
        free(cent); free(stack); free(out->prim_idx); free(out->nodes);
        memset(out, 0, sizeof(*out));
        return -1;
    }
    for (uint32_t i = 0; i < n; ++i) {
        out->prim_idx[i] = i;
        for (int a = 0; a < 3; ++a) cent[3 * i + a] = 0.5f * (boxes[i].lo[a] + boxes[i].hi[a]);
    }
    out->num_prims = n;
    out->num_nodes = 1;
    uint32_t sp = 0;
    stack[sp++] = (obuild_item){0, 0, n, 1};
    obox binb[64];
    uint32_t binc[64];
    obox rb[64];
    uint32_t rc[64];
    while (sp) {
        obuild_item it = stack[--sp];
        onode* nd = &out->nodes[it.node];
        obox bb, cb;
        obox_empty(&bb);
        obox_empty(&cb);
        for (uint32_t i = it.begin; i < it.end; ++i) {
            const uint32_t p = out->prim_idx[i];
            obox_grow(&bb, &boxes[p]);
            obox c = {{cent[3 * p], cent[3 * p + 1], cent[3 * p + 2]}, {cent[3 * p], cent[3 * p + 1], cent[3 * p + 2]}};
            obox_grow(&cb, &c);
        }
        memcpy(nd->lo, bb.lo, sizeof bb.lo);
        memcpy(nd->hi, bb.hi, sizeof bb.hi);
        if (it.depth > out->max_depth) out->max_depth = it.depth;
        const uint32_t cnt = it.end - it.begin;
        if (cnt <= (uint32_t)leaf_max) {
            nd->first = it.begin;
            nd->count = cnt;
            continue;
        }
        int best_axis = -1, best_split = -1;
        double best_cost = INFINITY;
        for (int a = 0; a < 3; ++a) {
            const float ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0.0f)) continue;
            const float scale = (float)bins / ext;
            for (int k = 0; k < bins; ++k) { obox_empty(&binb[k]); binc[k] = 0; }
            for (uint32_t i = it.begin; i < it.end; ++i) {
                const uint32_t p = out->prim_idx[i];
                int k = (int)((cent[3 * p + a] - cb.lo[a]) * scale);
                if (k < 0) k = 0;
                if (k >= bins) k = bins - 1;
                obox_grow(&binb[k], &boxes[p]);
                binc[k]++;
            }
            obox acc;
            obox_empty(&acc);
            uint32_t accn = 0;
            for (int k = bins - 1; k > 0; --k) {
                obox_grow(&acc, &binb[k]);
                accn += binc[k];
                rb[k] = acc;
                rc[k] = accn;
            }
            obox_empty(&acc);
            accn = 0;
            for (int k = 0; k < bins - 1; ++k) {
                obox_grow(&acc, &binb[k]);
                accn += binc[k];
                if (accn == 0 || rc[k + 1] == 0) continue;
                const double cost = obox_area(&acc) * accn + obox_area(&rb[k + 1]) * rc[k + 1];
                if (cost < best_cost) { best_cost = cost; best_axis = a; best_split = k; }
            }
        }
        uint32_t mid;
        if (best_axis < 0) {
            mid = it.begin + cnt / 2;
        } else {
            const float ext = cb.hi[best_axis] - cb.lo[best_axis];
            const float scale = (float)bins / ext;
            uint32_t i = it.begin, j = it.end;
            while (i < j) {
                const uint32_t p = out->prim_idx[i];
                int k = (int)((cent[3 * p + best_axis] - cb.lo[best_axis]) * scale);
                if (k < 0) k = 0;
                if (k >= bins) k = bins - 1;
                if (k <= best_split) {
                    ++i;
                } else {
                    --j;
                    const uint32_t tmp = out->prim_idx[i];
                    out->prim_idx[i] = out->prim_idx[j];
                    out->prim_idx[j] = tmp;
                }
            }
            mid = i;
            if (mid == it.begin || mid == it.end) mid = it.begin + cnt / 2;
        }
        const uint32_t left = out->num_nodes;
        out->num_nodes += 2;
        nd->first = left;
        nd->count = 0;
        stack[sp++] = (obuild_item){left + 1, mid, it.end, it.depth + 1};
        stack[sp++] = (obuild_item){left, it.begin, mid, it.depth + 1};
    }
    free(stack);
    free(cent);
    return 0;
}

static void obvh_free(obvh* b)
{
    free(b->nodes);
    free(b->prim_idx);
    memset(b, 0, sizeof(*b));
}

orc_scene* orc_scene_new(void)
{
    orc_scene* s = (orc_scene*)calloc(1, sizeof(orc_scene));
    return s;
}

void orc_scene_free(orc_scene* s)
{
    if (!s) return;
    for (uint32_t i = 0; i < s->num_meshes; ++i) {
        free(s->meshes[i].tris);
        free(s->meshes[i].mat);
        free(s->meshes[i].light);
        obvh_free(&s->meshes[i].bvh);
    }
    free(s->meshes);
    free(s->insts);
    free(s->materials);
    free(s->lights);
    obvh_free(&s->tlas);
    free(s);
}

int orc_scene_add_mesh(orc_scene* s, const float* tri_verts, uint32_t num_tris)
{
    if (!s || (!tri_verts && num_tris)) return -1;
    if (s->num_meshes == s->cap_meshes) {
        uint32_t nc = s->cap_meshes ? 2 * s->cap_meshes : 8;
        omesh* m = (omesh*)realloc(s->meshes, sizeof(omesh) * nc);
        if (!m) return -1;
        s->meshes = m;
        s->cap_meshes = nc;
    }
    omesh* m = &s->meshes[s->num_meshes];
    memset(m, 0, sizeof(*m));
    m->tris = (float*)malloc(sizeof(float) * 9 * (size_t)(num_tris ? num_tris : 1));
    if (!m->tris) return -1;
    if (num_tris) memcpy(m->tris, tri_verts, sizeof(float) * 9 * (size_t)num_tris);
    m->num_tris = num_tris;
    s->built = 0;
    return (int)s->num_meshes++;
}

int orc_scene_add_instance(orc_scene* s, uint32_t mesh, const float* o2w, const float* w2o)
{
    if (!s || mesh >= s->num_meshes || !o2w || !w2o) return -1;
    if (s->num_insts == s->cap_insts) {
        uint32_t nc = s->cap_insts ? 2 * s->cap_insts : 8;
        oinst* p = (oinst*)realloc(s->insts, sizeof(oinst) * nc);
        if (!p) return -1;
        s->insts = p;
        s->cap_insts = nc;
    }
    oinst* in = &s->insts[s->num_insts];
    memset(in, 0, sizeof(*in));
    in->mesh = mesh;
    memcpy(in->o2w, o2w, sizeof(in->o2w));
    memcpy(in->w2o, w2o, sizeof(in->w2o));
    s->built = 0;
    return (int)s->num_insts++;
}

static void tri_box(const float* t, obox* b)
{
    for (int a = 0; a < 3; ++a) {
        b->lo[a] = fminf(fminf(t[a], t[3 + a]), t[6 + a]);
        b->hi[a] = fmaxf(fmaxf(t[a], t[3 + a]), t[6 + a]);
    }
}

typedef struct build_job {
    orc_scene* s;
    atomic_uint next;
    atomic_int err;
    int leaf_max, bins;
} build_job;

static void* build_worker(void* arg)
{
    build_job* j = (build_job*)arg;
    for (;;) {
        uint32_t i = atomic_fetch_add(&j->next, 1u);
        if (i >= j->s->num_meshes) break;
        omesh* m = &j->s->meshes[i];
        obox* boxes = (obox*)malloc(sizeof(obox) * (size_t)(m->num_tris ? m->num_tris : 1));
        if (!boxes) { atomic_store(&j->err, -1); break; }
        for (uint32_t t = 0; t < m->num_tris; ++t) tri_box(m->tris + 9 * (size_t)t, &boxes[t]);
        obvh_free(&m->bvh);
        if (obvh_build(&m->bvh, boxes, m->num_tris, j->leaf_max, j->bins)) atomic_store(&j->err, -1);
        free(boxes);
    }
    return NULL;
}

static int online_cpus(void)
{
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

/* World box of an instance: the 8 corners of the mesh box mapped by
 * object_to_world in double, rounded outward to float, then widened by
 * 2^-12 of the extent so that float round-off of the object-space ray can
 * never leave it. */
static void inst_world_box(const oinst* in, const obvh* b, obox* out)
{
    obox_empty(out);
    if (b->num_nodes == 0) return;
    const onode* r = &b->nodes[0];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int c = 0; c < 8; ++c) {
        const double p[3] = {(c & 1) ? r->hi[0] : r->lo[0], (c & 2) ? r->hi[1] : r->lo[1], (c & 4) ? r->hi[2] : r->lo[2]};
        for (int i = 0; i < 3; ++i) {
            const float* row = in->o2w + 4 * i;
            const double w = (double)row[0] * p[0] + (double)row[1] * p[1] + (double)row[2] * p[2] + (double)row[3];
            if (w < lo[i]) lo[i] = w;
            if (w > hi[i]) hi[i] = w;
        }
    }
    for (int i = 0; i < 3; ++i) {
        const double pad = (hi[i] - lo[i]) * 0x1p-12 + 1e-30;
        out->lo[i] = nextafterf((float)(lo[i] - pad), -INFINITY);
        out->hi[i] = nextafterf((float)(hi[i] + pad), INFINITY);
    }
}

int orc_scene_build(orc_scene* s, int leaf_max, int bins, int nthreads)
{
    if (!s || s->num_meshes == 0) return -1;
    if (nthreads <= 0) nthreads = online_cpus();
    if ((uint32_t)nthreads > s->num_meshes) nthreads = (int)s->num_meshes;
    build_job job;
    job.s = s;
    atomic_init(&job.next, 0u);
    atomic_init(&job.err, 0);
    job.leaf_max = leaf_max;
    job.bins = bins;
    pthread_t th[256];
    if (nthreads > 256) nthreads = 256;
    for (int i = 1; i < nthreads; ++i) pthread_create(&th[i], NULL, build_worker, &job);
    build_worker(&job);
    for (int i = 1; i < nthreads; ++i) pthread_join(th[i], NULL);
    if (atomic_load(&job.err)) return -1;
    obvh_free(&s->tlas);
    if (s->num_insts) {
        obox* boxes = (obox*)malloc(sizeof(obox) * s->num_insts);
        if (!boxes) return -1;
        uint32_t base = 0;
        for (uint32_t i = 0; i < s->num_insts; ++i) {
            oinst* in = &s->insts[i];
            in->id_base = base;
            base += s->meshes[in->mesh].num_tris;
            inst_world_box(in, &s->meshes[in->mesh].bvh, &boxes[i]);
            memcpy(in->lo, boxes[i].lo, sizeof in->lo);
            memcpy(in->hi, boxes[i].hi, sizeof in->hi);
        }
        int rc = obvh_build(&s->tlas, boxes, s->num_insts, 1, bins);
        free(boxes);
        if (rc) return -1;
    }
    s->built = 1;
    return 0;
}

uint64_t orc_scene_num_prims(const orc_scene* s)
{
    if (!s) return 0;
    if (!s->num_insts) return s->num_meshes ? s->meshes[0].num_tris : 0;
    uint64_t n = 0;
    for (uint32_t i = 0; i < s->num_insts; ++i) n += s->meshes[s->insts[i].mesh].num_tris;
    return n;
}

static int box_contains(const float* lo, const float* hi, const obox* b)
{
    for (int a = 0; a < 3; ++a)
        if (b->lo[a] < lo[a] || b->hi[a] > hi[a]) return 0;
    return 1;
}

/* Codes: 1 leaf too large, 2 child box not inside parent, 3 primitive not in
 * exactly one leaf, 4 leaf box misses a primitive, 5 not built, 6 root box is
 * not the union of the primitives. */
static int obvh_check(const obvh* b, const obox* prim_boxes, uint32_t n, int leaf_max)
{
    if (n == 0) return b->num_nodes == 0 ? 0 : 3;
    uint8_t* seen = (uint8_t*)calloc(n, 1);
    if (!seen) return -1;
    int rc = 0;
    for (uint32_t i = 0; i < b->num_nodes && !rc; ++i) {
        const onode* nd = &b->nodes[i];
        if (nd->count) {
            if (nd->count > (uint32_t)leaf_max) rc = 1;
            for (uint32_t k = 0; k < nd->count && !rc; ++k) {
                const uint32_t p = b->prim_idx[nd->first + k];
                if (p >= n || seen[p]) { rc = 3; break; }
                seen[p] = 1;
                if (!box_contains(nd->lo, nd->hi, &prim_boxes[p])) rc = 4;
            }
        } else {
            for (int c = 0; c < 2 && !rc; ++c) {
                const onode* ch = &b->nodes[nd->first + c];
                obox cb;
                memcpy(cb.lo, ch->lo, sizeof cb.lo);
                memcpy(cb.hi, ch->hi, sizeof cb.hi);
                if (!box_contains(nd->lo, nd->hi, &cb)) rc = 2;
            }
        }
    }
    for (uint32_t p = 0; p < n && !rc; ++p)
        if (!seen[p]) rc = 3;
    if (!rc) {
        obox u;
        obox_empty(&u);
        for (uint32_t p = 0; p < n; ++p) obox_grow(&u, &prim_boxes[p]);
        for (int a = 0; a < 3; ++a)
            if (u.lo[a] != b->nodes[0].lo[a] || u.hi[a] != b->nodes[0].hi[a]) rc = 6;
    }
    free(seen);
    return rc;
}

int orc_scene_check(const orc_scene* s)
{
    if (!s || !s->built) return 5;
    for (uint32_t i = 0; i < s->num_meshes; ++i) {
        const omesh* m = &s->meshes[i];
        obox* boxes = (obox*)malloc(sizeof(obox) * (size_t)(m->num_tris ? m->num_tris : 1));
        if (!boxes) return -1;
        for (uint32_t t = 0; t < m->num_tris; ++t) tri_box(m->tris + 9 * (size_t)t, &boxes[t]);
        int rc = obvh_check(&m->bvh, boxes, m->num_tris, 64);
        free(boxes);
        if (rc) return rc;
    }
    if (s->num_insts) {
        obox* boxes = (obox*)malloc(sizeof(obox) * s->num_insts);
        if (!boxes) return -1;
        for (uint32_t i = 0; i < s->num_insts; ++i) {
            memcpy(boxes[i].lo, s->insts[i].lo, sizeof boxes[i].lo);
            memcpy(boxes[i].hi, s->insts[i].hi, sizeof boxes[i].hi);
        }
        int rc = obvh_check(&s->tlas, boxes, s->num_insts, 1);
        free(boxes);
        if (rc) return 10 + rc;
    }
    return 0;
}

/* ------------------------------------------------------------------ traversal */

typedef struct obest {
    float t, u, v;
    uint32_t id;
    int done;
    uint32_t nodes, tests;
} obest;

/* Closest-hit tie rule: the lexicographic minimum of (t, id) over hits with
 * tmin < t < tmax, so the answer does not depend on the visiting order. */
static inline void consider(const oray* r, obest* b, float t, float u, float v, uint32_t id, int anyhit)
{
    if (!(t > r->tmin && t < r->tmax)) return;
    if (anyhit) {
        b->t = t; b->u = u; b->v = v; b->id = id; b->done = 1;
        return;
    }
    if (t < b->t || (t == b->t && id < b->id)) {
        b->t = t; b->u = u; b->v = v; b->id = id;
    }
}

static void trace_mesh(const omesh* m, const oray* r, uint32_t id_base, obest* b, int anyhit, uint32_t* stack_n,
                       float* stack_t)
{
    const obvh* bv = &m->bvh;
    if (!bv->num_nodes) return;
    float tn;
    if (!oray_box(r, bv->nodes[0].lo, bv->nodes[0].hi, b->t, &tn)) return;
    uint32_t sp = 0, node = 0;
    for (;;) {
        const onode* nd = &bv->nodes[node];
        b->nodes++;
        if (nd->count) {
            for (uint32_t k = 0; k < nd->count; ++k) {
                const uint32_t p = bv->prim_idx[nd->first + k];
                const float* tri = m->tris + 9 * (size_t)p;
                float t, u, v;
                b->tests++;
                if (oray_isect(r, tri, tri + 3, tri + 6, &t, &u, &v)) {
                    consider(r, b, t, u, v, id_base + p, anyhit);
                    if (b->done) return;
                }
            }
        } else {
            float tl, tr;
            const int hl = oray_box(r, bv->nodes[nd->first].lo, bv->nodes[nd->first].hi, b->t, &tl);
            const int hr = oray_box(r, bv->nodes[nd->first + 1].lo, bv->nodes[nd->first + 1].hi, b->t, &tr);
            if (hl && hr) {
                uint32_t nearn = nd->first, farn = nd->first + 1;
                float tfar = tr;
                if (tr < tl) { nearn = nd->first + 1; farn = nd->first; tfar = tl; }
                stack_n[sp] = farn;
                stack_t[sp] = tfar;
                ++sp;
                node = nearn;
                continue;
            }
            if (hl) { node = nd->first; continue; }
            if (hr) { node = nd->first + 1; continue; }
        }
        /* pop, dropping entries that lie beyond the best hit so far */
        for (;;) {
            if (sp == 0) return;
            --sp;
            if (stack_t[sp] <= b->t * ORC_SLACK) { node = stack_n[sp]; break; }
        }
    }
}

static void trace_one(const orc_scene* s, const float* ray8, int anyhit, int brute, float* hit4, uint32_t* cnt2,
                      uint32_t* stack_n, float* stack_t, uint32_t* tstack_n, float* tstack_t)
{
    if (orc_ray_rejected(ray8)) {
        orc_miss_record(ray8, hit4);
        if (cnt2) cnt2[0] = cnt2[1] = 0;
        return;
    }
    oray r;
    oray_setup(&r, ray8, ray8 + 4, ray8[3], ray8[7]);
    obest b = {r.tmax, 0.0f, 0.0f, ORC_NO_HIT, 0, 0, 0};
    if (!s->num_insts) {
        const omesh* m = &s->meshes[0];
        if (brute) {
            for (uint32_t p = 0; p < m->num_tris && !b.done; ++p) {
                const float* tri = m->tris + 9 * (size_t)p;
                float t, u, v;
                b.tests++;
                if (oray_isect(&r, tri, tri + 3, tri + 6, &t, &u, &v)) consider(&r, &b, t, u, v, p, anyhit);
            }
        } else {
            trace_mesh(m, &r, 0, &b, anyhit, stack_n, stack_t);
        }
    } else if (brute) {
        for (uint32_t i = 0; i < s->num_insts && !b.done; ++i) {
            const oinst* in = &s->insts[i];
            const omesh* m = &s->meshes[in->mesh];
            float oo[3], od[3];
            xform_ray(in->w2o, r.o, r.d, oo, od);
            oray ro;
            oray_setup(&ro, oo, od, r.tmin, r.tmax);
            for (uint32_t p = 0; p < m->num_tris && !b.done; ++p) {
                const float* tri = m->tris + 9 * (size_t)p;
                float t, u, v;
                b.tests++;
                if (oray_isect(&ro, tri, tri + 3, tri + 6, &t, &u, &v)) consider(&ro, &b, t, u, v, in->id_base + p, anyhit);
            }
        }
    } else {
        const obvh* tl = &s->tlas;
        float tn;
        if (tl->num_nodes && oray_box(&r, tl->nodes[0].lo, tl->nodes[0].hi, b.t, &tn)) {
            uint32_t sp = 0, node = 0;
            for (;;) {
                const onode* nd = &tl->nodes[node];
                b.nodes++;
                int descended = 0;
                if (nd->count) {
                    for (uint32_t k = 0; k < nd->count; ++k) {
                        const oinst* in = &s->insts[tl->prim_idx[nd->first + k]];
                        float oo[3], od[3];
                        xform_ray(in->w2o, r.o, r.d, oo, od);
                        oray ro;
                        oray_setup(&ro, oo, od, r.tmin, r.tmax);
                        trace_mesh(&s->meshes[in->mesh], &ro, in->id_base, &b, anyhit, stack_n, stack_t);
                        if (b.done) break;
                    }
                    if (b.done) break;
                } else {
                    float ta, tb;
                    const int ha = oray_box(&r, tl->nodes[nd->first].lo, tl->nodes[nd->first].hi, b.t, &ta);
                    const int hb = oray_box(&r, tl->nodes[nd->first + 1].lo, tl->nodes[nd->first + 1].hi, b.t, &tb);
                    if (ha && hb) {
                        uint32_t nearn = nd->first, farn = nd->first + 1;
                        float tfar = tb;
                        if (tb < ta) { nearn = nd->first + 1; farn = nd->first; tfar = ta; }
                        tstack_n[sp] = farn;
                        tstack_t[sp] = tfar;
                        ++sp;
                        node = nearn;
                        descended = 1;
                    } else if (ha) { node = nd->first; descended = 1; }
                    else if (hb) { node = nd->first + 1; descended = 1; }
                }
                if (descended) continue;
                int found = 0;
                while (sp) {
                    --sp;
                    if (tstack_t[sp] <= b.t * ORC_SLACK) { node = tstack_n[sp]; found = 1; break; }
                }
                if (!found) break;
            }
        }
    }
    if (b.id == ORC_NO_HIT) {
        b.t = r.tmax;
        b.u = b.v = 0.0f;
    }
    hit4[0] = b.t;
    memcpy(&hit4[1], &b.id, 4);
    hit4[2] = b.u;
    hit4[3] = b.v;
    if (cnt2) { cnt2[0] = b.nodes; cnt2[1] = b.tests; }
}

typedef struct trace_job {
    const orc_scene* s;
    const float* rays;
    uint32_t n;
    int anyhit, brute;
    float* hits;
    uint32_t* counts;
    atomic_uint next;
    uint32_t stack_cap, tstack_cap;
} trace_job;

static void* trace_worker(void* arg)
{
    trace_job* j = (trace_job*)arg;
    uint32_t* sn = (uint32_t*)malloc(sizeof(uint32_t) * j->stack_cap);
    float* st = (float*)malloc(sizeof(float) * j->stack_cap);
    uint32_t* tn = (uint32_t*)malloc(sizeof(uint32_t) * j->tstack_cap);
    float* tt = (float*)malloc(sizeof(float) * j->tstack_cap);
    if (sn && st && tn && tt) {
        for (;;) {
            const uint32_t b = atomic_fetch_add(&j->next, 256u);
            if (b >= j->n) break;
            const uint32_t e = b + 256u < j->n ? b + 256u : j->n;
            for (uint32_t i = b; i < e; ++i)
                trace_one(j->s, j->rays + 8 * (size_t)i, j->anyhit, j->brute, j->hits + 4 * (size_t)i,
                          j->counts ? j->counts + 2 * (size_t)i : NULL, sn, st, tn, tt);
        }
    }
    free(sn); free(st); free(tn); free(tt);
    return NULL;
}

void orc_trace(const orc_scene* s, const float* rays, uint32_t n, int anyhit, int brute, int nthreads, float* hits,
               uint32_t* counts)
{
    if (!s || !s->built || !rays || !hits) return;
    trace_job j;
    j.s = s; j.rays = rays; j.n = n; j.anyhit = anyhit; j.brute = brute; j.hits = hits; j.counts = counts;
    atomic_init(&j.next, 0u);
    uint32_t md = 1;
    for (uint32_t i = 0; i < s->num_meshes; ++i)
        if (s->meshes[i].bvh.max_depth > md) md = s->meshes[i].bvh.max_depth;
    j.stack_cap = md + 4;
    j.tstack_cap = s->tlas.max_depth + 4;
    if (nthreads <= 0) nthreads = online_cpus();
    if (nthreads > 256) nthreads = 256;
    const uint32_t chunks = (n + 255u) / 256u;
    if ((uint32_t)nthreads > chunks) nthreads = chunks ? (int)chunks : 1;
    pthread_t th[256];
    for (int i = 1; i < nthreads; ++i) pthread_create(&th[i], NULL, trace_worker, &j);
    trace_worker(&j);
    for (int i = 1; i < nthreads; ++i) pthread_join(th[i], NULL);
}

/* ------------------------------------------------------------------ shading data */

int orc_scene_set_mesh_attrs(orc_scene* s, uint32_t mesh, const uint32_t* mat_ids, const int32_t* light_ids)
{
    if (!s || mesh >= s->num_meshes) return -1;
    omesh* m = &s->meshes[mesh];
    free(m->mat);
    free(m->light);
    m->mat = NULL;
    m->light = NULL;
    if (mat_ids) {
        m->mat = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(m->num_tris ? m->num_tris : 1));
        if (!m->mat) return -1;
        memcpy(m->mat, mat_ids, sizeof(uint32_t) * (size_t)m->num_tris);
    }
    if (light_ids) {
        m->light = (int32_t*)malloc(sizeof(int32_t) * (size_t)(m->num_tris ? m->num_tris : 1));
        if (!m->light) return -1;
        memcpy(m->light, light_ids, sizeof(int32_t) * (size_t)m->num_tris);
    }
    return 0;
}

/* 8 floats per material as fnd_material: type (as uint32 bits), albedo[3],
 * roughness, emission[3].  Roughness is clamped to [0.02, 1]. */
int orc_scene_set_materials(orc_scene* s, const float* mats8, uint32_t n)
{
    if (!s || (!mats8 && n)) return -1;
    free(s->materials);
    s->materials = (orc_material*)calloc(n ? n : 1, sizeof(orc_material));
    if (!s->materials) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        orc_material* m = &s->materials[i];
        memcpy(&m->type, &mats8[8 * (size_t)i], 4);
        for (int c = 0; c < 3; ++c) {
            m->albedo[c] = mats8[8 * (size_t)i + 1 + c];
            m->emission[c] = mats8[8 * (size_t)i + 5 + c];
        }
        m->roughness = fminf(1.0f, fmaxf(0.02f, mats8[8 * (size_t)i + 4]));
    }
    s->num_materials = n;
    return 0;
}

/* 12 floats per light as fnd_quad_light: corner, edge0, edge1, radiance.
 * Lights are chosen in proportion to mean radiance times area. */
int orc_scene_set_lights(orc_scene* s, const float* l12, uint32_t n)
{
    if (!s || (!l12 && n)) return -1;
    free(s->lights);
    s->lights = (orc_light*)calloc(n ? n : 1, sizeof(orc_light));
    if (!s->lights) return -1;
    double total = 0.0;
    double* pw = (double*)malloc(sizeof(double) * (n ? n : 1));
    if (!pw) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        orc_light* L = &s->lights[i];
        const float* q = l12 + 12 * (size_t)i;
        memcpy(L->corner, q, 12);
        memcpy(L->e0, q + 3, 12);
        memcpy(L->e1, q + 6, 12);
        memcpy(L->radiance, q + 9, 12);
        const double cx = (double)L->e0[1] * L->e1[2] - (double)L->e0[2] * L->e1[1];
        const double cy = (double)L->e0[2] * L->e1[0] - (double)L->e0[0] * L->e1[2];
        const double cz = (double)L->e0[0] * L->e1[1] - (double)L->e0[1] * L->e1[0];
        const double a = sqrt(cx * cx + cy * cy + cz * cz);
        if (!(a > 0.0)) { free(pw); return -1; }
        L->area = (float)a;
        L->n[0] = (float)(cx / a); L->n[1] = (float)(cy / a); L->n[2] = (float)(cz / a);
        pw[i] = (((double)L->radiance[0] + L->radiance[1] + L->radiance[2]) / 3.0) * a;
        if (pw[i] < 1e-12) pw[i] = 1e-12;
        total += pw[i];
    }
    double acc = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        acc += pw[i];
        s->lights[i].cdf = (i + 1 == n) ? 1.0f : (float)(acc / total);
        s->lights[i].pdf_select = (float)(pw[i] / total);
    }
    free(pw);
    s->num_lights = n;
    return 0;
}

void orc_scene_set_background(orc_scene* s, const float* rgb)
{
    if (s && rgb) memcpy(s->background, rgb, 12);
}

/* One ray through the trees with stacks on the C stack (depth-checked). */
static int trace_core(const orc_scene* s, const float* o, const float* d, float tmax, int anyhit, obest* b)
{
    uint32_t md = 1;
    for (uint32_t i = 0; i < s->num_meshes; ++i)
        if (s->meshes[i].bvh.max_depth > md) md = s->meshes[i].bvh.max_depth;
    if (md + 4 > 512 || s->tlas.max_depth + 4 > 512) return -1;
    uint32_t sn[512], tn[512];
    float st[512], tt[512];
    float ray8[8] = {o[0], o[1], o[2], 0.0f, d[0], d[1], d[2], tmax};
    float hit4[4];
    (void)b;
    trace_one(s, ray8, anyhit, 0, hit4, NULL, sn, st, tn, tt);
    uint32_t id;
    memcpy(&id, &hit4[1], 4);
    b->t = hit4[0];
    b->id = id;
    return id != ORC_NO_HIT;
}

int orc_closest_shade(const orc_scene* s, v3f_in o, v3f_in d, orc_shade_hit* out)
{
    obest b;
    const float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
    if (trace_core(s, of, df, 3.0e38f, 0, &b) != 1) return 0;
    uint32_t mesh = 0, local = b.id;
    const float* w2o = NULL;
    if (s->num_insts) {
        uint32_t lo = 0, hi = s->num_insts;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (s->insts[mid].id_base <= b.id) lo = mid; else hi = mid;
        }
        mesh = s->insts[lo].mesh;
        local = b.id - s->insts[lo].id_base;
        w2o = s->insts[lo].w2o;
    }
    const omesh* m = &s->meshes[mesh];
    const float* v = m->tris + 9 * (size_t)local;
    const float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
    const float e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
    float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    if (w2o) {
        const float nn[3] = {w2o[0] * n[0] + w2o[4] * n[1] + w2o[8] * n[2],
                             w2o[1] * n[0] + w2o[5] * n[1] + w2o[9] * n[2],
                             w2o[2] * n[0] + w2o[6] * n[1] + w2o[10] * n[2]};
        memcpy(n, nn, sizeof n);
    }
    out->t = b.t;
    memcpy(out->n, n, sizeof n);
    out->material = m->mat ? m->mat[local] : 0u;
    out->light = m->light ? m->light[local] : -1;
    return 1;
}

int orc_occluded(const orc_scene* s, v3f_in o, v3f_in d, float tmax)
{
    obest b;
    const float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
    return trace_core(s, of, df, tmax, 1, &b) == 1;
}
