/*
 * pt_oracle_wide.c — host transcription of the GPU traversal over the
 * encoded 64-byte wide-BVH layout.
 *
 * TEST INFRASTRUCTURE ONLY (see pt_oracle.h).  The input is the byte layout
 * that fnd_scene_encode produces (the same bytes fnd_pt_scene_create
 * uploads); the layout is restated below field by field.  The walk follows
 * the kernel step for step — the same quantised slab expressions, the same
 * sorting network for the visiting order, the same push/pop and cull rule —
 * so per-ray counts of node visits and triangle tests equal the kernel's, and
 * they are the "visited nodes" and "tested triangles" of the algorithmic byte
 * count (32 B ray + 16 B hit + 64 B per node + 48 B per triangle).
 *
 * Closest-hit records are independent of the visiting order (tie rule), so
 * this walk and the binary-BVH walk of pt_oracle.c must agree bit for bit;
 * the tests check that, which pins this transcription's decoding.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "pt_oracle.h"
#include "pt_oracle_internal.h"

#define W_SLACK (1.0f + 0x1p-16f)
#define W_TINY 0x1p-80f
#define W_ENTRY_TRI 0x80000000u
#define W_ENTRY_INST 0x40000000u
#define W_ENTRY_RETURN 0x7FFFFFFFu

typedef struct wnode {          /* 64 bytes */
    float p[3];
    int8_t e[3];
    uint8_t imask;
    uint32_t entry[4];          /* slot k's stack entry, tagged; 0 when empty */
    uint32_t qlo[3], qhi[3];
    uint32_t child_base, prim_base;
} wnode;

typedef struct wtri {           /* 48 bytes */
    float v0[3];
    uint32_t local_id;
    float v1[3];
    uint32_t material;
    float v2[3];
    int32_t light;
} wtri;

typedef struct winst {          /* 64 bytes */
    float w2o[12];
    uint32_t blas_root, id_base, mesh, orig_index;
} winst;

typedef struct wray {
    float o[3], d[3], invd[3];
    int kx, ky, kz;
    float Sx, Sy, Sz;
} wray;

static void wray_setup(wray* r, const float* o, const float* d)
{
    for (int i = 0; i < 3; ++i) {
        r->o[i] = o[i];
        r->d[i] = d[i];
        float di = d[i];
        if (fabsf(di) < W_TINY) di = copysignf(W_TINY, di);
        r->invd[i] = 1.0f / di;
    }
    int kz = 0;
    float m = fabsf(d[0]);
    if (fabsf(d[1]) > m) { kz = 1; m = fabsf(d[1]); }
    if (fabsf(d[2]) > m) kz = 2;
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    if (d[kz] < 0.0f) { int t = kx; kx = ky; ky = t; }
    r->kx = kx; r->ky = ky; r->kz = kz;
    /* one division fewer than Woop's d[kx] / d[kz]: Sz is the reciprocal
     * already formed for the slab tests and Sx, Sy are products with it; a
     * per-ray constant either way, so shared edges stay watertight */
    r->Sz = r->invd[kz];
    r->Sx = d[kx] * r->Sz;
    r->Sy = d[ky] * r->Sz;
}

static int wray_isect(const wray* r, const float* a, const float* b, const float* c, float* t, float* u, float* v)
{
    const int kx = r->kx, ky = r->ky, kz = r->kz;
    const float Akx = a[kx] - r->o[kx], Aky = a[ky] - r->o[ky], Akz = a[kz] - r->o[kz];
    const float Bkx = b[kx] - r->o[kx], Bky = b[ky] - r->o[ky], Bkz = b[kz] - r->o[kz];
    const float Ckx = c[kx] - r->o[kx], Cky = c[ky] - r->o[ky], Ckz = c[kz] - r->o[kz];
    const float Ax = fmaf(-r->Sx, Akz, Akx), Ay = fmaf(-r->Sy, Akz, Aky);
    const float Bx = fmaf(-r->Sx, Bkz, Bkx), By = fmaf(-r->Sy, Bkz, Bky);
    const float Cx = fmaf(-r->Sx, Ckz, Ckx), Cy = fmaf(-r->Sy, Ckz, Cky);
    /* plain products, as on the device (exact negation across a shared edge) */
    const float U = Cx * By - Cy * Bx;
    const float V = Ax * Cy - Ay * Cx;
    const float W = Bx * Ay - By * Ax;
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
    const float det = (U + V) + W;
    if (det == 0.0f) return 0;
    const float T = fmaf(U, r->Sz * Akz, fmaf(V, r->Sz * Bkz, W * (r->Sz * Ckz)));
    *t = T / det;
    *u = V / det;
    *v = W / det;
    return 1;
}

static inline float qbyte(uint32_t w, int k) { return (float)((w >> (8 * k)) & 0xffu); }

typedef struct wscene {
    const wnode* nodes;
    const wtri* tris;
    const winst* insts;
    uint32_t root;
    int two_level;
} wscene;

static atomic_uint g_max_stack;

/* Deepest stack any walk reached since the last call of this function. */
uint32_t orc_wide_max_stack(int reset)
{
    return reset ? atomic_exchange(&g_max_stack, 0u) : atomic_load(&g_max_stack);
}

static void note_stack(uint32_t sp)
{
    uint32_t cur = atomic_load(&g_max_stack);
    while (sp > cur && !atomic_compare_exchange_weak(&g_max_stack, &cur, sp)) {
    }
}

static void walk(const wscene* s, const float* ray8, int anyhit, uint32_t* stk_e, float* stk_t, float* hit4,
                 uint32_t* cnt3, uint32_t stack_cap)
{
    if (orc_ray_rejected(ray8)) {
        /* the device tests the root node's slots with an empty range, then stops */
        orc_miss_record(ray8, hit4);
        if (cnt3) { cnt3[0] = 1; cnt3[1] = 0; cnt3[2] = 0; }
        return;
    }
    uint32_t max_sp = 0;
    wray r;
    const float* wo = ray8;
    const float* wd = ray8 + 4;
    const float tmin = ray8[3], tmax = ray8[7];
    wray_setup(&r, wo, wd);
    float tbest = tmax, bu = 0.0f, bv = 0.0f;
    uint32_t best = ORC_NO_HIT, id_base = 0, sp = 0, cur = s->root;
    uint32_t c_nodes = 0, c_tris = 0, c_inst = 0;
    for (;;) {
        int pop = 0, finish = 0;
        if (cur < W_ENTRY_INST) {
            const wnode* n = &s->nodes[cur];
            c_nodes++;
            float sc[3], cc[3];
            uint32_t nw[3], fw[3];
            for (int a = 0; a < 3; ++a) {
                sc[a] = ldexpf(1.0f, n->e[a]) * r.invd[a];
                cc[a] = (n->p[a] - r.o[a]) * r.invd[a];
                const int pos = r.invd[a] >= 0.0f;
                nw[a] = pos ? n->qlo[a] : n->qhi[a];
                fw[a] = pos ? n->qhi[a] : n->qlo[a];
            }
            float tk[4];
            uint32_t ek[4];
            for (int k = 0; k < 4; ++k) {
                const float tn = fmaxf(fmaxf(fmaf(qbyte(nw[0], k), sc[0], cc[0]), fmaf(qbyte(nw[1], k), sc[1], cc[1])),
                                       fmaxf(fmaf(qbyte(nw[2], k), sc[2], cc[2]), tmin));
                const float tf = fminf(fminf(fmaf(qbyte(fw[0], k), sc[0], cc[0]), fmaf(qbyte(fw[1], k), sc[1], cc[1])),
                                       fminf(fmaf(qbyte(fw[2], k), sc[2], cc[2]), tbest));
                const int hit = ((n->imask >> (4 + k)) & 1u) && tn <= tf * W_SLACK;
                tk[k] = hit ? tn : INFINITY;
                ek[k] = n->entry[k];
            }
            if (!anyhit) {
                static const int net[5][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}};
                for (int q = 0; q < 5; ++q) {
                    const int a = net[q][0], b = net[q][1];
                    if (tk[b] < tk[a]) {
                        float tt = tk[a]; tk[a] = tk[b]; tk[b] = tt;
                        uint32_t ee = ek[a]; ek[a] = ek[b]; ek[b] = ee;
                    }
                }
                for (int k = 3; k >= 1; --k)
                    if (tk[k] != INFINITY) { stk_e[sp] = ek[k]; stk_t[sp] = tk[k]; ++sp; }
                if (tk[0] != INFINITY) cur = ek[0]; else pop = 1;
            } else {
                int have = 0;
                for (int k = 3; k >= 0; --k) {
                    if (tk[k] != INFINITY) {
                        if (have) { stk_e[sp] = cur; stk_t[sp] = 0.0f; ++sp; }
                        cur = ek[k];
                        have = 1;
                    }
                }
                if (!have) pop = 1;
            }
        } else if (cur & W_ENTRY_TRI) {
            c_tris++;
            const uint32_t ti = (cur >> 3) & 0x0FFFFFFFu;
            const wtri* tr = &s->tris[ti];
            float t, u, v;
            if (wray_isect(&r, tr->v0, tr->v1, tr->v2, &t, &u, &v) && t > tmin && t < tmax) {
                const uint32_t id = id_base + tr->local_id;
                if (anyhit) {
                    tbest = t; best = id; bu = u; bv = v;
                    finish = 1;
                } else if (t < tbest || (t == tbest && id < best)) {
                    tbest = t; best = id; bu = u; bv = v;
                }
            }
            if (!finish) {
                if ((cur & 7u) > 1u) cur += 7u; else pop = 1;
            }
        } else if (cur == W_ENTRY_RETURN) {
            wray_setup(&r, wo, wd);
            id_base = 0;
            pop = 1;
        } else {
            c_inst++;
            const winst* in = &s->insts[(cur >> 3) & 0x07FFFFFFu];
            const float* m = in->w2o;
            float oo[3], od[3];
            for (int i = 0; i < 3; ++i) {
                const float* row = m + 4 * i;
                oo[i] = fmaf(row[0], wo[0], fmaf(row[1], wo[1], fmaf(row[2], wo[2], row[3])));
                od[i] = fmaf(row[0], wd[0], fmaf(row[1], wd[1], row[2] * wd[2]));
            }
            wray_setup(&r, oo, od);
            stk_e[sp] = W_ENTRY_RETURN;
            stk_t[sp] = -INFINITY;
            ++sp;
            cur = in->blas_root;
            id_base = in->id_base;
        }
        if (sp > max_sp) max_sp = sp;
        if (sp + 4 >= stack_cap) { finish = 1; max_sp = 0xFFFFFFFFu; break; }
        if (pop) {
            for (;;) {
                if (sp == 0) { finish = 1; break; }
                --sp;
                if (anyhit || stk_t[sp] <= tbest * W_SLACK) { cur = stk_e[sp]; break; }
            }
        }
        if (finish) break;
    }
    note_stack(max_sp);
    const int miss = best == ORC_NO_HIT;
    hit4[0] = miss ? tmax : tbest;
    memcpy(&hit4[1], &best, 4);
    hit4[2] = miss ? 0.0f : bu;
    hit4[3] = miss ? 0.0f : bv;
    if (cnt3) { cnt3[0] = c_nodes; cnt3[1] = c_tris; cnt3[2] = c_inst; }
}

typedef struct wjob {
    wscene s;
    const float* rays;
    uint32_t n;
    int anyhit;
    float* hits;
    uint32_t* counts;
    uint32_t stack_cap;
    atomic_uint next;
} wjob;

static void* wworker(void* arg)
{
    wjob* j = (wjob*)arg;
    uint32_t* se = (uint32_t*)malloc(sizeof(uint32_t) * j->stack_cap);
    float* st = (float*)malloc(sizeof(float) * j->stack_cap);
    if (se && st) {
        for (;;) {
            const uint32_t b = atomic_fetch_add(&j->next, 256u);
            if (b >= j->n) break;
            const uint32_t e = b + 256u < j->n ? b + 256u : j->n;
            for (uint32_t i = b; i < e; ++i)
                walk(&j->s, j->rays + 8 * (size_t)i, j->anyhit, se, st, j->hits + 4 * (size_t)i,
                     j->counts ? j->counts + 3 * (size_t)i : NULL, j->stack_cap);
        }
    }
    free(se);
    free(st);
    return NULL;
}

void orc_trace_wide(const void* nodes, const void* tris, const void* instances, uint32_t root, int two_level,
                    uint32_t stack_entries, const float* rays, uint32_t n, int anyhit, int nthreads, float* hits,
                    uint32_t* counts)
{
    wjob j;
    j.s.nodes = (const wnode*)nodes;
    j.s.tris = (const wtri*)tris;
    j.s.insts = (const winst*)instances;
    j.s.root = root;
    j.s.two_level = two_level;
    j.rays = rays;
    j.n = n;
    j.anyhit = anyhit;
    j.hits = hits;
    j.counts = counts;
    j.stack_cap = stack_entries + 8;
    atomic_init(&j.next, 0u);
    if (nthreads <= 0) {
        long c = sysconf(_SC_NPROCESSORS_ONLN);
        nthreads = c > 0 ? (int)c : 1;
    }
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    for (int i = 1; i < nthreads; ++i) pthread_create(&th[i], NULL, wworker, &j);
    wworker(&j);
    for (int i = 1; i < nthreads; ++i) pthread_join(th[i], NULL);
}
