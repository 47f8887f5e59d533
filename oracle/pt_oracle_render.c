/*
 * pt_oracle_render.c — CPU oracle of the path tracer (estimator, BSDFs,
 * light sampling, accumulation).
 *
 * TEST INFRASTRUCTURE ONLY (see pt_oracle.h).  The reference has no renderer
 * beyond a textured-quad raster pass (src/Renderer/Renderer.cpp:312-366), so
 * this restates published methods:
 *   estimator        unidirectional path tracing with next-event estimation
 *                    and multiple importance sampling (power heuristic),
 *                    Veach 1997, ch. 9; depth convention of pbrt-v3
 *                    (PathIntegrator::Li): emission is added at every vertex,
 *                    the path stops at bounce == max_depth.
 *   camera           glm::lookAt + perspective, Z up, vertical fov
 *                    (src/Renderer/Renderer.cpp:373-380), jittered pixel.
 *   Lambert          f = albedo / pi, cosine sampling (Malley).
 *   GGX              D of Walter et al. 2007, Smith height-correlated G2,
 *                    Schlick F (F0 = albedo), alpha = roughness^2, visible
 *                    normal sampling of Heitz, JCGT 7(4), 2018.
 *   lights           parallelograms sampled uniformly by area; one-sided.
 *                    With at most 8 lights the light is chosen with
 *                    probability 0.9 w_l / sum(w) + 0.1 (power share), w_l =
 *                    power x max corner cosine at the receiver x distance
 *                    from the light plane / d^3 to the centre, evaluated at
 *                    the offset origin with the normal quantised to 16+16-bit
 *                    octahedral (Cigolle et al., JCGT 3(2), 2014) both at NEE
 *                    and for MIS at a BSDF-sampled light hit; otherwise in
 *                    proportion to power.
 *   RNG              PCG32 with one stream per (seed, sample, pixel, bounce).
 * Pinned by: the closed-form furnace (exact), agreement of the MIS estimator
 * with BSDF-only sampling (unbiasedness), and the pdf checks of the tests.
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

typedef struct { double x, y, z; } d3;  /* only for camera set-up */

typedef struct v3f { float x, y, z; } v3f;
static inline v3f mk(float x, float y, float z) { v3f r = {x, y, z}; return r; }
static inline v3f add(v3f a, v3f b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3f sub(v3f a, v3f b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3f scl(v3f a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3f mul(v3f a, v3f b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
/* fma nesting of the device's dot(): the oracle evaluates the device's
 * expressions in the same order, so a path's radiance has the same bits. */
static inline float dot3(v3f a, v3f b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
static inline v3f cross3(v3f a, v3f b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline v3f nrm(v3f a) { return scl(a, 1.0f / sqrtf(dot3(a, a))); }
static inline float maxc(v3f a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }

#define PI_F 3.14159265358979323846f
#define INV_PI_F 0.31830988618379067154f

/* sin and cos of 2 pi u, u in [0, 1) a multiple of 2^-24: exact reduction to
 * the nearest quarter turn (q = round(4u), r = u - q/4 exact, |r| <= 1/8),
 * x = 2 pi r, Taylor series of sin through x^11 and of cos through x^12 in
 * Horner form (first omitted terms below 2.2e-9 and 1.1e-10 relative), then
 * a rotation by q quarter turns. */
static void sincos_2pi(float u, float* s, float* c)
{
    const float qf = floorf(fmaf(u, 4.0f, 0.5f));
    const float r = fmaf(qf, -0.25f, u);
    const uint32_t q = (uint32_t)qf & 3u;
    const float x = r * 6.28318530717958647692f;
    const float x2 = x * x;
    float ps = fmaf(x2, 2.7557319223985893e-06f, -1.9841269841269841e-04f);
    ps = fmaf(x2, ps, 8.3333333333333333e-03f);
    ps = fmaf(x2, ps, -1.6666666666666667e-01f);
    const float sx = fmaf(x * x2, ps, x);
    float pc = fmaf(x2, 2.0876756987868099e-09f, -2.7557319223985891e-07f);
    pc = fmaf(x2, pc, 2.4801587301587302e-05f);
    pc = fmaf(x2, pc, -1.3888888888888889e-03f);
    pc = fmaf(x2, pc, 4.1666666666666667e-02f);
    pc = fmaf(x2, pc, -0.5f);
    const float cx = fmaf(x2, pc, 1.0f);
    float ss = (q & 1u) ? cx : sx, cc = (q & 1u) ? -sx : cx;
    if (q & 2u) { ss = -ss; cc = -cc; }
    *s = ss;
    *c = cc;
}

void orc_sincos_2pi(const float* u, uint32_t n, float* out2)
{
    for (uint32_t i = 0; i < n; ++i) sincos_2pi(u[i], &out2[2 * (size_t)i], &out2[2 * (size_t)i + 1]);
}

/* SplitMix64 output mixer (Steele, Lea, Flood, OOPSLA 2014). */
static uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* PCG32 stream of one (seed, sample, pixel, bounce): the key is hashed into
 * the initial state, the pixel selects the stream. */
static void rng_for(orc_pcg32* r, uint64_t seed, uint32_t sample, uint32_t pixel, uint32_t bounce)
{
    uint64_t k = mix64(seed + 0x9E3779B97F4A7C15ULL * ((uint64_t)sample + 1u));
    k = mix64(k ^ (0xD1B54A32D192ED03ULL * ((uint64_t)bounce + 1u)));
    k = mix64(k ^ (0x8CB92BA72F3D8DD7ULL * ((uint64_t)pixel + 1u)));
    orc_pcg32_srandom(r, k, pixel);
}

/* Frisvad-style basis made branchless: Duff et al., JCGT 6(1), 2017. */
static void basis(v3f n, v3f* t, v3f* b)
{
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float bb = n.x * n.y * a;
    *t = mk(1.0f + sign * n.x * n.x * a, sign * bb, -sign * n.x);
    *b = mk(bb, sign + n.y * n.y * a, -n.y);
}

static float ggx_D(float nh, float a2)
{
    const float d = fmaf(nh * nh, a2 - 1.0f, 1.0f);
    return a2 / (PI_F * d * d);
}

static float ggx_Lambda(float c, float a2)
{
    const float c2 = c * c;
    const float t2 = fmaxf(0.0f, 1.0f - c2) / fmaxf(c2, 1e-20f);
    return 0.5f * (-1.0f + sqrtf(fmaf(a2, t2, 1.0f)));
}

static v3f schlick(v3f f0, float c)
{
    const float m = 1.0f - fminf(fmaxf(c, 0.0f), 1.0f);
    const float m5 = (m * m) * (m * m) * m;
    return add(f0, scl(sub(mk(1, 1, 1), f0), m5));
}

/* f * cos(theta_i) and the BSDF-sampling pdf of wi, local frame (n = +z). */
static void bsdf_eval(int type, v3f albedo, float rough, v3f wo, v3f wi, v3f* fcos, float* pdf)
{
    *fcos = mk(0, 0, 0);
    *pdf = 0.0f;
    if (wo.z <= 0.0f || wi.z <= 0.0f) return;
    if (type == ORC_MAT_LAMBERT) {
        *fcos = scl(albedo, wi.z * INV_PI_F);
        *pdf = wi.z * INV_PI_F;
        return;
    }
    const float a = rough * rough, a2 = a * a;
    const v3f h = nrm(add(wo, wi));
    const float D = ggx_D(h.z, a2);
    const float Lo = ggx_Lambda(wo.z, a2), Li = ggx_Lambda(wi.z, a2);
    const float G2 = 1.0f / (1.0f + Lo + Li), G1o = 1.0f / (1.0f + Lo);
    const v3f F = schlick(albedo, dot3(wi, h));
    *fcos = scl(F, D * G2 / (4.0f * wo.z));
    *pdf = G1o * D / (4.0f * wo.z);
}

/* Heitz 2018 visible-normal sample for view wo. */
static v3f ggx_vndf(v3f wo, float a, float u1, float u2)
{
    const v3f vh = nrm(mk(a * wo.x, a * wo.y, wo.z));
    const float l2 = vh.x * vh.x + vh.y * vh.y;
    const v3f T1 = l2 > 0.0f ? scl(mk(-vh.y, vh.x, 0.0f), 1.0f / sqrtf(l2)) : mk(1, 0, 0);
    const v3f T2 = cross3(vh, T1);
    const float r = sqrtf(u1);
    float sn, cs;
    sincos_2pi(u2, &sn, &cs);
    const float t1 = r * cs;
    float t2 = r * sn;
    const float s = 0.5f * (1.0f + vh.z);
    t2 = (1.0f - s) * sqrtf(fmaxf(0.0f, 1.0f - t1 * t1)) + s * t2;
    const v3f nh = add(add(scl(T1, t1), scl(T2, t2)), scl(vh, sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2))));
    return nrm(mk(a * nh.x, a * nh.y, fmaxf(0.0f, nh.z)));
}

/* Returns 1 with (wi, weight = f cos / pdf, pdf). */
static int bsdf_sample(int type, v3f albedo, float rough, v3f wo, float u1, float u2, v3f* wi, v3f* w, float* pdf)
{
    if (wo.z <= 0.0f) return 0;
    if (type == ORC_MAT_LAMBERT) {
        const float r = sqrtf(u1);
        float sn, cs;
        sincos_2pi(u2, &sn, &cs);
        *wi = mk(r * cs, r * sn, sqrtf(fmaxf(0.0f, 1.0f - u1)));
        *pdf = wi->z * INV_PI_F;
        *w = albedo;
        return wi->z > 0.0f;
    }
    const float a = rough * rough, a2 = a * a;
    const v3f h = ggx_vndf(wo, a, u1, u2);
    const float oh = dot3(wo, h);
    *wi = sub(scl(h, 2.0f * oh), wo);
    if (wi->z <= 0.0f || oh <= 0.0f) return 0;
    const float Lo = ggx_Lambda(wo.z, a2), Li = ggx_Lambda(wi->z, a2);
    *pdf = ggx_D(h.z, a2) / (4.0f * wo.z * (1.0f + Lo));
    *w = scl(schlick(albedo, oh), (1.0f + Lo) / (1.0f + Lo + Li));
    return *pdf > 0.0f;
}

static float power_h(float a, float b)
{
    const float a2 = a * a, b2 = b * b;
    return a2 + b2 > 0.0f ? a2 / (a2 + b2) : 0.0f;
}

typedef struct rjob {
    const orc_scene* s;
    uint32_t width, height, spp, sample_offset, max_depth, rr_depth, mode;
    uint64_t seed;
    v3f eye, fwd, right, up;
    float tan_half, aspect;
    float* accum;
    double* accum64;
    atomic_uint next;
    atomic_ullong rays;
} rjob;

/* ---- light selection adapted to the shading point (see the header comment) */
#define ADAPTIVE_LIGHTS 8u

static uint32_t oct_encode(v3f n)
{
    const float s = fabsf(n.x) + fabsf(n.y) + fabsf(n.z);
    float u = n.x / s, v = n.y / s;
    if (n.z < 0.0f) {
        const float ou = u;
        u = (1.0f - fabsf(v)) * (ou >= 0.0f ? 1.0f : -1.0f);
        v = (1.0f - fabsf(ou)) * (v >= 0.0f ? 1.0f : -1.0f);
    }
    const uint32_t qu = (uint32_t)(fminf(fmaxf(u * 0.5f + 0.5f, 0.0f), 1.0f) * 65535.0f + 0.5f);
    const uint32_t qv = (uint32_t)(fminf(fmaxf(v * 0.5f + 0.5f, 0.0f), 1.0f) * 65535.0f + 0.5f);
    return qu | (qv << 16);
}

static v3f oct_decode(uint32_t c)
{
    float u = (float)(c & 0xFFFFu) * (2.0f / 65535.0f) - 1.0f;
    float v = (float)(c >> 16) * (2.0f / 65535.0f) - 1.0f;
    const float z = 1.0f - fabsf(u) - fabsf(v);
    if (z < 0.0f) {
        const float ou = u;
        u = (1.0f - fabsf(v)) * (ou >= 0.0f ? 1.0f : -1.0f);
        v = (1.0f - fabsf(ou)) * (v >= 0.0f ? 1.0f : -1.0f);
    }
    return nrm(mk(u, v, z));
}

/* Unit normal of a light, formed in FP32 from its edges as the device forms it. */
static v3f light_n(const orc_light* lt)
{
    return nrm(cross3(mk(lt->e0[0], lt->e0[1], lt->e0[2]), mk(lt->e1[0], lt->e1[1], lt->e1[2])));
}

static float light_weight(const orc_light* lt, v3f p, v3f n)
{
    const v3f c0 = mk(lt->corner[0], lt->corner[1], lt->corner[2]);
    const v3f e0 = mk(lt->e0[0], lt->e0[1], lt->e0[2]), e1 = mk(lt->e1[0], lt->e1[1], lt->e1[2]);
    const float side = dot3(light_n(lt), sub(p, c0));
    if (!(side > 0.0f)) return 0.0f;
    float cx = 0.0f;
    const v3f cs[4] = {c0, add(c0, e0), add(c0, e1), add(add(c0, e0), e1)};
    for (int k = 0; k < 4; ++k) {
        const v3f v = sub(cs[k], p);
        const float l2 = dot3(v, v);
        if (l2 > 0.0f) cx = fmaxf(cx, dot3(n, v) / sqrtf(l2));
    }
    if (!(cx > 0.0f)) return 0.0f;
    const v3f v = sub(add(c0, scl(add(e0, e1), 0.5f)), p);
    const float d2 = fmaxf(dot3(v, v), 0.25f * lt->area);
    const float power = (lt->radiance[0] + lt->radiance[1] + lt->radiance[2]) * lt->area;
    return power * cx * side / (d2 * sqrtf(d2));
}

/* Probability that NEE at (p, n) picks light li; with pick != NULL, also the
 * light chosen for uniform u (returned through *pick, its probability returned). */
static float light_select(const orc_scene* s, v3f p, v3f n, uint32_t li, float u, uint32_t* pick)
{
    const uint32_t nl = s->num_lights;
    if (nl > ADAPTIVE_LIGHTS) {
        if (pick) {
            uint32_t l = 0;
            while (l + 1 < nl && u >= s->lights[l].cdf) ++l;
            *pick = l;
            li = l;
        }
        return s->lights[li].pdf_select;
    }
    float w[ADAPTIVE_LIGHTS], total = 0.0f;
    for (uint32_t l = 0; l < nl; ++l) {
        w[l] = light_weight(&s->lights[l], p, n);
        total += w[l];
    }
    const float inv = total > 0.0f ? 0.9f / total : 0.0f;
    const float floor_share = total > 0.0f ? 0.1f : 1.0f;
    if (pick) {
        float acc = 0.0f;
        li = nl - 1;
        for (uint32_t l = 0; l < nl; ++l) {
            acc += w[l] * inv + floor_share * s->lights[l].pdf_select;
            if (u < acc) { li = l; break; }
        }
        *pick = li;
    }
    return w[li] * inv + floor_share * s->lights[li].pdf_select;
}

/* Camera of the reference's only view (src/Renderer/Renderer.cpp:373-380):
 * glm::lookAt with a Z-up world, forward = normalize(at - eye), right =
 * normalize(cross(forward, up)), up' = cross(right, forward), a vertical
 * field of view and aspect = width / height.  The basis is formed in double
 * and rounded once; tan(fov / 2) likewise. */
static void camera_init(rjob* j, const float* cam10, uint32_t width, uint32_t height)
{
    d3 e = {cam10[0], cam10[1], cam10[2]}, a = {cam10[3], cam10[4], cam10[5]}, u = {cam10[6], cam10[7], cam10[8]};
    d3 f = {a.x - e.x, a.y - e.y, a.z - e.z};
    double l = sqrt(f.x * f.x + f.y * f.y + f.z * f.z);
    f.x /= l; f.y /= l; f.z /= l;
    d3 r = {f.y * u.z - f.z * u.y, f.z * u.x - f.x * u.z, f.x * u.y - f.y * u.x};
    l = sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
    r.x /= l; r.y /= l; r.z /= l;
    d3 up = {r.y * f.z - r.z * f.y, r.z * f.x - r.x * f.z, r.x * f.y - r.y * f.x};
    j->width = width;
    j->height = height;
    j->eye = mk((float)e.x, (float)e.y, (float)e.z);
    j->fwd = mk((float)f.x, (float)f.y, (float)f.z);
    j->right = mk((float)r.x, (float)r.y, (float)r.z);
    j->up = mk((float)up.x, (float)up.y, (float)up.z);
    j->tan_half = (float)tan((double)cam10[9] * M_PI / 360.0);
    j->aspect = (float)((double)width / (double)height);
}

/* Unit direction through image point (x + jx, y + jy), row 0 at the top:
 * screen x in [-1, 1] scaled by tan(fov / 2) * aspect, screen y by tan(fov / 2). */
static v3f camera_dir(const rjob* j, uint32_t x, uint32_t y, float jx, float jy)
{
    const float sx = (2.0f * ((float)x + jx) / (float)j->width - 1.0f) * j->tan_half * j->aspect;
    const float sy = (1.0f - 2.0f * ((float)y + jy) / (float)j->height) * j->tan_half;
    return nrm(add(add(j->fwd, scl(j->right, sx)), scl(j->up, sy)));
}

/* Camera rays as the renderer starts its paths: out8 per point (origin, 0,
 * direction, 0).  With jitter NULL, point i is (xs[i] + jx[i], ys[i] + jy[i])
 * given by the caller (closed-form checks); otherwise every pixel of the
 * image with the jitter of sample `sample` drawn from rng_for(seed, sample,
 * pixel, 0), in pixel order. */
void orc_camera_rays(const float* cam10, uint32_t width, uint32_t height, const uint32_t* xs, const uint32_t* ys,
                     const float* jxy, uint32_t n, uint64_t seed, uint32_t sample, float* out8)
{
    rjob j;
    memset(&j, 0, sizeof j);
    camera_init(&j, cam10, width, height);
    const uint32_t count = xs ? n : width * height;
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t x, y;
        float jx, jy;
        if (xs) {
            x = xs[i]; y = ys[i]; jx = jxy[2 * (size_t)i]; jy = jxy[2 * (size_t)i + 1];
        } else {
            x = i % width; y = i / width;
            orc_pcg32 rng;
            rng_for(&rng, seed, sample, i, 0);
            jx = orc_pcg32_float(&rng);
            jy = orc_pcg32_float(&rng);
        }
        const v3f d = camera_dir(&j, x, y, jx, jy);
        float* o = out8 + 8 * (size_t)i;
        o[0] = j.eye.x; o[1] = j.eye.y; o[2] = j.eye.z; o[3] = 0.0f;
        o[4] = d.x; o[5] = d.y; o[6] = d.z; o[7] = 3.0e38f;
    }
}

/* `ndraws` outputs of the PCG32 stream of (seed, sample, pixel, bounce) for
 * pixels pixel0 .. pixel0 + npix - 1: out[(p - pixel0) * ndraws + k]. */
void orc_rng_draws(uint64_t seed, uint32_t sample, uint32_t pixel0, uint32_t npix, uint32_t bounce, uint32_t ndraws,
                   uint32_t* out)
{
    for (uint32_t p = 0; p < npix; ++p) {
        orc_pcg32 r;
        rng_for(&r, seed, sample, pixel0 + p, bounce);
        for (uint32_t k = 0; k < ndraws; ++k) out[(size_t)p * ndraws + k] = orc_pcg32_next(&r);
    }
}

/* mode: 0 = MIS (NEE + BSDF), 1 = BSDF sampling only, 2 = NEE only for direct light */
static void render_pixel(rjob* j, uint32_t pixel)
{
    const orc_scene* s = j->s;
    const uint32_t x = pixel % j->width, y = pixel / j->width;
    float acc[3] = {j->accum[3 * (size_t)pixel], j->accum[3 * (size_t)pixel + 1], j->accum[3 * (size_t)pixel + 2]};
    double acc64[3] = {0, 0, 0};
    if (j->accum64) for (int c = 0; c < 3; ++c) acc64[c] = j->accum64[3 * (size_t)pixel + c];
    uint64_t rays = 0;
    for (uint32_t si = 0; si < j->spp; ++si) {
        const uint32_t sample = j->sample_offset + si;
        orc_pcg32 rng;
        rng_for(&rng, j->seed, sample, pixel, 0);
        const float jx = orc_pcg32_float(&rng), jy = orc_pcg32_float(&rng);
        v3f o = j->eye;
        v3f d = camera_dir(j, x, y, jx, jy);
        v3f thr = mk(1, 1, 1), L = mk(0, 0, 0);
        float pdf_prev = 0.0f;
        uint32_t ncode_prev = 0;  /* quantised normal of the previous vertex (MIS of light hits) */
        for (uint32_t bounce = 0; bounce <= j->max_depth; ++bounce) {
            orc_shade_hit h;
            rays++;
            if (!orc_closest_shade(s, (v3f_in){o.x, o.y, o.z}, (v3f_in){d.x, d.y, d.z}, &h)) {
                L = add(L, mul(thr, mk(s->background[0], s->background[1], s->background[2])));
                break;
            }
            const orc_material* m = &s->materials[h.material];
            v3f ng = mk(h.n[0], h.n[1], h.n[2]);
            ng = nrm(ng);
            const int front = dot3(ng, d) < 0.0f;
            const v3f nf = front ? ng : scl(ng, -1.0f);
            const v3f p = add(o, scl(d, h.t));
            const v3f Le = mk(m->emission[0], m->emission[1], m->emission[2]);
            const int emissive = m->type == ORC_MAT_EMISSIVE;
            if ((Le.x > 0 || Le.y > 0 || Le.z > 0) && (!emissive || front)) {
                float w = 1.0f;
                if (bounce > 0 && pdf_prev > 0.0f && h.light >= 0) {
                    const orc_light* lt = &s->lights[h.light];
                    const float cl = fabsf(dot3(light_n(lt), d));
                    const float ps = light_select(s, o, oct_decode(ncode_prev), (uint32_t)h.light, 0.0f, NULL);
                    const float pl = cl > 0.0f ? ps * h.t * h.t / (lt->area * cl) : 0.0f;
                    w = j->mode == 1 ? 1.0f : (j->mode == 2 ? 0.0f : power_h(pdf_prev, pl));
                }
                L = add(L, scl(mul(thr, Le), w));
            }
            if (emissive || bounce >= j->max_depth) break;
            orc_pcg32 r2;
            rng_for(&r2, j->seed, sample, pixel, bounce + 1);
            v3f t, b;
            basis(nf, &t, &b);
            const v3f wo = mk(-dot3(d, t), -dot3(d, b), -dot3(d, nf));
            const v3f alb = mk(m->albedo[0], m->albedo[1], m->albedo[2]);
            const float eps = 1e-4f * fmaxf(1.0f, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
            const v3f org = add(p, scl(nf, eps));
            const float ul = orc_pcg32_float(&r2), ua = orc_pcg32_float(&r2), ub = orc_pcg32_float(&r2);
            const uint32_t ncode = oct_encode(nf);
            if (s->num_lights > 0 && j->mode != 1) {
                uint32_t li = 0;
                const float psel = light_select(s, org, oct_decode(ncode), 0, ul, &li);
                const orc_light* lt = &s->lights[li];
                const v3f xl = add(add(mk(lt->corner[0], lt->corner[1], lt->corner[2]),
                                       scl(mk(lt->e0[0], lt->e0[1], lt->e0[2]), ua)),
                                   scl(mk(lt->e1[0], lt->e1[1], lt->e1[2]), ub));
                v3f wi = sub(xl, p);
                const float d2 = dot3(wi, wi);
                wi = scl(wi, 1.0f / sqrtf(d2));
                const float cl = -dot3(light_n(lt), wi);
                if (cl > 0.0f && d2 > 0.0f) {
                    const float pl = psel * d2 / (lt->area * cl);
                    v3f fc;
                    float pb;
                    bsdf_eval((int)m->type, alb, m->roughness, wo, mk(dot3(wi, t), dot3(wi, b), dot3(wi, nf)), &fc, &pb);
                    if (pb > 0.0f) {
                        const float w = j->mode == 2 ? 1.0f : power_h(pl, pb);
                        const v3f c = scl(mul(mul(thr, fc), mk(lt->radiance[0], lt->radiance[1], lt->radiance[2])), w / pl);
                        if (c.x > 0 || c.y > 0 || c.z > 0) {
                            rays++;
                            const v3f to = sub(xl, org);
                            if (!orc_occluded(s, (v3f_in){org.x, org.y, org.z}, (v3f_in){to.x, to.y, to.z}, 1.0f - 1e-3f))
                                L = add(L, c);
                        }
                    }
                }
            }
            const float u1 = orc_pcg32_float(&r2), u2 = orc_pcg32_float(&r2), urr = orc_pcg32_float(&r2);
            v3f wl, wgt;
            float pdf;
            if (!bsdf_sample((int)m->type, alb, m->roughness, wo, u1, u2, &wl, &wgt, &pdf)) break;
            v3f nthr = mul(thr, wgt);
            if (!(maxc(nthr) > 0.0f)) break;
            if (bounce >= j->rr_depth) {
                const float q = fminf(maxc(nthr), 0.95f);
                if (urr >= q) break;
                nthr = scl(nthr, 1.0f / q);
            }
            thr = nthr;
            pdf_prev = pdf;
            ncode_prev = ncode;
            o = org;
            d = nrm(add(add(scl(t, wl.x), scl(b, wl.y)), scl(nf, wl.z)));
        }
        acc[0] += L.x; acc[1] += L.y; acc[2] += L.z;
        acc64[0] += L.x; acc64[1] += L.y; acc64[2] += L.z;
    }
    for (int c = 0; c < 3; ++c) j->accum[3 * (size_t)pixel + c] = acc[c];
    if (j->accum64) for (int c = 0; c < 3; ++c) j->accum64[3 * (size_t)pixel + c] = acc64[c];
    atomic_fetch_add(&j->rays, rays);
}

static void* render_worker(void* arg)
{
    rjob* j = (rjob*)arg;
    const uint32_t npix = j->width * j->height;
    for (;;) {
        const uint32_t b = atomic_fetch_add(&j->next, 64u);
        if (b >= npix) break;
        const uint32_t e = b + 64u < npix ? b + 64u : npix;
        for (uint32_t p = b; p < e; ++p) render_pixel(j, p);
    }
    return NULL;
}

uint64_t orc_render(const orc_scene* s, const float* cam10, uint32_t width, uint32_t height, uint32_t spp,
                    uint32_t sample_offset, uint32_t max_depth, uint32_t rr_depth, uint64_t seed, uint32_t mode,
                    int nthreads, float* accum, double* accum64)
{
    if (!s || !cam10 || !accum || !width || !height) return 0;
    rjob j;
    memset(&j, 0, sizeof j);
    j.s = s; j.width = width; j.height = height; j.spp = spp; j.sample_offset = sample_offset;
    j.max_depth = max_depth; j.rr_depth = rr_depth; j.mode = mode; j.seed = seed;
    j.accum = accum; j.accum64 = accum64;
    camera_init(&j, cam10, width, height);
    atomic_init(&j.next, 0u);
    atomic_init(&j.rays, 0ull);
    if (nthreads <= 0) {
        long c = sysconf(_SC_NPROCESSORS_ONLN);
        nthreads = c > 0 ? (int)c : 1;
    }
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    for (int i = 1; i < nthreads; ++i) pthread_create(&th[i], NULL, render_worker, &j);
    render_worker(&j);
    for (int i = 1; i < nthreads; ++i) pthread_join(th[i], NULL);
    return (uint64_t)atomic_load(&j.rays);
}

/* Exposed for the sampling tests: one BSDF sample / evaluation in the local
 * frame.  out7 = wi[3], weight[3], pdf (sample); out4 = fcos[3], pdf (eval). */
int orc_bsdf_sample(int type, const float* albedo3, float rough, const float* wo3, float u1, float u2, float* out7)
{
    v3f wi, w;
    float pdf = 0.0f;
    const int ok = bsdf_sample(type, mk(albedo3[0], albedo3[1], albedo3[2]), rough, mk(wo3[0], wo3[1], wo3[2]), u1, u2,
                               &wi, &w, &pdf);
    out7[0] = wi.x; out7[1] = wi.y; out7[2] = wi.z;
    out7[3] = w.x; out7[4] = w.y; out7[5] = w.z;
    out7[6] = pdf;
    return ok;
}

void orc_bsdf_eval(int type, const float* albedo3, float rough, const float* wo3, const float* wi3, float* out4)
{
    v3f fc;
    float pdf;
    bsdf_eval(type, mk(albedo3[0], albedo3[1], albedo3[2]), rough, mk(wo3[0], wo3[1], wo3[2]), mk(wi3[0], wi3[1], wi3[2]),
              &fc, &pdf);
    out4[0] = fc.x; out4[1] = fc.y; out4[2] = fc.z; out4[3] = pdf;
}
