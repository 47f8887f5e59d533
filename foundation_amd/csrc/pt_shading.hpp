// pt_shading.hpp — device sampling and BSDF functions of the path tracer.
//
// None of this exists in the reference (no BSDF, RNG or light sampling; its
// only shader is a textured-quad raster pass, src/Renderer/Triangle.slang:23-37).
// Each function restates a published method:
//  * PCG32 XSH-RR (O'Neill 2014); one generator per (seed, sample, pixel,
//    bounce), so a path's random numbers do not depend on where it sits in a
//    queue — the property that makes images independent of queue sizes.
//  * Orthonormal basis: Duff et al., "Building an Orthonormal Basis,
//    Revisited", JCGT 6(1), 2017.
//  * Cosine-weighted hemisphere: Malley's method (disk -> hemisphere).
//  * GGX / Trowbridge-Reitz D (Walter et al. 2007), Smith height-correlated
//    masking-shadowing G2 = 1 / (1 + L(wo) + L(wi)), Schlick Fresnel with
//    F0 = albedo (conductor), alpha = roughness^2, visible-normal sampling of
//    Heitz, "Sampling the GGX Distribution of Visible Normals", JCGT 7(4), 2018.
//  * Next-event estimation on parallelogram lights (uniform area sampling,
//    light chosen proportional to power) with the power heuristic
//    (Veach 1997) against BSDF sampling.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fnd {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;

struct Rng {
    uint64_t state;
    uint64_t inc;
};

__host__ __device__ __forceinline__ uint32_t rng_next(Rng& r)
{
    const uint64_t old = r.state;
    r.state = old * 6364136223846793005ULL + r.inc;
    const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

__host__ __device__ __forceinline__ float rng_float(Rng& r) { return float(rng_next(r) >> 8) * 0x1p-24f; }

// Output mixer of SplitMix64 (Steele, Lea, Flood, "Fast Splittable
// Pseudorandom Number Generators", OOPSLA 2014).
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// pcg32_srandom(initstate = hash(seed, sample, bounce, pixel), initseq = pixel).
// Hashing the key into the state matters: PCG streams that share a state and
// differ only in the increment are correlated.
__host__ __device__ __forceinline__ Rng rng_for(uint64_t seed, uint32_t sample, uint32_t pixel, uint32_t bounce)
{
    uint64_t k = mix64(seed + 0x9E3779B97F4A7C15ULL * (uint64_t(sample) + 1u));
    k = mix64(k ^ (0xD1B54A32D192ED03ULL * (uint64_t(bounce) + 1u)));
    k = mix64(k ^ (0x8CB92BA72F3D8DD7ULL * (uint64_t(pixel) + 1u)));
    Rng r;
    r.state = 0u;
    r.inc = (uint64_t(pixel) << 1u) | 1u;
    rng_next(r);
    r.state += k;
    rng_next(r);
    return r;
}

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) { return a * (1.0f / length(a)); }
__device__ __forceinline__ float max3(V3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }

struct Onb {
    V3 t, b, n;
};

__device__ __forceinline__ Onb onb_from(V3 n)
{
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    Onb o;
    o.t = v3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    o.b = v3(b, sign + n.y * n.y * a, -n.y);
    o.n = n;
    return o;
}
__device__ __forceinline__ V3 to_local(const Onb& o, V3 v) { return v3(dot(v, o.t), dot(v, o.b), dot(v, o.n)); }
__device__ __forceinline__ V3 to_world(const Onb& o, V3 v) { return o.t * v.x + o.b * v.y + o.n * v.z; }

// sin and cos of 2*pi*u for u in [0, 1), u a multiple of 2^-24 (the output
// of rng_float).  Written out so that any host evaluating the same fmaf
// sequence gets the same bits (a library sincosf differs between the device
// and a host libm by an ulp here and there):
//  * exact reduction to the nearest quarter turn: q = round(4u), r = u - q/4
//    is exact (both are multiples of 2^-24) with |r| <= 1/8;
//  * x = 2*pi*r in [-pi/4, pi/4], Taylor polynomials in Horner form through
//    x^11 (sin) and x^12 (cos); the first omitted terms are below 2.2e-9 and
//    1.1e-10 relative, so the results are within a few ulp of the true values;
//  * the quadrant q mod 4 rotates (sin x, cos x) by q quarter turns.
__host__ __device__ __forceinline__ void sincos_2pi(float u, float& s, float& c)
{
    const float qf = floorf(fmaf(u, 4.0f, 0.5f));
    const float r = fmaf(qf, -0.25f, u);
    const uint32_t q = uint32_t(qf) & 3u;
    const float x = r * 6.28318530717958647692f;
    const float x2 = x * x;
    float ps = fmaf(x2, 2.7557319223985893e-06f, -1.9841269841269841e-04f);  // 1/9!, -1/7!
    ps = fmaf(x2, ps, 8.3333333333333333e-03f);                               // 1/5!
    ps = fmaf(x2, ps, -1.6666666666666667e-01f);                              // -1/3!
    const float sx = fmaf(x * x2, ps, x);
    float pc = fmaf(x2, 2.0876756987868099e-09f, -2.7557319223985891e-07f);  // 1/12!, -1/10!
    pc = fmaf(x2, pc, 2.4801587301587302e-05f);                               // 1/8!
    pc = fmaf(x2, pc, -1.3888888888888889e-03f);                              // -1/6!
    pc = fmaf(x2, pc, 4.1666666666666667e-02f);                               // 1/4!
    pc = fmaf(x2, pc, -0.5f);                                                 // -1/2!
    const float cx = fmaf(x2, pc, 1.0f);
    s = (q & 1u) ? cx : sx;
    c = (q & 1u) ? -sx : cx;
    if (q & 2u) {
        s = -s;
        c = -c;
    }
}

// Malley: uniform disk point (polar), lifted to the hemisphere; pdf = cos / pi.
__device__ __forceinline__ V3 sample_cosine(float u1, float u2)
{
    const float r = sqrtf(u1);
    float s, c;
    sincos_2pi(u2, s, c);
    return v3(r * c, r * s, sqrtf(fmaxf(0.0f, 1.0f - u1)));
}

// ---- GGX (local frame, n = +z)

__device__ __forceinline__ float ggx_d(float nh, float a2)
{
    const float d = fmaf(nh * nh, a2 - 1.0f, 1.0f);
    return a2 / (kPi * d * d);
}

// Smith Lambda for GGX: (-1 + sqrt(1 + a^2 tan^2 theta)) / 2
__device__ __forceinline__ float ggx_lambda(float cos_t, float a2)
{
    const float c2 = cos_t * cos_t;
    const float t2 = fmaxf(0.0f, 1.0f - c2) / fmaxf(c2, 1e-20f);
    return 0.5f * (-1.0f + sqrtf(fmaf(a2, t2, 1.0f)));
}

__device__ __forceinline__ V3 schlick(V3 f0, float c)
{
    const float m = 1.0f - fminf(fmaxf(c, 0.0f), 1.0f);
    const float m5 = (m * m) * (m * m) * m;
    return f0 + (v3(1.0f, 1.0f, 1.0f) - f0) * m5;
}

// Heitz 2018, listing 1: visible normal for view direction wo (local, wo.z > 0).
__device__ __forceinline__ V3 ggx_sample_vndf(V3 wo, float alpha, float u1, float u2)
{
    const V3 vh = normalize(v3(alpha * wo.x, alpha * wo.y, wo.z));
    const float lensq = vh.x * vh.x + vh.y * vh.y;
    const V3 t1 = lensq > 0.0f ? v3(-vh.y, vh.x, 0.0f) * (1.0f / sqrtf(lensq)) : v3(1.0f, 0.0f, 0.0f);
    const V3 t2 = cross(vh, t1);
    const float r = sqrtf(u1);
    float s, c;
    sincos_2pi(u2, s, c);
    const float p1 = r * c;
    float p2 = r * s;
    const float sh = 0.5f * (1.0f + vh.z);
    p2 = (1.0f - sh) * sqrtf(fmaxf(0.0f, 1.0f - p1 * p1)) + sh * p2;
    const V3 nh = t1 * p1 + t2 * p2 + vh * sqrtf(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2));
    return normalize(v3(alpha * nh.x, alpha * nh.y, fmaxf(0.0f, nh.z)));
}

struct BsdfEval {
    V3 f_cos;   // f * cos(theta_i)
    float pdf;  // solid-angle pdf of BSDF sampling for wi
};

__device__ __forceinline__ BsdfEval bsdf_eval(uint32_t type, V3 albedo, float roughness, V3 wo, V3 wi)
{
    BsdfEval e{v3(0, 0, 0), 0.0f};
    if (wo.z <= 0.0f || wi.z <= 0.0f) return e;
    if (type == 0u) {  // Lambert
        e.f_cos = albedo * (wi.z * kInvPi);
        e.pdf = wi.z * kInvPi;
        return e;
    }
    const float a = roughness * roughness;
    const float a2 = a * a;
    const V3 h = normalize(wo + wi);
    const float D = ggx_d(h.z, a2);
    const float lo = ggx_lambda(wo.z, a2), li = ggx_lambda(wi.z, a2);
    const float G2 = 1.0f / (1.0f + lo + li);
    const float G1o = 1.0f / (1.0f + lo);
    const V3 F = schlick(albedo, dot(wi, h));
    e.f_cos = F * (D * G2 / (4.0f * wo.z));
    e.pdf = G1o * D / (4.0f * wo.z);
    return e;
}

struct BsdfSample {
    V3 wi;
    V3 weight;  // f * cos / pdf
    float pdf;
    bool ok;
};

__device__ __forceinline__ BsdfSample bsdf_sample(uint32_t type, V3 albedo, float roughness, V3 wo, float u1,
                                                  float u2)
{
    BsdfSample s{v3(0, 0, 1), v3(0, 0, 0), 0.0f, false};
    if (wo.z <= 0.0f) return s;
    if (type == 0u) {
        s.wi = sample_cosine(u1, u2);
        s.pdf = s.wi.z * kInvPi;
        s.weight = albedo;
        s.ok = s.wi.z > 0.0f;
        return s;
    }
    const float a = roughness * roughness;
    const float a2 = a * a;
    const V3 h = ggx_sample_vndf(wo, a, u1, u2);
    const float oh = dot(wo, h);
    s.wi = h * (2.0f * oh) - wo;
    if (s.wi.z <= 0.0f || oh <= 0.0f) return s;
    const float lo = ggx_lambda(wo.z, a2), li = ggx_lambda(s.wi.z, a2);
    const float D = ggx_d(h.z, a2);
    s.pdf = D / (4.0f * wo.z * (1.0f + lo));
    s.weight = schlick(albedo, oh) * ((1.0f + lo) / (1.0f + lo + li));
    s.ok = s.pdf > 0.0f;
    return s;
}

__device__ __forceinline__ float power_heuristic(float a, float b)
{
    const float a2 = a * a, b2 = b * b;
    return a2 + b2 > 0.0f ? a2 / (a2 + b2) : 0.0f;
}

}  // namespace fnd
