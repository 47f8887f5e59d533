// pt_device.hpp — device functions of the traversal: ray set-up, watertight
// ray/triangle test, quantised 4-wide slab test, instance transform.
//
// Every expression that decides a closest-hit record (ray constants,
// triangle test, instance transform) is written with explicit fmaf and this
// translation unit is compiled with -ffp-contract=off, so that a host
// evaluation of the same expressions in the same order gives the same bits.
// Box tests only need to be conservative: they carry a relative slack of
// 2^-16 and never decide which of two hits wins.
#pragma once
#include <hip/hip_runtime.h>

#include "pt_layout.hpp"

namespace fnd {

constexpr float kSlack = 1.0f + 0x1p-16f;
constexpr float kTinyDir = 0x1p-80f;

struct RayK {
    float ox, oy, oz;      // origin (current space)
    float dx, dy, dz;      // direction
    float ix, iy, iz;      // 1 / d with |d| clamped to >= 2^-80
    float Sx, Sy, Sz;      // shear constants
    uint32_t perm;         // kx | ky << 2 | kz << 4
};

__device__ __forceinline__ float sel3(uint32_t k, float x, float y, float z)
{
    return k == 0 ? x : (k == 1 ? y : z);
}

// Woop, Benthin, Wald, JCGT 2013, section 3: kz = axis of largest |d| (first
// on ties), kx, ky cyclic, swapped when d[kz] < 0.
__device__ __forceinline__ void ray_setup(RayK& r, float ox, float oy, float oz, float dx, float dy, float dz)
{
    r.ox = ox; r.oy = oy; r.oz = oz;
    r.dx = dx; r.dy = dy; r.dz = dz;
    const float cx = fabsf(dx) < kTinyDir ? copysignf(kTinyDir, dx) : dx;
    const float cy = fabsf(dy) < kTinyDir ? copysignf(kTinyDir, dy) : dy;
    const float cz = fabsf(dz) < kTinyDir ? copysignf(kTinyDir, dz) : dz;
    r.ix = 1.0f / cx;
    r.iy = 1.0f / cy;
    r.iz = 1.0f / cz;
    uint32_t kz = 0;
    float m = fabsf(dx);
    if (fabsf(dy) > m) { kz = 1; m = fabsf(dy); }
    if (fabsf(dz) > m) { kz = 2; }
    uint32_t kx = kz == 2 ? 0 : kz + 1;
    uint32_t ky = kx == 2 ? 0 : kx + 1;
    const float dkz = sel3(kz, dx, dy, dz);
    if (dkz < 0.0f) { const uint32_t t = kx; kx = ky; ky = t; }
    r.perm = kx | (ky << 2) | (kz << 4);
    // Sz is the slab reciprocal of axis kz and Sx, Sy are products with it
    // (Woop divides d[kx] / d[kz]; one per-ray constant is as watertight as the other)
    r.Sz = sel3(kz, r.ix, r.iy, r.iz);
    r.Sx = sel3(kx, dx, dy, dz) * r.Sz;
    r.Sy = sel3(ky, dx, dy, dz) * r.Sz;
}

// Row-major 3x4 world -> object, o' = M o + m, d' = M d, fma nesting fixed.
__device__ __forceinline__ void xform_ray(const float* m, float ox, float oy, float oz, float dx, float dy, float dz,
                                          float& rox, float& roy, float& roz, float& rdx, float& rdy, float& rdz)
{
    rox = fmaf(m[0], ox, fmaf(m[1], oy, fmaf(m[2], oz, m[3])));
    roy = fmaf(m[4], ox, fmaf(m[5], oy, fmaf(m[6], oz, m[7])));
    roz = fmaf(m[8], ox, fmaf(m[9], oy, fmaf(m[10], oz, m[11])));
    rdx = fmaf(m[0], dx, fmaf(m[1], dy, m[2] * dz));
    rdy = fmaf(m[4], dx, fmaf(m[5], dy, m[6] * dz));
    rdz = fmaf(m[8], dx, fmaf(m[9], dy, m[10] * dz));
}

// Woop et al. 2013, FP32 only: exact zeros of U, V, W count as inside.
// Returns true with t = T / det and the barycentric numerators V, W and
// det when the ray's line crosses the triangle; the caller range-checks t and
// divides u = V / det, v = W / det only for the hit it keeps (the same
// operations, so the same bits, as dividing here).
__device__ __forceinline__ bool isect_tri(const RayK& r, const float4 a, const float4 b, const float4 c, float& t,
                                          float& Vn, float& Wn, float& detn)
{
    const uint32_t kx = r.perm & 3u, ky = (r.perm >> 2) & 3u, kz = (r.perm >> 4) & 3u;
    // vertex - origin per axis, then permuted (the same differences as
    // permuting first; no permuted origin is kept per ray)
    const float ax = a.x - r.ox, ay = a.y - r.oy, az = a.z - r.oz;
    const float bx = b.x - r.ox, by = b.y - r.oy, bz = b.z - r.oz;
    const float cx = c.x - r.ox, cy = c.y - r.oy, cz = c.z - r.oz;
    const float Akx = sel3(kx, ax, ay, az);
    const float Aky = sel3(ky, ax, ay, az);
    const float Akz = sel3(kz, ax, ay, az);
    const float Bkx = sel3(kx, bx, by, bz);
    const float Bky = sel3(ky, bx, by, bz);
    const float Bkz = sel3(kz, bx, by, bz);
    const float Ckx = sel3(kx, cx, cy, cz);
    const float Cky = sel3(ky, cx, cy, cz);
    const float Ckz = sel3(kz, cx, cy, cz);
    const float Ax = fmaf(-r.Sx, Akz, Akx);
    const float Ay = fmaf(-r.Sy, Akz, Aky);
    const float Bx = fmaf(-r.Sx, Bkz, Bkx);
    const float By = fmaf(-r.Sy, Bkz, Bky);
    const float Cx = fmaf(-r.Sx, Ckz, Ckx);
    const float Cy = fmaf(-r.Sy, Ckz, Cky);
    // Plain products (no fma): the edge function of an edge seen from either of
    // its triangles is then the exact negation (products commute, x - y =
    // -(y - x)), which keeps shared edges watertight and gives det == 0 for a
    // triangle with two equal vertices.
    const float U = Cx * By - Cy * Bx;
    const float V = Ax * Cy - Ay * Cx;
    const float W = Bx * Ay - By * Ax;
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = (U + V) + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * Akz;
    const float Bz = r.Sz * Bkz;
    const float Cz = r.Sz * Ckz;
    const float T = fmaf(U, Az, fmaf(V, Bz, W * Cz));
    t = T / det;
    Vn = V;
    Wn = W;
    detn = det;
    return true;
}

// Zeroes the traversal work counter's shards; threads [0, kCounterShards) of
// a kernel that runs before the traversal on the same stream each clear one.
__device__ __forceinline__ void zero_trace_counter(uint32_t* counter, uint32_t t)
{
    if (t < kCounterShards) counter[t * kCounterStride] = 0u;
}

// x * 2^e, exact (v_ldexp_f32)
__device__ __forceinline__ float scale_exp2(float x, int e)
{
    return __builtin_amdgcn_ldexpf(x, e);
}

__device__ __forceinline__ float ubyte_f(uint32_t w, int k)
{
    return float((w >> (8 * k)) & 0xffu);
}

}  // namespace fnd
