/*
 * pt_oracle_cache.c — FP64 CPU oracle of the radiance cache.
 *
 * TEST INFRASTRUCTURE ONLY (see pt_oracle.h).  The reference has no learned
 * component.  Restated here:
 *   encoding  multiresolution hash grid, Mueller, Evans, Schied, Keller,
 *             "Instant Neural Graphics Primitives with a Multiresolution Hash
 *             Encoding", ACM TOG 41(4), 2022, section 3: L levels,
 *             N_l = floor(N_min * b^l), b = exp((ln N_max - ln N_min)/(L-1)),
 *             trilinear interpolation of the 8 corners of the cell of x*N_l,
 *             dense indexing while (N_l+1)^3 <= T, else
 *             (x * 1 xor y * 2654435761 xor z * 805459861) mod T.
 *   network   ReLU MLP without biases, Mueller, Rousselle, Novak, Keller,
 *             "Real-time Neural Radiance Caching for Path Tracing", ACM TOG
 *             40(4), 2021: input 48 (32 grid features; oriented normal, 1,
 *             normalised position, 0; outgoing direction, roughness, glossy
 *             flag = roughness > 0, zeros), 5 x 64 hidden, 16 outputs (RGB).
 *             The paper also feeds the direction and roughness to the
 *             network; a diffuse query passes direction 0 and roughness 0.
 *   loss      relative L2 of the same paper: (p - y)^2 / (sg(p)^2 + 0.01),
 *             summed over samples and RGB (sg: held constant when
 *             differentiating); plain_l2 = 1 drops the denominator, which
 *             makes the returned gradient the derivative of the returned loss
 *             (used to pin the backward pass by finite differences).
 * With round_fp16 = 1 the weights, grid values, encoded features and hidden
 * activations are rounded to FP16 as the GPU forward pass stores them; the
 * arithmetic stays in double.
 * Pinned by central finite differences of this same loss (tests) and by the
 * closed-form Adam step.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pt_oracle.h"

#define CL 16
#define CT (1u << 19)
#define CIN 48
#define CH 64
#define COUT 16

static const int W_OFF[6] = {0, 3072, 7168, 11264, 15360, 19456};
static const int W_IN[6] = {48, 64, 64, 64, 64, 64};
static const int W_OUT[6] = {64, 64, 64, 64, 64, 16};

/* IEEE binary16 round-to-nearest-even of a float, returned as double. */
static double to_half(double x)
{
    if (x != x) return x;
    const double ax = fabs(x);
    if (ax >= 65520.0) return x > 0 ? INFINITY : -INFINITY;
    if (ax < 0x1p-25) return copysign(0.0, x);
    int e;
    frexp(ax, &e);                 /* ax = m * 2^e, m in [0.5, 1) */
    int q = e - 11;                /* 11 significant bits */
    if (q < -24) q = -24;          /* subnormal spacing 2^-24 */
    const double scale = ldexp(1.0, q);
    return copysign(nearbyint(ax / scale) * scale, x);
}

void orc_cache_levels(int* res, uint32_t* dense_mask)
{
    const double b = exp((log(2048.0) - log(16.0)) / (double)(CL - 1));
    *dense_mask = 0;
    for (int l = 0; l < CL; ++l) {
        res[l] = (int)floor(16.0 * pow(b, (double)l) + 1e-9);
        const double cells = (double)(res[l] + 1) * (res[l] + 1) * (res[l] + 1);
        if (cells <= (double)CT) *dense_mask |= 1u << l;
    }
}

static void corners(double x, double y, double z, int res, int dense, uint32_t* idx, double* w)
{
    const double sx = x * res, sy = y * res, sz = z * res;
    int ix = (int)floor(sx), iy = (int)floor(sy), iz = (int)floor(sz);
    ix = ix < 0 ? 0 : (ix > res - 1 ? res - 1 : ix);
    iy = iy < 0 ? 0 : (iy > res - 1 ? res - 1 : iy);
    iz = iz < 0 ? 0 : (iz > res - 1 ? res - 1 : iz);
    const double fx = sx - ix, fy = sy - iy, fz = sz - iz;
    const uint32_t st = (uint32_t)(res + 1);
    for (int c = 0; c < 8; ++c) {
        const uint32_t cx = (uint32_t)(ix + (c & 1)), cy = (uint32_t)(iy + ((c >> 1) & 1)), cz = (uint32_t)(iz + (c >> 2));
        w[c] = ((c & 1) ? fx : 1 - fx) * (((c >> 1) & 1) ? fy : 1 - fy) * ((c >> 2) ? fz : 1 - fz);
        idx[c] = dense ? cx + cy * st + cz * st * st : ((cx) ^ (cy * 2654435761u) ^ (cz * 805459861u)) & (CT - 1u);
    }
}

/* Returns the loss summed over samples; pred (n*3), dmlp (20480) and dgrid
 * (16 * 2^19 * 2, or NULL) receive the prediction and the gradients.
 * queries: n * 10 (position, normal, outgoing direction, roughness);
 * lo3, inv_extent: the normalisation. */
double orc_cache_loss_grad(const float* mlp, const float* grid, const float* queries, const float* targets, uint32_t n,
                           const float* lo3, float inv_extent, int round_fp16, int plain_l2, float* pred, double* dmlp,
                           double* dgrid)
{
    int res[CL];
    uint32_t dense;
    orc_cache_levels(res, &dense);
    double* W = (double*)malloc(sizeof(double) * 20480);
    for (int i = 0; i < 20480; ++i) W[i] = round_fp16 ? to_half(mlp[i]) : mlp[i];
    if (dmlp) memset(dmlp, 0, sizeof(double) * 20480);
    double loss = 0.0;
    double act[7][CH];   /* act[0] = input (48), act[1..5] = hidden outputs, act[6] = output */
    double z[6][CH];
    for (uint32_t s = 0; s < n; ++s) {
        const float* q = queries + 10 * (size_t)s;
        double x[3];
        for (int a = 0; a < 3; ++a) {  /* normalised in float, as the device does */
            float vf = (q[a] - lo3[a]) * inv_extent;
            vf = vf < 0.f ? 0.f : (vf > 1.f ? 1.f : vf);
            x[a] = vf;
        }
        uint32_t idx[CL][8];
        double wc[CL][8];
        for (int l = 0; l < CL; ++l) {
            corners(x[0], x[1], x[2], res[l], (dense >> l) & 1u, idx[l], wc[l]);
            for (int f = 0; f < 2; ++f) {
                double v = 0.0;
                for (int c = 0; c < 8; ++c) {
                    double g = grid[((size_t)l * CT + idx[l][c]) * 2 + f];
                    if (round_fp16) g = to_half(g);
                    v += wc[l][c] * g;
                }
                act[0][2 * l + f] = round_fp16 ? to_half(v) : v;
            }
        }
        double ex[16] = {0};
        ex[0] = q[3]; ex[1] = q[4]; ex[2] = q[5]; ex[3] = 1.0;
        ex[4] = x[0]; ex[5] = x[1]; ex[6] = x[2];
        ex[8] = q[6]; ex[9] = q[7]; ex[10] = q[8]; ex[11] = q[9]; ex[12] = q[9] > 0.0f ? 1.0 : 0.0;
        for (int k = 0; k < 16; ++k) act[0][32 + k] = round_fp16 ? to_half(ex[k]) : ex[k];
        for (int L = 0; L < 6; ++L) {
            for (int o = 0; o < W_OUT[L]; ++o) {
                double acc = 0.0;
                for (int i = 0; i < W_IN[L]; ++i) acc += W[W_OFF[L] + i * W_OUT[L] + o] * act[L][i];
                z[L][o] = acc;
                if (L < 5) {
                    const double h = acc > 0 ? acc : 0.0;
                    act[L + 1][o] = round_fp16 ? to_half(h) : h;
                } else {
                    act[6][o] = acc;
                }
            }
        }
        double dz[6][CH];
        memset(dz, 0, sizeof dz);
        for (int c = 0; c < 3; ++c) {
            const double p = act[6][c], y = targets[3 * (size_t)s + c];
            const double den = plain_l2 ? 1.0 : p * p + 0.01;
            loss += (p - y) * (p - y) / den;
            dz[5][c] = 2.0 * (p - y) / den;
            if (pred) pred[3 * (size_t)s + c] = (float)p;
        }
        for (int L = 5; L >= 0; --L) {
            if (dmlp)
                for (int i = 0; i < W_IN[L]; ++i)
                    for (int o = 0; o < W_OUT[L]; ++o) dmlp[W_OFF[L] + i * W_OUT[L] + o] += act[L][i] * dz[L][o];
            double dx[CH];
            for (int i = 0; i < W_IN[L]; ++i) {
                double acc = 0.0;
                for (int o = 0; o < W_OUT[L]; ++o) acc += W[W_OFF[L] + i * W_OUT[L] + o] * dz[L][o];
                dx[i] = acc;
            }
            if (L > 0) {
                for (int i = 0; i < CH; ++i) dz[L - 1][i] = z[L - 1][i] > 0 ? dx[i] : 0.0;
            } else if (dgrid) {
                for (int l = 0; l < CL; ++l)
                    for (int f = 0; f < 2; ++f)
                        for (int c = 0; c < 8; ++c)
                            dgrid[((size_t)l * CT + idx[l][c]) * 2 + f] += wc[l][c] * dx[2 * l + f];
            }
        }
    }
    free(W);
    return loss;
}

/* One Adam step (Kingma & Ba 2015) in double on (p, m, v) with gradient g;
 * t is the step number after this update (bias corrections 1 - beta^t). */
void orc_adam_step(double* p, double* m, double* v, const double* g, uint64_t n, double lr, double b1, double b2,
                   double eps, uint32_t t)
{
    const double bc1 = 1.0 - pow(b1, (double)t), bc2 = 1.0 - pow(b2, (double)t);
    for (uint64_t i = 0; i < n; ++i) {
        m[i] = b1 * m[i] + (1.0 - b1) * g[i];
        v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
        p[i] -= lr * (m[i] / bc1) / (sqrt(v[i] / bc2) + eps);
    }
}
