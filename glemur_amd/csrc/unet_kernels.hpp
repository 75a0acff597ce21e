// Device kernels of the warm-start U-Net (bfloat16 activations and
// gradients, FP32 accumulation on the bf16 matrix cores of gfx950).
//
// Layout: point p = x + nx*(y + ny*z), channels of a point contiguous with
// pitch `ld` elements (a multiple of 8, slices start at multiples of 8, so
// every 8-channel group is one 16-byte load).  Weights W[co][k] with
// k = t*cin + ci, tap t = (dx+1) + 3(dy+1) + 9(dz+1), rows padded with zeros
// to a multiple of 16 and k padded with zeros to `kpad` (a multiple of 32).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gmu {

using bf16_t = uint16_t;

// out[p][co] = bias[co] + sum_k in[p + off(t)][ci] W[co][k] (zero padding),
// then optional ReLU, optional mask (value kept where mask[p][co] > 0),
// stored as bf16 (out) or fp32 (outf); channels co >= cout are not stored.
// With W the flipped transpose of a layer's weights this is the layer's
// input gradient.
struct ConvArgs {
    const bf16_t *in = nullptr;
    long long ldi = 0;
    const bf16_t *w = nullptr;
    int kpad = 0;
    const float *bias = nullptr;
    bf16_t *out = nullptr;
    long long ldo = 0;
    float *outf = nullptr;
    long long ldf = 0;
    const bf16_t *mask = nullptr;
    long long ldm = 0;
    int nx = 0, ny = 0, nz = 0;
    int cin = 0;      // multiple of 8
    int coutp = 0;    // rows of W: 16, 32, 64, 96, 128 or 192
    int cout = 0;     // channels stored (<= coutp)
    int relu = 0;
};
void launch_conv(hipStream_t s, const ConvArgs &a);

// Weight gradient partials: part[chunk][co][t][ci] = sum over the chunk's
// points p of g[p][co] in[p + off(t)][ci].
struct WgradArgs {
    const bf16_t *in = nullptr;
    long long ldi = 0;
    const bf16_t *g = nullptr;
    long long ldg = 0;
    int nx = 0, ny = 0, nz = 0;
    int cin = 0, cout = 0;    // cin multiple of 8, cout multiple of 16
    int chunk = 0;            // points per chunk (multiple of 64)
    int nchunks = 0;
    float *part = nullptr;
};
void launch_wgrad(hipStream_t s, const WgradArgs &a);
// grad[co][t*cin+ci] (pitch kpad) (+)= sum over chunks in order.
void launch_wgrad_reduce(hipStream_t s, const float *part, int nchunks, int cout, int cin, int kpad, float *grad,
                         int accumulate);
// bgrad[c] (+)= sum_p g[p][c], through per-chunk partials (fixed order).
void launch_bias_grad(hipStream_t s, const bf16_t *g, long long ldg, long long P, int c, float *part, float *bgrad,
                      int accumulate);
int bias_grad_chunks(long long P);

// 3x3x3 average, stride 2, padding 1, divisor 27 (nx.. fine sizes).
void launch_pool(hipStream_t s, const bf16_t *in, long long ldi, int nx, int ny, int nz, int c, bf16_t *out,
                 long long ldo);
// g[p] = mask(p) * (skip[p] + pool^T(dP)[p]) on the fine grid.
void launch_pool_bwd(hipStream_t s, const bf16_t *dP, long long ldp, int nx, int ny, int nz, int c,
                     const bf16_t *skip, long long lds, const bf16_t *mask, long long ldm, bf16_t *out, long long ldo);
// Trilinear x2 with aligned corners (nx.. coarse sizes).
void launch_up(hipStream_t s, const bf16_t *in, long long ldi, int nx, int ny, int nz, int c, bf16_t *out,
               long long ldo);
// g[q] = mask(q) * up^T(dF)[q] on the coarse grid.
void launch_up_bwd(hipStream_t s, const bf16_t *dF, long long ldf, int nx, int ny, int nz, int c,
                   const bf16_t *mask, long long ldm, bf16_t *out, long long ldo);

// Loss of one sample: d = y[p][c] - target[p][c] * inv_scale (c < 3);
// g[p][c] = bf16(2 d / (3P)) (0 for c >= 3, 16 channels); loss_acc += sum d^2 / (3P)
// in a fixed order.
void launch_loss(hipStream_t s, const float *y, const float *target, float inv_scale, long long P, bf16_t *g,
                 double *part, double *loss_acc);
int loss_blocks(long long P);

// Adam (Kingma & Ba 2015) on n master weights: m = b1 m + (1-b1) g',
// v = b2 v + (1-b2) g'^2, w -= lr mhat / (sqrt(vhat) + eps), with
// g' = g * gscale; w16 = bf16(w) when w16 != null.
void launch_adam(hipStream_t s, float *w, const float *g, float *m, float *v, long long n, float gscale, float lr,
                 float b1, float b2, float eps, int step, bf16_t *w16);
// wT[ci][t'*coutp + co] = w16[co][(26 - t')*cin + ci] (pitch kpadT).
void launch_flip(hipStream_t s, const bf16_t *w16, int coutp, int cin, int kpad, bf16_t *wT, int kpadT);
void launch_f32_to_bf16(hipStream_t s, const float *in, long long n, bf16_t *out);
void launch_zero(hipStream_t s, void *p, long long bytes);

// Features of a relaxation context's B0 (SoA, padded (nx+2)(ny+2)(nz+2),
// FP64 or FP32) at every `stride`-th inner point of the fine grid:
// (Bx, By, Bz - 1, 2i/(n-1) - 1 per axis, 0, 0) as bf16 [P][8].
void launch_features(hipStream_t s, const double *B0_64, const float *B0_32, long long Pf, int fsx, long long fsxy,
                     int stride, int nx, int ny, int nz, bf16_t *feat);
// Target of a relaxed context: positions minus the uniform grid
// (O + i h per axis) at the coarse points, fp32 [P][3].
void launch_target(hipStream_t s, const double *X64, const float *Xh, const float *Xl, long long Pf, int fsx,
                   long long fsxy, int nx, int ny, int nz, double Ox, double Oy, double Oz, double dx, double dy,
                   double dz, float *target);
// Warm start: positions of the fine grid (inner points) += scale * up(y)
// where y is the network output [Pc][16] (channels 0..2) of the coarse grid
// (stride 2: fine = 2 coarse - 1; stride 1: same grid).
void launch_apply(hipStream_t s, const float *y, int cnx, int cny, int cnz, int stride, double scale, double *X64,
                  float *Xh, float *Xl, long long Pf, int fsx, long long fsxy);
// out[0] = sum of v[i]^2 (fixed order; part holds 512 doubles).
void launch_sumsq(hipStream_t s, const float *v, long long n, double *part, double *out);
// min over inner points of det (FP64, padded layout) and a NaN flag:
// out[0] = min det, out[1] = 1 if any det is NaN.
void launch_min_det(hipStream_t s, const double *det, int nx, int ny, int nz, int sx, long long sxy, double *part,
                    double *out);   // part holds 1024 doubles

}  // namespace gmu
