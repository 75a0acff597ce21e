/* CPU oracle of the warm-start U-Net (test infrastructure only).
 *
 * The U-Net has no counterpart in the reference: it is pinned by PyTorch's
 * CPU conv3d / avg_pool3d / trilinear interpolate (an independent library)
 * and by finite differences of the loss.  Arithmetic is FP64; with
 * round_bf16 set, every stored activation and gradient tensor is rounded to
 * bfloat16 (nearest even) at the points where the GPU path stores bfloat16,
 * and weights enter the convolutions rounded to bfloat16.
 *
 * Layout: point p = x + nx*(y + ny*z) (x fastest), channels of a point
 * contiguous with pitch ld; weights W[co][t][ci] with tap
 * t = (dx+1) + 3(dy+1) + 9(dz+1); zero padding 1; the coarse level of an
 * n-point axis has (n+1)/2 points (fine = 2*coarse - 1).
 */
#ifndef GM_ORACLE_UNET_H
#define GM_ORACLE_UNET_H

#ifdef __cplusplus
extern "C" {
#endif

double ou_bf16(double x);

/* out[p][co] = bias[co] + sum_t sum_ci in[p+off(t)][ci] W[co][t][ci], then
 * max(0, .) when relu. */
void ou_conv3d(const double *in, int ldi, int nx, int ny, int nz, int cin, const double *w,
               const double *bias, int cout, int relu, double *out, int ldo);
/* Adjoint in the input: gx[p][ci] = sum_t sum_co gy[p-off(t)][co] W[co][t][ci]. */
void ou_conv3d_bwd_data(const double *gy, int ldg, int nx, int ny, int nz, int cout, const double *w,
                        int cin, double *gx, int ldx);
/* gw[co][t][ci] = sum_p gy[p][co] in[p+off(t)][ci]; gb[co] = sum_p gy[p][co]. */
void ou_conv3d_bwd_weight(const double *in, int ldi, const double *gy, int ldg, int nx, int ny, int nz,
                          int cin, int cout, double *gw, double *gb);
/* 3x3x3 average, stride 2, padding 1, divisor 27; nx.. are the fine sizes. */
void ou_pool(const double *in, int ldi, int nx, int ny, int nz, int c, double *out, int ldo);
void ou_pool_bwd(const double *gout, int ldg, int nx, int ny, int nz, int c, double *gin, int ldi);
/* Trilinear x2 with aligned corners; nx.. are the coarse sizes. */
void ou_up(const double *in, int ldi, int nx, int ny, int nz, int c, double *out, int ldo);
void ou_up_bwd(const double *gout, int ldg, int nx, int ny, int nz, int c, double *gin, int ldi);

/* Whole network. */
#define OU_LAYERS 11
long ou_unet_nparams(void);
/* Input features [P][8] of a grid: B0 (AoS, padded (nx+2)(ny+2)(nz+2)
 * points of the fine grid fnx..) sampled at every `stride`-th inner point:
 * (Bx, By, Bz - 1, 2i/(nx-1) - 1, 2j/(ny-1) - 1, 2k/(nz-1) - 1, 0, 0). */
void ou_unet_features(const double *B0pad, int fnx, int fny, int fnz, int stride, int nx, int ny, int nz,
                      double *feat);
/* Forward (y = scale * head output, [P][3]), loss = mean over P and 3 of
 * (head - target/scale)^2, and, when grad != NULL, its gradient in the
 * packed parameters (nparams - 1 entries; the scale is not trained). */
double ou_unet_loss(const double *params, const double *feat, const double *target, int nx, int ny, int nz,
                    int round_bf16, double *y, double *grad);

#ifdef __cplusplus
}
#endif
#endif
