/* CPU oracle of the warm-start U-Net: FP64 restatement of the network the
 * GPU path runs (test infrastructure; the product never calls it).
 * No reference counterpart (the reference has no learned warm start); the
 * operators are pinned against PyTorch's CPU conv3d, avg_pool3d and
 * trilinear interpolation and the gradient against finite differences. */
#include "unet.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const int LCIN[OU_LAYERS] = {8, 32, 32, 64, 64, 128, 192, 64, 96, 32, 32};
static const int LCOUT[OU_LAYERS] = {32, 32, 64, 64, 128, 128, 64, 64, 32, 32, 3};

double ou_bf16(double x)
{
    float f = (float)x;
    unsigned u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;   /* inf / nan unchanged */
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return f;
}

static inline int inb(int v, int n) { return v >= 0 && v < n; }

void ou_conv3d(const double *in, int ldi, int nx, int ny, int nz, int cin, const double *w,
               const double *bias, int cout, int relu, double *out, int ldo)
{
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                long p = x + (long)nx * (y + (long)ny * z);
                for (int co = 0; co < cout; co++) {
                    double s = bias ? bias[co] : 0.0;
                    for (int t = 0; t < 27; t++) {
                        int qx = x + t % 3 - 1, qy = y + (t / 3) % 3 - 1, qz = z + t / 9 - 1;
                        if (!inb(qx, nx) || !inb(qy, ny) || !inb(qz, nz)) continue;
                        const double *a = in + (qx + (long)nx * (qy + (long)ny * qz)) * ldi;
                        const double *wt = w + ((long)co * 27 + t) * cin;
                        for (int ci = 0; ci < cin; ci++) s += a[ci] * wt[ci];
                    }
                    out[p * ldo + co] = relu && s < 0 ? 0.0 : s;
                }
            }
}

void ou_conv3d_bwd_data(const double *gy, int ldg, int nx, int ny, int nz, int cout, const double *w,
                        int cin, double *gx, int ldx)
{
    /* gather form of the adjoint: the output point q = p - off(t) used
     * input p through tap t */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                long p = x + (long)nx * (y + (long)ny * z);
                for (int ci = 0; ci < cin; ci++) gx[p * ldx + ci] = 0.0;
                for (int t = 0; t < 27; t++) {
                    int qx = x - (t % 3 - 1), qy = y - ((t / 3) % 3 - 1), qz = z - (t / 9 - 1);
                    if (!inb(qx, nx) || !inb(qy, ny) || !inb(qz, nz)) continue;
                    const double *g = gy + (qx + (long)nx * (qy + (long)ny * qz)) * ldg;
                    for (int co = 0; co < cout; co++) {
                        const double *wt = w + ((long)co * 27 + t) * cin;
                        for (int ci = 0; ci < cin; ci++) gx[p * ldx + ci] += g[co] * wt[ci];
                    }
                }
            }
}

void ou_conv3d_bwd_weight(const double *in, int ldi, const double *gy, int ldg, int nx, int ny, int nz,
                          int cin, int cout, double *gw, double *gb)
{
    const long P = (long)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (int co = 0; co < cout; co++) {
        double *row = gw + (long)co * 27 * cin;
        for (long i = 0; i < 27L * cin; i++) row[i] = 0.0;
        double sb = 0.0;
        for (long p = 0; p < P; p++) {
            int x = (int)(p % nx), y = (int)((p / nx) % ny), z = (int)(p / ((long)nx * ny));
            double g = gy[p * ldg + co];
            sb += g;
            if (g == 0.0) continue;
            for (int t = 0; t < 27; t++) {
                int qx = x + t % 3 - 1, qy = y + (t / 3) % 3 - 1, qz = z + t / 9 - 1;
                if (!inb(qx, nx) || !inb(qy, ny) || !inb(qz, nz)) continue;
                const double *a = in + (qx + (long)nx * (qy + (long)ny * qz)) * ldi;
                for (int ci = 0; ci < cin; ci++) row[t * cin + ci] += g * a[ci];
            }
        }
        if (gb) gb[co] = sb;
    }
}

void ou_pool(const double *in, int ldi, int nx, int ny, int nz, int c, double *out, int ldo)
{
    const int cx = (nx + 1) / 2, cy = (ny + 1) / 2, cz = (nz + 1) / 2;
    for (int Z = 0; Z < cz; Z++)
        for (int Y = 0; Y < cy; Y++)
            for (int X = 0; X < cx; X++) {
                long q = X + (long)cx * (Y + (long)cy * Z);
                for (int ch = 0; ch < c; ch++) {
                    double s = 0;
                    for (int t = 0; t < 27; t++) {
                        int x = 2 * X + t % 3 - 1, y = 2 * Y + (t / 3) % 3 - 1, z = 2 * Z + t / 9 - 1;
                        if (inb(x, nx) && inb(y, ny) && inb(z, nz))
                            s += in[(x + (long)nx * (y + (long)ny * z)) * ldi + ch];
                    }
                    out[q * ldo + ch] = s / 27.0;
                }
            }
}

void ou_pool_bwd(const double *gout, int ldg, int nx, int ny, int nz, int c, double *gin, int ldi)
{
    const int cx = (nx + 1) / 2, cy = (ny + 1) / 2, cz = (nz + 1) / 2;
    const long P = (long)nx * ny * nz;
    for (long p = 0; p < P; p++)
        for (int ch = 0; ch < c; ch++) gin[p * ldi + ch] = 0.0;
    /* scatter form: each coarse value spreads to its 27 window points */
    for (int Z = 0; Z < cz; Z++)
        for (int Y = 0; Y < cy; Y++)
            for (int X = 0; X < cx; X++) {
                const double *g = gout + (X + (long)cx * (Y + (long)cy * Z)) * ldg;
                for (int t = 0; t < 27; t++) {
                    int x = 2 * X + t % 3 - 1, y = 2 * Y + (t / 3) % 3 - 1, z = 2 * Z + t / 9 - 1;
                    if (!(inb(x, nx) && inb(y, ny) && inb(z, nz))) continue;
                    double *d = gin + (x + (long)nx * (y + (long)ny * z)) * ldi;
                    for (int ch = 0; ch < c; ch++) d[ch] += g[ch] / 27.0;
                }
            }
}

/* weights of fine index f on coarse indices: f even -> f/2 (1); f odd ->
 * (f-1)/2 and (f+1)/2 (1/2 each) */
static void lin1(int f, int *i0, int *i1, double *w0, double *w1)
{
    if (f % 2 == 0) { *i0 = *i1 = f / 2; *w0 = 1.0; *w1 = 0.0; }
    else { *i0 = (f - 1) / 2; *i1 = *i0 + 1; *w0 = 0.5; *w1 = 0.5; }
}

void ou_up(const double *in, int ldi, int nx, int ny, int nz, int c, double *out, int ldo)
{
    const int fx = 2 * nx - 1, fy = 2 * ny - 1, fz = 2 * nz - 1;
    for (int z = 0; z < fz; z++)
        for (int y = 0; y < fy; y++)
            for (int x = 0; x < fx; x++) {
                int ix[2], iy[2], iz[2];
                double wx[2], wy[2], wz[2];
                lin1(x, &ix[0], &ix[1], &wx[0], &wx[1]);
                lin1(y, &iy[0], &iy[1], &wy[0], &wy[1]);
                lin1(z, &iz[0], &iz[1], &wz[0], &wz[1]);
                double *o = out + (x + (long)fx * (y + (long)fy * z)) * ldo;
                for (int ch = 0; ch < c; ch++) o[ch] = 0.0;
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++)
                        for (int e = 0; e < 2; e++) {
                            double w = wx[e] * wy[b] * wz[a];
                            if (w == 0.0) continue;
                            const double *s = in + (ix[e] + (long)nx * (iy[b] + (long)ny * iz[a])) * ldi;
                            for (int ch = 0; ch < c; ch++) o[ch] += w * s[ch];
                        }
            }
}

void ou_up_bwd(const double *gout, int ldg, int nx, int ny, int nz, int c, double *gin, int ldi)
{
    const int fx = 2 * nx - 1, fy = 2 * ny - 1, fz = 2 * nz - 1;
    const long P = (long)nx * ny * nz;
    for (long p = 0; p < P; p++)
        for (int ch = 0; ch < c; ch++) gin[p * ldi + ch] = 0.0;
    /* scatter form of the transpose of ou_up */
    for (int z = 0; z < fz; z++)
        for (int y = 0; y < fy; y++)
            for (int x = 0; x < fx; x++) {
                int ix[2], iy[2], iz[2];
                double wx[2], wy[2], wz[2];
                lin1(x, &ix[0], &ix[1], &wx[0], &wx[1]);
                lin1(y, &iy[0], &iy[1], &wy[0], &wy[1]);
                lin1(z, &iz[0], &iz[1], &wz[0], &wz[1]);
                const double *g = gout + (x + (long)fx * (y + (long)fy * z)) * ldg;
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++)
                        for (int e = 0; e < 2; e++) {
                            double w = wx[e] * wy[b] * wz[a];
                            if (w == 0.0) continue;
                            double *d = gin + (ix[e] + (long)nx * (iy[b] + (long)ny * iz[a])) * ldi;
                            for (int ch = 0; ch < c; ch++) d[ch] += w * g[ch];
                        }
            }
}

long ou_unet_nparams(void)
{
    long n = 0;
    for (int l = 0; l < OU_LAYERS; l++) n += (long)LCOUT[l] * 27 * LCIN[l] + LCOUT[l];
    return n + 1;   /* + output scale */
}

void ou_unet_features(const double *B0pad, int fnx, int fny, int fnz, int stride, int nx, int ny, int nz,
                      double *feat)
{
    const long sx = fnx + 2, sxy = (long)(fnx + 2) * (fny + 2);
    (void)fnz;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                long g = (stride * i + 1) + sx * (stride * j + 1) + sxy * (stride * k + 1);
                double *f = feat + 8 * (i + (long)nx * (j + (long)ny * k));
                f[0] = B0pad[3 * g];
                f[1] = B0pad[3 * g + 1];
                f[2] = B0pad[3 * g + 2] - 1.0;
                f[3] = 2.0 * i / (nx - 1) - 1.0;
                f[4] = 2.0 * j / (ny - 1) - 1.0;
                f[5] = 2.0 * k / (nz - 1) - 1.0;
                f[6] = f[7] = 0.0;
            }
}

static void rnd(double *a, long n, int on)
{
    if (!on) return;
    for (long i = 0; i < n; i++) a[i] = ou_bf16(a[i]);
}

/* channels [0, c) of n points with pitch ld */
static void rnd_slice(double *a, long n, int ld, int c, int on)
{
    if (!on) return;
    for (long i = 0; i < n; i++)
        for (int ch = 0; ch < c; ch++) a[i * ld + ch] = ou_bf16(a[i * ld + ch]);
}

double ou_unet_loss(const double *params, const double *feat, const double *target, int nx, int ny, int nz,
                    int round_bf16, double *y, double *grad)
{
    const int R = round_bf16;
    const int n2x = (nx + 1) / 2, n2y = (ny + 1) / 2, n2z = (nz + 1) / 2;
    const int n3x = (n2x + 1) / 2, n3y = (n2y + 1) / 2, n3z = (n2z + 1) / 2;
    const long P1 = (long)nx * ny * nz, P2 = (long)n2x * n2y * n2z, P3 = (long)n3x * n3y * n3z;

    /* parameters: W rounded to bf16 when R, biases as given */
    const double *W[OU_LAYERS], *Bv[OU_LAYERS];
    double *Wr[OU_LAYERS];
    long off[OU_LAYERS];
    long o = 0;
    for (int l = 0; l < OU_LAYERS; l++) {
        long nw = (long)LCOUT[l] * 27 * LCIN[l];
        off[l] = o;
        W[l] = params + o;
        Bv[l] = params + o + nw;
        Wr[l] = (double *)malloc(nw * sizeof(double));
        memcpy(Wr[l], W[l], nw * sizeof(double));
        rnd(Wr[l], nw, R);
        o += nw + LCOUT[l];
    }
    const double scale = params[o];

#define ALLOC(n) ((double *)calloc((size_t)(n), sizeof(double)))
    double *x0 = ALLOC(P1 * 8), *a1 = ALLOC(P1 * 32), *cat1 = ALLOC(P1 * 96), *p2 = ALLOC(P2 * 32);
    double *a3 = ALLOC(P2 * 64), *cat2 = ALLOC(P2 * 192), *p3 = ALLOC(P3 * 64), *a5 = ALLOC(P3 * 128);
    double *a6 = ALLOC(P3 * 128), *a7 = ALLOC(P2 * 64), *a8 = ALLOC(P2 * 64), *a9 = ALLOC(P1 * 32);
    double *a10 = ALLOC(P1 * 32), *h = ALLOC(P1 * 3);
    memcpy(x0, feat, P1 * 8 * sizeof(double));
    rnd(x0, P1 * 8, R);
    ou_conv3d(x0, 8, nx, ny, nz, 8, Wr[0], Bv[0], 32, 1, a1, 32);
    rnd(a1, P1 * 32, R);
    ou_conv3d(a1, 32, nx, ny, nz, 32, Wr[1], Bv[1], 32, 1, cat1 + 64, 96);   /* a2 */
    rnd_slice(cat1 + 64, P1, 96, 32, R);
    ou_pool(cat1 + 64, 96, nx, ny, nz, 32, p2, 32);
    rnd(p2, P2 * 32, R);
    ou_conv3d(p2, 32, n2x, n2y, n2z, 32, Wr[2], Bv[2], 64, 1, a3, 64);
    rnd(a3, P2 * 64, R);
    ou_conv3d(a3, 64, n2x, n2y, n2z, 64, Wr[3], Bv[3], 64, 1, cat2 + 128, 192);   /* a4 */
    rnd_slice(cat2 + 128, P2, 192, 64, R);
    ou_pool(cat2 + 128, 192, n2x, n2y, n2z, 64, p3, 64);
    rnd(p3, P3 * 64, R);
    ou_conv3d(p3, 64, n3x, n3y, n3z, 64, Wr[4], Bv[4], 128, 1, a5, 128);
    rnd(a5, P3 * 128, R);
    ou_conv3d(a5, 128, n3x, n3y, n3z, 128, Wr[5], Bv[5], 128, 1, a6, 128);
    rnd(a6, P3 * 128, R);
    ou_up(a6, 128, n3x, n3y, n3z, 128, cat2, 192);   /* u2 */
    rnd_slice(cat2, P2, 192, 128, R);
    ou_conv3d(cat2, 192, n2x, n2y, n2z, 192, Wr[6], Bv[6], 64, 1, a7, 64);
    rnd(a7, P2 * 64, R);
    ou_conv3d(a7, 64, n2x, n2y, n2z, 64, Wr[7], Bv[7], 64, 1, a8, 64);
    rnd(a8, P2 * 64, R);
    ou_up(a8, 64, n2x, n2y, n2z, 64, cat1, 96);      /* u1 */
    rnd_slice(cat1, P1, 96, 64, R);
    ou_conv3d(cat1, 96, nx, ny, nz, 96, Wr[8], Bv[8], 32, 1, a9, 32);
    rnd(a9, P1 * 32, R);
    ou_conv3d(a9, 32, nx, ny, nz, 32, Wr[9], Bv[9], 32, 1, a10, 32);
    rnd(a10, P1 * 32, R);
    ou_conv3d(a10, 32, nx, ny, nz, 32, Wr[10], Bv[10], 3, 0, h, 3);

    double loss = 0;
    for (long i = 0; i < P1 * 3; i++) {
        double d = h[i] - target[i] / scale;
        loss += d * d;
    }
    loss /= (double)(P1 * 3);
    if (y)
        for (long i = 0; i < P1 * 3; i++) y[i] = scale * h[i];

    if (grad) {
        double *g11 = ALLOC(P1 * 3), *g10 = ALLOC(P1 * 32), *g9 = ALLOC(P1 * 32), *dcat1 = ALLOC(P1 * 96);
        double *g8 = ALLOC(P2 * 64), *g7 = ALLOC(P2 * 64), *dcat2 = ALLOC(P2 * 192), *g6 = ALLOC(P3 * 128);
        double *g5 = ALLOC(P3 * 128), *dp3 = ALLOC(P3 * 64), *g4 = ALLOC(P2 * 64), *g3 = ALLOC(P2 * 64);
        double *dp2 = ALLOC(P2 * 32), *g2 = ALLOC(P1 * 32), *g1 = ALLOC(P1 * 32);
        for (long i = 0; i < P1 * 3; i++) g11[i] = 2.0 * (h[i] - target[i] / scale) / (double)(P1 * 3);
        rnd(g11, P1 * 3, R);
#define WG(l, in, ldi, gy, ldg, sx, sy, sz) \
    ou_conv3d_bwd_weight(in, ldi, gy, ldg, sx, sy, sz, LCIN[l], LCOUT[l], grad + off[l], \
                         grad + off[l] + (long)LCOUT[l] * 27 * LCIN[l])
#define MASK(g, a, lda, n, c)                                      \
    for (long i_ = 0; i_ < (n); i_++)                              \
        for (int c_ = 0; c_ < (c); c_++)                           \
            if (!((a)[i_ * (lda) + c_] > 0)) (g)[i_ * (c) + c_] = 0.0;
        WG(10, a10, 32, g11, 3, nx, ny, nz);
        ou_conv3d_bwd_data(g11, 3, nx, ny, nz, 3, Wr[10], 32, g10, 32);
        MASK(g10, a10, 32, P1, 32);
        rnd(g10, P1 * 32, R);
        WG(9, a9, 32, g10, 32, nx, ny, nz);
        ou_conv3d_bwd_data(g10, 32, nx, ny, nz, 32, Wr[9], 32, g9, 32);
        MASK(g9, a9, 32, P1, 32);
        rnd(g9, P1 * 32, R);
        WG(8, cat1, 96, g9, 32, nx, ny, nz);
        ou_conv3d_bwd_data(g9, 32, nx, ny, nz, 32, Wr[8], 96, dcat1, 96);
        rnd(dcat1, P1 * 96, R);
        ou_up_bwd(dcat1, 96, n2x, n2y, n2z, 64, g8, 64);
        MASK(g8, a8, 64, P2, 64);
        rnd(g8, P2 * 64, R);
        WG(7, a7, 64, g8, 64, n2x, n2y, n2z);
        ou_conv3d_bwd_data(g8, 64, n2x, n2y, n2z, 64, Wr[7], 64, g7, 64);
        MASK(g7, a7, 64, P2, 64);
        rnd(g7, P2 * 64, R);
        WG(6, cat2, 192, g7, 64, n2x, n2y, n2z);
        ou_conv3d_bwd_data(g7, 64, n2x, n2y, n2z, 64, Wr[6], 192, dcat2, 192);
        rnd(dcat2, P2 * 192, R);
        ou_up_bwd(dcat2, 192, n3x, n3y, n3z, 128, g6, 128);
        MASK(g6, a6, 128, P3, 128);
        rnd(g6, P3 * 128, R);
        WG(5, a5, 128, g6, 128, n3x, n3y, n3z);
        ou_conv3d_bwd_data(g6, 128, n3x, n3y, n3z, 128, Wr[5], 128, g5, 128);
        MASK(g5, a5, 128, P3, 128);
        rnd(g5, P3 * 128, R);
        WG(4, p3, 64, g5, 128, n3x, n3y, n3z);
        ou_conv3d_bwd_data(g5, 128, n3x, n3y, n3z, 128, Wr[4], 64, dp3, 64);
        rnd(dp3, P3 * 64, R);
        ou_pool_bwd(dp3, 64, n2x, n2y, n2z, 64, g4, 64);
        for (long i = 0; i < P2; i++)
            for (int c = 0; c < 64; c++) g4[i * 64 + c] += dcat2[i * 192 + 128 + c];
        MASK(g4, cat2 + 128, 192, P2, 64);
        rnd(g4, P2 * 64, R);
        WG(3, a3, 64, g4, 64, n2x, n2y, n2z);
        ou_conv3d_bwd_data(g4, 64, n2x, n2y, n2z, 64, Wr[3], 64, g3, 64);
        MASK(g3, a3, 64, P2, 64);
        rnd(g3, P2 * 64, R);
        WG(2, p2, 32, g3, 64, n2x, n2y, n2z);
        ou_conv3d_bwd_data(g3, 64, n2x, n2y, n2z, 64, Wr[2], 32, dp2, 32);
        rnd(dp2, P2 * 32, R);
        ou_pool_bwd(dp2, 32, nx, ny, nz, 32, g2, 32);
        for (long i = 0; i < P1; i++)
            for (int c = 0; c < 32; c++) g2[i * 32 + c] += dcat1[i * 96 + 64 + c];
        MASK(g2, cat1 + 64, 96, P1, 32);
        rnd(g2, P1 * 32, R);
        WG(1, a1, 32, g2, 32, nx, ny, nz);
        ou_conv3d_bwd_data(g2, 32, nx, ny, nz, 32, Wr[1], 32, g1, 32);
        MASK(g1, a1, 32, P1, 32);
        rnd(g1, P1 * 32, R);
        WG(0, x0, 8, g1, 32, nx, ny, nz);
#undef WG
#undef MASK
        free(g11); free(g10); free(g9); free(dcat1); free(g8); free(g7); free(dcat2); free(g6);
        free(g5); free(dp3); free(g4); free(g3); free(dp2); free(g2); free(g1);
    }
#undef ALLOC
    free(x0); free(a1); free(cat1); free(p2); free(a3); free(cat2); free(p3); free(a5);
    free(a6); free(a7); free(a8); free(a9); free(a10); free(h);
    for (int l = 0; l < OU_LAYERS; l++) free(Wr[l]);
    return loss;
}
