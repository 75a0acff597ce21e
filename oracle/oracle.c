/*
 * GLEMuR CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Each function restates one reference kernel or host routine in plain C on
 * the reference's AoS padded layout, point by point, in the reference's
 * operation order.  Citations give the reference file:line followed.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.141592653589793

/* padded point index (i,j,k in 0..n+1) */
static inline long pid(const or_params *p, long i, long j, long k)
{
    return i + (long)(p->nx + 2) * (j + (long)(p->ny + 2) * k);
}
/* inner point index (i,j,k in 0..n-1) */
static inline long iid(const or_params *p, long i, long j, long k)
{
    return i + (long)p->nx * (j + (long)p->ny * k);
}
static inline long npad(const or_params *p)
{
    return (long)(p->nx + 2) * (p->ny + 2) * (p->nz + 2);
}
static inline long ninner(const or_params *p)
{
    return (long)p->nx * p->ny * p->nz;
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* src/rungeKutta.cu:10-25 */
void or_assign_coefficients(double a[6], double b[42], double c[6], double cs[6])
{
    static const double A[6] = {0, 0.2, 0.3, 0.6, 1, 0.875};
    const double B[7][6] = {
        {0, 0, 0, 0, 0, 0},
        {0.2, 0, 0, 0, 0, 0},
        {3 / 40., 9 / 40., 0, 0, 0, 0},
        {0.3, -0.9, 1.2, 0, 0, 0},
        {-11 / 54., 2.5, -70 / 27., 35 / 27., 0, 0},
        {1631 / 55296., 175 / 512., 575 / 13824., 44275 / 110592., 253 / 4096., 0},
        {0, 0, 0, 0, 0, 0}};
    const double C[6] = {37 / 378., 0, 250 / 621., 125 / 594., 0, 512 / 1771.};
    const double CS[6] = {2825 / 27648., 0, 18575 / 48384., 13525 / 55296., 277 / 14336., 0.25};
    memcpy(a, A, sizeof(A));
    memcpy(b, B, sizeof(B));
    memcpy(c, C, sizeof(C));
    memcpy(cs, CS, sizeof(CS));
}

static int starts(const char *s, const char *w)
{
    return strncmp(s, w, strlen(w)) == 0;
}

/* src/init.cu:13-26 */
double or_init_residual(const or_params *p)
{
    double r = 0;
    if (starts(p->bInit, "Pontin09 ") || starts(p->bInit, "analytic ") ||
        starts(p->bInit, "homZ ") || starts(p->bInit, "blobs "))
        r = 1 * p->Lx * p->Ly * p->Lz * p->ampl;
    if (starts(p->bInit, "Borromean "))
        r = 1 * p->Lx * p->Ly * p->Lz * p->bGround;
    return r;
}

/* Periodic copies on the host, src/boundaries.cu:492-528. */
static void set_peri_planes(const or_params *p, double *f)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    if (p->zPeri)
        for (int j = 0; j < ny + 2; j++)
            for (int i = 0; i < nx + 2; i++)
                for (int l = 0; l < 3; l++) {
                    f[l + 3 * pid(p, i, j, 0)] = f[l + 3 * pid(p, i, j, nz)];
                    f[l + 3 * pid(p, i, j, nz + 1)] = f[l + 3 * pid(p, i, j, 1)];
                }
    if (p->yPeri)
        for (int k = 0; k < nz + 2; k++)
            for (int i = 0; i < nx + 2; i++)
                for (int l = 0; l < 3; l++) {
                    f[l + 3 * pid(p, i, 0, k)] = f[l + 3 * pid(p, i, ny, k)];
                    f[l + 3 * pid(p, i, ny + 1, k)] = f[l + 3 * pid(p, i, 1, k)];
                }
    if (p->xPeri)
        for (int k = 0; k < nz + 2; k++)
            for (int j = 0; j < ny + 2; j++)
                for (int l = 0; l < 3; l++) {
                    f[l + 3 * pid(p, 0, j, k)] = f[l + 3 * pid(p, nx, j, k)];
                    f[l + 3 * pid(p, nx + 1, j, k)] = f[l + 3 * pid(p, 1, j, k)];
                }
}

void or_set_peri(const or_params *p, double *field)
{
    set_peri_planes(p, field);
}

/* Periodic grid copies with the +-(L+h) shift, src/boundaries.cu:532-582.
 * The host twin (src/boundaries.cu:586-634) applies the shift outside its
 * loops (brace scoping), so at initialisation the reference leaves the
 * periodic coordinate of the ghost planes unshifted; this restatement applies
 * the shift as the device kernel does, at initialisation too. */
void or_set_grid_peri(const or_params *p, double *xb)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    if (p->zPeri)
        for (int j = 0; j < ny + 2; j++)
            for (int i = 0; i < nx + 2; i++) {
                for (int l = 0; l < 2; l++) {
                    xb[l + 3 * pid(p, i, j, 0)] = xb[l + 3 * pid(p, i, j, nz)];
                    xb[l + 3 * pid(p, i, j, nz + 1)] = xb[l + 3 * pid(p, i, j, 1)];
                }
                xb[2 + 3 * pid(p, i, j, 0)] = xb[2 + 3 * pid(p, i, j, nz)] - p->Lz - p->dz;
                xb[2 + 3 * pid(p, i, j, nz + 1)] = xb[2 + 3 * pid(p, i, j, 1)] + p->Lz + p->dz;
            }
    if (p->yPeri)
        for (int k = 0; k < nz + 2; k++)
            for (int i = 0; i < nx + 2; i++) {
                for (int l = 0; l < 3; l += 2) {
                    xb[l + 3 * pid(p, i, 0, k)] = xb[l + 3 * pid(p, i, ny, k)];
                    xb[l + 3 * pid(p, i, ny + 1, k)] = xb[l + 3 * pid(p, i, 1, k)];
                }
                xb[1 + 3 * pid(p, i, 0, k)] = xb[1 + 3 * pid(p, i, ny, k)] - p->Ly - p->dy;
                xb[1 + 3 * pid(p, i, ny + 1, k)] = xb[1 + 3 * pid(p, i, 1, k)] + p->Ly + p->dy;
            }
    if (p->xPeri)
        for (int k = 0; k < nz + 2; k++)
            for (int j = 0; j < ny + 2; j++) {
                for (int l = 1; l < 3; l++) {
                    xb[l + 3 * pid(p, 0, j, k)] = xb[l + 3 * pid(p, nx, j, k)];
                    xb[l + 3 * pid(p, nx + 1, j, k)] = xb[l + 3 * pid(p, 1, j, k)];
                }
                xb[0 + 3 * pid(p, 0, j, k)] = xb[0 + 3 * pid(p, nx, j, k)] - p->Lx - p->dx;
                xb[0 + 3 * pid(p, nx + 1, j, k)] = xb[0 + 3 * pid(p, 1, j, k)] + p->Lx + p->dx;
            }
}

/* src/init.cu:30-226 (the closed-form fields the configurations use) and
 * src/init.cu:388-419 (initDistortion), then the periodic host copies
 * (src/main.cu:166-169). */
int or_init_state(const or_params *p, double *B0, double *xb)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const char *bi = p->bInit;
    int known = starts(bi, "Pontin09 ") || starts(bi, "analytic ") || starts(bi, "analyticX ") ||
                starts(bi, "analyticY ") || starts(bi, "homZ ") || starts(bi, "blobs ");
    if (!known) return -1;
    for (int k = 0; k < nz + 2; k++) {
        double z = p->dz * (k - 1) + p->Oz;
        for (int j = 0; j < ny + 2; j++) {
            double y = p->dy * (j - 1) + p->Oy;
            for (int i = 0; i < nx + 2; i++) {
                double x = p->dx * (i - 1) + p->Ox;
                double *b = B0 + 3 * pid(p, i, j, k);
                double *g = xb + 3 * pid(p, i, j, k);
                if (starts(bi, "Pontin09 ")) {
                    double e1 = exp(-(x * x + y * y) / (p->ar * p->ar) - (z - p->L1) * (z - p->L1) / (p->az * p->az));
                    double e2 = exp(-(x * x + y * y) / (p->ar * p->ar) - (z - p->L2) * (z - p->L2) / (p->az * p->az));
                    b[0] = -2 * p->ampl / (OR_PI * p->ar) * y * (p->phi1 * e1 + p->phi2 * e2);
                    b[1] = 2 * p->ampl / (OR_PI * p->ar) * x * (p->phi1 * e1 + p->phi2 * e2);
                    b[2] = p->ampl;
                }
                if (starts(bi, "analytic ")) {
                    double e = exp(-(x * x + y * y) / (p->ar * p->ar) - z * z / (p->az * p->az));
                    b[0] = 2 * p->ampl * z * e * p->phi1 / (p->az * p->az) * y;
                    b[1] = -2 * p->ampl * z * e * p->phi1 / (p->az * p->az) * x;
                    b[2] = p->ampl;
                }
                if (starts(bi, "analyticX ")) {
                    double e = exp(-(z * z + y * y) / (p->ar * p->ar) - x * x / (p->az * p->az));
                    b[0] = p->ampl;
                    b[1] = -2 * p->ampl * x * e * p->phi1 / (p->az * p->az) * z;
                    b[2] = 2 * p->ampl * x * e * p->phi1 / (p->az * p->az) * y;
                }
                if (starts(bi, "analyticY ")) {
                    double e = exp(-(x * x + z * z) / (p->ar * p->ar) - y * y / (p->az * p->az));
                    b[0] = 2 * p->ampl * y * e * p->phi1 / (p->az * p->az) * z;
                    b[1] = p->ampl;
                    b[2] = -2 * p->ampl * y * e * p->phi1 / (p->az * p->az) * x;
                }
                if (starts(bi, "homZ ")) {
                    b[0] = 0.;
                    b[1] = 0.;
                    b[2] = p->ampl;
                }
                if (starts(bi, "blobs ")) {
                    b[0] = 0.;
                    b[1] = 0.;
                    b[2] = 0.;
                    for (int m = 0; m < p->nBlobs; m++) {
                        double dxc = x - p->blobXc[m], dyc = y - p->blobYc[m], dzc = z - p->blobZc[m];
                        double e = exp((-dxc * dxc - dyc * dyc) / (p->blobScale[m] * p->blobScale[m]) -
                                       dzc * dzc / (p->blobZl[m] * p->blobZl[m]));
                        double amp = 2 * p->ampl * p->blobTwist[m] / p->blobScale[m];
                        b[0] += amp * e * (-dyc);
                        b[1] += amp * e * dxc;
                    }
                    b[2] = p->ampl;
                }
                g[0] = x;
                g[1] = y;
                g[2] = z;
            }
        }
    }
    /* initDistortion, src/init.cu:388-419 */
    if (!starts(p->initDist, "none ")) {
        for (int k = 0; k < nz + 2; k++) {
            double z = p->dz * (k - 1) + p->Oz;
            for (int j = 0; j < ny + 2; j++) {
                double y = p->dy * (j - 1) + p->Oy;
                for (int i = 0; i < nx + 2; i++) {
                    double x = p->dx * (i - 1) + p->Ox;
                    double *g = xb + 3 * pid(p, i, j, k);
                    if (starts(p->initDist, "initShearX ")) {
                        g[1] -= p->initShearA * sin(p->initShearK * 2 * OR_PI * (x + p->Ox - p->dx / 2) / (p->Lx + p->dx)) * z;
                        double yy = g[1];
                        g[0] -= p->initShearB * sin(p->initShearK * 2 * OR_PI * (yy + p->Oy - p->dy / 2) / (p->Ly + p->dy)) * z;
                    }
                    if (starts(p->initDist, "centerShift ")) {
                        g[1] += p->initShearA * exp(-p->initShearK * x * x) * (1 - y * y) * (exp(-1 * z + 1) - 5.0804508111226507);
                    }
                }
            }
        }
    }
    set_peri_planes(p, B0);
    or_set_grid_peri(p, xb);
    return 0;
}

/* src/rungeKutta.cu:28-108 */
void or_b_jacb0(const or_params *p, const double *xb, const double *B0, double *B, double *detJac)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                const double *xm = xb + 3 * pid(p, i - 1, j, k), *xp = xb + 3 * pid(p, i + 1, j, k);
                const double *ym = xb + 3 * pid(p, i, j - 1, k), *yp = xb + 3 * pid(p, i, j + 1, k);
                const double *zm = xb + 3 * pid(p, i, j, k - 1), *zp = xb + 3 * pid(p, i, j, k + 1);
                double jac[3][3];
                for (int l = 0; l < 3; l++) {
                    jac[l][0] = (xp[l] - xm[l]) * p->dx1 / 2;
                    jac[l][1] = (yp[l] - ym[l]) * p->dy1 / 2;
                    jac[l][2] = (zp[l] - zm[l]) * p->dz1 / 2;
                }
                double detJac1 = 1. / (jac[0][0] * jac[1][1] * jac[2][2] + jac[0][1] * jac[1][2] * jac[2][0] +
                                       jac[0][2] * jac[1][0] * jac[2][1] - jac[0][0] * jac[1][2] * jac[2][1] -
                                       jac[0][1] * jac[1][0] * jac[2][2] - jac[0][2] * jac[1][1] * jac[2][0]);
                long c = pid(p, i, j, k);
                const double *b0 = B0 + 3 * c;
                for (int l = 0; l < 3; l++)
                    B[l + 3 * c] = (jac[l][0] * b0[0] + jac[l][1] * b0[1] + jac[l][2] * b0[2]) * detJac1;
                detJac[c] = 1. / detJac1;
            }
}

/* addDist, src/boundaries.cu:11-35 (x = dx*i + Ox from the inner index). */
static double add_dist(const or_params *p, int i, int j, double z, int l)
{
    double x = p->dx * i + p->Ox;
    double y = p->dy * j + p->Oy;
    double distX = 0, distY = 0;
    if (p->initDistCode == 0) {
        distY = -p->initShearA * sin(p->initShearK * 2 * OR_PI * (x + p->Ox - p->dx / 2) / (p->Lx + p->dx)) * z;
        distX = -p->initShearB * sin(p->initShearK * 2 * OR_PI * (distY + y + p->Oy - p->dy / 2) / (p->Ly + p->dy)) * z;
    }
    if (p->initDistCode == 1) {
        distX = 0;
        distY = p->initShearA * exp(-p->initShearK * x * x) * (1 - y * y) * (exp(-1 * z + 1) - 0.15366);
    }
    return l == 0 ? distX : distY;
}

/* jac, 1/det and B with the normal component held (normal < 0: none). */
static void bound_b(const or_params *p, double jac[3][3], const double *b0, int normal,
                    double *Bout, double *detOut)
{
    double detJac1 = 1 / (jac[0][0] * jac[1][1] * jac[2][2] + jac[0][1] * jac[1][2] * jac[2][0] +
                          jac[0][2] * jac[1][0] * jac[2][1] - jac[0][0] * jac[1][2] * jac[2][1] -
                          jac[0][1] * jac[1][0] * jac[2][2] - jac[0][2] * jac[1][1] * jac[2][0]);
    (void)p;
    for (int l = 0; l < 3; l++) {
        if (l == normal)
            Bout[l] = b0[l];
        else
            Bout[l] = (jac[l][0] * b0[0] + jac[l][1] * b0[1] + jac[l][2] * b0[2]) * detJac1;
    }
    *detOut = 1 / detJac1;
}


/* Periodic image coordinate as the reference forms it: a value read by a
 * neighbouring thread of the launch tile is shifted as (a -+ L) -+ h, a value
 * read at a tile edge as a + (-+(L + h)) (src/boundaries.cu:112-160). */
static double pshift(double a, int side, double L, double h, int edge)
{
    if (edge) return a + (2 * side - 1) * (L + h);
    return side ? a + L + h : a - L - h;
}

/* updateBbound, src/boundaries.cu:39-446, faces 2 (z), 1 (y), 0 (x).
 * Line-tied faces: the outer ghost is the ghost point moved one cell along
 * the outward normal (plus addDist on z faces); in-plane neighbours at the
 * edges of the reference's 16x16 launch tiles and at the last row are
 * rebuilt as x +- h instead of being read (src/boundaries.cu:72-109). */
void or_update_bbound(const or_params *p, const double *xb, const double *B0, double *B, double *detJac)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const double dx = p->dx, dy = p->dy, dz = p->dz;

    /* ---- face 2: z = const planes ---- */
    for (int side = 0; side < 2; side++) {
        int kg = side ? nz + 1 : 0;
#pragma omp parallel for schedule(static)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                double L0[3], L2[3], xl[3], xr[3], yd[3], yu[3], jac[3][3];
                long g = pid(p, i + 1, j + 1, kg);
                const double *b0;
                if (!p->zPeri) {
                    const double *xg = xb + 3 * g;
                    const double *xin = xb + 3 * pid(p, i + 1, j + 1, side ? nz : 1);
                    double x2[3];
                    double zsec = 2 * dz * (2 * side - 1) + p->Oz + p->Lz * side;
                    double zg = dz * (2 * side - 1) + p->Oz + p->Lz * side;
                    for (int l = 0; l < 2; l++)
                        x2[l] = xg[l] - add_dist(p, i, j, xg[2], l) + add_dist(p, i, j, zsec, l);
                    x2[2] = xg[2] + (2 * side - 1) * dz;
                    for (int l = 0; l < 3; l++) {
                        L0[l] = side ? xin[l] : x2[l];
                        L2[l] = side ? x2[l] : xin[l];
                    }
                    if (i % 16 == 0) {
                        xl[0] = xg[0] - dx - add_dist(p, i, j, zg, 0) + add_dist(p, i - 1, j, zg, 0);
                        xl[1] = xg[1] - add_dist(p, i, j, zg, 1) + add_dist(p, i - 1, j, zg, 1);
                        xl[2] = xg[2];
                    } else
                        memcpy(xl, xb + 3 * pid(p, i, j + 1, kg), sizeof xl);
                    if (i % 16 == 15 || i == nx - 1) {
                        xr[0] = xg[0] + dx - add_dist(p, i, j, zg, 0) + add_dist(p, i + 1, j, zg, 0);
                        xr[1] = xg[1] - add_dist(p, i, j, zg, 1) + add_dist(p, i + 1, j, zg, 1);
                        xr[2] = xg[2];
                    } else
                        memcpy(xr, xb + 3 * pid(p, i + 2, j + 1, kg), sizeof xr);
                    if (j % 16 == 0) {
                        yd[0] = xg[0] - add_dist(p, i, j, zg, 0) + add_dist(p, i, j - 1, zg, 0);
                        yd[1] = xg[1] - dy - add_dist(p, i, j, zg, 1) + add_dist(p, i, j - 1, zg, 1);
                        yd[2] = xg[2];
                    } else
                        memcpy(yd, xb + 3 * pid(p, i + 1, j, kg), sizeof yd);
                    if (j % 16 == 15 || j == ny - 1) {
                        yu[0] = xg[0] - add_dist(p, i, j, zg, 0) + add_dist(p, i, j + 1, zg, 0);
                        yu[1] = xg[1] + dy - add_dist(p, i, j, zg, 1) + add_dist(p, i, j + 1, zg, 1);
                        yu[2] = xg[2];
                    } else
                        memcpy(yu, xb + 3 * pid(p, i + 1, j + 2, kg), sizeof yu);
                    b0 = B0 + 3 * g;
                } else {
                    /* periodic: the ghost is evaluated at the image plane
                     * with the shifted coordinate (src/boundaries.cu:112-162) */
                    int k0 = side ? nz : nz - 1, k2 = side ? 2 : 1;
                    int ke = side ? 1 : nz;          /* image plane of the ghost */
                    const double *a0 = xb + 3 * pid(p, i + 1, j + 1, k0);
                    const double *a2 = xb + 3 * pid(p, i + 1, j + 1, k2);
                    for (int l = 0; l < 2; l++) {
                        L0[l] = a0[l];
                        L2[l] = a2[l];
                    }
                    L0[2] = side ? a0[2] : a0[2] - p->Lz - p->dz;
                    L2[2] = side ? a2[2] + p->Lz + p->dz : a2[2];
                    const double *ql = xb + 3 * pid(p, i, j + 1, ke), *qr = xb + 3 * pid(p, i + 2, j + 1, ke);
                    const double *qd = xb + 3 * pid(p, i + 1, j, ke), *qu = xb + 3 * pid(p, i + 1, j + 2, ke);
                    for (int l = 0; l < 2; l++) {
                        xl[l] = ql[l];
                        xr[l] = qr[l];
                        yd[l] = qd[l];
                        yu[l] = qu[l];
                    }
                    xl[2] = pshift(ql[2], side, p->Lz, p->dz, i % 16 == 0);
                    xr[2] = pshift(qr[2], side, p->Lz, p->dz, i % 16 == 15 || i == nx - 1);
                    yd[2] = pshift(qd[2], side, p->Lz, p->dz, j % 16 == 0);
                    yu[2] = pshift(qu[2], side, p->Lz, p->dz, j % 16 == 15 || j == ny - 1);
                    b0 = B0 + 3 * pid(p, i + 1, j + 1, ke);
                }
                for (int l = 0; l < 3; l++) {
                    jac[l][0] = (xr[l] - xl[l]) * p->dx1 / 2;
                    jac[l][1] = (yu[l] - yd[l]) * p->dy1 / 2;
                    jac[l][2] = (L2[l] - L0[l]) * p->dz1 / 2;
                }
                bound_b(p, jac, b0, p->zPeri ? -1 : 2, B + 3 * g, detJac + g);
            }
    }

    /* ---- face 1: y = const planes ---- */
    for (int side = 0; side < 2; side++) {
        int jg = side ? ny + 1 : 0;
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < nx; i++) {
                double L0[3], L2[3], xl[3], xr[3], zd[3], zu[3], jac[3][3];
                long g = pid(p, i + 1, jg, k + 1);
                const double *b0;
                if (!p->yPeri) {
                    const double *xg = xb + 3 * g;
                    const double *xin = xb + 3 * pid(p, i + 1, side ? ny : 1, k + 1);
                    double x2[3] = {xg[0], xg[1] + (2 * side - 1) * dy, xg[2]};
                    for (int l = 0; l < 3; l++) {
                        L0[l] = side ? xin[l] : x2[l];
                        L2[l] = side ? x2[l] : xin[l];
                    }
                    if (i % 16 == 0) {
                        xl[0] = xg[0] - dx; xl[1] = xg[1]; xl[2] = xg[2];
                    } else
                        memcpy(xl, xb + 3 * pid(p, i, jg, k + 1), sizeof xl);
                    if (i % 16 == 15 || i == nx - 1) {
                        xr[0] = xg[0] + dx; xr[1] = xg[1]; xr[2] = xg[2];
                    } else
                        memcpy(xr, xb + 3 * pid(p, i + 2, jg, k + 1), sizeof xr);
                    if (k % 16 == 0) {
                        zd[0] = xg[0]; zd[1] = xg[1]; zd[2] = xg[2] - dz;
                    } else
                        memcpy(zd, xb + 3 * pid(p, i + 1, jg, k), sizeof zd);
                    if (k % 16 == 15 || k == nz - 1) {
                        zu[0] = xg[0]; zu[1] = xg[1]; zu[2] = xg[2] + dz;
                    } else
                        memcpy(zu, xb + 3 * pid(p, i + 1, jg, k + 2), sizeof zu);
                    b0 = B0 + 3 * g;
                } else {
                    int j0 = side ? ny : ny - 1, j2 = side ? 2 : 1;
                    int je = side ? 1 : ny;
                    const double *a0 = xb + 3 * pid(p, i + 1, j0, k + 1);
                    const double *a2 = xb + 3 * pid(p, i + 1, j2, k + 1);
                    for (int l = 0; l < 3; l += 2) {
                        L0[l] = a0[l];
                        L2[l] = a2[l];
                    }
                    L0[1] = side ? a0[1] : a0[1] - p->Ly - p->dy;
                    L2[1] = side ? a2[1] + p->Ly + p->dy : a2[1];
                    const double *ql = xb + 3 * pid(p, i, je, k + 1), *qr = xb + 3 * pid(p, i + 2, je, k + 1);
                    const double *qd = xb + 3 * pid(p, i + 1, je, k), *qu = xb + 3 * pid(p, i + 1, je, k + 2);
                    for (int l = 0; l < 3; l += 2) {
                        xl[l] = ql[l];
                        xr[l] = qr[l];
                        zd[l] = qd[l];
                        zu[l] = qu[l];
                    }
                    xl[1] = pshift(ql[1], side, p->Ly, p->dy, i % 16 == 0);
                    xr[1] = pshift(qr[1], side, p->Ly, p->dy, i % 16 == 15 || i == nx - 1);
                    zd[1] = pshift(qd[1], side, p->Ly, p->dy, k % 16 == 0);
                    zu[1] = pshift(qu[1], side, p->Ly, p->dy, k % 16 == 15 || k == nz - 1);
                    b0 = B0 + 3 * pid(p, i + 1, je, k + 1);
                }
                for (int l = 0; l < 3; l++) {
                    jac[l][0] = (xr[l] - xl[l]) * p->dx1 / 2;
                    jac[l][1] = (L2[l] - L0[l]) * p->dy1 / 2;
                    jac[l][2] = (zu[l] - zd[l]) * p->dz1 / 2;
                }
                bound_b(p, jac, b0, p->yPeri ? -1 : 1, B + 3 * g, detJac + g);
            }
    }

    /* ---- face 0: x = const planes ---- */
    for (int side = 0; side < 2; side++) {
        int ig = side ? nx + 1 : 0;
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++) {
                double L0[3], L2[3], yl[3], yr[3], zd[3], zu[3], jac[3][3];
                long g = pid(p, ig, j + 1, k + 1);
                const double *b0;
                if (!p->xPeri) {
                    const double *xg = xb + 3 * g;
                    const double *xin = xb + 3 * pid(p, side ? nx : 1, j + 1, k + 1);
                    double x2[3] = {xg[0] + (2 * side - 1) * dx, xg[1], xg[2]};
                    for (int l = 0; l < 3; l++) {
                        L0[l] = side ? xin[l] : x2[l];
                        L2[l] = side ? x2[l] : xin[l];
                    }
                    if (j % 16 == 0) {
                        yl[0] = xg[0]; yl[1] = xg[1] - dy; yl[2] = xg[2];
                    } else
                        memcpy(yl, xb + 3 * pid(p, ig, j, k + 1), sizeof yl);
                    if (j % 16 == 15 || j == ny - 1) {
                        yr[0] = xg[0]; yr[1] = xg[1] + dy; yr[2] = xg[2];
                    } else
                        memcpy(yr, xb + 3 * pid(p, ig, j + 2, k + 1), sizeof yr);
                    if (k % 16 == 0) {
                        zd[0] = xg[0]; zd[1] = xg[1]; zd[2] = xg[2] - dz;
                    } else
                        memcpy(zd, xb + 3 * pid(p, ig, j + 1, k), sizeof zd);
                    if (k % 16 == 15 || k == nz - 1) {
                        zu[0] = xg[0]; zu[1] = xg[1]; zu[2] = xg[2] + dz;
                    } else
                        memcpy(zu, xb + 3 * pid(p, ig, j + 1, k + 2), sizeof zu);
                    b0 = B0 + 3 * g;
                } else {
                    int i0 = side ? nx : nx - 1, i2 = side ? 2 : 1;
                    int ie = side ? 1 : nx;
                    const double *a0 = xb + 3 * pid(p, i0, j + 1, k + 1);
                    const double *a2 = xb + 3 * pid(p, i2, j + 1, k + 1);
                    for (int l = 1; l < 3; l++) {
                        L0[l] = a0[l];
                        L2[l] = a2[l];
                    }
                    L0[0] = side ? a0[0] : a0[0] - p->Lx - p->dx;
                    L2[0] = side ? a2[0] + p->Lx + p->dx : a2[0];
                    const double *ql = xb + 3 * pid(p, ie, j, k + 1), *qr = xb + 3 * pid(p, ie, j + 2, k + 1);
                    const double *qd = xb + 3 * pid(p, ie, j + 1, k), *qu = xb + 3 * pid(p, ie, j + 1, k + 2);
                    for (int l = 1; l < 3; l++) {
                        yl[l] = ql[l];
                        yr[l] = qr[l];
                        zd[l] = qd[l];
                        zu[l] = qu[l];
                    }
                    yl[0] = pshift(ql[0], side, p->Lx, p->dx, j % 16 == 0);
                    yr[0] = pshift(qr[0], side, p->Lx, p->dx, j % 16 == 15 || j == ny - 1);
                    zd[0] = pshift(qd[0], side, p->Lx, p->dx, k % 16 == 0);
                    zu[0] = pshift(qu[0], side, p->Lx, p->dx, k % 16 == 15 || k == nz - 1);
                    b0 = B0 + 3 * pid(p, ie, j + 1, k + 1);
                }
                for (int l = 0; l < 3; l++) {
                    jac[l][0] = (L2[l] - L0[l]) * p->dx1 / 2;
                    jac[l][1] = (yr[l] - yl[l]) * p->dy1 / 2;
                    jac[l][2] = (zu[l] - zd[l]) * p->dz1 / 2;
                }
                bound_b(p, jac, b0, p->xPeri ? -1 : 0, B + 3 * g, detJac + g);
            }
    }
}

static inline void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* One diamond of the Stokes current (src/current.cu:291-320): vertices in
 * circulation order v1..v4; returns the circulation and the mean area vector. */
static double diamond(const double *x[4], const double *b[4], double na[3])
{
    double d[4][3], c[4][3];
    double I = 0;
    for (int l = 0; l < 3; l++) {
        d[0][l] = x[1][l] - x[0][l];
        d[1][l] = x[2][l] - x[1][l];
        d[2][l] = x[3][l] - x[2][l];
        d[3][l] = x[0][l] - x[3][l];
        I += (b[1][l] + b[0][l]) / 2 * d[0][l] + (b[2][l] + b[1][l]) / 2 * d[1][l] +
             (b[3][l] + b[2][l]) / 2 * d[2][l] + (b[0][l] + b[3][l]) / 2 * d[3][l];
    }
    cross3(d[0], d[1], c[0]);
    cross3(d[1], d[2], c[1]);
    cross3(d[2], d[3], c[2]);
    cross3(d[3], d[0], c[3]);
    for (int l = 0; l < 3; l++)
        na[l] = (c[0][l] + c[1][l] + c[2][l] + c[3][l]) / 4;
    return I;
}

/* src/current.cu:222-409 */
void or_jstokes(const or_params *p, const double *xb, const double *B, double *J)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                long im = pid(p, i - 1, j, k), ip = pid(p, i + 1, j, k);
                long jm = pid(p, i, j - 1, k), jp = pid(p, i, j + 1, k);
                long km = pid(p, i, j, k - 1), kp = pid(p, i, j, k + 1);
                double NA[3][3], II[3];
                {   /* quadrilateral 0: W, S, E, N */
                    const double *x[4] = {xb + 3 * im, xb + 3 * jm, xb + 3 * ip, xb + 3 * jp};
                    const double *b[4] = {B + 3 * im, B + 3 * jm, B + 3 * ip, B + 3 * jp};
                    II[0] = diamond(x, b, NA[0]);
                }
                {   /* quadrilateral 1: y-, z-, y+, z+ */
                    const double *x[4] = {xb + 3 * jm, xb + 3 * km, xb + 3 * jp, xb + 3 * kp};
                    const double *b[4] = {B + 3 * jm, B + 3 * km, B + 3 * jp, B + 3 * kp};
                    II[1] = diamond(x, b, NA[1]);
                }
                {   /* quadrilateral 2: z-, x-, z+, x+ */
                    const double *x[4] = {xb + 3 * km, xb + 3 * im, xb + 3 * kp, xb + 3 * ip};
                    const double *b[4] = {B + 3 * km, B + 3 * im, B + 3 * kp, B + 3 * ip};
                    II[2] = diamond(x, b, NA[2]);
                }
                double detN = NA[0][0] * NA[1][1] * NA[2][2] + NA[0][1] * NA[1][2] * NA[2][0] +
                              NA[0][2] * NA[1][0] * NA[2][1] - NA[0][0] * NA[1][2] * NA[2][1] -
                              NA[0][1] * NA[1][0] * NA[2][2] - NA[0][2] * NA[1][1] * NA[2][0];
                double detN1 = 1 / detN;
                double A1[3][3];
                A1[0][0] = (NA[1][1] * NA[2][2] - NA[1][2] * NA[2][1]) * detN1;
                A1[0][1] = (NA[0][2] * NA[2][1] - NA[0][1] * NA[2][2]) * detN1;
                A1[0][2] = (NA[0][1] * NA[1][2] - NA[0][2] * NA[1][1]) * detN1;
                A1[1][0] = (NA[1][2] * NA[2][0] - NA[1][0] * NA[2][2]) * detN1;
                A1[1][1] = (NA[0][0] * NA[2][2] - NA[0][2] * NA[2][0]) * detN1;
                A1[1][2] = (NA[0][2] * NA[1][0] - NA[0][0] * NA[1][2]) * detN1;
                A1[2][0] = (NA[1][0] * NA[2][1] - NA[1][1] * NA[2][0]) * detN1;
                A1[2][1] = (NA[0][1] * NA[2][0] - NA[0][0] * NA[2][1]) * detN1;
                A1[2][2] = (NA[0][0] * NA[1][1] - NA[0][1] * NA[1][0]) * detN1;
                double *Jo = J + 3 * iid(p, i - 1, j - 1, k - 1);
                for (int l = 0; l < 3; l++) {
                    double s = 0;
                    for (int m = 0; m < 3; m++)
                        s += A1[l][m] * II[m];
                    Jo[l] = s;
                }
            }
}

/* src/gradP.cu:10-156 */
void or_gradp(const or_params *p, const double *xb, const double *detJac, double *gradP)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
#define X(di, dj, dk, l) xb[(l) + 3 * pid(p, i + (di), j + (dj), k + (dk))]
                double det1 = 1 / detJac[pid(p, i, j, k)];
                double jac[3][3], jac1[3][3], d2[3][3][3];
                for (int l = 0; l < 3; l++) {
                    jac[l][0] = (X(1, 0, 0, l) - X(-1, 0, 0, l)) * p->dx1 / 2;
                    jac[l][1] = (X(0, 1, 0, l) - X(0, -1, 0, l)) * p->dy1 / 2;
                    jac[l][2] = (X(0, 0, 1, l) - X(0, 0, -1, l)) * p->dz1 / 2;
                }
                jac1[0][0] = (jac[1][1] * jac[2][2] - jac[1][2] * jac[2][1]) * det1;
                jac1[0][1] = (jac[0][2] * jac[2][1] - jac[0][1] * jac[2][2]) * det1;
                jac1[0][2] = (jac[0][1] * jac[1][2] - jac[0][2] * jac[1][1]) * det1;
                jac1[1][0] = (jac[1][2] * jac[2][0] - jac[1][0] * jac[2][2]) * det1;
                jac1[1][1] = (jac[0][0] * jac[2][2] - jac[0][2] * jac[2][0]) * det1;
                jac1[1][2] = (jac[0][2] * jac[1][0] - jac[0][0] * jac[1][2]) * det1;
                jac1[2][0] = (jac[1][0] * jac[2][1] - jac[1][1] * jac[2][0]) * det1;
                jac1[2][1] = (jac[0][1] * jac[2][0] - jac[0][0] * jac[2][1]) * det1;
                jac1[2][2] = (jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0]) * det1;
                for (int l = 0; l < 3; l++) {
                    d2[l][0][0] = (X(1, 0, 0, l) - 2 * X(0, 0, 0, l) + X(-1, 0, 0, l)) * p->dx1 * p->dx1;
                    d2[l][1][1] = (X(0, 1, 0, l) - 2 * X(0, 0, 0, l) + X(0, -1, 0, l)) * p->dy1 * p->dy1;
                    d2[l][2][2] = (X(0, 0, 1, l) - 2 * X(0, 0, 0, l) + X(0, 0, -1, l)) * p->dz1 * p->dz1;
                    d2[l][0][1] = (X(1, 1, 0, l) - X(1, -1, 0, l) - X(-1, 1, 0, l) + X(-1, -1, 0, l)) * p->dx1 * p->dy1 / 4;
                    d2[l][0][2] = (X(1, 0, 1, l) - X(1, 0, -1, l) - X(-1, 0, 1, l) + X(-1, 0, -1, l)) * p->dx1 * p->dz1 / 4;
                    d2[l][1][2] = (X(0, 1, 1, l) - X(0, -1, 1, l) - X(0, 1, -1, l) + X(0, -1, -1, l)) * p->dy1 * p->dz1 / 4;
                    d2[l][1][0] = d2[l][0][1];
                    d2[l][2][0] = d2[l][0][2];
                    d2[l][2][1] = d2[l][1][2];
                }
#undef X
                double g[3] = {0, 0, 0};
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++)
                        for (int l = 0; l < 3; l++) {
                            g[0] -= d2[a][b][l] * jac1[b][a] * jac1[l][0];
                            g[1] -= d2[a][b][l] * jac1[b][a] * jac1[l][1];
                            g[2] -= d2[a][b][l] * jac1[b][a] * jac1[l][2];
                        }
                double *go = gradP + 3 * iid(p, i - 1, j - 1, k - 1);
                for (int l = 0; l < 3; l++)
                    go[l] = g[l] * (p->beta * det1);
            }
}

/* epsilon profile of src/rungeKutta.cu:180-185 (inner indices). */
static double eps_prof(const or_params *p, int i, int j, int k)
{
    if (!p->epsilonProf) return 1;
    return (-(i - (p->nx) / 2.) / (p->nx) * (i + 1 - (p->nx) / 2.) / (p->nx) * 4 + 1) *
           (-(j - (p->ny) / 2.) / (p->ny) * (j + 1 - (p->ny) / 2.) / (p->ny) * 4 + 1) *
           (-(k - (p->nz) / 2.) / (p->nz) * (k + 1 - (p->nz) / 2.) / (p->nz) * 4 + 1);
}

/* src/rungeKutta.cu:112-247, magnetofrictional branch (inertia and jxbAver
 * off). */
void or_kk(const or_params *p, const double b[42], int n, double dt, const double *xb, const double *B,
           const double *J, const double *gradP, double *kk, double *xb_tmp)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const long N = ninner(p);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                long c = pid(p, i + 1, j + 1, k + 1), q = iid(p, i, j, k);
                double kv[3];
                cross3(J + 3 * q, B + 3 * c, kv);
                if (p->jxbAver) {
                    /* src/rungeKutta.cu:172-177, 357-443, with the +z
                     * neighbour's B at its own padded index */
                    const double w6 = p->jxbAverWeight / 6;
                    for (int l = 0; l < 3; l++) kv[l] *= (1 - p->jxbAverWeight);
                    const int id[3] = {i, j, k}, nn[3] = {nx, ny, nz};
                    for (int m = 0; m < 6; m++) {
                        int d = m / 2, up = m % 2, o[3] = {0, 0, 0};
                        double jm[3] = {0, 0, 0}, bm[3] = {0, 0, 0}, cc[3];
                        o[d] = up ? 1 : -1;
                        if (up ? id[d] < nn[d] - 1 : id[d] > 0) {
                            const double *js = J + 3 * iid(p, i + o[0], j + o[1], k + o[2]);
                            const double *bs = B + 3 * pid(p, i + 1 + o[0], j + 1 + o[1], k + 1 + o[2]);
                            for (int l = 0; l < 3; l++) { jm[l] = js[l]; bm[l] = bs[l]; }
                        }
                        cross3(jm, bm, cc);
                        for (int l = 0; l < 3; l++) kv[l] += cc[l] * w6;
                    }
                }
                double eps = eps_prof(p, i, j, k);
                for (int l = 0; l < 3; l++) {
                    kv[l] *= dt * eps;
                    if (p->pressure) kv[l] += -dt * gradP[l + 3 * q];
                }
                for (int l = 0; l < 3; l++)
                    kk[l + 3 * q + 3 * N * n] = kv[l];
                for (int l = 0; l < 3; l++) {
                    double x = xb[l + 3 * c];
                    for (int m = 0; m < n + 1; m++)
                        x += (b[m + (n + 1) * 6] - b[m + n * 6]) * kk[l + 3 * q + 3 * N * m];
                    xb_tmp[l + 3 * c] = x;
                }
            }
}

/* src/rungeKutta.cu:251-353 with the block max of src/maths.cu:144-153 and
 * the grid max of src/maths.cu:58-130 (both skip NaN: they start at 0 and
 * keep a value only when '>' holds). */
double or_xnewstar(const or_params *p, const double c[6], const double cs[6], const double *xb,
                   const double *kk, double *xb_new, int *nan_seen)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const long N = ninner(p);
    double mx = 0;
    int nan = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                long g = pid(p, i + 1, j + 1, k + 1), q = iid(p, i, j, k);
                for (int l = 0; l < 3; l++) {
                    double xn = xb[l + 3 * g], xs = xb[l + 3 * g];
                    if (l < 2 || p->zUpdate)
                        for (int m = 0; m < 6; m++) {
                            xn += c[m] * kk[l + 3 * q + 3 * N * m];
                            xs += cs[m] * kk[l + 3 * q + 3 * N * m];
                        }
                    xb_new[l + 3 * g] = xn;
                    double d = fabs(xn - xs);
                    if (d != d) nan = 1;
                    if (d > mx) mx = d;
                }
            }
    if (nan_seen) *nan_seen = nan;
    return mx;
}

/* src/rungeKutta.cu:112-247, inertia branch (:195-212, 216-246): the
 * position slope k_x = dt u and the velocity slope
 * k_u = dt det (J x B) - dt nu u det - dt gradP, six values per point in
 * kk[n] ([stage][point][x0 x1 x2 u0 u1 u2]); position and velocity are
 * advanced to the next stage in xb_tmp and uu_tmp.  No friction profile
 * and no J x B averaging in this branch. */
void or_kk_inertia(const or_params *p, const double b[42], int n, double dt, const double *xb, const double *B,
                   const double *J, const double *gradP, const double *detJac, const double *uu, double *kk,
                   double *xb_tmp, double *uu_tmp)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const long N = ninner(p);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                long c = pid(p, i + 1, j + 1, k + 1), q = iid(p, i, j, k);
                const double *u = uu + 3 * q;
                double ku[3], *ks = kk + 6 * q + 6 * N * n;
                cross3(J + 3 * q, B + 3 * c, ku);
                for (int l = 0; l < 3; l++) {
                    ks[l] = dt * u[l];
                    ku[l] *= dt * detJac[c];
                    ku[l] -= dt * p->nu * u[l] * detJac[c];
                    if (p->pressure) ku[l] += -dt * gradP[l + 3 * q];
                    ks[3 + l] = ku[l];
                }
                for (int l = 0; l < 3; l++) {
                    double x = xb[l + 3 * c], v = u[l];
                    for (int m = 0; m < n + 1; m++) {
                        const double w = b[m + (n + 1) * 6] - b[m + n * 6];
                        x += w * kk[l + 6 * q + 6 * N * m];
                        v += w * kk[3 + l + 6 * q + 6 * N * m];
                    }
                    xb_tmp[l + 3 * c] = x;
                    uu_tmp[l + 3 * q] = v;
                }
            }
}

/* src/rungeKutta.cu:251-353 with inertia: x5 / x4 from the position slopes
 * (the error is measured on the positions only), u5 from the velocity
 * slopes into uu_new. */
double or_xnewstar_inertia(const or_params *p, const double c[6], const double cs[6], const double *xb,
                           const double *uu, const double *kk, double *xb_new, double *uu_new, int *nan_seen)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const long N = ninner(p);
    double mx = 0;
    int nan = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                long g = pid(p, i + 1, j + 1, k + 1), q = iid(p, i, j, k);
                for (int l = 0; l < 3; l++) {
                    double xn = xb[l + 3 * g], xs = xb[l + 3 * g];
                    if (l < 2 || p->zUpdate)
                        for (int m = 0; m < 6; m++) {
                            xn += c[m] * kk[l + 6 * q + 6 * N * m];
                            xs += cs[m] * kk[l + 6 * q + 6 * N * m];
                        }
                    double un = uu[l + 3 * q];
                    for (int m = 0; m < 6; m++) un += c[m] * kk[3 + l + 6 * q + 6 * N * m];
                    xb_new[l + 3 * g] = xn;
                    uu_new[l + 3 * q] = un;
                    double d = fabs(xn - xs);
                    if (d != d) nan = 1;
                    if (d > mx) mx = d;
                }
            }
    if (nan_seen) *nan_seen = nan;
    return mx;
}

/* U2_det + sumGlobal, src/maths.cu:683-717, src/diagnostics.cu:114-122:
 * sum of |u|^2 det with the boundary weights 1/2^(#faces), times
 * dx dy dz.  Summed serially in (k,j,i) order. */
double or_u2(const or_params *p, const double *uu, const double *detJac)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    double s = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const double *u = uu + 3 * iid(p, i, j, k);
                double v = (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * detJac[pid(p, i + 1, j + 1, k + 1)];
                int bound = (i == 0) + (j == 0) + (k == 0) + (i == nx - 1) + (j == ny - 1) + (k == nz - 1);
                bound = 1 << bound;
                v /= bound;
                s += v;
            }
    return s * p->dx * p->dy * p->dz;
}

/* mag + findDevMax, src/maths.cu:157-164, src/diagnostics.cu:124-130:
 * NaN-blind max of |u| over inner points. */
double or_umax(const or_params *p, const double *uu)
{
    return or_jmax(p, uu);
}

/* B2_det + sumGlobal, src/maths.cu:647-680, src/diagnostics.cu:73-82.
 * Summed serially in (k,j,i) order. */
double or_b2(const or_params *p, const double *B, const double *detJac)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    double s = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                long g = pid(p, i + 1, j + 1, k + 1);
                const double *b = B + 3 * g;
                double v = (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]) * detJac[g];
                int bound = (i == 0) + (j == 0) + (k == 0) + (i == nx - 1) + (j == ny - 1) + (k == nz - 1);
                bound = 1 << bound;
                v /= bound;
                s += v;
            }
    return s * p->dx * p->dy * p->dz;
}

/* mag + findDevMax, src/maths.cu:157-164, src/diagnostics.cu:15-22 */
double or_jmax(const or_params *p, const double *J)
{
    const long N = ninner(p);
    double mx = 0;
    for (long q = 0; q < N; q++) {
        const double *j = J + 3 * q;
        double v = sqrt(j[0] * j[0] + j[1] * j[1] + j[2] * j[2]);
        if (v > mx) mx = v;
    }
    return mx;
}

/* JxB_B2 + findDevMax, src/maths.cu:168-200, src/diagnostics.cu:23-32 */
double or_jxb_b2_max(const or_params *p, const double *B, const double *J)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    double mx = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const double *b = B + 3 * pid(p, i + 1, j + 1, k + 1);
                const double *jj = J + 3 * iid(p, i, j, k);
                double c[3];
                cross3(jj, b, c);
                double b2 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
                double v = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) / b2;
                if (v > mx) mx = v;
            }
    return mx;
}

/* max |JxB|/(|J||B|), 0/0 := 0.  No reference counterpart: the relaxation
 * criterion of the workload. */
double or_jxb_jb_max(const or_params *p, const double *B, const double *J)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    double mx = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const double *b = B + 3 * pid(p, i + 1, j + 1, k + 1);
                const double *jj = J + 3 * iid(p, i, j, k);
                double c[3];
                cross3(jj, b, c);
                double nc = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
                double nj = sqrt(jj[0] * jj[0] + jj[1] * jj[1] + jj[2] * jj[2]);
                double nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
                double den = nj * nb;
                double v = den > 0 ? nc / den : 0;
                if (v > mx) mx = v;
            }
    return mx;
}

/* max |JxB| / max (|J||B|).  No reference counterpart: the relaxation
 * criterion of the workload. */
double or_jxbmax_jbmax(const or_params *p, const double *B, const double *J)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    double mc = 0, md = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const double *b = B + 3 * pid(p, i + 1, j + 1, k + 1);
                const double *jj = J + 3 * iid(p, i, j, k);
                double c[3];
                cross3(jj, b, c);
                double nc = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
                double den = sqrt(jj[0] * jj[0] + jj[1] * jj[1] + jj[2] * jj[2]) *
                             sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
                if (nc > mc) mc = nc;
                if (den > md) md = den;
            }
    return md > 0 ? mc / md : 0;
}

int or_state_create(const or_params *p, const double *B0, const double *xb, or_state *s)
{
    long P3 = 3 * npad(p), N3 = 3 * ninner(p);
    memset(s, 0, sizeof *s);
    s->B0 = (double *)malloc(P3 * sizeof(double));
    s->xb = (double *)malloc(P3 * sizeof(double));
    s->B = (double *)malloc(P3 * sizeof(double));
    s->detJac = (double *)calloc(npad(p), sizeof(double));
    s->J = (double *)calloc(N3, sizeof(double));
    s->gradP = (double *)calloc(N3, sizeof(double));
    s->kk = (double *)calloc((p->inertia ? 12 : 6) * N3, sizeof(double));
    s->xb_tmp = (double *)malloc(P3 * sizeof(double));
    s->xb_new = (double *)malloc(P3 * sizeof(double));
    if (p->inertia) {
        /* the velocity starts at rest (uInit = 'nil'; the reference leaves
         * h.uu as allocated, src/init.cu:358-378, src/memory.cu:34) */
        s->uu = (double *)calloc(N3, sizeof(double));
        s->uu_tmp = (double *)calloc(N3, sizeof(double));
        s->uu_new = (double *)calloc(N3, sizeof(double));
    }
    if (!s->B0 || !s->xb || !s->B || !s->detJac || !s->J || !s->gradP || !s->kk || !s->xb_tmp || !s->xb_new ||
        (p->inertia && (!s->uu || !s->uu_tmp || !s->uu_new))) {
        or_state_free(s);
        return -1;
    }
    memcpy(s->B0, B0, P3 * sizeof(double));
    memcpy(s->xb, xb, P3 * sizeof(double));
    memcpy(s->B, B0, P3 * sizeof(double));
    memcpy(s->xb_tmp, xb, P3 * sizeof(double));
    memcpy(s->xb_new, xb, P3 * sizeof(double));
    s->t = 0;
    s->dt = p->dt0;
    or_assign_coefficients(s->a, s->b, s->c, s->cs);
    return 0;
}

void or_state_free(or_state *s)
{
    free(s->B0); free(s->xb); free(s->B); free(s->detJac); free(s->J);
    free(s->gradP); free(s->kk); free(s->xb_tmp); free(s->xb_new);
    free(s->uu); free(s->uu_tmp); free(s->uu_new);
    memset(s, 0, sizeof *s);
}

void or_jclassic(const or_params *p, const double *xb, const double *B0, const double *detJac, double *J);
void or_jstokes4th(const or_params *p, const double *xb, const double *B, double *J);

static int any_peri(const or_params *p)
{
    return p->xPeri || p->yPeri || p->zPeri;
}

void or_fields(const or_params *p, or_state *s)
{
    or_b_jacb0(p, s->xb, s->B0, s->B, s->detJac);
    or_update_bbound(p, s->xb, s->B0, s->B, s->detJac);
    if (any_peri(p)) or_set_peri(p, s->B);
    if (strncmp(p->jMethod, "Classic ", 8) == 0)
        or_jclassic(p, s->xb, s->B0, s->detJac, s->J);
    else if (strncmp(p->jMethod, "Stokes4th ", 10) == 0)
        or_jstokes4th(p, s->xb, s->B, s->J);
    else
        or_jstokes(p, s->xb, s->B, s->J);
}

/* src/main.cu:260-319 */
double or_attempt(const or_params *p, or_state *s, int *nan_seen)
{
    long P3 = 3 * npad(p);
    for (int n = 0; n < 6; n++) {
        or_fields(p, s);
        if (p->pressure) or_gradp(p, s->xb, s->detJac, s->gradP);
        if (p->inertia)
            or_kk_inertia(p, s->b, n, s->dt, s->xb, s->B, s->J, s->gradP, s->detJac, s->uu, s->kk, s->xb_tmp,
                          s->uu_tmp);
        else
            or_kk(p, s->b, n, s->dt, s->xb, s->B, s->J, s->gradP, s->kk, s->xb_tmp);
        if (any_peri(p)) or_set_grid_peri(p, s->xb_tmp);
        memcpy(s->xb, s->xb_tmp, P3 * sizeof(double));
        if (p->inertia) memcpy(s->uu, s->uu_tmp, 3 * ninner(p) * sizeof(double));
    }
    double md = p->inertia ? or_xnewstar_inertia(p, s->c, s->cs, s->xb, s->uu, s->kk, s->xb_new, s->uu_new, nan_seen)
                           : or_xnewstar(p, s->c, s->cs, s->xb, s->kk, s->xb_new, nan_seen);
    if (any_peri(p)) or_set_grid_peri(p, s->xb_new);
    return md;
}

/* src/main.cu:259-383 */
int or_step(const or_params *p, or_state *s, double *dt_used, double *maxDelta, int *nan_seen)
{
    long P3 = 3 * npad(p);
    double dt = s->dt;
    double md = or_attempt(p, s, nan_seen);
    if (dt_used) *dt_used = dt;
    if (maxDelta) *maxDelta = md;
    int acc = md < p->maxError;
    if (acc) {
        s->t += dt;
        memcpy(s->xb, s->xb_new, P3 * sizeof(double));
        if (p->inertia) memcpy(s->uu, s->uu_new, 3 * ninner(p) * sizeof(double));
        or_fields(p, s);
    }
    if (md > 0) s->dt = dt * pow(0.9 * p->maxError / md, 0.2);
    return acc;
}

/* ---- mesh and error diagnostics ------------------------------------ */

/* dirCellVol, gridWedgeMin, gridConvexity (src/maths.cu:443-640) per inner
 * point (x fastest, any array may be NULL), and the grid minima of
 * convexity and wedgeMin (src/diagnostics.cu:84-110: -max(-v)). */
void or_mesh(const or_params *p, const double *xb, double *cellVol, double *convex, double *wedge,
             double *convex_min, double *wedge_min)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    double cmin = 0, wmin = 0;
    int first = 1;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const double *x = xb + 3 * pid(p, i + 1, j + 1, k + 1);
                double wm = 0, cv = 1;
                for (int u = 0; u <= 2; u += 2)
                    for (int v = 0; v <= 2; v += 2)
                        for (int w = 0; w <= 2; w += 2) {
                            const double *a = xb + 3 * pid(p, i + u, j + 1, k + 1);
                            const double *b = xb + 3 * pid(p, i + 1, j + v, k + 1);
                            const double *c = xb + 3 * pid(p, i + 1, j + 1, k + w);
                            double d1[3], d2[3], d3[3], cr[3];
                            for (int l = 0; l < 3; l++) {
                                d1[l] = a[l] - x[l];
                                d2[l] = b[l] - x[l];
                                d3[l] = c[l] - x[l];
                            }
                            cross3(d1, d2, cr);
                            double wc = cr[0] * d3[0] + cr[1] * d3[1] + cr[2] * d3[2];
                            if ((wc < 0) && (u + v + w == 2 || u + v + w == 6)) cv = -1;
                            if ((wc > 0) && (u + v + w == 0 || u + v + w == 4)) cv = -1;
                            if (u + v + w == 0) wm = -wc;
                            if (u + v + w == 0 || u + v + w == 4) wc = -wc;
                            if (wc < wm) wm = wc;
                        }
                double d1[3], d2[3], d3[3], cr[3];
                const double *a = xb + 3 * pid(p, i + 2, j + 1, k + 1);
                const double *b = xb + 3 * pid(p, i + 1, j + 2, k + 1);
                const double *c = xb + 3 * pid(p, i + 1, j + 1, k + 2);
                for (int l = 0; l < 3; l++) {
                    d1[l] = a[l] - x[l];
                    d2[l] = b[l] - x[l];
                    d3[l] = c[l] - x[l];
                }
                cross3(d1, d2, cr);
                long q = iid(p, i, j, k);
                if (cellVol) cellVol[q] = cr[0] * d3[0] + cr[1] * d3[1] + cr[2] * d3[2];
                if (convex) convex[q] = cv;
                if (wedge) wedge[q] = wm;
                if (first || cv < cmin) cmin = cv;
                if (first || wm < wmin) wmin = wm;
                first = 0;
            }
    if (convex_min) *convex_min = cmin;
    if (wedge_min) *wedge_min = wmin;
}

/* B_1ez + sumGlobal (src/maths.cu:294-330, src/diagnostics.cu:51-60) */
double or_err_b_1ez(const or_params *p, const double *B)
{
    double s = 0;
    for (int k = 0; k < p->nz; k++)
        for (int j = 0; j < p->ny; j++)
            for (int i = 0; i < p->nx; i++) {
                const double *b = B + 3 * pid(p, i + 1, j + 1, k + 1);
                s += b[0] * b[0] + b[1] * b[1] + (b[2] - 1.) * (b[2] - 1.);
            }
    return sqrt(s) / sqrt((double)p->nx * p->ny * p->nz);
}

/* xb_XbAn + sumGlobal (src/maths.cu:334-377, src/diagnostics.cu:62-71):
 * distance to the analytic twist x_an of the "analytic" initial field */
double or_err_xb_xban(const or_params *p, const double *xb)
{
    double s = 0;
    for (int k = 0; k < p->nz; k++)
        for (int j = 0; j < p->ny; j++)
            for (int i = 0; i < p->nx; i++) {
                const double *v = xb + 3 * pid(p, i + 1, j + 1, k + 1);
                double x = i * p->dx + p->Ox, y = j * p->dy + p->Oy, z = k * p->dz + p->Oz;
                double e = exp(-(x * x + y * y) / (p->ar * p->ar) - z * z / (p->az * p->az));
                double d0 = v[0] - cos(e * p->phi1) * x - sin(e * p->phi1) * y;
                double d1 = v[1] + sin(e * p->phi1) * x - cos(e * p->phi1) * y;
                double d2 = v[2] - z;
                s += d0 * d0 + d1 * d1 + d2 * d2;
            }
    return sqrt(s) / sqrt((double)p->nx * p->ny * p->nz);
}

/* JB_B2 + epsilonStar + findDevMax (src/maths.cu:204-290,
 * src/diagnostics.cu:33-49): along the column i = nx/2, j = ny/2, the
 * largest sqrt2 |a_n - a_k| / |z_n - z_k| over all pairs (0 where z is
 * equal), a = J.B/B^2, z the position's z.  For nz <= 256 the reference
 * covers exactly these pairs; above, its tiling leaves entries unwritten. */
double or_epsilon_star(const or_params *p, const double *xb, const double *B, const double *J)
{
    const int nz = p->nz, i = p->nx / 2, j = p->ny / 2;
    const double SQRT2 = 1.414213562373095;   /* src/global.h:18 */
    double *a = (double *)malloc(sizeof(double) * nz), *z = (double *)malloc(sizeof(double) * nz);
    for (int k = 0; k < nz; k++) {
        const double *b = B + 3 * pid(p, i + 1, j + 1, k + 1);
        const double *jj = J + 3 * iid(p, i, j, k);
        a[k] = (jj[0] * b[0] + jj[1] * b[1] + jj[2] * b[2]) / (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        z[k] = xb[2 + 3 * pid(p, i + 1, j + 1, k + 1)];
    }
    double m = 0;
    for (int k = 0; k < nz; k++)
        for (int n = 0; n < nz; n++) {
            double v = z[n] != z[k] ? SQRT2 * fabs((a[n] - a[k]) / (z[n] - z[k])) : 0;
            if (m < v) m = v;
        }
    free(a);
    free(z);
    return m;
}

/* JClassic, src/current.cu:10-219: curl of B = jac B0/det by the chain
 * rule with central differences (reference operation order). */
void or_jclassic(const or_params *p, const double *xb, const double *B0, const double *detJac, double *J)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const double h1[3] = {p->dx1, p->dy1, p->dz1};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const int c0[3] = {i + 1, j + 1, k + 1};
                double jac[3][3], d2[3][3][3], dB0[3][3], dD1[3], jac1[3][3];
#define XP(di, dj, dk) (xb + 3 * pid(p, c0[0] + (di), c0[1] + (dj), c0[2] + (dk)))
#define BP(di, dj, dk) (B0 + 3 * pid(p, c0[0] + (di), c0[1] + (dj), c0[2] + (dk)))
#define DP(di, dj, dk) (detJac[pid(p, c0[0] + (di), c0[1] + (dj), c0[2] + (dk))])
                const int e[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
                const double *x0 = XP(0, 0, 0), *b0 = BP(0, 0, 0);
                for (int m = 0; m < 3; m++) {
                    const double *xp = XP(e[m][0], e[m][1], e[m][2]), *xm = XP(-e[m][0], -e[m][1], -e[m][2]);
                    const double *bp = BP(e[m][0], e[m][1], e[m][2]), *bm = BP(-e[m][0], -e[m][1], -e[m][2]);
                    for (int l = 0; l < 3; l++) {
                        jac[l][m] = (xp[l] - xm[l]) * h1[m] / 2;
                        d2[l][m][m] = (xp[l] - 2 * x0[l] + xm[l]) * h1[m] * h1[m];
                        dB0[l][m] = (bp[l] - bm[l]) * h1[m] / 2;
                    }
                    dD1[m] = (1 / DP(e[m][0], e[m][1], e[m][2]) - 1 / DP(-e[m][0], -e[m][1], -e[m][2])) * h1[m] / 2;
                }
                /* mixed derivatives in the reference's term order */
                for (int l = 0; l < 3; l++) {
                    d2[l][0][1] = (XP(1, 1, 0)[l] - XP(1, -1, 0)[l] - XP(-1, 1, 0)[l] + XP(-1, -1, 0)[l]) * h1[0] * h1[1] / 4;
                    d2[l][0][2] = (XP(1, 0, 1)[l] - XP(1, 0, -1)[l] - XP(-1, 0, 1)[l] + XP(-1, 0, -1)[l]) * h1[0] * h1[2] / 4;
                    d2[l][1][2] = (XP(0, 1, 1)[l] - XP(0, -1, 1)[l] - XP(0, 1, -1)[l] + XP(0, -1, -1)[l]) * h1[1] * h1[2] / 4;
                    d2[l][1][0] = d2[l][0][1];
                    d2[l][2][0] = d2[l][0][2];
                    d2[l][2][1] = d2[l][1][2];
                }
#undef XP
#undef BP
#undef DP
                const double det1 = 1 / detJac[pid(p, c0[0], c0[1], c0[2])];
                jac1[0][0] = (jac[1][1] * jac[2][2] - jac[1][2] * jac[2][1]) * det1;
                jac1[0][1] = (jac[0][2] * jac[2][1] - jac[0][1] * jac[2][2]) * det1;
                jac1[0][2] = (jac[0][1] * jac[1][2] - jac[0][2] * jac[1][1]) * det1;
                jac1[1][0] = (jac[1][2] * jac[2][0] - jac[1][0] * jac[2][2]) * det1;
                jac1[1][1] = (jac[0][0] * jac[2][2] - jac[0][2] * jac[2][0]) * det1;
                jac1[1][2] = (jac[0][2] * jac[1][0] - jac[0][0] * jac[1][2]) * det1;
                jac1[2][0] = (jac[1][0] * jac[2][1] - jac[1][1] * jac[2][0]) * det1;
                jac1[2][1] = (jac[0][1] * jac[2][0] - jac[0][0] * jac[2][1]) * det1;
                jac1[2][2] = (jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0]) * det1;
                double *Jq = J + 3 * iid(p, i, j, k);
                Jq[0] = Jq[1] = Jq[2] = 0;
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) {
                        double T[3];
                        for (int l = 0; l < 3; l++)
                            T[l] = d2[l][a][b] * b0[b] * det1 + jac[l][b] * dB0[b][a] * det1 + jac[l][b] * b0[b] * dD1[a];
                        Jq[0] += jac1[a][1] * T[2] - jac1[a][2] * T[1];
                        Jq[1] += jac1[a][2] * T[0] - jac1[a][0] * T[2];
                        Jq[2] += jac1[a][0] * T[1] - jac1[a][1] * T[0];
                    }
            }
}

/* JStokes4th, src/current.cu:413-735: the octilateral through the 4 face and
 * 4 edge neighbours in each coordinate plane (rows xy, yz, zx of NA), the
 * trapezoid circulation, the area vector (half the corner cross products
 * plus a quarter of the inner diamond's), J = NA^-1 I with each cofactor
 * divided by det NA (reference operation order). */
void or_jstokes4th(const or_params *p, const double *xb, const double *B, double *J)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    static const int ax[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    static const int V[8][2] = {{0, 0}, {1, 0}, {2, 0}, {2, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                double NA[3][3], II[3];
                for (int q = 0; q < 3; q++) {
                    const double *x[8], *b[8];
                    for (int v = 0; v < 8; v++) {
                        int o[3] = {0, 0, 0};
                        o[ax[q][0]] = V[v][0] - 1;
                        o[ax[q][1]] = V[v][1] - 1;
                        long g = pid(p, i + o[0], j + o[1], k + o[2]);
                        x[v] = xb + 3 * g;
                        b[v] = B + 3 * g;
                    }
                    double dx[8][3], dA[3], dB[3], dC[3], dD[3], I = 0;
                    for (int l = 0; l < 3; l++) {
                        for (int e = 1; e <= 8; e++) dx[e - 1][l] = x[e % 8][l] - x[e - 1][l];
                        I += (b[1][l] + b[0][l]) / 2 * dx[0][l] + (b[2][l] + b[1][l]) / 2 * dx[1][l] +
                             (b[3][l] + b[2][l]) / 2 * dx[2][l] + (b[4][l] + b[3][l]) / 2 * dx[3][l] +
                             (b[5][l] + b[4][l]) / 2 * dx[4][l] + (b[6][l] + b[5][l]) / 2 * dx[5][l] +
                             (b[7][l] + b[6][l]) / 2 * dx[6][l] + (b[0][l] + b[7][l]) / 2 * dx[7][l];
                        dA[l] = x[1][l] - x[7][l];
                        dB[l] = x[3][l] - x[1][l];
                        dC[l] = x[5][l] - x[3][l];
                        dD[l] = x[7][l] - x[5][l];
                    }
                    II[q] = I;
                    double c1[3], c2[3], c3[3], c4[3];
                    cross3(dx[1], dx[2], c1);
                    cross3(dx[3], dx[4], c2);
                    cross3(dx[5], dx[6], c3);
                    cross3(dx[7], dx[0], c4);
                    for (int l = 0; l < 3; l++) NA[q][l] = (c1[l] + c2[l] + c3[l] + c4[l]) / 2;
                    cross3(dA, dB, c1);
                    cross3(dB, dC, c2);
                    cross3(dC, dD, c3);
                    cross3(dD, dA, c4);
                    for (int l = 0; l < 3; l++) NA[q][l] += (c1[l] + c2[l] + c3[l] + c4[l]) / 4;
                }
                double dN = NA[0][0] * NA[1][1] * NA[2][2] + NA[0][1] * NA[1][2] * NA[2][0] +
                            NA[0][2] * NA[1][0] * NA[2][1] - NA[0][0] * NA[1][2] * NA[2][1] -
                            NA[0][1] * NA[1][0] * NA[2][2] - NA[0][2] * NA[1][1] * NA[2][0];
                double A1[3][3];
                A1[0][0] = (NA[1][1] * NA[2][2] - NA[1][2] * NA[2][1]) / dN;
                A1[0][1] = (NA[0][2] * NA[2][1] - NA[0][1] * NA[2][2]) / dN;
                A1[0][2] = (NA[0][1] * NA[1][2] - NA[0][2] * NA[1][1]) / dN;
                A1[1][0] = (NA[1][2] * NA[2][0] - NA[1][0] * NA[2][2]) / dN;
                A1[1][1] = (NA[0][0] * NA[2][2] - NA[0][2] * NA[2][0]) / dN;
                A1[1][2] = (NA[0][2] * NA[1][0] - NA[0][0] * NA[1][2]) / dN;
                A1[2][0] = (NA[1][0] * NA[2][1] - NA[1][1] * NA[2][0]) / dN;
                A1[2][1] = (NA[0][1] * NA[2][0] - NA[0][0] * NA[2][1]) / dN;
                A1[2][2] = (NA[0][0] * NA[1][1] - NA[0][1] * NA[1][0]) / dN;
                double *Jo = J + 3 * iid(p, i - 1, j - 1, k - 1);
                for (int l = 0; l < 3; l++) {
                    double s = 0;
                    for (int m = 0; m < 3; m++) s += A1[l][m] * II[m];
                    Jo[l] = s;
                }
            }
}
