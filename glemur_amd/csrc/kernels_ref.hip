// Per-kernel path: one HIP kernel per step of the reference's call
// structure (B_JacB0, updateBbound, setPeri, setGridPeri, JStokes,
// gradPClassic, kk, xNewStar, reductions), FP64, structure-of-arrays on the
// padded grid, one thread per point.  Not the fast path -- the fused stage
// kernels are -- but it keeps the reference's call-by-call boundary and
// checks each piece of the arithmetic on its own.
#include "gm_device.hpp"

#include <mutex>
#include <vector>

namespace gmd {

namespace {

constexpr int kBlock = 256;

inline int blocks_for(long long n, int bs = kBlock)
{
    long long b = (n + bs - 1) / bs;
    return (int)(b < 1 ? 1 : b);
}

// position of padded point a, component-major SoA
__device__ __forceinline__ void ld3(const double *F, long long P, long long a, double v[3])
{
    v[0] = F[a];
    v[1] = F[P + a];
    v[2] = F[2 * P + a];
}

// ---------------------------------------------------------------- B_JacB0
__global__ void __launch_bounds__(kBlock) k_b_jacb0(Geo g, const double *X, const double *B0, double *B, double *det)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long N = (long long)g.nx * g.ny * g.nz;
    if (t >= N) return;
    int i = (int)(t % g.nx) + 1, j = (int)((t / g.nx) % g.ny) + 1, k = (int)(t / ((long long)g.nx * g.ny)) + 1;
    long long a = gidx(g, i, j, k);
    double xm[3], xp[3], ym[3], yp[3], zm[3], zp[3], b0[3], d[3], jac[3][3], b[3];
    ld3(X, g.P, a - 1, xm);
    ld3(X, g.P, a + 1, xp);
    ld3(X, g.P, a - g.sx, ym);
    ld3(X, g.P, a + g.sx, yp);
    ld3(X, g.P, a - g.sxy, zm);
    ld3(X, g.P, a + g.sxy, zp);
    for (int l = 0; l < 3; ++l) d[l] = xp[l] - xm[l];
    jac_col(jac, 0, d, g.dx1);
    for (int l = 0; l < 3; ++l) d[l] = yp[l] - ym[l];
    jac_col(jac, 1, d, g.dy1);
    for (int l = 0; l < 3; ++l) d[l] = zp[l] - zm[l];
    jac_col(jac, 2, d, g.dz1);
    ld3(B0, g.P, a, b0);
    double detJac1 = cauchy_b(jac, b0, -1, b);
    B[a] = b[0];
    B[g.P + a] = b[1];
    B[2 * g.P + a] = b[2];
    det[a] = 1. / detJac1;
}

// ----------------------------------------------------------- updateBbound
// Periodic image coordinate (a -+ L) -+ h for values read by a neighbouring
// thread of the reference's launch tile, a + (-+(L+h)) at tile edges.
__device__ __forceinline__ double pshift(double a, int side, double L, double h, bool edge)
{
    if (edge) return a + (2 * side - 1) * (L + h);
    return side ? a + L + h : a - L - h;
}

// One ghost-face point.  face: 0 x, 1 y, 2 z; (u, v): inner in-plane
// indices along the face's first and second tangential axes (x,y for z
// faces, x,z for y faces, y,z for x faces).
__device__ void bbound_point(const Geo &g, const double *X, const double *B0, double *B, double *det, int face,
                             int side, int u, int v)
{
    // axis numbers: normal n, tangential t0 (u), t1 (v)
    const int n = face, t0 = face == 0 ? 1 : 0, t1 = face == 2 ? 1 : 2;
    int nn[3] = {g.nx, g.ny, g.nz};
    double hh[3] = {g.dx, g.dy, g.dz}, h1[3] = {g.dx1, g.dy1, g.dz1}, LL[3] = {g.Lx, g.Ly, g.Lz};
    int peri = face == 0 ? g.xPeri : face == 1 ? g.yPeri : g.zPeri;
    long long st[3] = {1, g.sx, g.sxy};
    int c[3];
    c[n] = side ? nn[n] + 1 : 0;
    c[t0] = u + 1;
    c[t1] = v + 1;
    long long a = c[0] * st[0] + c[1] * st[1] + c[2] * st[2];
    double L0[3], L2[3], um[3], up[3], vm[3], vp[3];
    bool eum = u % 16 == 0, eup = (u % 16 == 15) || (u == nn[t0] - 1);
    bool evm = v % 16 == 0, evp = (v % 16 == 15) || (v == nn[t1] - 1);
    long long ab;  // B0 source
    if (!peri) {
        double xg[3], xin[3];
        ld3(X, g.P, a, xg);
        ld3(X, g.P, a + (side ? -st[n] : st[n]), xin);
        double x2[3] = {xg[0], xg[1], xg[2]};
        if (face == 2) {
            // addDist on the outer ghost of z faces (src/boundaries.cu:63-67)
            double zsec = 2 * g.dz * (2 * side - 1) + g.Oz + g.Lz * side;
            for (int l = 0; l < 2; ++l)
                x2[l] = xg[l] - add_dist(g.dist, u, v, xg[2], l) + add_dist(g.dist, u, v, zsec, l);
            x2[2] = xg[2] + (2 * side - 1) * g.dz;
        } else {
            x2[n] = xg[n] + (2 * side - 1) * hh[n];
        }
        for (int l = 0; l < 3; ++l) {
            L0[l] = side ? xin[l] : x2[l];
            L2[l] = side ? x2[l] : xin[l];
        }
        // in-plane neighbours: read, or rebuilt as x +- h at tile edges
        auto nb = [&](double out[3], int axis, int sgn, bool edge, long long off, int du, int dv) {
            if (edge) {
                for (int l = 0; l < 3; ++l) out[l] = xg[l];
                if (face == 2) {
                    double zg = g.dz * (2 * side - 1) + g.Oz + g.Lz * side;
                    double hx = axis == 0 ? sgn * g.dx : 0.0, hy = axis == 1 ? sgn * g.dy : 0.0;
                    out[0] = (axis == 0 ? xg[0] + hx : xg[0]) - add_dist(g.dist, u, v, zg, 0) +
                             add_dist(g.dist, u + du, v + dv, zg, 0);
                    out[1] = (axis == 1 ? xg[1] + hy : xg[1]) - add_dist(g.dist, u, v, zg, 1) +
                             add_dist(g.dist, u + du, v + dv, zg, 1);
                } else {
                    out[axis] = xg[axis] + sgn * hh[axis];
                }
            } else {
                ld3(X, g.P, a + off, out);
            }
        };
        nb(um, t0, -1, eum, -st[t0], -1, 0);
        nb(up, t0, +1, eup, st[t0], +1, 0);
        nb(vm, t1, -1, evm, -st[t1], 0, -1);
        nb(vp, t1, +1, evp, st[t1], 0, +1);
        ab = a;
    } else {
        // periodic: evaluated at the image plane (src/boundaries.cu:112-162)
        int i0 = side ? nn[n] : nn[n] - 1, i2 = side ? 2 : 1, ie = side ? 1 : nn[n];
        long long base = a - c[n] * st[n];
        double a0[3], a2[3];
        ld3(X, g.P, base + i0 * st[n], a0);
        ld3(X, g.P, base + i2 * st[n], a2);
        for (int l = 0; l < 3; ++l) {
            L0[l] = a0[l];
            L2[l] = a2[l];
        }
        L0[n] = side ? a0[n] : a0[n] - LL[n] - hh[n];
        L2[n] = side ? a2[n] + LL[n] + hh[n] : a2[n];
        long long ai = base + ie * st[n];
        ld3(X, g.P, ai - st[t0], um);
        ld3(X, g.P, ai + st[t0], up);
        ld3(X, g.P, ai - st[t1], vm);
        ld3(X, g.P, ai + st[t1], vp);
        um[n] = pshift(um[n], side, LL[n], hh[n], eum);
        up[n] = pshift(up[n], side, LL[n], hh[n], eup);
        vm[n] = pshift(vm[n], side, LL[n], hh[n], evm);
        vp[n] = pshift(vp[n], side, LL[n], hh[n], evp);
        ab = ai;
    }
    double jac[3][3], d[3], b0[3], b[3];
    for (int l = 0; l < 3; ++l) d[l] = L2[l] - L0[l];
    jac_col(jac, n, d, h1[n]);
    for (int l = 0; l < 3; ++l) d[l] = up[l] - um[l];
    jac_col(jac, t0, d, h1[t0]);
    for (int l = 0; l < 3; ++l) d[l] = vp[l] - vm[l];
    jac_col(jac, t1, d, h1[t1]);
    ld3(B0, g.P, ab, b0);
    double detJac1 = cauchy_b(jac, b0, peri ? -1 : n, b);
    B[a] = b[0];
    B[g.P + a] = b[1];
    B[2 * g.P + a] = b[2];
    det[a] = 1 / detJac1;
}

__global__ void __launch_bounds__(kBlock) k_bbound(Geo g, const double *X, const double *B0, double *B, double *det, int face)
{
    int nu = face == 0 ? g.ny : g.nx, nv = face == 2 ? g.ny : g.nz;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long per = (long long)nu * nv;
    if (t >= 2 * per) return;
    int side = (int)(t / per);
    long long r = t % per;
    bbound_point(g, X, B0, B, det, face, side, (int)(r % nu), (int)(r / nu));
}

// ------------------------------------------------------ periodic copies
__global__ void k_set_peri(Geo g, double *F, int face)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int na = face == 0 ? g.ny + 2 : g.nx + 2, nb = face == 2 ? g.ny + 2 : g.nz + 2;
    if (t >= (long long)na * nb) return;
    int u = (int)(t % na), v = (int)(t / na);
    long long lo, hi, srclo, srchi;
    if (face == 2) {
        lo = gidx(g, u, v, 0); hi = gidx(g, u, v, g.nz + 1);
        srclo = gidx(g, u, v, g.nz); srchi = gidx(g, u, v, 1);
    } else if (face == 1) {
        lo = gidx(g, u, 0, v); hi = gidx(g, u, g.ny + 1, v);
        srclo = gidx(g, u, g.ny, v); srchi = gidx(g, u, 1, v);
    } else {
        lo = gidx(g, 0, u, v); hi = gidx(g, g.nx + 1, u, v);
        srclo = gidx(g, g.nx, u, v); srchi = gidx(g, 1, u, v);
    }
    for (int l = 0; l < 3; ++l) {
        F[l * g.P + lo] = F[l * g.P + srclo];
        F[l * g.P + hi] = F[l * g.P + srchi];
    }
}

__global__ void k_set_grid_peri(Geo g, double *X, int face)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int na = face == 0 ? g.ny + 2 : g.nx + 2, nb = face == 2 ? g.ny + 2 : g.nz + 2;
    if (t >= (long long)na * nb) return;
    int u = (int)(t % na), v = (int)(t / na);
    long long lo, hi, srclo, srchi;
    double L, h;
    if (face == 2) {
        lo = gidx(g, u, v, 0); hi = gidx(g, u, v, g.nz + 1);
        srclo = gidx(g, u, v, g.nz); srchi = gidx(g, u, v, 1);
        L = g.Lz; h = g.dz;
    } else if (face == 1) {
        lo = gidx(g, u, 0, v); hi = gidx(g, u, g.ny + 1, v);
        srclo = gidx(g, u, g.ny, v); srchi = gidx(g, u, 1, v);
        L = g.Ly; h = g.dy;
    } else {
        lo = gidx(g, 0, u, v); hi = gidx(g, g.nx + 1, u, v);
        srclo = gidx(g, g.nx, u, v); srchi = gidx(g, 1, u, v);
        L = g.Lx; h = g.dx;
    }
    for (int l = 0; l < 3; ++l) {
        double vlo = X[l * g.P + srclo], vhi = X[l * g.P + srchi];
        if (l == face) {
            vlo = vlo - L - h;
            vhi = vhi + L + h;
        }
        X[l * g.P + lo] = vlo;
        X[l * g.P + hi] = vhi;
    }
}

// ---------------------------------------------------------------- JStokes
// JStokes4th (src/current.cu:413-735): per coordinate plane through the
// point (xy, yz, zx: rows 0, 1, 2 of NA), the octilateral through the 4
// face and 4 edge neighbours, vertices (a, b) = (-1,-1), (0,-1), (1,-1),
// (1,0), (1,1), (0,1), (-1,1), (-1,0) along the plane's axes: trapezoid
// circulation I = sum (B_k + B_k-1)/2 . (x_k - x_k-1), area vector = half
// the corner cross products plus a quarter of the inner diamond's; then
// J = NA^-1 I with every cofactor divided by det NA.
__global__ void __launch_bounds__(kBlock) k_jstokes4th(Geo g, const double *X, const double *B, double *J)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long N = (long long)g.nx * g.ny * g.nz;
    if (t >= N) return;
    int i = (int)(t % g.nx) + 1, j = (int)((t / g.nx) % g.ny) + 1, k = (int)(t / ((long long)g.nx * g.ny)) + 1;
    const long long a0 = gidx(g, i, j, k);
    const long long st[3] = {1, g.sx, g.sxy};
    const int ax[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    const int V[8][2] = {{0, 0}, {1, 0}, {2, 0}, {2, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
    double NA[3][3], II[3];
    for (int q = 0; q < 3; ++q) {
        double x[8][3], b[8][3];
        for (int v = 0; v < 8; ++v) {
            const long long o = a0 + (V[v][0] - 1) * st[ax[q][0]] + (V[v][1] - 1) * st[ax[q][1]];
            ld3(X, g.P, o, x[v]);
            ld3(B, g.P, o, b[v]);
        }
        double dx[8][3], dA[3], dB[3], dC[3], dD[3];   // dx[k-1] = x_k - x_(k-1), k = 1..8
        double I = 0;
        for (int l = 0; l < 3; ++l) {
            for (int e = 1; e <= 8; ++e) dx[e - 1][l] = x[e % 8][l] - x[e - 1][l];
            double sI = (b[1][l] + b[0][l]) / 2 * dx[0][l];
            for (int e = 2; e <= 8; ++e) sI = sI + (b[e % 8][l] + b[e - 1][l]) / 2 * dx[e - 1][l];
            I += sI;
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
        for (int l = 0; l < 3; ++l) NA[q][l] = (c1[l] + c2[l] + c3[l] + c4[l]) / 2;
        cross3(dA, dB, c1);
        cross3(dB, dC, c2);
        cross3(dC, dD, c3);
        cross3(dD, dA, c4);
        for (int l = 0; l < 3; ++l) NA[q][l] += (c1[l] + c2[l] + c3[l] + c4[l]) / 4;
    }
    const double dN = det3(NA);
    double A[3][3];
    A[0][0] = (NA[1][1] * NA[2][2] - NA[1][2] * NA[2][1]) / dN;
    A[0][1] = (NA[0][2] * NA[2][1] - NA[0][1] * NA[2][2]) / dN;
    A[0][2] = (NA[0][1] * NA[1][2] - NA[0][2] * NA[1][1]) / dN;
    A[1][0] = (NA[1][2] * NA[2][0] - NA[1][0] * NA[2][2]) / dN;
    A[1][1] = (NA[0][0] * NA[2][2] - NA[0][2] * NA[2][0]) / dN;
    A[1][2] = (NA[0][2] * NA[1][0] - NA[0][0] * NA[1][2]) / dN;
    A[2][0] = (NA[1][0] * NA[2][1] - NA[1][1] * NA[2][0]) / dN;
    A[2][1] = (NA[0][1] * NA[2][0] - NA[0][0] * NA[2][1]) / dN;
    A[2][2] = (NA[0][0] * NA[1][1] - NA[0][1] * NA[1][0]) / dN;
    for (int l = 0; l < 3; ++l) {
        double sJ = 0;
        for (int m = 0; m < 3; ++m) sJ += A[l][m] * II[m];
        J[l * g.P + a0] = sJ;
    }
}

// JClassic (src/current.cu:10-219): J = curl_x of B = jac B0 / det by the
// chain rule, with central differences on the initial grid X:
// dB_l/dX_a = sum_b d2x_l/dX_a dX_b B0_b / det + jac_lb dB0_b/dX_a / det
// + jac_lb B0_b d(1/det)/dX_a, then dB/dx = dB/dX jac^-1.  The reference's
// operation order is kept (bitwise FP64 parity with the oracle).
__global__ void __launch_bounds__(kBlock) k_jclassic(Geo g, const double *X, const double *B0, const double *det,
                                                     double *J)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long N = (long long)g.nx * g.ny * g.nz;
    if (t >= N) return;
    int i = (int)(t % g.nx) + 1, j = (int)((t / g.nx) % g.ny) + 1, k = (int)(t / ((long long)g.nx * g.ny)) + 1;
    const long long a0 = gidx(g, i, j, k);
    const long long st[3] = {1, g.sx, g.sxy};
    const double h1[3] = {g.dx1, g.dy1, g.dz1};
    double x0[3], b0[3], jac[3][3], d2[3][3][3], dB0[3][3], dD1[3];
    ld3(X, g.P, a0, x0);
    ld3(B0, g.P, a0, b0);
    for (int m = 0; m < 3; ++m) {
        double xp[3], xm[3], bp[3], bm[3];
        ld3(X, g.P, a0 + st[m], xp);
        ld3(X, g.P, a0 - st[m], xm);
        ld3(B0, g.P, a0 + st[m], bp);
        ld3(B0, g.P, a0 - st[m], bm);
        for (int l = 0; l < 3; ++l) {
            jac[l][m] = (xp[l] - xm[l]) * h1[m] / 2;
            d2[l][m][m] = (xp[l] - 2 * x0[l] + xm[l]) * h1[m] * h1[m];
            dB0[l][m] = (bp[l] - bm[l]) * h1[m] / 2;
        }
        dD1[m] = (1 / det[a0 + st[m]] - 1 / det[a0 - st[m]]) * h1[m] / 2;
    }
    for (int m = 0; m < 3; ++m)
        for (int n = m + 1; n < 3; ++n) {
            double pp[3], pm[3], mp[3], mm[3];
            ld3(X, g.P, a0 + st[m] + st[n], pp);
            ld3(X, g.P, a0 + st[m] - st[n], pm);
            ld3(X, g.P, a0 - st[m] + st[n], mp);
            ld3(X, g.P, a0 - st[m] - st[n], mm);
            for (int l = 0; l < 3; ++l) {
                // the y-z term subtracts its middle corners in the other order
                d2[l][m][n] = (m == 1 ? pp[l] - mp[l] - pm[l] + mm[l] : pp[l] - pm[l] - mp[l] + mm[l]) * h1[m] *
                              h1[n] / 4;
                d2[l][n][m] = d2[l][m][n];
            }
        }
    const double det1 = 1 / det[a0];
    double jac1[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            const int c1 = (c + 1) % 3, c2 = (c + 2) % 3, r1 = (r + 1) % 3, r2 = (r + 2) % 3;
            jac1[r][c] = (jac[c1][r1] * jac[c2][r2] - jac[c1][r2] * jac[c2][r1]) * det1;
        }
    double Jv[3] = {0, 0, 0};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double T[3];
            for (int l = 0; l < 3; ++l)
                T[l] = d2[l][a][b] * b0[b] * det1 + jac[l][b] * dB0[b][a] * det1 + jac[l][b] * b0[b] * dD1[a];
            for (int c = 0; c < 3; ++c) {
                const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                Jv[c] += jac1[a][c1] * T[c2] - jac1[a][c2] * T[c1];
            }
        }
    for (int c = 0; c < 3; ++c) J[c * g.P + a0] = Jv[c];
}

__global__ void __launch_bounds__(kBlock) k_jstokes(Geo g, const double *X, const double *B, double *J)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long N = (long long)g.nx * g.ny * g.nz;
    if (t >= N) return;
    int i = (int)(t % g.nx) + 1, j = (int)((t / g.nx) % g.ny) + 1, k = (int)(t / ((long long)g.nx * g.ny)) + 1;
    long long a = gidx(g, i, j, k);
    long long im = a - 1, ip = a + 1, jm = a - g.sx, jp = a + g.sx, km = a - g.sxy, kp = a + g.sxy;
    // diamonds in circulation order (src/current.cu:291-382)
    const long long quad[3][4] = {{im, jm, ip, jp}, {jm, km, jp, kp}, {km, im, kp, ip}};
    double NA[3][3], II[3];
    for (int q = 0; q < 3; ++q) {
        double x[4][3], bv[4][3], d[4][3];
        for (int e = 0; e < 4; ++e) {
            ld3(X, g.P, quad[q][e], x[e]);
            ld3(B, g.P, quad[q][e], bv[e]);
        }
        for (int l = 0; l < 3; ++l) {
            d[0][l] = x[1][l] - x[0][l];
            d[1][l] = x[2][l] - x[1][l];
            d[2][l] = x[3][l] - x[2][l];
            d[3][l] = x[0][l] - x[3][l];
        }
        II[q] = diamond(d, bv, NA[q]);
    }
    double Jv[3];
    solve_na(NA, II, Jv);
    J[a] = Jv[0];
    J[g.P + a] = Jv[1];
    J[2 * g.P + a] = Jv[2];
}

// ------------------------------------------------------------ gradPClassic
__global__ void __launch_bounds__(kBlock) k_gradp(Geo g, const double *X, const double *det, double *G)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long N = (long long)g.nx * g.ny * g.nz;
    if (t >= N) return;
    int i = (int)(t % g.nx) + 1, j = (int)((t / g.nx) % g.ny) + 1, k = (int)(t / ((long long)g.nx * g.ny)) + 1;
    long long a = gidx(g, i, j, k);
    auto x = [&](int di, int dj, int dk, int l) { return X[l * g.P + a + di + dj * g.sx + dk * g.sxy]; };
    double det1 = 1 / det[a];
    double jac[3][3], d2[3][3][3];
    for (int l = 0; l < 3; ++l) {
        jac[l][0] = (x(1, 0, 0, l) - x(-1, 0, 0, l)) * g.dx1 / 2;
        jac[l][1] = (x(0, 1, 0, l) - x(0, -1, 0, l)) * g.dy1 / 2;
        jac[l][2] = (x(0, 0, 1, l) - x(0, 0, -1, l)) * g.dz1 / 2;
    }
    for (int l = 0; l < 3; ++l) {
        d2[l][0][0] = (x(1, 0, 0, l) - 2 * x(0, 0, 0, l) + x(-1, 0, 0, l)) * g.dx1 * g.dx1;
        d2[l][1][1] = (x(0, 1, 0, l) - 2 * x(0, 0, 0, l) + x(0, -1, 0, l)) * g.dy1 * g.dy1;
        d2[l][2][2] = (x(0, 0, 1, l) - 2 * x(0, 0, 0, l) + x(0, 0, -1, l)) * g.dz1 * g.dz1;
        d2[l][0][1] = (x(1, 1, 0, l) - x(1, -1, 0, l) - x(-1, 1, 0, l) + x(-1, -1, 0, l)) * g.dx1 * g.dy1 / 4;
        d2[l][0][2] = (x(1, 0, 1, l) - x(1, 0, -1, l) - x(-1, 0, 1, l) + x(-1, 0, -1, l)) * g.dx1 * g.dz1 / 4;
        d2[l][1][2] = (x(0, 1, 1, l) - x(0, -1, 1, l) - x(0, 1, -1, l) + x(0, -1, -1, l)) * g.dy1 * g.dz1 / 4;
        d2[l][1][0] = d2[l][0][1];
        d2[l][2][0] = d2[l][0][2];
        d2[l][2][1] = d2[l][1][2];
    }
    double gv[3];
    gradp_core(jac, d2, det1, g.beta, gv);
    G[a] = gv[0];
    G[g.P + a] = gv[1];
    G[2 * g.P + a] = gv[2];
}

// --------------------------------------------------------------------- kk
__global__ void __launch_bounds__(kBlock) k_kk(Geo g, const Tableau *tab, int n, double dt, const double *X,
                                               const double *B, const double *J, const double *G, double *K,
                                               double *Xtmp, const double *det, const double *U, double *KU,
                                               double *Utmp)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long N = (long long)g.nx * g.ny * g.nz;
    if (t >= N) return;
    int i = (int)(t % g.nx), j = (int)((t / g.nx) % g.ny), k = (int)(t / ((long long)g.nx * g.ny));
    long long a = gidx(g, i + 1, j + 1, k + 1);
    const long long P3 = 3 * g.P;
    double jv[3], bv[3], kv[3];
    ld3(J, g.P, a, jv);
    ld3(B, g.P, a, bv);
    cross3(jv, bv, kv);
    if (g.inertia) {
        // inertia branch (src/rungeKutta.cu:195-212, 216-246): position
        // slope dt u, velocity slope dt det (J x B) - dt nu u det - dt gradP;
        // no friction profile and no averaging here
        const double dj = det[a];
        for (int l = 0; l < 3; ++l) {
            const double u = U[l * g.P + a];
            K[n * P3 + l * g.P + a] = dt * u;
            kv[l] *= dt * dj;
            kv[l] -= dt * g.nu * u * dj;
            if (g.pressure) kv[l] += -dt * G[l * g.P + a];
            KU[n * P3 + l * g.P + a] = kv[l];
        }
        for (int l = 0; l < 3; ++l) {
            double x = X[l * g.P + a], v = U[l * g.P + a];
            for (int m = 0; m < n + 1; ++m) {
                const double w = tab->b[m + (n + 1) * 6] - tab->b[m + n * 6];
                x += w * K[m * P3 + l * g.P + a];
                v += w * KU[m * P3 + l * g.P + a];
            }
            Xtmp[l * g.P + a] = x;
            Utmp[l * g.P + a] = v;
        }
        return;
    }
    if (g.jxbAver) {
        // jxbAver (src/rungeKutta.cu:172-177, 357-443): (1 - w) J x B plus
        // w/6 J x B of each face neighbour that is an inner point, in the
        // order -x, +x, -y, +y, -z, +z; the +z neighbour's B is read at its
        // own padded index (the reference's (nx+1)(ny+1) pitch there is a
        // known defect, not reproduced)
        const double w6 = g.jxbAverWeight / 6;
        for (int l = 0; l < 3; ++l) kv[l] *= (1 - g.jxbAverWeight);
        const long long st[3] = {1, g.sx, g.sxy};
        const int id[3] = {i, j, k}, nn[3] = {g.nx, g.ny, g.nz};
        for (int m = 0; m < 6; ++m) {
            const int d = m / 2, up = m % 2;
            double jm[3] = {0, 0, 0}, bm[3] = {0, 0, 0}, c[3];
            if (up ? id[d] < nn[d] - 1 : id[d] > 0) {
                const long long o = a + (up ? st[d] : -st[d]);
                ld3(J, g.P, o, jm);
                ld3(B, g.P, o, bm);
            }
            cross3(jm, bm, c);
            for (int l = 0; l < 3; ++l) kv[l] += c[l] * w6;
        }
    }
    double eps = g.epsilonProf ? eps_axis(i, g.nx) * eps_axis(j, g.ny) * eps_axis(k, g.nz) : 1.0;
    for (int l = 0; l < 3; ++l) {
        kv[l] *= dt * eps;
        if (g.pressure) kv[l] += -dt * G[l * g.P + a];
        K[n * P3 + l * g.P + a] = kv[l];
    }
    for (int l = 0; l < 3; ++l) {
        double x = X[l * g.P + a];
        for (int m = 0; m < n + 1; ++m)
            x += (tab->b[m + (n + 1) * 6] - tab->b[m + n * 6]) * K[m * P3 + l * g.P + a];
        Xtmp[l * g.P + a] = x;
    }
}

// --------------------------------------------------------------- xNewStar
__global__ void __launch_bounds__(kBlock) k_xnewstar(Geo g, const Tableau *tab, const double *X, const double *K,
                                                     double *Xnew, double *partials, int *nanflag, const double *U,
                                                     const double *KU, double *Unew)
{
    __shared__ double wmax[kBlock / 64];
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long N = (long long)g.nx * g.ny * g.nz;
    double m = 0;
    if (t < N) {
        int i = (int)(t % g.nx), j = (int)((t / g.nx) % g.ny), k = (int)(t / ((long long)g.nx * g.ny));
        long long a = gidx(g, i + 1, j + 1, k + 1);
        const long long P3 = 3 * g.P;
        bool nan = false;
        for (int l = 0; l < 3; ++l) {
            double xn = X[l * g.P + a], xs = xn;
            if (l < 2 || g.zUpdate)
                for (int s = 0; s < 6; ++s) {
                    double kv = K[s * P3 + l * g.P + a];
                    xn += tab->c[s] * kv;
                    xs += tab->cs[s] * kv;
                }
            Xnew[l * g.P + a] = xn;
            double d = fabs(xn - xs);
            if (d != d) nan = true;
            m = nb_max(m, d);
        }
        if (g.inertia)   // u5 (src/rungeKutta.cu:311-322); no error term
            for (int l = 0; l < 3; ++l) {
                double un = U[l * g.P + a];
                for (int s = 0; s < 6; ++s) un += tab->c[s] * KU[s * P3 + l * g.P + a];
                Unew[l * g.P + a] = un;
            }
        if (nan) atomicOr(nanflag, 1);
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int w = 0; w < kBlock / 64; ++w) r = nb_max(r, wmax[w]);
        partials[blockIdx.x] = r;
    }
}

__global__ void __launch_bounds__(1024) k_max_final(const double *partials, int n, double *out)
{
    __shared__ double wmax[16];
    double m = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) m = nb_max(m, partials[i]);
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r = nb_max(r, wmax[w]);
        out[0] = r;
    }
}

// ------------------------------------------------------- velocity diagnostics
// U2_det (src/maths.cu:683-717): |u|^2 det / 2^(#boundary faces) summed,
// and mag + findDevMax (src/maths.cu:157-164, 58-130): NaN-blind max |u|.
// 512 grid-stride blocks write partials; one thread combines them in order.
constexpr int kUBlocks = 512;

__global__ void __launch_bounds__(kBlock) k_u_diag(Geo g, const double *U, const double *det, double *part)
{
    __shared__ double ss[kBlock], sm[kBlock];
    const long long N = (long long)g.nx * g.ny * g.nz;
    double e = 0, m = 0;
    for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < N; t += (long long)gridDim.x * kBlock) {
        const int i = (int)(t % g.nx), j = (int)((t / g.nx) % g.ny), k = (int)(t / ((long long)g.nx * g.ny));
        const long long a = gidx(g, i + 1, j + 1, k + 1);
        const double u0 = U[a], u1 = U[g.P + a], u2 = U[2 * g.P + a];
        double v = (u0 * u0 + u1 * u1 + u2 * u2) * det[a];
        int bound = (i == 0) + (j == 0) + (k == 0) + (i == g.nx - 1) + (j == g.ny - 1) + (k == g.nz - 1);
        bound = 1 << bound;
        v /= bound;
        e += v;
        m = nb_max(m, sqrt(u0 * u0 + u1 * u1 + u2 * u2));
    }
    ss[threadIdx.x] = e;
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            ss[threadIdx.x] += ss[threadIdx.x + s];
            sm[threadIdx.x] = nb_max(sm[threadIdx.x], sm[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = ss[0];
        part[kUBlocks + blockIdx.x] = sm[0];
    }
}

__global__ void k_u_final(Geo g, const double *part, double *out)
{
    if (threadIdx.x != 0) return;
    double e = 0, m = 0;
    for (int b = 0; b < kUBlocks; ++b) {
        e += part[b];
        m = nb_max(m, part[kUBlocks + b]);
    }
    out[0] = e * g.dx * g.dy * g.dz;
    out[1] = m;
}

__global__ void __launch_bounds__(kBlock) k_inner_aos_to_padded(Geo g, const double *src, double *dst)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long N = (long long)g.nx * g.ny * g.nz;
    if (t >= N) return;
    const int i = (int)(t % g.nx), j = (int)((t / g.nx) % g.ny), k = (int)(t / ((long long)g.nx * g.ny));
    const long long a = gidx(g, i + 1, j + 1, k + 1);
    for (int l = 0; l < 3; ++l) dst[l * g.P + a] = src[3 * t + l];
}

// ------------------------------------------------------------ diagnostics
// Per point: energy density |B|^2 det / 2^(#boundary faces)
// (src/maths.cu:647-680), |J|, |JxB|/|B|^2 (src/maths.cu:168-200) and
// |JxB|/(|J||B|) (0/0 := 0), and the maxima of |JxB| and |J||B| (their
// quotient is the relaxation criterion); grid-stride, per-block partials.
template <class R>
__global__ void __launch_bounds__(kBlock) k_diag(Geo g, const R *B, const R *J, const double *det, double *part,
                                                 int nblk)
{
    __shared__ double red[6][kBlock / 64];
    long long N = (long long)g.nx * g.ny * g.nz;
    double e = 0, jm = 0, r1 = 0, r2 = 0, r3 = 0, r4 = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < N; t += (long long)gridDim.x * blockDim.x) {
        int i = (int)(t % g.nx), j = (int)((t / g.nx) % g.ny), k = (int)(t / ((long long)g.nx * g.ny));
        long long a = gidx(g, i + 1, j + 1, k + 1);
        double b[3] = {(double)B[a], (double)B[g.P + a], (double)B[2 * g.P + a]};
        double jj[3] = {(double)J[a], (double)J[g.P + a], (double)J[2 * g.P + a]};
        double v = (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]) * det[a];
        int bound = (i == 0) + (j == 0) + (k == 0) + (i == g.nx - 1) + (j == g.ny - 1) + (k == g.nz - 1);
        bound = 1 << bound;
        v /= bound;
        e += v;
        double c[3];
        cross3(jj, b, c);
        double nc = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        double nj = sqrt(jj[0] * jj[0] + jj[1] * jj[1] + jj[2] * jj[2]);
        double b2 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
        double nb = sqrt(b2);
        jm = nb_max(jm, nj);
        r1 = nb_max(r1, nc / b2);
        double den = nj * nb;
        r2 = nb_max(r2, den > 0 ? nc / den : 0.0);
        r3 = nb_max(r3, nc);
        r4 = nb_max(r4, den);
    }
    e = wave_sum(e);
    jm = wave_max(jm);
    r1 = wave_max(r1);
    r2 = wave_max(r2);
    r3 = wave_max(r3);
    r4 = wave_max(r4);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = e;
        red[1][w] = jm;
        red[2][w] = r1;
        red[3][w] = r2;
        red[4][w] = r3;
        red[5][w] = r4;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0, a[5] = {0, 0, 0, 0, 0};
        for (int q = 0; q < kBlock / 64; ++q) {
            s += red[0][q];
            for (int m = 0; m < 5; ++m) a[m] = nb_max(a[m], red[m + 1][q]);
        }
        part[blockIdx.x] = s;
        for (int m = 0; m < 5; ++m) part[(m + 1) * nblk + blockIdx.x] = a[m];
    }
}

__global__ void __launch_bounds__(1024) k_diag_final(Geo g, const double *part, int nblk, double *out)
{
    // one wave per quantity, fixed order
    int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (q >= 6) return;
    double v = 0;
    for (int i = lane; i < nblk; i += 64) v = q == 0 ? v + part[i] : nb_max(v, part[q * nblk + i]);
    v = q == 0 ? wave_sum(v) : wave_max(v);
    if (lane == 0) out[q] = q == 0 ? v * g.dx * g.dy * g.dz : v;
}

// ------------------------------------------------------------ layouts
__global__ void k_aos_to_soa_f64(long long P, const double *src, double *dst)
{
    long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P) return;
    for (int l = 0; l < 3; ++l) dst[l * P + a] = src[3 * a + l];
}
__global__ void k_soa_to_aos_f64(long long P, const double *src, double *dst)
{
    long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P) return;
    for (int l = 0; l < 3; ++l) dst[3 * a + l] = src[l * P + a];
}
// F32K positions are (hi, lo) triples: x, y, z of padded point a at 3a
__global__ void k_aos_to_soa_f32k(long long P, const double *src, float *hi, float *lo)
{
    long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P) return;
    for (int l = 0; l < 3; ++l) {
        double v = src[3 * a + l];
        float h = (float)v;
        hi[3 * a + l] = h;
        lo[3 * a + l] = (float)(v - (double)h);
    }
}
__global__ void k_soa_to_aos_f32k(long long P, const float *hi, const float *lo, double *dst)
{
    long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P) return;
    for (int l = 0; l < 3; ++l) dst[3 * a + l] = (double)hi[3 * a + l] + (double)lo[3 * a + l];
}
// FP32 arrays in component order (triples = false) or as triples
__global__ void k_soa_to_aos_f32(long long P, const float *src, double *dst, bool triples)
{
    long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P) return;
    for (int l = 0; l < 3; ++l) dst[3 * a + l] = (double)(triples ? src[3 * a + l] : src[l * P + a]);
}
__global__ void k_aos_to_soa_f32(long long P, const double *src, float *dst, bool triples)
{
    long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P) return;
    for (int l = 0; l < 3; ++l) dst[triples ? 3 * a + l : l * P + a] = (float)src[3 * a + l];
}
template <class R>
__global__ void k_padded_to_inner_aos(Geo g, const R *src, double *dst, bool triples)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long N = (long long)g.nx * g.ny * g.nz;
    if (t >= N) return;
    int i = (int)(t % g.nx), j = (int)((t / g.nx) % g.ny), k = (int)(t / ((long long)g.nx * g.ny));
    long long a = gidx(g, i + 1, j + 1, k + 1);
    for (int l = 0; l < 3; ++l) dst[3 * t + l] = (double)(triples ? src[3 * a + l] : src[l * g.P + a]);
}
// Plain device copy: the reference kernel for HBM byte counters and the
// achievable-bandwidth probe.
// HBM yardsticks.  Each workgroup owns kProbeU * 256 consecutive 16-byte
// words; a lane issues its kProbeU loads (one per 4 KB slice, coalesced
// across the wave) before any store, so kProbeU * 16 bytes per lane are in
// flight, and the grid holds one workgroup per slice set (no grid-stride
// loop whose trips would serialise on the load latency).
constexpr int kProbeU = 8;
__global__ void __launch_bounds__(256) k_copy_probe(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                    long long n)
{
    const long long base = (long long)blockIdx.x * (kProbeU * 256) + threadIdx.x;
    float4 v[kProbeU];
#pragma unroll
    for (int u = 0; u < kProbeU; ++u) {
        const long long i = base + u * 256;
        v[u] = i < n ? src[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < kProbeU; ++u) {
        const long long i = base + u * 256;
        if (i < n) dst[i] = v[u];
    }
}
// read only: the loaded words are summed; the sum is stored only if it is
// NaN (never, for the zero-filled source), which keeps the loads live
__global__ void __launch_bounds__(256) k_read_probe(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                    long long n)
{
    const long long base = (long long)blockIdx.x * (kProbeU * 256) + threadIdx.x;
    float4 v[kProbeU];
#pragma unroll
    for (int u = 0; u < kProbeU; ++u) {
        const long long i = base + u * 256;
        v[u] = i < n ? src[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    float a = 0.f;
#pragma unroll
    for (int u = 0; u < kProbeU; ++u) a += (v[u].x + v[u].y) + (v[u].z + v[u].w);
    if (a != a) dst[base] = make_float4(a, a, a, a);
}
__global__ void __launch_bounds__(256) k_write_probe(float4 *__restrict__ dst, long long n)
{
    const long long base = (long long)blockIdx.x * (kProbeU * 256) + threadIdx.x;
#pragma unroll
    for (int u = 0; u < kProbeU; ++u) {
        const long long i = base + u * 256;
        if (i < n) dst[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}
__global__ void k_fill_f64(double *p, long long n, double v)
{
    long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n) p[a] = v;
}
__global__ void k_fill_f32(float *p, long long n, float v)
{
    long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n) p[a] = v;
}

// ------------------------------------------- allocation-free entry point
// Scratch of gm_rk_xnewstar_noalloc: slot 0 the grid maximum, slot 1 the NaN
// flag (an int), slots 8.. the per-block maxima of k_xnewstar (one per 256
// inner points).  One copy per device, kept until the process ends; every
// call on a device shares it, so calls that overlap on one device race.
constexpr long long kNoAllocScratch = 1LL << 20;   // doubles: 8 MiB
static __device__ __align__(256) double scratch[kNoAllocScratch];

}  // namespace

long long noalloc_scratch_doubles() { return kNoAllocScratch; }

// Device address of `scratch` on `device`, which must be current: looked up
// once per device by hipGetSymbolAddress (host code may not use the array's
// name as a device address); nullptr if the lookup fails.
double *noalloc_scratch(int device)
{
    static std::mutex mu;
    static std::vector<double *> ptr;
    std::lock_guard<std::mutex> lock(mu);
    if (device < 0) return nullptr;
    if ((int)ptr.size() <= device) ptr.resize(device + 1, nullptr);
    if (!ptr[device]) {
        void *a = nullptr;
        if (hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        ptr[device] = static_cast<double *>(a);
    }
    return ptr[device];
}

// ---------------------------------------------------------------- wrappers
static long long ninner(const Geo &g) { return (long long)g.nx * g.ny * g.nz; }

void launch_b_jacb0(hipStream_t s, const Geo &g, const double *X, const double *B0, double *B, double *det)
{
    hipLaunchKernelGGL(k_b_jacb0, dim3(blocks_for(ninner(g))), dim3(kBlock), 0, s, g, X, B0, B, det);
}
void launch_bbound(hipStream_t s, const Geo &g, const double *X, const double *B0, double *B, double *det)
{
    // faces 2, 1, 0 as setBbound (src/boundaries.cu:637-642); each writes
    // disjoint ghost points and reads only X and B0
    const long long per[3] = {2LL * g.ny * g.nz, 2LL * g.nx * g.nz, 2LL * g.nx * g.ny};
    for (int face = 2; face >= 0; --face)
        hipLaunchKernelGGL(k_bbound, dim3(blocks_for(per[face])), dim3(kBlock), 0, s, g, X, B0, B, det, face);
}
void launch_set_peri(hipStream_t s, const Geo &g, double *F)
{
    if (g.zPeri) hipLaunchKernelGGL(k_set_peri, dim3(blocks_for((long long)(g.nx + 2) * (g.ny + 2))), dim3(kBlock), 0, s, g, F, 2);
    if (g.yPeri) hipLaunchKernelGGL(k_set_peri, dim3(blocks_for((long long)(g.nx + 2) * (g.nz + 2))), dim3(kBlock), 0, s, g, F, 1);
    if (g.xPeri) hipLaunchKernelGGL(k_set_peri, dim3(blocks_for((long long)(g.ny + 2) * (g.nz + 2))), dim3(kBlock), 0, s, g, F, 0);
}
void launch_set_grid_peri(hipStream_t s, const Geo &g, double *X)
{
    if (g.zPeri) hipLaunchKernelGGL(k_set_grid_peri, dim3(blocks_for((long long)(g.nx + 2) * (g.ny + 2))), dim3(kBlock), 0, s, g, X, 2);
    if (g.yPeri) hipLaunchKernelGGL(k_set_grid_peri, dim3(blocks_for((long long)(g.nx + 2) * (g.nz + 2))), dim3(kBlock), 0, s, g, X, 1);
    if (g.xPeri) hipLaunchKernelGGL(k_set_grid_peri, dim3(blocks_for((long long)(g.ny + 2) * (g.nz + 2))), dim3(kBlock), 0, s, g, X, 0);
}
void launch_jstokes4th(hipStream_t s, const Geo &g, const double *X, const double *B, double *J)
{
    hipLaunchKernelGGL(k_jstokes4th, dim3(blocks_for(ninner(g))), dim3(kBlock), 0, s, g, X, B, J);
}
void launch_jclassic(hipStream_t s, const Geo &g, const double *X, const double *B0, const double *det, double *J)
{
    hipLaunchKernelGGL(k_jclassic, dim3(blocks_for(ninner(g))), dim3(kBlock), 0, s, g, X, B0, det, J);
}
void launch_jstokes(hipStream_t s, const Geo &g, const double *X, const double *B, double *J)
{
    hipLaunchKernelGGL(k_jstokes, dim3(blocks_for(ninner(g))), dim3(kBlock), 0, s, g, X, B, J);
}
void launch_gradp(hipStream_t s, const Geo &g, const double *X, const double *det, double *G)
{
    hipLaunchKernelGGL(k_gradp, dim3(blocks_for(ninner(g))), dim3(kBlock), 0, s, g, X, det, G);
}
void launch_kk(hipStream_t s, const Geo &g, const Tableau *tab, int n, double dt, const double *X, const double *B,
               const double *J, const double *G, double *K, double *Xtmp, const double *det, const double *U,
               double *KU, double *Utmp)
{
    hipLaunchKernelGGL(k_kk, dim3(blocks_for(ninner(g))), dim3(kBlock), 0, s, g, tab, n, dt, X, B, J, G, K, Xtmp, det,
                       U, KU, Utmp);
}
int launch_xnewstar(hipStream_t s, const Geo &g, const Tableau *tab, const double *X, const double *K, double *Xnew,
                    double *partials, int *nanflag, const double *U, const double *KU, double *Unew)
{
    int nb = blocks_for(ninner(g));
    hipLaunchKernelGGL(k_xnewstar, dim3(nb), dim3(kBlock), 0, s, g, tab, X, K, Xnew, partials, nanflag, U, KU, Unew);
    return nb;
}
void launch_u_diag(hipStream_t s, const Geo &g, const double *U, const double *det, double *part, double *out)
{
    hipLaunchKernelGGL(k_u_diag, dim3(kUBlocks), dim3(kBlock), 0, s, g, U, det, part);
    hipLaunchKernelGGL(k_u_final, dim3(1), dim3(64), 0, s, g, part, out);
}
void launch_inner_aos_to_padded(hipStream_t s, const Geo &g, const double *src, double *dst)
{
    hipLaunchKernelGGL(k_inner_aos_to_padded, dim3(blocks_for(ninner(g))), dim3(kBlock), 0, s, g, src, dst);
}
void launch_max_final(hipStream_t s, const double *partials, int n, double *out)
{
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(1024), 0, s, partials, n, out);
}
template <class R>
void launch_diag(hipStream_t s, const Geo &g, const R *B, const R *J, const double *det, double *scratch,
                 int max_blocks, double *out)
{
    int nb = blocks_for(ninner(g));
    if (nb > max_blocks) nb = max_blocks;
    hipLaunchKernelGGL(k_diag<R>, dim3(nb), dim3(kBlock), 0, s, g, B, J, det, scratch, nb);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(6 * 64), 0, s, g, scratch, nb, out);  // one wave per quantity
}
template void launch_diag<double>(hipStream_t, const Geo &, const double *, const double *, const double *, double *, int, double *);
template void launch_diag<float>(hipStream_t, const Geo &, const float *, const float *, const double *, double *, int, double *);

void launch_aos_to_soa_f64(hipStream_t s, long long P, const double *src, double *dst)
{
    hipLaunchKernelGGL(k_aos_to_soa_f64, dim3(blocks_for(P)), dim3(kBlock), 0, s, P, src, dst);
}
void launch_soa_to_aos_f64(hipStream_t s, long long P, const double *src, double *dst)
{
    hipLaunchKernelGGL(k_soa_to_aos_f64, dim3(blocks_for(P)), dim3(kBlock), 0, s, P, src, dst);
}
void launch_aos_to_soa_f32k(hipStream_t s, long long P, const double *src, float *hi, float *lo)
{
    hipLaunchKernelGGL(k_aos_to_soa_f32k, dim3(blocks_for(P)), dim3(kBlock), 0, s, P, src, hi, lo);
}
void launch_soa_to_aos_f32k(hipStream_t s, long long P, const float *hi, const float *lo, double *dst)
{
    hipLaunchKernelGGL(k_soa_to_aos_f32k, dim3(blocks_for(P)), dim3(kBlock), 0, s, P, hi, lo, dst);
}
void launch_soa_to_aos_f32(hipStream_t s, long long P, const float *src, double *dst, bool triples)
{
    hipLaunchKernelGGL(k_soa_to_aos_f32, dim3(blocks_for(P)), dim3(kBlock), 0, s, P, src, dst, triples);
}
void launch_aos_to_soa_f32(hipStream_t s, long long P, const double *src, float *dst, bool triples)
{
    hipLaunchKernelGGL(k_aos_to_soa_f32, dim3(blocks_for(P)), dim3(kBlock), 0, s, P, src, dst, triples);
}
template <class R>
void launch_padded_to_inner_aos(hipStream_t s, const Geo &g, const R *src, double *dst, bool triples)
{
    hipLaunchKernelGGL(k_padded_to_inner_aos<R>, dim3(blocks_for(ninner(g))), dim3(kBlock), 0, s, g, src, dst, triples);
}
template void launch_padded_to_inner_aos<double>(hipStream_t, const Geo &, const double *, double *, bool);
template void launch_padded_to_inner_aos<float>(hipStream_t, const Geo &, const float *, double *, bool);
void launch_hbm_probe(hipStream_t s, int mode, const void *src, void *dst, long long bytes)
{
    const long long n = bytes / 16;
    const long long blocks = (n + kProbeU * 256 - 1) / (kProbeU * 256);
    if (mode == 1)
        hipLaunchKernelGGL(k_read_probe, dim3((unsigned)blocks), dim3(256), 0, s, (const float4 *)src, (float4 *)dst, n);
    else if (mode == 2)
        hipLaunchKernelGGL(k_write_probe, dim3((unsigned)blocks), dim3(256), 0, s, (float4 *)dst, n);
    else
        hipLaunchKernelGGL(k_copy_probe, dim3((unsigned)blocks), dim3(256), 0, s, (const float4 *)src, (float4 *)dst, n);
}
void launch_fill_f64(hipStream_t s, double *p, long long n, double v)
{
    hipLaunchKernelGGL(k_fill_f64, dim3(blocks_for(n)), dim3(kBlock), 0, s, p, n, v);
}
void launch_fill_f32(hipStream_t s, float *p, long long n, float v)
{
    hipLaunchKernelGGL(k_fill_f32, dim3(blocks_for(n)), dim3(kBlock), 0, s, p, n, v);
}

}  // namespace gmd
