// Fused Cash-Karp stage kernel for gfx950 and the device-side step
// controller.
//
// One launch computes one stage n of one attempted step over the whole
// grid.  A workgroup owns an x-y tile of TX x TY inner points and marches
// through a z-chunk of CZ planes.  Per plane it
//   (b) forms the stage positions x_n = x0 + sum_{m<n} b_nm k_m on the tile
//       plus a 2-point halo (loaded once from HBM, kept in an LDS ring of
//       four planes),
//   (d) evaluates B = jac.B0/det on the tile plus a 1-point face halo (LDS
//       ring of three planes), with the line-tied boundary formula on
//       ghost faces,
//   (a) evaluates the Stokes current J, the force and the slope k_n on the
//       tile points of the previous plane and writes k_n (or, in the last
//       stage, the new positions and the error maximum).
// So per point and stage HBM sees x0, B0 and the earlier slopes once and the
// new slope once; B, J and x_n never go to HBM.
//
// Two arithmetics: F64 follows the reference's operation order (incremental
// stage positions of src/rungeKutta.cu:216-227 included) so it matches the
// FP64 oracle bit for bit; F32K keeps positions as FP32 hi + lo words, forms
// every position difference from hi and lo separately and centres the
// Stokes circulation on the point's own field, with the determinant, error
// norm, energy and dt in FP64.
#include "gm_device.hpp"
#include "gm_fused.hpp"

namespace gmd {

namespace {

__device__ __forceinline__ int ring4(int p) { return (p + 8) & 3; }
__device__ __forceinline__ int ring3(int p) { return (p + 9) % 3; }

// Workgroup id -> logical tile so that the blocks one XCD receives
// (hardware deals ids round-robin over the 8 XCDs) are a contiguous range of
// tiles and share halos in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
    int q = nb >> 3, r = nb & 7;
    int x = bid & 7, w = bid >> 3;
    int base = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
    return base + w;
}

// Two-sum (Knuth): s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e)
{
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// A ghost-face point belongs to the tile that owns its nearest inner point.
__device__ __forceinline__ bool owns(const Geo &g, int gi, int gj, int gk, int ix0, int ixe, int iy0, int iye, int Ks,
                                     int Ke)
{
    int ci = min(max(gi, 1), g.nx), cj = min(max(gj, 1), g.ny), ck = min(max(gk, 1), g.nz);
    return ci - 1 >= ix0 && ci - 1 < ixe && cj - 1 >= iy0 && cj - 1 < iye && ck >= Ks && ck <= Ke;
}

// ghost classification of a padded point: number of ghost coordinates and
// the axis/side of the single one
__device__ __forceinline__ int ghost_axis(const Geo &g, int gi, int gj, int gk, int &side)
{
    int cnt = 0, ax = -1;
    side = 0;
    if (gi == 0 || gi == g.nx + 1) { ++cnt; ax = 0; side = gi != 0; }
    if (gj == 0 || gj == g.ny + 1) { ++cnt; ax = 1; side = gj != 0; }
    if (gk == 0 || gk == g.nz + 1) { ++cnt; ax = 2; side = gk != 0; }
    return cnt == 1 ? ax : (cnt == 0 ? -1 : -2);
}

// Line-tied ghost-face B (src/boundaries.cu:50-110,164-186 and the y/x face
// twins), from double positions: xg ghost, xin first inner point along the
// normal, in-plane stored neighbours um/up (first tangential axis) and vm/vp
// (second).  (u, v) are the inner in-plane indices; tile-edge neighbours are
// rebuilt as x +- h as the reference's 16x16 launch tiles do.
__device__ void ghost_b(const Geo &g, int face, int side, int u, int v, const double xg[3], const double xin[3],
                        const double um_s[3], const double up_s[3], const double vm_s[3], const double vp_s[3],
                        const double b0[3], double B[3], double &det)
{
    const int n = face, t0 = face == 0 ? 1 : 0, t1 = face == 2 ? 1 : 2;
    const int nn[3] = {g.nx, g.ny, g.nz};
    const double hh[3] = {g.dx, g.dy, g.dz}, h1[3] = {g.dx1, g.dy1, g.dz1};
    double x2[3] = {xg[0], xg[1], xg[2]};
    double zg = 0;
    if (face == 2) {
        double zsec = 2 * g.dz * (2 * side - 1) + g.Oz + g.Lz * side;
        for (int l = 0; l < 2; ++l)
            x2[l] = xg[l] - add_dist(g.dist, u, v, xg[2], l) + add_dist(g.dist, u, v, zsec, l);
        x2[2] = xg[2] + (2 * side - 1) * g.dz;
        zg = g.dz * (2 * side - 1) + g.Oz + g.Lz * side;
    } else {
        x2[n] = xg[n] + (2 * side - 1) * hh[n];
    }
    double L0[3], L2[3];
    for (int l = 0; l < 3; ++l) {
        L0[l] = side ? xin[l] : x2[l];
        L2[l] = side ? x2[l] : xin[l];
    }
    auto nb = [&](double out[3], const double st[3], int axis, int sgn, bool edge, int du, int dv) {
        if (!edge) {
            out[0] = st[0]; out[1] = st[1]; out[2] = st[2];
            return;
        }
        out[0] = xg[0]; out[1] = xg[1]; out[2] = xg[2];
        if (face == 2) {
            out[0] = (axis == 0 ? xg[0] + sgn * g.dx : xg[0]) - add_dist(g.dist, u, v, zg, 0) +
                     add_dist(g.dist, u + du, v + dv, zg, 0);
            out[1] = (axis == 1 ? xg[1] + sgn * g.dy : xg[1]) - add_dist(g.dist, u, v, zg, 1) +
                     add_dist(g.dist, u + du, v + dv, zg, 1);
        } else {
            out[axis] = xg[axis] + sgn * hh[axis];
        }
    };
    double um[3], up[3], vm[3], vp[3];
    nb(um, um_s, t0, -1, u % 16 == 0, -1, 0);
    nb(up, up_s, t0, +1, (u % 16 == 15) || (u == nn[t0] - 1), +1, 0);
    nb(vm, vm_s, t1, -1, v % 16 == 0, 0, -1);
    nb(vp, vp_s, t1, +1, (v % 16 == 15) || (v == nn[t1] - 1), 0, +1);
    double jac[3][3], d[3];
    for (int l = 0; l < 3; ++l) d[l] = L2[l] - L0[l];
    jac_col(jac, n, d, h1[n]);
    for (int l = 0; l < 3; ++l) d[l] = up[l] - um[l];
    jac_col(jac, t0, d, h1[t0]);
    for (int l = 0; l < 3; ++l) d[l] = vp[l] - vm[l];
    jac_col(jac, t1, d, h1[t1]);
    double detJac1 = cauchy_b(jac, b0, n, B);
    det = 1 / detJac1;
}

}  // namespace

// ============================================================== F64 kernel
template <int NT, int NPTS>
__global__ void __launch_bounds__(NT) k_fused_f64(FusedArgs A)
{
    extern __shared__ __align__(16) unsigned char smem[];
    double *XL = reinterpret_cast<double *>(smem);            // [4][3][NPTS]
    double *BL = XL + 4 * 3 * NPTS;                            // [3][3][NPTS]
    __shared__ double red[5][NT / 64];

    const Geo &g = A.g;
    Ctl *ctl = A.ctl;
    if (A.check_done && ctl->done) return;
    const int cur = ctl->cur;
    const double dt = ctl->dt;
    const int stage = A.stage;              // -1: fields only
    const int n = stage < 0 ? 0 : stage;
    const double *X = A.X64[cur];
    const double *B0 = A.B0_64;
    const double *K = A.K64;
    const long long P = g.P, P3 = 3 * P;

    int L = xcd_remap(blockIdx.x, A.nblocks);
    int tx = L % A.tiles_x, ty = (L / A.tiles_x) % A.tiles_y, ch = L / (A.tiles_x * A.tiles_y);
    const int ix0 = tx * A.TX, iy0 = ty * A.TY;
    const int Ks = ch * A.CZ + 1, Ke = min(ch * A.CZ + A.CZ, g.nz);
    const int RW = A.TX + 4, RH = A.TY + 4, npts = RW * RH;
    const int ixe = min(ix0 + A.TX, g.nx), iye = min(iy0 + A.TY, g.ny);  // exclusive inner ends

    double dmax = 0, e_sum = 0, jmax = 0, r1 = 0, r2 = 0;
    int nan_seen = 0;

    // stage positions of one rect point on plane q into LDS slot ring4(q)
    auto xn_plane = [&](int q) {
        int s = ring4(q);
        for (int pt = threadIdx.x; pt < npts; pt += NT) {
            int c = pt % RW, r = pt / RW;
            int gi = ix0 - 1 + c, gj = iy0 - 1 + r;
            if (gi < 0 || gi > g.nx + 1 || gj < 0 || gj > g.ny + 1) continue;
            long long a = gidx(g, gi, gj, q);
            bool ghost = gi == 0 || gi == g.nx + 1 || gj == 0 || gj == g.ny + 1 || q == 0 || q == g.nz + 1;
            for (int l = 0; l < 3; ++l) {
                double x = X[l * P + a];
                if (!ghost) {
                    // x_n by the reference's incremental stage updates
                    double kv[5];
                    for (int m = 0; m < n; ++m) kv[m] = K[m * P3 + l * P + a];
                    for (int s2 = 0; s2 < n; ++s2)
                        for (int m = 0; m <= s2; ++m) x += A.tab.db[s2 * 6 + m] * kv[m];
                }
                XL[(s * 3 + l) * NPTS + pt] = x;
            }
        }
    };

    auto XLv = [&](int q, int l, int pt) { return XL[(ring4(q) * 3 + l) * NPTS + pt]; };
    auto BLv = [&](int q, int l, int pt) { return BL[(ring3(q) * 3 + l) * NPTS + pt]; };

    // prologue: planes Ks-2, Ks-1
    for (int q = Ks - 2; q <= Ks - 1; ++q)
        if (q >= 0) xn_plane(q);
    __syncthreads();

    for (int p = Ks - 1; p <= Ke + 1; ++p) {
        // (b) positions two planes ahead of the current output plane
        if (p + 1 <= g.nz + 1 && p + 1 <= Ke + 2) xn_plane(p + 1);
        __syncthreads();

        // (d) B on plane p over tile + face halo
        if (p >= 0 && p <= g.nz + 1) {
            int sb = ring3(p);
            for (int pt = threadIdx.x; pt < npts; pt += NT) {
                int c = pt % RW, r = pt / RW;
                if (c < 1 || c > RW - 2 || r < 1 || r > RH - 2) continue;
                if ((c == 1 || c == RW - 2) && (r == 1 || r == RH - 2)) continue;
                int gi = ix0 - 1 + c, gj = iy0 - 1 + r;
                if (gi < 0 || gi > g.nx + 1 || gj < 0 || gj > g.ny + 1) continue;
                int side;
                int ax = ghost_axis(g, gi, gj, p, side);
                if (ax == -2) continue;
                long long a = gidx(g, gi, gj, p);
                double b0[3] = {B0[a], B0[P + a], B0[2 * P + a]};
                double Bv[3];
                if (ax == -1) {
                    double jac[3][3], d[3];
                    for (int l = 0; l < 3; ++l) d[l] = XLv(p, l, pt + 1) - XLv(p, l, pt - 1);
                    jac_col(jac, 0, d, g.dx1);
                    for (int l = 0; l < 3; ++l) d[l] = XLv(p, l, pt + RW) - XLv(p, l, pt - RW);
                    jac_col(jac, 1, d, g.dy1);
                    for (int l = 0; l < 3; ++l) d[l] = XLv(p + 1, l, pt) - XLv(p - 1, l, pt);
                    jac_col(jac, 2, d, g.dz1);
                    cauchy_b(jac, b0, -1, Bv);
                } else {
                    double xg[3], xin[3], um[3], up[3], vm[3], vp[3], det;
                    int pin, pum, pup, pvm, pvp, qin, qum, qup, qvm, qvp;
                    // neighbour rect offsets and planes by face
                    if (ax == 0) {          // x face: tangents y (rows), z (planes)
                        pin = side ? pt - 1 : pt + 1; qin = p;
                        pum = pt - RW; qum = p; pup = pt + RW; qup = p;
                        pvm = pt; qvm = p - 1; pvp = pt; qvp = p + 1;
                    } else if (ax == 1) {   // y face: tangents x (cols), z (planes)
                        pin = side ? pt - RW : pt + RW; qin = p;
                        pum = pt - 1; qum = p; pup = pt + 1; qup = p;
                        pvm = pt; qvm = p - 1; pvp = pt; qvp = p + 1;
                    } else {                // z face: tangents x, y
                        pin = pt; qin = side ? p - 1 : p + 1;
                        pum = pt - 1; qum = p; pup = pt + 1; qup = p;
                        pvm = pt - RW; qvm = p; pvp = pt + RW; qvp = p;
                    }
                    int u = ax == 0 ? gj - 1 : gi - 1, v = ax == 2 ? gj - 1 : p - 1;
                    bool eum = u % 16 == 0, eup = (u % 16 == 15) || (u == (ax == 0 ? g.ny : g.nx) - 1);
                    int nv = ax == 2 ? g.ny : g.nz;
                    bool evm = v % 16 == 0, evp = (v % 16 == 15) || (v == nv - 1);
                    for (int l = 0; l < 3; ++l) {
                        xg[l] = XLv(p, l, pt);
                        xin[l] = XLv(qin, l, pin);
                        um[l] = eum ? 0 : XLv(qum, l, pum);
                        up[l] = eup ? 0 : XLv(qup, l, pup);
                        vm[l] = evm ? 0 : XLv(qvm, l, pvm);
                        vp[l] = evp ? 0 : XLv(qvp, l, pvp);
                    }
                    ghost_b(g, ax, side, u, v, xg, xin, um, up, vm, vp, b0, Bv, det);
                    if (A.emit && owns(g, gi, gj, p, ix0, ixe, iy0, iye, Ks, Ke)) {
                        for (int l = 0; l < 3; ++l) A.eB64[l * P + a] = Bv[l];
                        A.edet[a] = det;
                    }
                }
                for (int l = 0; l < 3; ++l) BL[(sb * 3 + l) * NPTS + pt] = Bv[l];
            }
        }
        __syncthreads();

        // (a) J, force and slope on plane pj = p - 1
        const int pj = p - 1;
        if (pj >= Ks && pj <= Ke) {
            for (int pt = threadIdx.x; pt < npts; pt += NT) {
                int c = pt % RW, r = pt / RW;
                int gi = ix0 - 1 + c, gj = iy0 - 1 + r;
                if (c < 2 || c > RW - 3 || r < 2 || r > RH - 3) continue;
                if (gi < 1 || gi > ixe || gj < 1 || gj > iye) continue;
                long long a = gidx(g, gi, gj, pj);
                const int nb4[3][4][2] = {
                    // {rect offset kind, plane offset}: 0:-x 1:-y 2:+x 3:+y 4:-z 5:+z
                    {{0, 0}, {1, 0}, {2, 0}, {3, 0}},
                    {{1, 0}, {4, -1}, {3, 0}, {5, 1}},
                    {{4, -1}, {0, 0}, {5, 1}, {2, 0}}};
                const int off[6] = {-1, -RW, +1, +RW, 0, 0};
                double NA[3][3], II[3];
                for (int q = 0; q < 3; ++q) {
                    double x[4][3], bv[4][3], d[4][3];
                    for (int e = 0; e < 4; ++e) {
                        int kind = nb4[q][e][0], dq = nb4[q][e][1];
                        int pp = pt + off[kind];
                        for (int l = 0; l < 3; ++l) {
                            x[e][l] = XLv(pj + dq, l, pp);
                            bv[e][l] = BLv(pj + dq, l, pp);
                        }
                    }
                    for (int l = 0; l < 3; ++l) {
                        d[0][l] = x[1][l] - x[0][l];
                        d[1][l] = x[2][l] - x[1][l];
                        d[2][l] = x[3][l] - x[2][l];
                        d[3][l] = x[0][l] - x[3][l];
                    }
                    II[q] = diamond(d, bv, NA[q]);
                }
                double Jv[3], Bc[3], kv[3];
                solve_na(NA, II, Jv);
                for (int l = 0; l < 3; ++l) Bc[l] = BLv(pj, l, pt);
                cross3(Jv, Bc, kv);
                int ii = gi - 1, jj = gj - 1, kk = pj - 1;
                double eps = g.epsilonProf ? A.epsx[ii] * A.epsy[jj] * A.epsz[kk] : 1.0;
                double gv[3] = {0, 0, 0};
                double detv = 0;
                if (g.pressure || A.emit) {
                    auto x = [&](int di, int dj, int dk, int l) { return XLv(pj + dk, l, pt + di + dj * RW); };
                    double jac[3][3];
                    for (int l = 0; l < 3; ++l) {
                        jac[l][0] = (x(1, 0, 0, l) - x(-1, 0, 0, l)) * g.dx1 / 2;
                        jac[l][1] = (x(0, 1, 0, l) - x(0, -1, 0, l)) * g.dy1 / 2;
                        jac[l][2] = (x(0, 0, 1, l) - x(0, 0, -1, l)) * g.dz1 / 2;
                    }
                    double detJac1 = 1. / det3(jac);
                    detv = 1. / detJac1;
                    if (g.pressure) {
                        double d2[3][3][3];
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
                        double det1 = 1 / detv;
                        gradp_core(jac, d2, det1, g.beta, gv);
                    }
                }
                for (int l = 0; l < 3; ++l) {
                    kv[l] *= dt * eps;
                    if (g.pressure) kv[l] += -dt * gv[l];
                }
                if (A.emit) {
                    for (int l = 0; l < 3; ++l) {
                        A.eB64[l * P + a] = Bc[l];
                        A.eJ64[l * P + a] = Jv[l];
                        if (g.pressure) A.eG64[l * P + a] = gv[l];
                    }
                    A.edet[a] = detv;
                }
                if (A.diag) {
                    double jxb[3];
                    cross3(Jv, Bc, jxb);
                    double b2 = Bc[0] * Bc[0] + Bc[1] * Bc[1] + Bc[2] * Bc[2];
                    double nc = sqrt(jxb[0] * jxb[0] + jxb[1] * jxb[1] + jxb[2] * jxb[2]);
                    double nj = sqrt(Jv[0] * Jv[0] + Jv[1] * Jv[1] + Jv[2] * Jv[2]);
                    double dv = detv;
                    if (!(g.pressure || A.emit)) {
                        // det at the point for the energy weight
                        auto x = [&](int di, int dj, int dk, int l) { return XLv(pj + dk, l, pt + di + dj * RW); };
                        double jac[3][3];
                        for (int l = 0; l < 3; ++l) {
                            jac[l][0] = (x(1, 0, 0, l) - x(-1, 0, 0, l)) * g.dx1 / 2;
                            jac[l][1] = (x(0, 1, 0, l) - x(0, -1, 0, l)) * g.dy1 / 2;
                            jac[l][2] = (x(0, 0, 1, l) - x(0, 0, -1, l)) * g.dz1 / 2;
                        }
                        dv = 1. / (1. / det3(jac));
                    }
                    double v = b2 * dv;
                    int bound = (ii == 0) + (jj == 0) + (kk == 0) + (ii == g.nx - 1) + (jj == g.ny - 1) + (kk == g.nz - 1);
                    v /= (1 << bound);
                    e_sum += v;
                    jmax = nb_max(jmax, nj);
                    r1 = nb_max(r1, nc / b2);
                    double den = nj * sqrt(b2);
                    r2 = nb_max(r2, den > 0 ? nc / den : 0.0);
                }
                if (stage >= 0 && stage < 5) {
                    for (int l = 0; l < 3; ++l) A.K64[stage * P3 + l * P + a] = kv[l];
                } else if (stage == 5) {
                    // x_6 = x_5 + sum_m (b6m - b5m) k_m, then x5/x4 weights
                    // (src/rungeKutta.cu:216-227, 296-309)
                    // x_new to buffer cur+1; the stage-6 position x_6 (where a
                    // rejected attempt leaves xb, src/main.cu:288,378-383) to cur+2
                    double *Xd = A.X64[(cur + 1) % 3], *X6 = A.X64[(cur + 2) % 3];
                    for (int l = 0; l < 3; ++l) {
                        double k6[6];
                        for (int m = 0; m < 5; ++m) k6[m] = K[m * P3 + l * P + a];
                        k6[5] = kv[l];
                        double x6 = XLv(pj, l, pt);
                        for (int m = 0; m < 6; ++m) x6 += A.tab.db[5 * 6 + m] * k6[m];
                        X6[l * P + a] = x6;
                        double xn = x6, xs = x6;
                        if (l < 2 || g.zUpdate)
                            for (int m = 0; m < 6; ++m) {
                                xn += A.tab.c[m] * k6[m];
                                xs += A.tab.cs[m] * k6[m];
                            }
                        Xd[l * P + a] = xn;
                        double dd = fabs(xn - xs);
                        if (dd != dd) nan_seen = 1;
                        dmax = nb_max(dmax, dd);
                    }
                }
            }
        }
        __syncthreads();
    }

    // block partials
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (stage == 5) {
        dmax = wave_max(dmax);
        if (lane == 0) red[0][w] = dmax;
        if (nan_seen) atomicOr(&ctl->nan_seen, 1);
    }
    if (A.diag) {
        e_sum = wave_sum(e_sum);
        jmax = wave_max(jmax);
        r1 = wave_max(r1);
        r2 = wave_max(r2);
        if (lane == 0) {
            red[1][w] = e_sum;
            red[2][w] = jmax;
            red[3][w] = r1;
            red[4][w] = r2;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (stage == 5) {
            double m = 0;
            for (int q = 0; q < NT / 64; ++q) m = nb_max(m, red[0][q]);
            A.part_delta[blockIdx.x] = m;
        }
        if (A.diag) {
            double s = 0, a1 = 0, a2 = 0, a3 = 0;
            for (int q = 0; q < NT / 64; ++q) {
                s += red[1][q];
                a1 = nb_max(a1, red[2][q]);
                a2 = nb_max(a2, red[3][q]);
                a3 = nb_max(a3, red[4][q]);
            }
            A.part_diag[blockIdx.x] = s;
            A.part_diag[A.nblocks + blockIdx.x] = a1;
            A.part_diag[2 * A.nblocks + blockIdx.x] = a2;
            A.part_diag[3 * A.nblocks + blockIdx.x] = a3;
        }
    }
}

// ============================================================= F32K kernel
template <int NT, int NPTS>
__global__ void __launch_bounds__(NT) k_fused_f32k(FusedArgs A)
{
    extern __shared__ __align__(16) unsigned char smem[];
    float *XH = reinterpret_cast<float *>(smem);     // [4][3][NPTS]
    float *XS = XH + 4 * 3 * NPTS;                    // [4][3][NPTS]
    float *BL = XS + 4 * 3 * NPTS;                    // [3][3][NPTS]
    __shared__ double red[5][NT / 64];

    const Geo &g = A.g;
    Ctl *ctl = A.ctl;
    if (A.check_done && ctl->done) return;
    const int cur = ctl->cur;
    const float dt = (float)ctl->dt;
    const int stage = A.stage;
    const int n = stage < 0 ? 0 : stage;
    const float *XHg = A.Xh[cur], *XLg = A.Xl[cur];
    const float *B0 = A.B0_32;
    const float *K = A.K32;
    const long long P = g.P, P3 = 3 * P;
    const float h1x = (float)g.dx1, h1y = (float)g.dy1, h1z = (float)g.dz1;

    int L = xcd_remap(blockIdx.x, A.nblocks);
    int tx = L % A.tiles_x, ty = (L / A.tiles_x) % A.tiles_y, ch = L / (A.tiles_x * A.tiles_y);
    const int ix0 = tx * A.TX, iy0 = ty * A.TY;
    const int Ks = ch * A.CZ + 1, Ke = min(ch * A.CZ + A.CZ, g.nz);
    const int RW = A.TX + 4, RH = A.TY + 4, npts = RW * RH;
    const int ixe = min(ix0 + A.TX, g.nx), iye = min(iy0 + A.TY, g.ny);

    float bn[5];
    for (int m = 0; m < 5; ++m) bn[m] = (float)A.tab.b[n * 6 + m];

    double dmax = 0, e_sum = 0, jmax = 0, r1 = 0, r2 = 0;
    int nan_seen = 0;

    auto xn_plane = [&](int q) {
        int s = ring4(q);
        for (int pt = threadIdx.x; pt < npts; pt += NT) {
            int c = pt % RW, r = pt / RW;
            int gi = ix0 - 1 + c, gj = iy0 - 1 + r;
            if (gi < 0 || gi > g.nx + 1 || gj < 0 || gj > g.ny + 1) continue;
            long long a = gidx(g, gi, gj, q);
            bool ghost = gi == 0 || gi == g.nx + 1 || gj == 0 || gj == g.ny + 1 || q == 0 || q == g.nz + 1;
            for (int l = 0; l < 3; ++l) {
                float sv = XLg[l * P + a];
                if (!ghost)
                    for (int m = 0; m < n; ++m) sv += bn[m] * K[m * P3 + l * P + a];
                XH[(s * 3 + l) * NPTS + pt] = XHg[l * P + a];
                XS[(s * 3 + l) * NPTS + pt] = sv;
            }
        }
    };
    auto xh = [&](int q, int l, int pt) { return XH[(ring4(q) * 3 + l) * NPTS + pt]; };
    auto xs = [&](int q, int l, int pt) { return XS[(ring4(q) * 3 + l) * NPTS + pt]; };
    // compensated difference x(qa,pa) - x(qb,pb)
    auto dif = [&](int qa, int pa, int qb, int pb, int l) {
        return (xh(qa, l, pa) - xh(qb, l, pb)) + (xs(qa, l, pa) - xs(qb, l, pb));
    };
    auto xd = [&](int q, int l, int pt) { return (double)xh(q, l, pt) + (double)xs(q, l, pt); };
    auto BLv = [&](int q, int l, int pt) { return BL[(ring3(q) * 3 + l) * NPTS + pt]; };

    for (int q = Ks - 2; q <= Ks - 1; ++q)
        if (q >= 0) xn_plane(q);
    __syncthreads();

    for (int p = Ks - 1; p <= Ke + 1; ++p) {
        if (p + 1 <= g.nz + 1 && p + 1 <= Ke + 2) xn_plane(p + 1);
        __syncthreads();

        if (p >= 0 && p <= g.nz + 1) {
            int sb = ring3(p);
            for (int pt = threadIdx.x; pt < npts; pt += NT) {
                int c = pt % RW, r = pt / RW;
                if (c < 1 || c > RW - 2 || r < 1 || r > RH - 2) continue;
                if ((c == 1 || c == RW - 2) && (r == 1 || r == RH - 2)) continue;
                int gi = ix0 - 1 + c, gj = iy0 - 1 + r;
                if (gi < 0 || gi > g.nx + 1 || gj < 0 || gj > g.ny + 1) continue;
                int side;
                int ax = ghost_axis(g, gi, gj, p, side);
                if (ax == -2) continue;
                long long a = gidx(g, gi, gj, p);
                float Bv[3];
                if (ax == -1) {
                    float jac[3][3], d[3];
                    for (int l = 0; l < 3; ++l) d[l] = dif(p, pt + 1, p, pt - 1, l);
                    jac_col(jac, 0, d, h1x);
                    for (int l = 0; l < 3; ++l) d[l] = dif(p, pt + RW, p, pt - RW, l);
                    jac_col(jac, 1, d, h1y);
                    for (int l = 0; l < 3; ++l) d[l] = dif(p + 1, pt, p - 1, pt, l);
                    jac_col(jac, 2, d, h1z);
                    double jd[3][3];
                    for (int u = 0; u < 3; ++u)
                        for (int v = 0; v < 3; ++v) jd[u][v] = jac[u][v];
                    double detJac1 = 1.0 / det3(jd);
                    float b0[3] = {B0[a], B0[P + a], B0[2 * P + a]};
                    for (int l = 0; l < 3; ++l)
                        Bv[l] = (float)((double)(jac[l][0] * b0[0] + jac[l][1] * b0[1] + jac[l][2] * b0[2]) * detJac1);
                } else {
                    double xg[3], xin[3], um[3], up[3], vm[3], vp[3], b0[3], Bd[3], det;
                    int pin, pum, pup, pvm, pvp, qin, qum, qup, qvm, qvp;
                    if (ax == 0) {
                        pin = side ? pt - 1 : pt + 1; qin = p;
                        pum = pt - RW; qum = p; pup = pt + RW; qup = p;
                        pvm = pt; qvm = p - 1; pvp = pt; qvp = p + 1;
                    } else if (ax == 1) {
                        pin = side ? pt - RW : pt + RW; qin = p;
                        pum = pt - 1; qum = p; pup = pt + 1; qup = p;
                        pvm = pt; qvm = p - 1; pvp = pt; qvp = p + 1;
                    } else {
                        pin = pt; qin = side ? p - 1 : p + 1;
                        pum = pt - 1; qum = p; pup = pt + 1; qup = p;
                        pvm = pt - RW; qvm = p; pvp = pt + RW; qvp = p;
                    }
                    int u = ax == 0 ? gj - 1 : gi - 1, v = ax == 2 ? gj - 1 : p - 1;
                    bool eum = u % 16 == 0, eup = (u % 16 == 15) || (u == (ax == 0 ? g.ny : g.nx) - 1);
                    int nv = ax == 2 ? g.ny : g.nz;
                    bool evm = v % 16 == 0, evp = (v % 16 == 15) || (v == nv - 1);
                    for (int l = 0; l < 3; ++l) {
                        xg[l] = xd(p, l, pt);
                        xin[l] = xd(qin, l, pin);
                        um[l] = eum ? 0 : xd(qum, l, pum);
                        up[l] = eup ? 0 : xd(qup, l, pup);
                        vm[l] = evm ? 0 : xd(qvm, l, pvm);
                        vp[l] = evp ? 0 : xd(qvp, l, pvp);
                        b0[l] = B0[l * P + a];
                    }
                    ghost_b(g, ax, side, u, v, xg, xin, um, up, vm, vp, b0, Bd, det);
                    for (int l = 0; l < 3; ++l) Bv[l] = (float)Bd[l];
                    if (A.emit && owns(g, gi, gj, p, ix0, ixe, iy0, iye, Ks, Ke)) {
                        for (int l = 0; l < 3; ++l) A.eB32[l * P + a] = Bv[l];
                        A.edet[a] = det;
                    }
                }
                for (int l = 0; l < 3; ++l) BL[(sb * 3 + l) * NPTS + pt] = Bv[l];
            }
        }
        __syncthreads();

        const int pj = p - 1;
        if (pj >= Ks && pj <= Ke) {
            for (int pt = threadIdx.x; pt < npts; pt += NT) {
                int c = pt % RW, r = pt / RW;
                int gi = ix0 - 1 + c, gj = iy0 - 1 + r;
                if (c < 2 || c > RW - 3 || r < 2 || r > RH - 3) continue;
                if (gi < 1 || gi > ixe || gj < 1 || gj > iye) continue;
                long long a = gidx(g, gi, gj, pj);
                const int nb4[3][4][2] = {{{0, 0}, {1, 0}, {2, 0}, {3, 0}},
                                          {{1, 0}, {4, -1}, {3, 0}, {5, 1}},
                                          {{4, -1}, {0, 0}, {5, 1}, {2, 0}}};
                const int off[6] = {-1, -RW, +1, +RW, 0, 0};
                float Bc[3];
                for (int l = 0; l < 3; ++l) Bc[l] = BLv(pj, l, pt);
                float NA[3][3], II[3];
                for (int q = 0; q < 3; ++q) {
                    int pv[4], qv[4];
                    for (int e = 0; e < 4; ++e) {
                        pv[e] = pt + off[nb4[q][e][0]];
                        qv[e] = pj + nb4[q][e][1];
                    }
                    float d[4][3], bv[4][3];
                    for (int l = 0; l < 3; ++l) {
                        for (int e = 0; e < 4; ++e) {
                            int e1 = (e + 1) & 3;
                            d[e][l] = dif(qv[e1], pv[e1], qv[e], pv[e], l);
                            // centred on the point's own field: exact in real
                            // arithmetic because the edges close
                            bv[e][l] = BLv(qv[e], l, pv[e]) - Bc[l];
                        }
                    }
                    II[q] = diamond(d, bv, NA[q]);
                }
                float Jv[3], kv[3];
                solve_na(NA, II, Jv);
                cross3(Jv, Bc, kv);
                int ii = gi - 1, jj = gj - 1, kk = pj - 1;
                float eps = g.epsilonProf ? (float)(A.epsx[ii] * A.epsy[jj] * A.epsz[kk]) : 1.0f;
                float gv[3] = {0, 0, 0};
                double detv = 0;
                if (g.pressure || A.emit || A.diag) {
                    float jac[3][3], d[3];
                    for (int l = 0; l < 3; ++l) d[l] = dif(pj, pt + 1, pj, pt - 1, l);
                    jac_col(jac, 0, d, h1x);
                    for (int l = 0; l < 3; ++l) d[l] = dif(pj, pt + RW, pj, pt - RW, l);
                    jac_col(jac, 1, d, h1y);
                    for (int l = 0; l < 3; ++l) d[l] = dif(pj + 1, pt, pj - 1, pt, l);
                    jac_col(jac, 2, d, h1z);
                    double jd[3][3];
                    for (int u = 0; u < 3; ++u)
                        for (int v = 0; v < 3; ++v) jd[u][v] = jac[u][v];
                    detv = det3(jd);
                    if (g.pressure) {
                        // second differences as differences of differences
                        auto dd2 = [&](int l, int dia, int dja, int dka, int dib, int djb, int dkb) {
                            return dif(pj + dka, pt + dia + dja * RW, pj + dkb, pt + dib + djb * RW, l);
                        };
                        float d2[3][3][3];
                        for (int l = 0; l < 3; ++l) {
                            d2[l][0][0] = (dd2(l, 1, 0, 0, 0, 0, 0) - dd2(l, 0, 0, 0, -1, 0, 0)) * h1x * h1x;
                            d2[l][1][1] = (dd2(l, 0, 1, 0, 0, 0, 0) - dd2(l, 0, 0, 0, 0, -1, 0)) * h1y * h1y;
                            d2[l][2][2] = (dd2(l, 0, 0, 1, 0, 0, 0) - dd2(l, 0, 0, 0, 0, 0, -1)) * h1z * h1z;
                            d2[l][0][1] = (dd2(l, 1, 1, 0, 1, -1, 0) - dd2(l, -1, 1, 0, -1, -1, 0)) * h1x * h1y / 4;
                            d2[l][0][2] = (dd2(l, 1, 0, 1, 1, 0, -1) - dd2(l, -1, 0, 1, -1, 0, -1)) * h1x * h1z / 4;
                            d2[l][1][2] = (dd2(l, 0, 1, 1, 0, -1, 1) - dd2(l, 0, 1, -1, 0, -1, -1)) * h1y * h1z / 4;
                            d2[l][1][0] = d2[l][0][1];
                            d2[l][2][0] = d2[l][0][2];
                            d2[l][2][1] = d2[l][1][2];
                        }
                        gradp_core(jac, d2, (float)(1.0 / detv), (float)g.beta, gv);
                    }
                }
                for (int l = 0; l < 3; ++l) {
                    kv[l] *= dt * eps;
                    if (g.pressure) kv[l] += -dt * gv[l];
                }
                if (A.emit) {
                    for (int l = 0; l < 3; ++l) {
                        A.eB32[l * P + a] = Bc[l];
                        A.eJ32[l * P + a] = Jv[l];
                        if (g.pressure) A.eG32[l * P + a] = gv[l];
                    }
                    A.edet[a] = detv;
                }
                if (A.diag) {
                    double Bd[3] = {Bc[0], Bc[1], Bc[2]}, Jd[3] = {Jv[0], Jv[1], Jv[2]}, jxb[3];
                    cross3(Jd, Bd, jxb);
                    double b2 = Bd[0] * Bd[0] + Bd[1] * Bd[1] + Bd[2] * Bd[2];
                    double nc = sqrt(jxb[0] * jxb[0] + jxb[1] * jxb[1] + jxb[2] * jxb[2]);
                    double nj = sqrt(Jd[0] * Jd[0] + Jd[1] * Jd[1] + Jd[2] * Jd[2]);
                    double v = b2 * detv;
                    int bound = (ii == 0) + (jj == 0) + (kk == 0) + (ii == g.nx - 1) + (jj == g.ny - 1) + (kk == g.nz - 1);
                    v /= (1 << bound);
                    e_sum += v;
                    jmax = nb_max(jmax, nj);
                    r1 = nb_max(r1, nc / b2);
                    double den = nj * sqrt(b2);
                    r2 = nb_max(r2, den > 0 ? nc / den : 0.0);
                }
                if (stage >= 0 && stage < 5) {
                    for (int l = 0; l < 3; ++l) A.K32[stage * P3 + l * P + a] = kv[l];
                } else if (stage == 5) {
                    float *Xhd = A.Xh[cur ^ 1], *Xld = A.Xl[cur ^ 1];
                    for (int l = 0; l < 3; ++l) {
                        float hi = XHg[l * P + a], lo = XLg[l * P + a];
                        if (l < 2 || g.zUpdate) {
                            float k6[6];
                            for (int m = 0; m < 5; ++m) k6[m] = K[m * P3 + l * P + a];
                            k6[5] = kv[l];
                            float dC = 0;
                            double err = 0;
                            for (int m = 0; m < 6; ++m) {
                                dC += (float)A.tab.c[m] * k6[m];
                                err += A.tab.ce[m] * (double)k6[m];
                            }
                            // (hi, lo) + dC with the rounding error of hi + dC kept
                            float s, e;
                            two_sum(hi, dC, s, e);
                            float lo2 = lo + e;
                            float hi2 = s + lo2;
                            lo = lo2 - (hi2 - s);
                            hi = hi2;
                            double dd = fabs(err);
                            if (dd != dd) nan_seen = 1;
                            dmax = nb_max(dmax, dd);
                        }
                        Xhd[l * P + a] = hi;
                        Xld[l * P + a] = lo;
                    }
                }
            }
        }
        __syncthreads();
    }

    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (stage == 5) {
        dmax = wave_max(dmax);
        if (lane == 0) red[0][w] = dmax;
        if (nan_seen) atomicOr(&ctl->nan_seen, 1);
    }
    if (A.diag) {
        e_sum = wave_sum(e_sum);
        jmax = wave_max(jmax);
        r1 = wave_max(r1);
        r2 = wave_max(r2);
        if (lane == 0) {
            red[1][w] = e_sum;
            red[2][w] = jmax;
            red[3][w] = r1;
            red[4][w] = r2;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (stage == 5) {
            double m = 0;
            for (int q = 0; q < NT / 64; ++q) m = nb_max(m, red[0][q]);
            A.part_delta[blockIdx.x] = m;
        }
        if (A.diag) {
            double s = 0, a1 = 0, a2 = 0, a3 = 0;
            for (int q = 0; q < NT / 64; ++q) {
                s += red[1][q];
                a1 = nb_max(a1, red[2][q]);
                a2 = nb_max(a2, red[3][q]);
                a3 = nb_max(a3, red[4][q]);
            }
            A.part_diag[blockIdx.x] = s;
            A.part_diag[A.nblocks + blockIdx.x] = a1;
            A.part_diag[2 * A.nblocks + blockIdx.x] = a2;
            A.part_diag[3 * A.nblocks + blockIdx.x] = a3;
        }
    }
}

// ============================================================ controller
// One workgroup: reduces the stage partials, applies accept/reject and the
// dt update of src/main.cu:321-383 and logs the attempt.
__global__ void __launch_bounds__(1024) k_controller(CtlArgs A)
{
    __shared__ double red[5][16];
    Ctl *ctl = A.ctl;
    if (ctl->done) return;
    const int nb = A.nblocks;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double m = 0, e = 0, jm = 0, r1 = 0, r2 = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        m = nb_max(m, A.part_delta[i]);
        if (A.diag) {
            e += A.part_diag[i];
            jm = nb_max(jm, A.part_diag[nb + i]);
            r1 = nb_max(r1, A.part_diag[2 * nb + i]);
            r2 = nb_max(r2, A.part_diag[3 * nb + i]);
        }
    }
    m = wave_max(m);
    e = wave_sum(e);
    jm = wave_max(jm);
    r1 = wave_max(r1);
    r2 = wave_max(r2);
    if (lane == 0) {
        red[0][w] = m;
        red[1][w] = e;
        red[2][w] = jm;
        red[3][w] = r1;
        red[4][w] = r2;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double md = 0, es = 0, jmx = 0, rr1 = 0, rr2 = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
        md = nb_max(md, red[0][q]);
        es += red[1][q];
        jmx = nb_max(jmx, red[2][q]);
        rr1 = nb_max(rr1, red[3][q]);
        rr2 = nb_max(rr2, red[4][q]);
    }
    StepRecord rec;
    rec.t = ctl->t;
    rec.dt_used = ctl->dt;
    rec.maxDelta = md;
    rec.nan_seen = ctl->nan_seen;
    if (A.diag) {
        ctl->energy = es * A.cell;
        ctl->jmax = jmx;
        ctl->jxb_b2 = rr1;
        ctl->ratio = rr2;
        ctl->diag_valid = 1;
    }
    rec.ratio = A.diag ? rr2 : -1;
    rec.energy = A.diag ? es * A.cell : -1;
    rec.jxb_b2 = A.diag ? rr1 : -1;
    rec.jmax = A.diag ? jmx : -1;
    // the state at the start of this attempt already meets the criterion:
    // stop there and drop the attempt
    if (A.diag && ctl->stop_ratio > 0 && rr2 <= ctl->stop_ratio) {
        ctl->converged = 1;
        ctl->done = 1;
        ctl->stop_reason = 3;
        ctl->accepted = 0;
        rec.accepted = -1;
        if (A.log) A.log[ctl->n_attempt % A.log_cap] = rec;
        return;
    }
    double dt = ctl->dt;
    ctl->dt_used = dt;
    ctl->maxDelta = md;
    int acc = md < A.maxError;
    ctl->accepted = acc;
    if (acc) {
        ctl->t += dt;
        ctl->cur = A.nbuf == 3 ? (ctl->cur + 1) % 3 : ctl->cur ^ 1;
        ctl->n_accept += 1;
    } else if (A.nbuf == 3) {
        ctl->cur = (ctl->cur + 2) % 3;
    }
    if (md > 0) ctl->dt = dt * pow(0.9 * A.maxError / md, 0.2);
    rec.accepted = acc;
    rec.t = ctl->t;
    if (A.log) A.log[ctl->n_attempt % A.log_cap] = rec;
    ctl->n_attempt += 1;
    if (ctl->n_accept >= ctl->n_cycle) { ctl->done = 1; ctl->stop_reason = 1; }
    else if (!(ctl->dt > A.dtMin)) { ctl->done = 1; ctl->stop_reason = 2; }
    else if (ctl->n_attempt >= ctl->max_attempt) { ctl->done = 1; ctl->stop_reason = 4; }
}

// ================================================================ launch
static bool g_attr_done[8];

template <int NT, int NPTS>
static int launch_fused_t(hipStream_t s, const FusedArgs &A, int f32k)
{
    size_t lds = f32k ? (size_t)(4 * 3 * NPTS * 2 + 3 * 3 * NPTS) * sizeof(float)
                      : (size_t)(4 * 3 * NPTS + 3 * 3 * NPTS) * sizeof(double);
    if (f32k) {
        if (!g_attr_done[1]) {
            (void)hipFuncSetAttribute((const void *)k_fused_f32k<NT, NPTS>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            g_attr_done[1] = true;
        }
        hipLaunchKernelGGL((k_fused_f32k<NT, NPTS>), dim3(A.nblocks), dim3(NT), lds, s, A);
    } else {
        if (!g_attr_done[0]) {
            (void)hipFuncSetAttribute((const void *)k_fused_f64<NT, NPTS>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            g_attr_done[0] = true;
        }
        hipLaunchKernelGGL((k_fused_f64<NT, NPTS>), dim3(A.nblocks), dim3(NT), lds, s, A);
    }
    return 0;
}

int fused_max_points() { return kFusedPts; }
size_t fused_lds_bytes(int f32k)
{
    return f32k ? (size_t)(4 * 3 * kFusedPts * 2 + 3 * 3 * kFusedPts) * sizeof(float)
                : (size_t)(4 * 3 * kFusedPts + 3 * 3 * kFusedPts) * sizeof(double);
}

void launch_fused(hipStream_t s, const FusedArgs &A, int f32k)
{
    launch_fused_t<kFusedThreads, kFusedPts>(s, A, f32k);
}

void launch_controller(hipStream_t s, const CtlArgs &A)
{
    hipLaunchKernelGGL(k_controller, dim3(1), dim3(1024), 0, s, A);
}

}  // namespace gmd
