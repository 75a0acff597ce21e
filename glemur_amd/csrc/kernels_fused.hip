// Fused Cash-Karp stage kernel for gfx950 and the device-side step
// controller.
//
// One launch computes one stage n of one attempted step over the whole
// grid.  A workgroup owns an x-y tile of TX x TY inner points and marches
// through a z-chunk of CZ planes; each thread owns one column of the
// (TX+4) x (TY+4) rect (tile plus a 2-point halo) for the whole march.  Per
// plane p a thread
//   (1) forms the stage position x_n(p+1) = x0 + sum_m b_nm k_m of its column
//       from inputs loaded one iteration earlier and publishes it in an LDS
//       ring of planes,
//   (0) issues the HBM loads of plane p+2 (x0, earlier slopes, B0) into
//       registers; they complete while the rest of the iteration computes,
//   (2) evaluates B = jac.B0/det at (column, p) on the tile plus a 1-point
//       face halo -- in-plane neighbours from LDS, z neighbours from its own
//       register queue -- with the line-tied formula on ghost faces, and
//       publishes it in an LDS ring of three planes,
//   -- one workgroup barrier --
//   (3) evaluates the Stokes current, the force and the slope k_n at
//       (column, p-1) on tile points and writes k_n (the last stage writes
//       the candidate positions and the error maximum instead).
// So per point and stage HBM sees x0, B0 and the earlier slopes once and the
// new slope once; B, J and x_n stay on chip.
//
// This file holds the FP64 kernel, which follows the reference's operation
// order (incremental stage positions of src/rungeKutta.cu:216-227 included)
// and matches the FP64 oracle bit for bit; the FP32 kernel with compensated
// positions is in kernels_f32.hip.
#include "gm_device.hpp"
#include "gm_fused.hpp"
#include "fused_common.hpp"

namespace gmd {

__constant__ TableauF c_tab;

namespace {

// ----------------------------------------------------------- ghost setup
// Tangential Jacobian columns of a ghost-face point: in-plane neighbours are
// read from the ghost layer, or rebuilt as x +- h (plus addDist on z faces)
// at the edges of the reference's 16x16 launch tiles and at the last row
// (src/boundaries.cu:70-110, 210-238, 338-366).
template <bool F32>
__global__ void k_ghost_setup(FusedArgs A, GhostPt *out)
{
    const Geo &g = A.g;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long cnt[6];
    long long total = 0;
    for (int fs = 0; fs < 6; ++fs) {
        int f = fs >> 1;
        cnt[fs] = f == 0 ? (long long)g.ny * g.nz : f == 1 ? (long long)g.nx * g.nz : (long long)g.nx * g.ny;
        total += cnt[fs];
    }
    if (t >= total) return;
    int fs = 0;
    long long r = t;
    while (r >= cnt[fs]) r -= cnt[fs++];
    const int face = fs >> 1, side = fs & 1;
    const int nu = nface(g, face);
    const int u = (int)(r % nu), v = (int)(r / nu);
    const int n = face, t0 = face == 0 ? 1 : 0, t1 = face == 2 ? 1 : 2;
    const int nn[3] = {g.nx, g.ny, g.nz};
    const long long st[3] = {1, g.sx, g.sxy};
    const double hh[3] = {g.dx, g.dy, g.dz}, h1[3] = {g.dx1, g.dy1, g.dz1};
    int c[3];
    c[n] = side ? nn[n] + 1 : 0;
    c[t0] = u + 1;
    c[t1] = v + 1;
    const long long a = c[0] * st[0] + c[1] * st[1] + c[2] * st[2];
    const int cur = A.ctl->cur;
    auto pos = [&](long long q, double x[3]) {
        for (int l = 0; l < 3; ++l) {
            if (F32)
                x[l] = (double)A.Xh[cur][3 * q + l] + (double)A.Xl[cur][3 * q + l];   // triples
            else
                x[l] = A.X64[cur][l * g.P + q];
        }
    };
    double xg[3];
    pos(a, xg);
    double x2[3] = {xg[0], xg[1], xg[2]}, zg = 0;
    if (face == 2) {
        double zsec = 2 * g.dz * (2 * side - 1) + g.Oz + g.Lz * side;
        for (int l = 0; l < 2; ++l)
            x2[l] = xg[l] - add_dist(g.dist, u, v, xg[2], l) + add_dist(g.dist, u, v, zsec, l);
        x2[2] = xg[2] + (2 * side - 1) * g.dz;
        zg = g.dz * (2 * side - 1) + g.Oz + g.Lz * side;
    } else {
        x2[n] = xg[n] + (2 * side - 1) * hh[n];
    }
    auto nb = [&](double o[3], int axis, int sgn, bool edge, long long off, int du, int dv) {
        if (!edge) {
            pos(a + off, o);
            return;
        }
        o[0] = xg[0]; o[1] = xg[1]; o[2] = xg[2];
        if (face == 2) {
            o[0] = (axis == 0 ? xg[0] + sgn * g.dx : xg[0]) - add_dist(g.dist, u, v, zg, 0) +
                   add_dist(g.dist, u + du, v + dv, zg, 0);
            o[1] = (axis == 1 ? xg[1] + sgn * g.dy : xg[1]) - add_dist(g.dist, u, v, zg, 1) +
                   add_dist(g.dist, u + du, v + dv, zg, 1);
        } else {
            o[axis] = xg[axis] + sgn * hh[axis];
        }
    };
    double um[3], up[3], vm[3], vp[3];
    nb(um, t0, -1, u % 16 == 0, -st[t0], -1, 0);
    nb(up, t0, +1, (u % 16 == 15) || (u == nn[t0] - 1), st[t0], +1, 0);
    nb(vm, t1, -1, v % 16 == 0, -st[t1], 0, -1);
    nb(vp, t1, +1, (v % 16 == 15) || (v == nn[t1] - 1), st[t1], 0, +1);
    GhostPt G;
    double jac[3][3], d[3];
    for (int l = 0; l < 3; ++l) d[l] = up[l] - um[l];
    jac_col(jac, t0, d, h1[t0]);
    for (int l = 0; l < 3; ++l) d[l] = vp[l] - vm[l];
    jac_col(jac, t1, d, h1[t1]);
    for (int l = 0; l < 3; ++l) {
        G.tc0[l] = jac[l][t0];
        G.tc1[l] = jac[l][t1];
        G.x2[l] = x2[l];
        G.b0[l] = F32 ? (double)A.B0_32[3 * a + l] : A.B0_64[l * g.P + a];
    }
    out[A.ghost_off[fs] + u + (long long)nu * v] = G;
}

struct Pos {
    double x[3];
};
__device__ __forceinline__ double pdiff(const Pos &a, const Pos &b, int l) { return a.x[l] - b.x[l]; }


}  // namespace

// =============================================================== kernel
// NT: threads = largest rect.  STAGE: 0..5, or 6 for the fields of the
// current positions only (no slope written).  PRESS: pressure force on.
// PERI: periodic faces (the plain kernel is built without their code).
// One workgroup's march: `L` is its logical index in the instance (after
// the XCD-aware remap), which picks the tile and z chunk.  Its partial
// results go to slot xcd_unremap(L), the hardware workgroup index of a
// standalone launch, so that the controller sums the partials in one order
// whether the instance runs alone or in an ensemble.
template <int NT, int STAGE, bool PRESS, bool PERI, bool KPRE = true>
__device__ __forceinline__ void stage64_body(const FusedArgs &A, const int L)
{
    constexpr bool FINAL = STAGE == 5;
    using R = double;
    constexpr int NXL = 5;  // planes of positions kept in LDS (ring)

    extern __shared__ __align__(16) unsigned char smem[];
    // LDS rings: positions [NXL][3][NT], B [3][3][NT]
    R *XLa = reinterpret_cast<R *>(smem);
    R *BL = XLa + NXL * 3 * NT;
    __shared__ double red[7][NT / 64];

    const Geo &g = A.g;
    Ctl *ctl = A.ctl;
    if (A.check_done && ctl->done) return;
    const int cur = ctl->cur;
    const R dt = (R)ctl->dt;
    constexpr int stage = STAGE == 6 ? -1 : STAGE;
    constexpr int n = STAGE == 6 ? 0 : STAGE;
    const long long P = g.P, P3 = 3 * P;

    const int tx = L % A.tiles_x, ty = (L / A.tiles_x) % A.tiles_y, ch = L / (A.tiles_x * A.tiles_y);
    const int ix0 = tx * A.TX, iy0 = ty * A.TY;
    const int Ks = ch * A.CZ + 1, Ke = min(ch * A.CZ + A.CZ, g.nz);
    const int RW = A.TX + 4, RH = A.TY + 4;
    const int ixe = min(ix0 + A.TX, g.nx), iye = min(iy0 + A.TY, g.ny);

    // ---- this thread's column
    const int t = threadIdx.x;
    const int c = t % RW, r = t / RW;
    const int gi = ix0 - 1 + c, gj = iy0 - 1 + r;
    // Periodic faces: a column beyond a periodic face stands for the inner
    // column the reference copies into it (its source si, sj: periodic
    // coordinates wrapped; src/boundaries.cu:450-582).  It holds that
    // column's positions unshifted and computes that column's B, so the
    // ghost's B is the copy, bit for bit.  A reader applies the period shift
    // of the reference's ghost positions, (x - L) - h or (x + L) + h, to a
    // neighbour that lies beyond a periodic face in the reader's own frame.
    const bool px = PERI && g.xPeri, py = PERI && g.yPeri, pz = PERI && g.zPeri;
    const int si = px ? (gi <= 0 ? gi + g.nx : gi >= g.nx + 1 ? gi - g.nx : gi) : gi;
    const int sj = py ? (gj <= 0 ? gj + g.ny : gj >= g.ny + 1 ? gj - g.ny : gj) : gj;
    const bool col = t < RW * RH && gi >= (px ? -1 : 0) && gi <= g.nx + (px ? 2 : 1) && gj >= (py ? -1 : 0) &&
                     gj <= g.ny + (py ? 2 : 1);
    // ghost columns / rows of line-tied faces (source frame)
    const bool gx = si == 0 || si == g.nx + 1, gy = sj == 0 || sj == g.ny + 1;
    const bool own_in = !PERI || (gi >= 0 && gi <= g.nx + 1 && gj >= 0 && gj <= g.ny + 1);   // own point padded
    const bool inB = col && own_in && c >= 1 && c <= RW - 2 && r >= 1 && r <= RH - 2 &&
                     !((c == 1 || c == RW - 2) && (r == 1 || r == RH - 2)) && !(gx && gy);
    const bool inJ = col && c >= 2 && c <= RW - 3 && r >= 2 && r <= RH - 3 && gi >= 1 && gi <= ixe && gj >= 1 &&
                     gj <= iye;
    const long long abase = (long long)gi + g.sx * (long long)gj;   // own point: stores
    const long long asrc = (long long)si + g.sx * (long long)sj;    // source column: loads
    const int ii = gi - 1, jj = gj - 1;
    // ghost face of an x/y ghost column (source coordinates)
    const int cside = gx ? (si != 0) : (sj != 0);
    const int cu = gx ? sj - 1 : si - 1;
    // period shifts (-1, 0, +1) of the in-plane neighbour reads
    const int shxm = px && si == 1 ? -1 : 0, shxp = px && si == g.nx ? 1 : 0;
    const int shym = py && sj == 1 ? -1 : 0, shyp = py && sj == g.ny ? 1 : 0;
    const double eps_xy = (g.epsilonProf && inJ) ? A.epsx[ii] * A.epsy[jj] : 1.0;
    const R *Xh0 = A.X64[cur];
    const R *B0 = A.B0_64;
    const R *K = A.K64;

    // ---- register state of the column
    Pos xq[4];       // x_n at planes p-2 .. p+1
    R bq[3][3];      // B at planes p-2 .. p
    double dq[2];    // det at planes p-1, p
    R b0q[2][3];     // B0 at planes p, p+1
    // prefetched inputs of one plane
    struct In {
        R x[3], k[5][3], b0[3];
    } pf;

    double dmax = 0, e_sum = 0, jmax = 0, r1 = 0, r2 = 0, r3 = 0, r4 = 0;
    int nan_seen = 0;

    // planes a chunk touches (one more beyond each end of a periodic z axis)
    auto plane_ok = [&](int q) { return q >= max(Ks - 2, pz ? -1 : 0) && q <= min(Ke + 2, g.nz + (pz ? 2 : 1)); };
    auto zgl = [&](int q) { return !pz && (q == 0 || q == g.nz + 1); };   // line-tied z ghost plane
    auto ghost_pt = [&](int q) { return gx || gy || zgl(q); };
    // source plane of a plane beyond a periodic z face, and the z shifts of
    // reading plane q - 1 or q + 1 in the frame of plane q
    auto zsrc = [&](int q) { return pz ? (q <= 0 ? q + g.nz : q >= g.nz + 1 ? q - g.nz : q) : q; };
    auto zsh_m = [&](int q) { return pz && zsrc(q) == 1 ? -1 : 0; };
    auto zsh_p = [&](int q) { return pz && zsrc(q) == g.nz ? 1 : 0; };
    // a position read across a periodic face: component `axis` shifted by
    // -+(L + h) in the reference's operation order
    auto shift = [&](Pos &v, int axis, int sgn) {
        if (PERI && sgn != 0) {
            const double Lp = axis == 0 ? g.Lx : axis == 1 ? g.Ly : g.Lz;
            const double h = axis == 0 ? g.dx : axis == 1 ? g.dy : g.dz;
            v.x[axis] = sgn < 0 ? (v.x[axis] - Lp) - h : (v.x[axis] + Lp) + h;
        }
    };
    (void)shxm; (void)shxp; (void)shym; (void)shyp; (void)zsh_m; (void)zsh_p; (void)shift;
    // ring slot of plane q (q > -NXL * 4 always holds here)
    auto XLslot = [&](int q) { return (unsigned)(q + NXL * 4) % (unsigned)NXL; };
    auto ld_x = [&](int q, int pt, Pos &o) {
        int s = XLslot(q);
#pragma unroll
        for (int l = 0; l < 3; ++l) o.x[l] = XLa[(s * 3 + l) * NT + pt];
    };
    auto st_x = [&](int q, const Pos &v) {
        int s = XLslot(q);
#pragma unroll
        for (int l = 0; l < 3; ++l) XLa[(s * 3 + l) * NT + t] = v.x[l];
    };
    auto ld_b = [&](int q, int pt, R o[3]) {
        int s = (unsigned)(q + 12) % 3u;
#pragma unroll
        for (int l = 0; l < 3; ++l) o[l] = BL[(s * 3 + l) * NT + pt];
    };

    // (0) loads of plane q into `in`
    auto load_plane = [&](int q, In &in) {
        long long a = asrc + g.sxy * (long long)zsrc(q);
        bool gh = ghost_pt(q);
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            in.x[l] = Xh0[l * P + a];
            in.b0[l] = B0[l * P + a];
        }
        if (!gh) {
#pragma unroll
            for (int m = 0; m < 5; ++m)
                if (m < n)
#pragma unroll
                    for (int l = 0; l < 3; ++l) in.k[m][l] = K[m * P3 + l * P + a];
        }
    };
    // (1) stage position from the loaded inputs
    auto form_x = [&](int q, const In &in, Pos &o) {
        bool gh = ghost_pt(q);
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            // x_n by the reference's incremental stage updates
            double x = in.x[l];
            if (!gh)
#pragma unroll
                for (int s2 = 0; s2 < 5; ++s2)
                    if (s2 < n)
#pragma unroll
                        for (int m = 0; m <= s2; ++m) x += c_tab.db[s2 * 6 + m] * in.k[m][l];
            o.x[l] = x;
        }
    };

    // ---- march
    const int p_first = Ks - 4, p_last = Ke + 1;
    for (int p = p_first; p <= p_last; ++p) {
        // (1) x_n(p+1) from the inputs loaded one iteration ago
        if (col && plane_ok(p + 1)) {
            form_x(p + 1, pf, xq[3]);
            st_x(p + 1, xq[3]);
#pragma unroll
            for (int l = 0; l < 3; ++l) b0q[1][l] = pf.b0[l];
        }
        // (0) loads of plane p+2, in flight while B and J are computed
        if (col && plane_ok(p + 2)) load_plane(p + 2, pf);
        // final stage: the five slopes of the output point (plane p-1),
        // issued here so that the B section covers their latency (KPRE; the
        // ensemble build, whose registers do not take them, loads them
        // where they are used)
        double kpre[5][3];
        (void)kpre;
        if constexpr (FINAL && KPRE) {
            if (inJ && p - 1 >= Ks && p - 1 <= Ke) {
                const long long a = abase + g.sxy * (long long)(p - 1);
#pragma unroll
                for (int m = 0; m < 5; ++m)
#pragma unroll
                    for (int l = 0; l < 3; ++l) kpre[m][l] = A.K64[m * P3 + l * P + a];
            }
        }
        // (2) B(p); ghost edges (two ghost coordinates) keep B0, as in the
        // reference, and are never read by J
        if (inB && p >= Ks - 1 && p <= Ke + 1 && p >= 0 && p <= g.nz + 1 && !((gx || gy) && zgl(p))) {
            const bool zg = zgl(p);
            R Bv[3];
            double detv = 0;
            if (!zg) {
                // The interior formula, which line-tied x and y ghost points
                // share: such a point's face formula (src/boundaries.cu:
                // 164-186, 287-335) is the same Jacobian with the outer
                // neighbour along the normal replaced by the outer ghost
                // x_g + h n and the tangential ones by x_g -+ h where the
                // reference's 16-wide launch tiles synthesise them (their
                // edges and the last row), B's normal component held at B0.
                // Taking one path keeps their waves from running two
                Pos xm, xp, ym, yp;
                ld_x(p, t - 1, xm);
                ld_x(p, t + 1, xp);
                ld_x(p, t - RW, ym);
                ld_x(p, t + RW, yp);
                Pos zp1 = xq[3], zm1 = xq[1];
                if constexpr (PERI) {
                    shift(xm, 0, shxm);
                    shift(xp, 0, shxp);
                    shift(ym, 1, shym);
                    shift(yp, 1, shyp);
                    shift(zp1, 2, zsh_p(p));
                    shift(zm1, 2, zsh_m(p));
                }
                if (gx || gy) {
                    const Pos &xg = xq[2];
                    const int v = zsrc(p) - 1;
                    if (gx) {
                        Pos x2 = xg;
                        x2.x[0] = xg.x[0] + (2 * cside - 1) * g.dx;
                        if (cside) xp = x2;
                        else xm = x2;
                        if (cu % 16 == 0) {
                            ym = xg;
                            ym.x[1] = xg.x[1] + (double)(-1) * g.dy;
                        }
                        if (cu % 16 == 15 || cu == g.ny - 1) {
                            yp = xg;
                            yp.x[1] = xg.x[1] + (double)(+1) * g.dy;
                        }
                    } else {
                        Pos x2 = xg;
                        x2.x[1] = xg.x[1] + (2 * cside - 1) * g.dy;
                        if (cside) yp = x2;
                        else ym = x2;
                        if (cu % 16 == 0) {
                            xm = xg;
                            xm.x[0] = xg.x[0] + (double)(-1) * g.dx;
                        }
                        if (cu % 16 == 15 || cu == g.nx - 1) {
                            xp = xg;
                            xp.x[0] = xg.x[0] + (double)(+1) * g.dx;
                        }
                    }
                    if (v % 16 == 0) {
                        zm1 = xg;
                        zm1.x[2] = xg.x[2] + (double)(-1) * g.dz;
                    }
                    if (v % 16 == 15 || v == g.nz - 1) {
                        zp1 = xg;
                        zp1.x[2] = xg.x[2] + (double)(+1) * g.dz;
                    }
                }
                R jac[3][3], d[3];
#pragma unroll
                for (int l = 0; l < 3; ++l) d[l] = pdiff(xp, xm, l);
                jac_col(jac, 0, d, (R)g.dx1);
#pragma unroll
                for (int l = 0; l < 3; ++l) d[l] = pdiff(yp, ym, l);
                jac_col(jac, 1, d, (R)g.dy1);
#pragma unroll
                for (int l = 0; l < 3; ++l) d[l] = pdiff(zp1, zm1, l);
                jac_col(jac, 2, d, (R)g.dz1);
                double detJac1 = cauchy_b(jac, b0q[0], gx ? 0 : gy ? 1 : -1, Bv);
                detv = 1. / detJac1;
                // a ghost point's B (and a line-tied one's det) for the
                // fields of a dump; a periodic ghost's det stays as the
                // reference's updateBbound leaves it
                if (A.emit && (gx || gy || (PERI && (gi == 0 || gi == g.nx + 1 || gj == 0 || gj == g.ny + 1 || p == 0 ||
                                                     p == g.nz + 1)))) {
                    int ci = min(max(gi, 1), g.nx) - 1, cj = min(max(gj, 1), g.ny) - 1, ck = min(max(p, 1), g.nz);
                    if (ci >= ix0 && ci < ixe && cj >= iy0 && cj < iye && ck >= Ks && ck <= Ke) {
                        long long a = abase + g.sxy * (long long)p;
#pragma unroll
                        for (int l = 0; l < 3; ++l) A.eB64[l * P + a] = Bv[l];
                        if (gx || gy) A.edet[a] = detv;
                    }
                }
            } else {
                // z faces: the tabulated face constants (the initial
                // distortion enters through them)
                const int face = 2, side = p != 0, u = si - 1, v = sj - 1;
                double xin[3];
                const Pos &pin = side ? xq[1] : xq[3];
#pragma unroll
                for (int l = 0; l < 3; ++l) xin[l] = pin.x[l];
                double Bd[3];
                const GhostPt G = A.ghost[A.ghost_off[face * 2 + side] + u + (long long)nface(g, face) * v];
                ghost_b_f<2>(g, G, side, xin, Bd, detv);
#pragma unroll
                for (int l = 0; l < 3; ++l) Bv[l] = (R)Bd[l];
                if (A.emit) {
                    int ci = min(max(gi, 1), g.nx) - 1, cj = min(max(gj, 1), g.ny) - 1, ck = min(max(p, 1), g.nz);
                    if (ci >= ix0 && ci < ixe && cj >= iy0 && cj < iye && ck >= Ks && ck <= Ke) {
                        long long a = abase + g.sxy * (long long)p;
#pragma unroll
                        for (int l = 0; l < 3; ++l) A.eB64[l * P + a] = Bv[l];
                        A.edet[a] = detv;
                    }
                }
            }
            int sb = (unsigned)(p + 12) % 3u;
#pragma unroll
            for (int l = 0; l < 3; ++l) {
                bq[2][l] = Bv[l];
                BL[(sb * 3 + l) * NT + t] = Bv[l];
            }
            dq[1] = detv;
        }

        // (3) J, force and slope at (column, p-1): in-plane data of plane p-1
        // was published before the previous iteration's barrier, B(p) is this
        // thread's own
        const int pj = p - 1;
        if (inJ && pj >= Ks && pj <= Ke) {
            const long long a = abase + g.sxy * (long long)pj;
            Pos xm, xp, ym, yp;
            ld_x(pj, t - 1, xm);
            ld_x(pj, t + 1, xp);
            ld_x(pj, t - RW, ym);
            ld_x(pj, t + RW, yp);
            Pos zm = xq[0], zp = xq[2];
            if constexpr (PERI) {
                shift(xm, 0, shxm);
                shift(xp, 0, shxp);
                shift(ym, 1, shym);
                shift(yp, 1, shyp);
                shift(zm, 2, zsh_m(pj));
                shift(zp, 2, zsh_p(pj));
            }
            R Bxm[3], Bxp[3], Bym[3], Byp[3];
            ld_b(pj, t - 1, Bxm);
            ld_b(pj, t + 1, Bxp);
            ld_b(pj, t - RW, Bym);
            ld_b(pj, t + RW, Byp);
            const R *Bzm = bq[0], *Bzp = bq[2], *Bc = bq[1];
            R Jv[3], kv[3];
            {
                // diamonds in circulation order (src/current.cu:291-382):
                // xy: W S E N; yz: y- z- y+ z+; xz: z- x- z+ x+
                const Pos *V[3][4] = {{&xm, &ym, &xp, &yp}, {&ym, &zm, &yp, &zp}, {&zm, &xm, &zp, &xp}};
                const R *W[3][4] = {{Bxm, Bym, Bxp, Byp}, {Bym, Bzm, Byp, Bzp}, {Bzm, Bxm, Bzp, Bxp}};
                R NA[3][3], II[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    R d[4][3], bv[4][3];
#pragma unroll
                    for (int e = 0; e < 4; ++e)
#pragma unroll
                        for (int l = 0; l < 3; ++l) {
                            d[e][l] = pdiff(*V[q][(e + 1) & 3], *V[q][e], l);
                            bv[e][l] = W[q][e][l];
                        }
                    II[q] = diamond(d, bv, NA[q]);
                }
                solve_na(NA, II, Jv);
            }
            cross3(Jv, Bc, kv);
            const int kk = pj - 1;
            R eps = (R)1;
            if (g.epsilonProf) eps = (R)(eps_xy * A.epsz[kk]);
            R gv[3] = {0, 0, 0};
            const double detv = dq[0];
            if constexpr (PRESS) {
                // gradP from second differences (src/gradP.cu:93-154)
                auto xat = [&](int di, int dj, int dk, Pos &o) {
                    if (di == 0 && dj == 0) {
                        o = dk < 0 ? xq[0] : dk > 0 ? xq[2] : xq[1];
                    } else {
                        ld_x(pj + dk, t + di + dj * RW, o);
                    }
                    if constexpr (PERI) {
                        shift(o, 0, di < 0 ? shxm : di > 0 ? shxp : 0);
                        shift(o, 1, dj < 0 ? shym : dj > 0 ? shyp : 0);
                        shift(o, 2, dk < 0 ? zsh_m(pj) : dk > 0 ? zsh_p(pj) : 0);
                    }
                };
                Pos pa, pb, pc, pd;
                R jac[3][3], d2[3][3][3], d[3];
#pragma unroll
                for (int l = 0; l < 3; ++l) d[l] = pdiff(xp, xm, l);
                jac_col(jac, 0, d, (R)g.dx1);
#pragma unroll
                for (int l = 0; l < 3; ++l) d[l] = pdiff(yp, ym, l);
                jac_col(jac, 1, d, (R)g.dy1);
#pragma unroll
                for (int l = 0; l < 3; ++l) d[l] = pdiff(zp, zm, l);
                jac_col(jac, 2, d, (R)g.dz1);
                const Pos &x0c = xq[1];
                const R h1[3] = {(R)g.dx1, (R)g.dy1, (R)g.dz1};
                const Pos *ax[3][2] = {{&xm, &xp}, {&ym, &yp}, {&zm, &zp}};
#pragma unroll
                for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
                    for (int l = 0; l < 3; ++l)
                        d2[l][a2][a2] = (ax[a2][1]->x[l] - 2 * x0c.x[l] + ax[a2][0]->x[l]) * h1[a2] * h1[a2];
                // mixed: (x(+a+b) - x(+a-b)) - (x(-a+b) - x(-a-b)) for (a,b) = (0,1),(0,2),(1,2)
                const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    int aa = pr[m][0], bb = pr[m][1];
                    int o1[3] = {0, 0, 0}, o2[3] = {0, 0, 0};
                    o1[aa] = 1;
                    o2[bb] = 1;
                    xat(o1[0] + o2[0], o1[1] + o2[1], o1[2] + o2[2], pa);
                    xat(o1[0] - o2[0], o1[1] - o2[1], o1[2] - o2[2], pb);
                    xat(-o1[0] + o2[0], -o1[1] + o2[1], -o1[2] + o2[2], pc);
                    xat(-o1[0] - o2[0], -o1[1] - o2[1], -o1[2] - o2[2], pd);
#pragma unroll
                    for (int l = 0; l < 3; ++l) {
                        R v;
                        if (m == 2)  // the reference's term order for d2/dy dz
                            v = (pa.x[l] - pc.x[l] - pb.x[l] + pd.x[l]) * h1[aa] * h1[bb] / 4;
                        else
                            v = (pa.x[l] - pb.x[l] - pc.x[l] + pd.x[l]) * h1[aa] * h1[bb] / 4;
                        d2[l][aa][bb] = v;
                        d2[l][bb][aa] = v;
                    }
                }
                R det1 = (R)(1 / detv);
                gradp_core(jac, d2, det1, (R)g.beta, gv);
            }
#pragma unroll
            for (int l = 0; l < 3; ++l) {
                kv[l] *= dt * eps;
                if (PRESS) kv[l] += -dt * gv[l];
            }
            if (A.emit) {
#pragma unroll
                for (int l = 0; l < 3; ++l) {
                    A.eB64[l * P + a] = Bc[l];
                    A.eJ64[l * P + a] = Jv[l];
                    if (PRESS) A.eG64[l * P + a] = gv[l];
                }
                A.edet[a] = detv;
            }
            if (A.diag) {
                double Bd[3] = {(double)Bc[0], (double)Bc[1], (double)Bc[2]};
                double Jd[3] = {(double)Jv[0], (double)Jv[1], (double)Jv[2]}, jxb[3];
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
                r3 = nb_max(r3, nc);
                r4 = nb_max(r4, den);
            }
            if constexpr (!FINAL) {
                if constexpr (stage >= 0) {
                    R *Ko = A.K64;
#pragma unroll
                    for (int l = 0; l < 3; ++l) Ko[stage * P3 + l * P + a] = kv[l];
                }
            } else {
                // x_6 = x_5 + sum_m (b6m - b5m) k_m, then the x5/x4 weights
                // (src/rungeKutta.cu:216-227, 296-309); x_new to buffer cur+1,
                // x_6 (where a rejected attempt leaves xb) to cur+2
                double *Xd = A.X64[(cur + 1) % 3], *X6 = A.X64[(cur + 2) % 3];
#pragma unroll
                for (int l = 0; l < 3; ++l) {
                    double k6[6];
#pragma unroll
                    for (int m = 0; m < 5; ++m) k6[m] = KPRE ? kpre[m][l] : A.K64[m * P3 + l * P + a];
                    k6[5] = kv[l];
                    double x6 = xq[1].x[l];
#pragma unroll
                    for (int m = 0; m < 6; ++m) x6 += c_tab.db[5 * 6 + m] * k6[m];
                    X6[l * P + a] = x6;
                    double xn = x6, xs = x6;
                    if (l < 2 || g.zUpdate)
#pragma unroll
                        for (int m = 0; m < 6; ++m) {
                            xn += c_tab.c[m] * k6[m];
                            xs += c_tab.cs[m] * k6[m];
                        }
                    Xd[l * P + a] = xn;
                    double dd = fabs(xn - xs);
                    if (dd != dd) nan_seen = 1;
                    dmax = nb_max(dmax, dd);
                }
            }
        }

        // one barrier per plane: LDS writes of this iteration become visible,
        // and no slot still read here is overwritten before everyone passed
        __syncthreads();

        // rotate the column queues
        xq[0] = xq[1];
        xq[1] = xq[2];
        xq[2] = xq[3];
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            bq[0][l] = bq[1][l];
            bq[1][l] = bq[2][l];
            b0q[0][l] = b0q[1][l];
        }
        dq[0] = dq[1];
    }

    // ---- block partials (fixed order: deterministic)
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (FINAL) {
        dmax = wave_max(dmax);
        if (lane == 0) red[0][w] = dmax;
        if (nan_seen) atomicOr(&ctl->nan_seen, 1);
    }
    if (A.diag) {
        e_sum = wave_sum(e_sum);
        jmax = wave_max(jmax);
        r1 = wave_max(r1);
        r2 = wave_max(r2);
        r3 = wave_max(r3);
        r4 = wave_max(r4);
        if (lane == 0) {
            red[1][w] = e_sum;
            red[2][w] = jmax;
            red[3][w] = r1;
            red[4][w] = r2;
            red[5][w] = r3;
            red[6][w] = r4;
        }
    }
    __syncthreads();
    const int pb = xcd_unremap(L, A.nblocks);
    if (threadIdx.x == 0) {
        if (FINAL) {
            double m = 0;
            for (int q = 0; q < NT / 64; ++q) m = nb_max(m, red[0][q]);
            A.part_delta[pb] = m;
        }
        if (A.diag) {
            double s = 0, a[5] = {0, 0, 0, 0, 0};
            for (int q = 0; q < NT / 64; ++q) {
                s += red[1][q];
                for (int m = 0; m < 5; ++m) a[m] = nb_max(a[m], red[m + 2][q]);
            }
            A.part_diag[pb] = s;
            for (int m = 0; m < 5; ++m) A.part_diag[(m + 1) * A.nblocks + pb] = a[m];
        }
    }
}

template <int NT, int STAGE, bool PRESS, bool PERI = false>
__global__ void __launch_bounds__(NT) k_stage64(FusedArgs A)
{
    stage64_body<NT, STAGE, PRESS, PERI>(A, xcd_remap(blockIdx.x, A.nblocks));
}

// Ensemble: one launch covers the workgroups of n instances; start[i] is
// the first (remapped) workgroup of instance i, start[n] the total.  Each
// instance's partials go to the slots its standalone launch would use.
template <int NT, int STAGE, bool PRESS>
__global__ void __launch_bounds__(NT) k_stage64_batch(const FusedArgs *__restrict__ args, const int *__restrict__ start,
                                                      int n)
{
    const int L = xcd_remap(blockIdx.x, start[n]);
    int lo = 0, hi = n - 1;
    while (lo < hi) {   // last instance whose first workgroup is <= L
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= L) lo = mid;
        else hi = mid - 1;
    }
    stage64_body<NT, STAGE, PRESS, false, false>(args[lo], L - start[lo]);
}

// ============================================================ controller
// One workgroup: reduces the stage partials, applies accept/reject and the
// dt update of src/main.cu:321-383 and logs the attempt.
// Controller workgroup: four waves.  The partials are a few thousand at
// most; at 129 x 129 x 385 the controller takes 9.1 us per attempt in four
// waves, 11.4 us in sixteen and in one
constexpr int kCtlThreads = 256;
__device__ __forceinline__ void controller_body(const CtlArgs &A)
{
    __shared__ double red[7][kCtlThreads / 64];
    Ctl *ctl = A.ctl;
    // thread 0 copies the control block into registers while the block
    // reduces the partials, and writes it back once at the end: its loads
    // are issued together instead of one after each dependent store.  (A
    // finished run still reduces: its partials are read and dropped.)
    Ctl h;
    if (threadIdx.x == 0) h = *ctl;
    const int nb = A.nb_diag;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double m = 0, e = 0, jm = 0, r1 = 0, r2 = 0, r3 = 0, r4 = 0;
    for (int i = threadIdx.x; i < A.nb_delta; i += blockDim.x) m = nb_max(m, A.part_delta[i]);
    if (A.diag)
        for (int i = threadIdx.x; i < nb; i += blockDim.x) {
            e += A.part_diag[i];
            jm = nb_max(jm, A.part_diag[nb + i]);
            r1 = nb_max(r1, A.part_diag[2 * nb + i]);
            r2 = nb_max(r2, A.part_diag[3 * nb + i]);
            r3 = nb_max(r3, A.part_diag[4 * nb + i]);
            r4 = nb_max(r4, A.part_diag[5 * nb + i]);
        }
    m = wave_max(m);
    e = wave_sum(e);
    jm = wave_max(jm);
    r1 = wave_max(r1);
    r2 = wave_max(r2);
    r3 = wave_max(r3);
    r4 = wave_max(r4);
    if (lane == 0) {
        red[0][w] = m;
        red[1][w] = e;
        red[2][w] = jm;
        red[3][w] = r1;
        red[4][w] = r2;
        red[5][w] = r3;
        red[6][w] = r4;
    }
    __syncthreads();
    if (threadIdx.x != 0 || h.done) return;
    double md = 0, es = 0, jmx = 0, rr1 = 0, rr2 = 0, rr3 = 0, rr4 = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
        md = nb_max(md, red[0][q]);
        es += red[1][q];
        jmx = nb_max(jmx, red[2][q]);
        rr1 = nb_max(rr1, red[3][q]);
        rr2 = nb_max(rr2, red[4][q]);
        rr3 = nb_max(rr3, red[5][q]);
        rr4 = nb_max(rr4, red[6][q]);
    }
    // relaxation criterion: max|JxB| / max(|J||B|) (0 when J vanishes)
    const double crit = rr4 > 0 ? rr3 / rr4 : 0.0;
    StepRecord rec;
    rec.t = h.t;
    rec.dt_used = h.dt;
    rec.maxDelta = md;
    rec.nan_seen = h.nan_seen;
    if (A.diag) {
        h.energy = es * A.cell;
        h.jmax = jmx;
        h.jxb_b2 = rr1;
        h.ratio = crit;
        h.ratio_pt = rr2;
        h.diag_valid = 1;
    }
    rec.ratio = A.diag ? crit : -1;
    rec.ratio_pt = A.diag ? rr2 : -1;
    rec.energy = A.diag ? es * A.cell : -1;
    rec.jxb_b2 = A.diag ? rr1 : -1;
    rec.jmax = A.diag ? jmx : -1;
    // the state at the start of this attempt already meets the criterion:
    // stop there and drop the attempt
    if (A.diag && h.stop_ratio > 0 && crit <= h.stop_ratio) {
        h.converged = 1;
        h.done = 1;
        h.stop_reason = 3;
        h.accepted = 0;
        rec.accepted = -1;
        if (A.log) A.log[h.n_attempt % A.log_cap] = rec;
        *ctl = h;
        return;
    }
    const double dt = h.dt;
    h.dt_used = dt;
    h.maxDelta = md;
    const int acc = md < A.maxError;
    h.accepted = acc;
    if (acc) {
        h.t += dt;
        h.cur = A.nbuf == 3 ? (h.cur + 1) % 3 : h.cur ^ 1;
        h.n_accept += 1;
    } else if (A.nbuf == 3) {
        h.cur = (h.cur + 2) % 3;
    }
    if (md > 0) h.dt = dt * pow(0.9 * A.maxError / md, 0.2);
    rec.accepted = acc;
    rec.t = h.t;
    if (A.log) A.log[h.n_attempt % A.log_cap] = rec;
    h.n_attempt += 1;
    h.event = acc && (interval_crossed(h.t, h.dt, h.ev_dump) || interval_crossed(h.t, h.dt, h.ev_save));
    if (h.n_accept >= h.n_cycle) { h.done = 1; h.stop_reason = 1; }
    else if (!(h.dt > A.dtMin)) { h.done = 1; h.stop_reason = 2; }
    else if (h.n_attempt >= h.max_attempt) { h.done = 1; h.stop_reason = 4; }
    else if (h.event) { h.done = 1; h.stop_reason = 5; }
    *ctl = h;
}

__global__ void __launch_bounds__(kCtlThreads) k_controller(CtlArgs A) { controller_body(A); }

// Ensemble: one workgroup per instance.
__global__ void __launch_bounds__(kCtlThreads) k_controller_batch(const CtlArgs *__restrict__ args)
{
    controller_body(args[blockIdx.x]);
}

// ================================================================ launch
namespace {
constexpr int kNT64 = 512;   // F64

template <int NT>
constexpr size_t lds_bytes64()
{
    return (size_t)(5 * 3 * NT + 3 * 3 * NT) * sizeof(double);
}

template <int NT, int STAGE, bool PRESS, bool PERI>
void launch_one(hipStream_t s, const FusedArgs &A)
{
    static bool attr = false;
    constexpr size_t lds = lds_bytes64<NT>();
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)k_stage64<NT, STAGE, PRESS, PERI>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        attr = true;
    }
    hipLaunchKernelGGL((k_stage64<NT, STAGE, PRESS, PERI>), dim3(A.nblocks), dim3(NT), lds, s, A);
}

template <int NT, int STAGE, bool PRESS>
void launch_one_batch(hipStream_t s, const FusedArgs *args, const int *start, int n, int total)
{
    static bool attr = false;
    constexpr size_t lds = lds_bytes64<NT>();
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)k_stage64_batch<NT, STAGE, PRESS>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        attr = true;
    }
    hipLaunchKernelGGL((k_stage64_batch<NT, STAGE, PRESS>), dim3(total), dim3(NT), lds, s, args, start, n);
}

template <int NT, bool PRESS>
void launch_stage_batch(hipStream_t s, int stage, const FusedArgs *args, const int *start, int n, int total)
{
    switch (stage) {
    case 0: launch_one_batch<NT, 0, PRESS>(s, args, start, n, total); break;
    case 1: launch_one_batch<NT, 1, PRESS>(s, args, start, n, total); break;
    case 2: launch_one_batch<NT, 2, PRESS>(s, args, start, n, total); break;
    case 3: launch_one_batch<NT, 3, PRESS>(s, args, start, n, total); break;
    case 4: launch_one_batch<NT, 4, PRESS>(s, args, start, n, total); break;
    default: launch_one_batch<NT, 5, PRESS>(s, args, start, n, total); break;
    }
}

template <int NT, bool PRESS, bool PERI>
void launch_stage(hipStream_t s, const FusedArgs &A)
{
    switch (A.stage) {
    case 0: launch_one<NT, 0, PRESS, PERI>(s, A); break;
    case 1: launch_one<NT, 1, PRESS, PERI>(s, A); break;
    case 2: launch_one<NT, 2, PRESS, PERI>(s, A); break;
    case 3: launch_one<NT, 3, PRESS, PERI>(s, A); break;
    case 4: launch_one<NT, 4, PRESS, PERI>(s, A); break;
    case 5: launch_one<NT, 5, PRESS, PERI>(s, A); break;
    default: launch_one<NT, 6, PRESS, PERI>(s, A); break;
    }
}

// Periodic ghost layers of the FP64 positions the controller left current
// (buffer ctl->cur, read on the device so that graph replays follow the
// controller's choice): the reference's setGridPeri (src/boundaries.cu:
// 532-582), faces z, y, x in that order, as the per-kernel path launches it
__global__ void k_set_grid_peri64(FusedArgs A, int face)
{
    const Geo &g = A.g;
    double *X = A.X64[A.ctl->cur];
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int na = face == 0 ? g.ny + 2 : g.nx + 2, nb = face == 2 ? g.ny + 2 : g.nz + 2;
    if (t >= (long long)na * nb) return;
    const int u = (int)(t % na), v = (int)(t / na);
    const long long st[3] = {1, g.sx, g.sxy};
    const int nn[3] = {g.nx, g.ny, g.nz};
    int lo3[3], s3[3];
    const int t0 = face == 0 ? 1 : 0, t1 = face == 2 ? 1 : 2;
    lo3[t0] = s3[t0] = u;
    lo3[t1] = s3[t1] = v;
    long long lo = 0, hi = 0, srclo = 0, srchi = 0;
    for (int a = 0; a < 3; ++a) {
        if (a == face) continue;
        lo += lo3[a] * st[a];
        srclo += s3[a] * st[a];
    }
    hi = lo + (long long)(nn[face] + 1) * st[face];
    srchi = srclo + st[face];
    srclo += (long long)nn[face] * st[face];
    const double Lp = face == 0 ? g.Lx : face == 1 ? g.Ly : g.Lz, h = face == 0 ? g.dx : face == 1 ? g.dy : g.dz;
    for (int l = 0; l < 3; ++l) {
        double vlo = X[l * g.P + srclo], vhi = X[l * g.P + srchi];
        if (l == face) {
            vlo = vlo - Lp - h;
            vhi = vhi + Lp + h;
        }
        X[l * g.P + lo] = vlo;
        X[l * g.P + hi] = vhi;
    }
}
}  // namespace

// both stage kernels hold up to 512 rect points (one per thread)
int fused_max_points(int) { return kNT64; }

void set_tableau(const TableauF &t)
{
    (void)hipMemcpyToSymbol(HIP_SYMBOL(c_tab), &t, sizeof t);
    set_tableau32(t);
}

long long ghost_points(const Geo &g, long long off[6])
{
    long long tot = 0;
    for (int fs = 0; fs < 6; ++fs) {
        int f = fs >> 1;
        off[fs] = tot;
        tot += f == 0 ? (long long)g.ny * g.nz : f == 1 ? (long long)g.nx * g.nz : (long long)g.nx * g.ny;
    }
    return tot;
}

void launch_ghost_setup(hipStream_t s, const FusedArgs &A, int f32k, GhostPt *out)
{
    long long off[6];
    long long tot = ghost_points(A.g, off);
    unsigned nb = (unsigned)((tot + 255) / 256);
    if (f32k)
        hipLaunchKernelGGL(k_ghost_setup<true>, dim3(nb), dim3(256), 0, s, A, out);
    else
        hipLaunchKernelGGL(k_ghost_setup<false>, dim3(nb), dim3(256), 0, s, A, out);
}

void launch_fused(hipStream_t s, const FusedArgs &A, int f32k)
{
    bool pr = A.g.pressure != 0;
    if (f32k) {
        launch_fused32(s, A);
    } else {
        const bool peri = A.g.xPeri || A.g.yPeri || A.g.zPeri;
        if (pr) peri ? launch_stage<kNT64, true, true>(s, A) : launch_stage<kNT64, true, false>(s, A);
        else peri ? launch_stage<kNT64, false, true>(s, A) : launch_stage<kNT64, false, false>(s, A);
    }
}

void launch_fused64_batch(hipStream_t s, int stage, int press, const FusedArgs *args, const int *start, int n,
                          int total)
{
    if (press) launch_stage_batch<kNT64, true>(s, stage, args, start, n, total);
    else launch_stage_batch<kNT64, false>(s, stage, args, start, n, total);
}

void launch_set_grid_peri64(hipStream_t s, const FusedArgs &A)
{
    const Geo &g = A.g;
    const int on[3] = {g.xPeri, g.yPeri, g.zPeri};
    for (int face = 2; face >= 0; --face) {
        if (!on[face]) continue;
        const long long n = face == 0 ? (long long)(g.ny + 2) * (g.nz + 2)
                            : face == 1 ? (long long)(g.nx + 2) * (g.nz + 2)
                                        : (long long)(g.nx + 2) * (g.ny + 2);
        hipLaunchKernelGGL(k_set_grid_peri64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, A, face);
    }
}

void launch_controller(hipStream_t s, const CtlArgs &A)
{
    hipLaunchKernelGGL(k_controller, dim3(1), dim3(kCtlThreads), 0, s, A);
}

void launch_controller_batch(hipStream_t s, const CtlArgs *args, int n)
{
    hipLaunchKernelGGL(k_controller_batch, dim3(n), dim3(kCtlThreads), 0, s, args);
}

}  // namespace gmd
