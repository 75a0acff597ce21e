// Device helpers shared by the fused stage kernels (FP64 and FP32).
#pragma once

#include "gm_device.hpp"
#include "gm_fused.hpp"

namespace gmd {
namespace {

// Workgroup id -> logical tile so that the blocks one XCD receives
// (hardware deals ids round-robin over the 8 XCDs) form a contiguous range
// of tiles and share halos in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
    int q = nb >> 3, r = nb & 7;
    int x = bid & 7, w = bid >> 3;
    int base = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
    return base + w;
}

// The inverse of xcd_remap: the workgroup index that xcd_remap sends to
// logical index L among nb.
__device__ __forceinline__ int xcd_unremap(int L, int nb)
{
    const int q = nb >> 3, r = nb & 7, head = r * (q + 1);
    int x, w;
    if (L < head) {
        x = L / (q + 1);
        w = L - x * (q + 1);
    } else {   // q > 0 here: L < nb = 8q + r
        x = r + (L - head) / q;
        w = L - head - (x - r) * q;
    }
    return w * 8 + x;
}

// Two-sum (Knuth): s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e)
{
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ int nface(const Geo &g, int face) { return face == 0 ? g.ny : g.nx; }

// B and det at a line-tied ghost-face point from its first inner neighbour
// x_in along the normal and the precomputed constants (src/boundaries.cu:
// 164-186): jac = [normal column from x_in and the outer ghost | tangential
// columns], B_normal = B0_normal.
// (FACE: the face as a compile-time constant, so that no array is indexed
// at run time)
template <int FACE>
__device__ __forceinline__ void ghost_b_f(const Geo &g, const GhostPt &G, int side, const double xin[3],
                                         double B[3], double &det)
{
    const double h1 = FACE == 0 ? g.dx1 : FACE == 1 ? g.dy1 : g.dz1;
    constexpr int t0 = FACE == 0 ? 1 : 0, t1 = FACE == 2 ? 1 : 2;
    double jac[3][3], d[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        d[l] = side ? G.x2[l] - xin[l] : xin[l] - G.x2[l];
        jac[l][t0] = G.tc0[l];
        jac[l][t1] = G.tc1[l];
    }
    jac_col(jac, FACE, d, h1);
    double detJac1 = cauchy_b(jac, G.b0, FACE, B);
    det = 1 / detJac1;
}
__device__ __forceinline__ void ghost_b(const Geo &g, const GhostPt &G, int face, int side, const double xin[3],
                                       double B[3], double &det)
{
    if (face == 0) ghost_b_f<0>(g, G, side, xin, B, det);
    else if (face == 1) ghost_b_f<1>(g, G, side, xin, B, det);
    else ghost_b_f<2>(g, G, side, xin, B, det);
}

}  // namespace
}  // namespace gmd
