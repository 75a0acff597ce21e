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
__device__ __forceinline__ void ghost_b(const Geo &g, const GhostPt &G, int face, int side, const double xin[3],
                                       double B[3], double &det)
{
    const double h1 = face == 0 ? g.dx1 : face == 1 ? g.dy1 : g.dz1;
    const int t0 = face == 0 ? 1 : 0, t1 = face == 2 ? 1 : 2;
    double jac[3][3], d[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        d[l] = side ? G.x2[l] - xin[l] : xin[l] - G.x2[l];
        jac[l][t0] = G.tc0[l];
        jac[l][t1] = G.tc1[l];
    }
    jac_col(jac, face, d, h1);
    double detJac1 = cauchy_b(jac, G.b0, face, B);
    det = 1 / detJac1;
}

// FP32 Stokes current from the six face neighbours, positions and fields
// taken relative to the point itself (r_v = x_v - x_c, b_v = B_v - B_c).
// For a closed quadrilateral v0..v3 the reference's circulation
// sum_e (b_e+1 + b_e)/2 . (x_e+1 - x_e) equals 1/2 sum_e (b_e . r_e+1 - b_e+1 . r_e)
// and its mean area vector 1/4 sum_e d_e x d_e+1 equals 1/2 (r2 - r0) x (r3 - r1);
// both identities are exact in real arithmetic and take half the operations.
// Order of the vertices as src/current.cu:291-382: xy W S E N, yz y- z- y+ z+,
// xz z- x- z+ x+ (indices into xm, ym, xp, yp, zm, zp = 0..5).
__device__ __forceinline__ void stokes_f32(const float r[6][3], const float b[6][3], float J[3])
{
#pragma clang fp contract(fast)
    const int Q[3][4] = {{0, 1, 2, 3}, {1, 4, 3, 5}, {4, 0, 5, 2}};
    float NA[3][3], II[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        float I = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float *be = b[Q[q][e]], *bn = b[Q[q][(e + 1) & 3]];
            const float *re = r[Q[q][e]], *rn = r[Q[q][(e + 1) & 3]];
            I += (be[0] * rn[0] + be[1] * rn[1] + be[2] * rn[2]) - (bn[0] * re[0] + bn[1] * re[1] + bn[2] * re[2]);
        }
        II[q] = 0.5f * I;
        const float *r0 = r[Q[q][0]], *r1 = r[Q[q][1]], *r2 = r[Q[q][2]], *r3 = r[Q[q][3]];
        float u[3] = {r2[0] - r0[0], r2[1] - r0[1], r2[2] - r0[2]};
        float v[3] = {r3[0] - r1[0], r3[1] - r1[1], r3[2] - r1[2]};
        NA[q][0] = 0.5f * (u[1] * v[2] - u[2] * v[1]);
        NA[q][1] = 0.5f * (u[2] * v[0] - u[0] * v[2]);
        NA[q][2] = 0.5f * (u[0] * v[1] - u[1] * v[0]);
    }
    // J = NA^-1 II by the adjugate
    float A00 = NA[1][1] * NA[2][2] - NA[1][2] * NA[2][1];
    float A01 = NA[0][2] * NA[2][1] - NA[0][1] * NA[2][2];
    float A02 = NA[0][1] * NA[1][2] - NA[0][2] * NA[1][1];
    float A10 = NA[1][2] * NA[2][0] - NA[1][0] * NA[2][2];
    float A11 = NA[0][0] * NA[2][2] - NA[0][2] * NA[2][0];
    float A12 = NA[0][2] * NA[1][0] - NA[0][0] * NA[1][2];
    float A20 = NA[1][0] * NA[2][1] - NA[1][1] * NA[2][0];
    float A21 = NA[0][1] * NA[2][0] - NA[0][0] * NA[2][1];
    float A22 = NA[0][0] * NA[1][1] - NA[0][1] * NA[1][0];
    float det = NA[0][0] * A00 + NA[0][1] * A10 + NA[0][2] * A20;
    float d1 = 1.0f / det;
    J[0] = (A00 * II[0] + A01 * II[1] + A02 * II[2]) * d1;
    J[1] = (A10 * II[0] + A11 * II[1] + A12 * II[2]) * d1;
    J[2] = (A20 * II[0] + A21 * II[1] + A22 * II[2]) * d1;
}
}  // namespace
}  // namespace gmd
