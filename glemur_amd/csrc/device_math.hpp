// Point-wise device math of the relaxation step, shared by the per-kernel
// path and the fused stage kernels.  The FP64 forms keep the reference's
// operation order (the library is built with -ffp-contract=off) so they
// reproduce it bit for bit; the *_c forms are the compensated FP32 variants.
#pragma once

#include <hip/hip_runtime.h>

namespace gmd {

// Cash-Karp tableau (src/rungeKutta.cu:10-25), b as 7x6 row-major.
struct Tableau {
    double b[42];
    double c[6];
    double cs[6];
};

template <class R>
__device__ __forceinline__ void cross3(const R a[3], const R b[3], R c[3])
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// det of the 3x3 grid Jacobian in the reference's 6-term order
// (src/rungeKutta.cu:89-90).
template <class R>
__device__ __forceinline__ R det3(const R j[3][3])
{
    return j[0][0] * j[1][1] * j[2][2] + j[0][1] * j[1][2] * j[2][0] + j[0][2] * j[1][0] * j[2][1] -
           j[0][0] * j[1][2] * j[2][1] - j[0][1] * j[1][0] * j[2][2] - j[0][2] * j[1][1] * j[2][0];
}

// Jacobian column m from the central difference d = x(X+e_m) - x(X-e_m):
// jac[l][m] = (d[l] * h1_m) / 2 (src/rungeKutta.cu:82-87).
template <class R>
__device__ __forceinline__ void jac_col(R jac[3][3], int m, const R d[3], R h1)
{
    jac[0][m] = d[0] * h1 / 2;
    jac[1][m] = d[1] * h1 / 2;
    jac[2][m] = d[2] * h1 / 2;
}

// B = jac.B0 / det with component `normal` held at B0 (normal < 0: none);
// returns 1/det as the reference's detJac1 (src/rungeKutta.cu:89-99,
// src/boundaries.cu:170-179).
template <class R>
__device__ __forceinline__ R cauchy_b(const R jac[3][3], const R b0[3], int normal, R B[3])
{
    R detJac1 = R(1) / det3(jac);
#pragma unroll
    for (int l = 0; l < 3; ++l)
        B[l] = (l == normal) ? b0[l] : (jac[l][0] * b0[0] + jac[l][1] * b0[1] + jac[l][2] * b0[2]) * detJac1;
    return detJac1;
}

// One Stokes diamond (src/current.cu:291-320): edge vectors d[e] of the
// closed quadrilateral v0->v1->v2->v3->v0 and the field values bv at the
// vertices; returns the circulation sum_e (b_e + b_e+1)/2 . d_e and the mean
// area vector na = (d0xd1 + d1xd2 + d2xd3 + d3xd0)/4.
template <class R>
__device__ __forceinline__ R diamond(const R d[4][3], const R bv[4][3], R na[3])
{
    R I = 0;
#pragma unroll
    for (int l = 0; l < 3; ++l)
        I += (bv[1][l] + bv[0][l]) / 2 * d[0][l] + (bv[2][l] + bv[1][l]) / 2 * d[1][l] +
             (bv[3][l] + bv[2][l]) / 2 * d[2][l] + (bv[0][l] + bv[3][l]) / 2 * d[3][l];
    R c[4][3];
    cross3(d[0], d[1], c[0]);
    cross3(d[1], d[2], c[1]);
    cross3(d[2], d[3], c[2]);
    cross3(d[3], d[0], c[3]);
#pragma unroll
    for (int l = 0; l < 3; ++l) na[l] = (c[0][l] + c[1][l] + c[2][l] + c[3][l]) / 4;
    return I;
}

// J = NA^-1 . II by the adjugate (src/current.cu:385-406).
template <class R>
__device__ __forceinline__ void solve_na(const R NA[3][3], const R II[3], R J[3])
{
    R detN = det3(NA);
    R d1 = R(1) / detN;
    R A[3][3];
    A[0][0] = (NA[1][1] * NA[2][2] - NA[1][2] * NA[2][1]) * d1;
    A[0][1] = (NA[0][2] * NA[2][1] - NA[0][1] * NA[2][2]) * d1;
    A[0][2] = (NA[0][1] * NA[1][2] - NA[0][2] * NA[1][1]) * d1;
    A[1][0] = (NA[1][2] * NA[2][0] - NA[1][0] * NA[2][2]) * d1;
    A[1][1] = (NA[0][0] * NA[2][2] - NA[0][2] * NA[2][0]) * d1;
    A[1][2] = (NA[0][2] * NA[1][0] - NA[0][0] * NA[1][2]) * d1;
    A[2][0] = (NA[1][0] * NA[2][1] - NA[1][1] * NA[2][0]) * d1;
    A[2][1] = (NA[0][1] * NA[2][0] - NA[0][0] * NA[2][1]) * d1;
    A[2][2] = (NA[0][0] * NA[1][1] - NA[0][1] * NA[1][0]) * d1;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        R s = 0;
        s += A[l][0] * II[0];
        s += A[l][1] * II[1];
        s += A[l][2] * II[2];
        J[l] = s;
    }
}

// Pressure force from the second derivatives of the map
// (src/gradP.cu:93-154): g_c = -(sum_{a,b,l} d2[a][b][l] jac1[b][a] jac1[l][c])
// * (beta * det1), jac1 = adj(jac) * det1.
template <class R>
__device__ __forceinline__ void gradp_core(const R jac[3][3], const R d2[3][3][3], R det1, R beta, R g[3])
{
    R jac1[3][3];
    jac1[0][0] = (jac[1][1] * jac[2][2] - jac[1][2] * jac[2][1]) * det1;
    jac1[0][1] = (jac[0][2] * jac[2][1] - jac[0][1] * jac[2][2]) * det1;
    jac1[0][2] = (jac[0][1] * jac[1][2] - jac[0][2] * jac[1][1]) * det1;
    jac1[1][0] = (jac[1][2] * jac[2][0] - jac[1][0] * jac[2][2]) * det1;
    jac1[1][1] = (jac[0][0] * jac[2][2] - jac[0][2] * jac[2][0]) * det1;
    jac1[1][2] = (jac[0][2] * jac[1][0] - jac[0][0] * jac[1][2]) * det1;
    jac1[2][0] = (jac[1][0] * jac[2][1] - jac[1][1] * jac[2][0]) * det1;
    jac1[2][1] = (jac[0][1] * jac[2][0] - jac[0][0] * jac[2][1]) * det1;
    jac1[2][2] = (jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0]) * det1;
    R s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int l = 0; l < 3; ++l) {
                s0 -= d2[a][b][l] * jac1[b][a] * jac1[l][0];
                s1 -= d2[a][b][l] * jac1[b][a] * jac1[l][1];
                s2 -= d2[a][b][l] * jac1[b][a] * jac1[l][2];
            }
    R w = beta * det1;
    g[0] = s0 * w;
    g[1] = s1 * w;
    g[2] = s2 * w;
}

// Friction profile epsilon of src/rungeKutta.cu:180-185 along one axis
// (inner index i of n points); the full factor is ex * ey * ez.
__host__ __device__ inline double eps_axis(int i, int n)
{
    return -(i - (n) / 2.) / (n) * (i + 1 - (n) / 2.) / (n) * 4 + 1;
}

// addDist (src/boundaries.cu:11-35): extra offset of the outer ghosts on
// z faces for an initially distorted grid; x = dx*i + Ox from the inner index.
struct DistParams {
    int code;
    double shearA, shearB, shearK, dx, dy, Ox, Oy, Lx, Ly;
};
__device__ inline double add_dist(const DistParams &q, int i, int j, double z, int l)
{
    const double pi = 3.141592653589793;
    double x = q.dx * i + q.Ox;
    double y = q.dy * j + q.Oy;
    double distX = 0, distY = 0;
    if (q.code == 0) {
        distY = -q.shearA * sin(q.shearK * 2 * pi * (x + q.Ox - q.dx / 2) / (q.Lx + q.dx)) * z;
        distX = -q.shearB * sin(q.shearK * 2 * pi * (distY + y + q.Oy - q.dy / 2) / (q.Ly + q.dy)) * z;
    }
    if (q.code == 1) {
        distX = 0;
        distY = q.shearA * exp(-q.shearK * x * x) * (1 - y * y) * (exp(-1 * z + 1) - 0.15366);
    }
    return l == 0 ? distX : distY;
}

// ---- wave64 / block reductions (deterministic order) ------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// NaN-blind max starting from 0, as the reference's maxValue
// (src/maths.cu:144-153): a value is taken only when '>' holds.
__device__ __forceinline__ double nb_max(double m, double v) { return v > m ? v : m; }
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = nb_max(v, __shfl_xor(v, o, 64));
    return v;
}

}  // namespace gmd
