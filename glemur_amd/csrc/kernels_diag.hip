// Mesh and error diagnostics of src/diagnostics.cu:33-110 and
// src/maths.cu:204-380,443-640: signed cell volume, octant wedge minimum,
// convexity, |B - ez| and |x - x_analytic| RMS errors, and the force-free
// parameter spread epsilon* along the central field line.  Positions are
// read in either precision (FP64, or FP32 hi + lo); per-point values are
// optionally stored for dumps, and the reductions are block partials summed
// or minimised in a fixed order.
#include "gm_device.hpp"

namespace gmd {
namespace {

constexpr int kB = 256;

struct PosRef {
    const double *x64;
    const float *xh, *xl;
    long long P;
    __device__ void at(long long a, double v[3]) const
    {
        for (int c = 0; c < 3; ++c)
            v[c] = x64 ? x64[c * P + a] : (double)xh[3 * a + c] + (double)xl[3 * a + c];   // F32K: triples
    }
};

__device__ inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ inline void cross3d(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// One inner point (i, j, k): the 8 octant wedge products of the edges to
// the face neighbours (src/maths.cu:497-570), oriented so that an
// undistorted cell is positive; convexity -1 when any is negative
// (src/maths.cu:573-640); the forward-difference cell volume
// (src/maths.cu:443-494).
__device__ void mesh_point(const Geo &g, const PosRef &X, int i, int j, int k, double &wmin, double &conv,
                           double &vol)
{
    const long long a = gidx(g, i + 1, j + 1, k + 1);
    double x0[3], xm[3][2][3];
    X.at(a, x0);
    const long long st[3] = {1, g.sx, g.sxy};
    for (int d = 0; d < 3; ++d)
        for (int s = 0; s < 2; ++s) {
            double v[3];
            X.at(a + (s ? st[d] : -st[d]), v);
            for (int c = 0; c < 3; ++c) xm[d][s][c] = v[c] - x0[c];
        }
    wmin = 0;
    conv = 1;
    for (int u = 0; u < 2; ++u)
        for (int v = 0; v < 2; ++v)
            for (int w = 0; w < 2; ++w) {
                double cr[3];
                cross3d(xm[0][u], xm[1][v], cr);
                double wc = dot3(cr, xm[2][w]);
                const int s = 2 * (u + v + w);       // the reference's u+v+w
                if (s == 0 || s == 4) wc = -wc;
                if (s == 0 || wc < wmin) wmin = wc;
                if (wc < 0) conv = -1;
            }
    double cr[3];
    cross3d(xm[0][1], xm[1][1], cr);
    vol = dot3(cr, xm[2][1]);
}

__global__ void __launch_bounds__(kB) k_mesh(Geo g, PosRef X, double *cellVol, double *convex, double *wedge,
                                             double *part, int nb)
{
    __shared__ double sw[kB], sc[kB];
    const long long N = (long long)g.nx * g.ny * g.nz;
    double wm = 1e300, cm = 1;
    for (long long t = (long long)blockIdx.x * kB + threadIdx.x; t < N; t += (long long)gridDim.x * kB) {
        const int i = (int)(t % g.nx), j = (int)((t / g.nx) % g.ny), k = (int)(t / ((long long)g.nx * g.ny));
        double w, c, v;
        mesh_point(g, X, i, j, k, w, c, v);
        if (cellVol) cellVol[t] = v;
        if (convex) convex[t] = c;
        if (wedge) wedge[t] = w;
        wm = w < wm ? w : wm;
        cm = c < cm ? c : cm;
    }
    sw[threadIdx.x] = wm;
    sc[threadIdx.x] = cm;
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sw[threadIdx.x] = fmin(sw[threadIdx.x], sw[threadIdx.x + s]);
            sc[threadIdx.x] = fmin(sc[threadIdx.x], sc[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sc[0];
        part[nb + blockIdx.x] = sw[0];
    }
}

// sums of |B - ez|^2 (src/maths.cu:294-330) and |x - x_an|^2
// (src/maths.cu:334-377) over inner points
template <class R>
__global__ void __launch_bounds__(kB) k_errs(Geo g, PosRef X, const R *B, double ar, double az, double phi1,
                                             double *part, int nb)
{
    __shared__ double s1[kB], s2[kB];
    const long long N = (long long)g.nx * g.ny * g.nz;
    double e1 = 0, e2 = 0;
    for (long long t = (long long)blockIdx.x * kB + threadIdx.x; t < N; t += (long long)gridDim.x * kB) {
        const int i = (int)(t % g.nx), j = (int)((t / g.nx) % g.ny), k = (int)(t / ((long long)g.nx * g.ny));
        const long long a = gidx(g, i + 1, j + 1, k + 1);
        const double b0 = B[a], b1 = B[g.P + a], b2 = (double)B[2 * g.P + a] - 1.;
        e1 += b0 * b0 + b1 * b1 + b2 * b2;
        const double x = i * g.dx + g.Ox, y = j * g.dy + g.Oy, z = k * g.dz + g.Oz;
        const double f = exp(-(x * x + y * y) / (ar * ar) - z * z / (az * az)) * phi1;
        double p[3];
        X.at(a, p);
        const double d0 = p[0] - cos(f) * x - sin(f) * y;
        const double d1 = p[1] + sin(f) * x - cos(f) * y;
        const double d2 = p[2] - z;
        e2 += d0 * d0 + d1 * d1 + d2 * d2;
    }
    s1[threadIdx.x] = e1;
    s2[threadIdx.x] = e2;
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            s1[threadIdx.x] += s1[threadIdx.x + s];
            s2[threadIdx.x] += s2[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * nb + blockIdx.x] = s1[0];
        part[3 * nb + blockIdx.x] = s2[0];
    }
}

// epsilon* (src/maths.cu:204-290): along the column i = nx/2, j = ny/2,
// a_k = J.B / B^2 and z_k = z of the position; the maximum over all pairs
// of sqrt2 |a_n - a_k| / |z_n - z_k| (pairs with equal z count 0).
template <class R>
__global__ void __launch_bounds__(kB) k_eps_star(Geo g, PosRef X, const R *B, const R *J, double *part, int nb)
{
    extern __shared__ double sh[];
    double *za = sh, *aa = sh + g.nz;
    __shared__ double red[kB];
    const int i = g.nx / 2, j = g.ny / 2;
    for (int k = threadIdx.x; k < g.nz; k += kB) {
        const long long a = gidx(g, i + 1, j + 1, k + 1);
        const double b[3] = {(double)B[a], (double)B[g.P + a], (double)B[2 * g.P + a]};
        const double jj[3] = {(double)J[a], (double)J[g.P + a], (double)J[2 * g.P + a]};
        aa[k] = dot3(jj, b) / dot3(b, b);
        double p[3];
        X.at(a, p);
        za[k] = p[2];
    }
    __syncthreads();
    const double SQRT2 = 1.414213562373095;   // the reference's literal (src/global.h:18)
    double m = 0;
    for (int k = threadIdx.x; k < g.nz; k += kB)
        for (int n = 0; n < g.nz; ++n)
            if (za[n] != za[k]) {
                const double v = SQRT2 * fabs((aa[n] - aa[k]) / (za[n] - za[k]));
                if (v > m) m = v;
            }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s && red[threadIdx.x + s] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[4 * nb] = red[0];
}

// out[0] = min convexity, out[1] = min wedge, out[2] = errB_1ez,
// out[3] = errXb_XbAn, out[4] = epsilon*
__global__ void k_extra_final(const double *part, int nb, long long N, int mesh, int errs, int eps, double *out)
{
    if (threadIdx.x != 0) return;
    if (mesh) {
        double c = 1, w = 1e300;
        for (int b = 0; b < nb; ++b) {
            c = fmin(c, part[b]);
            w = fmin(w, part[nb + b]);
        }
        out[0] = c;
        out[1] = w;
    }
    if (errs) {
        double s1 = 0, s2 = 0;
        for (int b = 0; b < nb; ++b) {
            s1 += part[2 * nb + b];
            s2 += part[3 * nb + b];
        }
        out[2] = sqrt(s1) / sqrt((double)N);
        out[3] = sqrt(s2) / sqrt((double)N);
    }
    if (eps) out[4] = part[4 * nb];
}

}  // namespace

int extra_diag_parts() { return 4 * 512 + 1; }

template <class R>
int launch_extra_diag(hipStream_t s, const Geo &g, const double *x64, const float *xh, const float *xl, const R *B,
                      const R *J, double ar, double az, double phi1, int mesh, int errs, int eps, double *cellVol,
                      double *convex, double *wedge, double *part, double *out)
{
    const int nb = 512;
    const PosRef X{x64, xh, xl, g.P};
    if (mesh || cellVol || convex || wedge)
        hipLaunchKernelGGL(k_mesh, dim3(nb), dim3(kB), 0, s, g, X, cellVol, convex, wedge, part, nb);
    if (errs) hipLaunchKernelGGL(k_errs<R>, dim3(nb), dim3(kB), 0, s, g, X, B, ar, az, phi1, part, nb);
    if (eps) {
        const size_t lds = 2 * (size_t)g.nz * sizeof(double);
        if (lds > 60 * 1024) return -1;
        hipLaunchKernelGGL(k_eps_star<R>, dim3(1), dim3(kB), lds, s, g, X, B, J, part, nb);
    }
    hipLaunchKernelGGL(k_extra_final, dim3(1), dim3(64), 0, s, part, nb, (long long)g.nx * g.ny * g.nz, mesh, errs,
                       eps, out);
    return 0;
}
template int launch_extra_diag<double>(hipStream_t, const Geo &, const double *, const float *, const float *,
                                       const double *, const double *, double, double, double, int, int, int,
                                       double *, double *, double *, double *, double *);
template int launch_extra_diag<float>(hipStream_t, const Geo &, const double *, const float *, const float *,
                                      const float *, const float *, double, double, double, int, int, int, double *,
                                      double *, double *, double *, double *);

}  // namespace gmd
