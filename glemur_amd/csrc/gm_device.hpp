// Device-side declarations shared by the HIP translation units and the
// host runtime (which launches through these wrappers).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "device_math.hpp"

namespace gmd {

// Grid geometry and constants passed by value to every kernel.
struct Geo {
    int nx, ny, nz;
    long long sx, sxy, P;     // padded strides and point count
    double dx, dy, dz, dx1, dy1, dz1;
    double Lx, Ly, Lz, Ox, Oy, Oz;
    int xPeri, yPeri, zPeri;
    int epsilonProf, zUpdate, pressure;
    double beta, maxError, dtMin;
    int jxbAver;              // J x B averaged with the face neighbours (per-kernel path)
    double jxbAverWeight;
    int inertia;              // velocity equation with friction nu (per-kernel path)
    double nu;
    DistParams dist;
};

__host__ __device__ inline long long gidx(const Geo &g, long long i, long long j, long long k)
{
    return i + g.sx * j + g.sxy * k;
}

// Device control block of a context: the step controller state lives in
// device memory so steps can be enqueued without host round trips.
struct Ctl {
    double t, dt;           // time and dt of the next attempt
    double dt_used, maxDelta;
    double ratio, energy, jxb_b2, jmax;  // diagnostics of the last evaluated state
    double ratio_pt;        // pointwise max |JxB|/(|J||B|) of that state
    int cur;                // index of the current position buffer
    int accepted, nan_seen, done, converged, stop_reason;
    long long n_attempt, n_accept;
    long long max_attempt, n_cycle;      // run limits (n_cycle counts accepted steps)
    double stop_ratio;
    int diag_valid;
    int event;              // the last accepted step is due a dump or a save
    double ev_dump, ev_save;  // dump / save intervals for `event` (<= 0: off)
};

// src/main.cu:362-376: after an accepted step a dump (save) is due when
// (t - dt, t] holds a multiple of the interval, dt being the already
// adapted step of the next attempt
__host__ __device__ inline bool interval_crossed(double t, double dt, double every)
{
    return every > 0 && ceil((t - dt) / every) != ceil(t / every);
}

struct StepRecord {
    double t, dt_used, maxDelta, ratio, ratio_pt, energy, jxb_b2, jmax;
    int accepted, nan_seen;
};

// ---- per-kernel path (FP64, reference call structure) -----------------
void launch_b_jacb0(hipStream_t s, const Geo &g, const double *X, const double *B0, double *B, double *det);
void launch_bbound(hipStream_t s, const Geo &g, const double *X, const double *B0, double *B, double *det);
void launch_set_peri(hipStream_t s, const Geo &g, double *F);
void launch_set_grid_peri(hipStream_t s, const Geo &g, double *X);
void launch_jstokes(hipStream_t s, const Geo &g, const double *X, const double *B, double *J);
void launch_jclassic(hipStream_t s, const Geo &g, const double *X, const double *B0, const double *det, double *J);
void launch_jstokes4th(hipStream_t s, const Geo &g, const double *X, const double *B, double *J);
void launch_gradp(hipStream_t s, const Geo &g, const double *X, const double *det, double *G);
// With g.inertia, K holds the position slopes dt u, KU the velocity slopes
// and U / Utmp the stage velocity (padded SoA, like the positions); det is
// read only then.  Without inertia U, KU, Utmp and det may be null.
void launch_kk(hipStream_t s, const Geo &g, const Tableau *tab, int n, double dt, const double *X,
               const double *B, const double *J, const double *G, double *K, double *Xtmp, const double *det,
               const double *U, double *KU, double *Utmp);
// x_new (and u_new with inertia) and per-block error maxima of the
// positions; returns the number of partials written.
int launch_xnewstar(hipStream_t s, const Geo &g, const Tableau *tab, const double *X, const double *K,
                    double *Xnew, double *partials, int *nanflag, const double *U, const double *KU, double *Unew);
// U2 (incl. dx dy dz, boundary weights 1/2^faces) into out[0] and the
// NaN-blind max |u| into out[1]; part holds 2 * 512 doubles.
void launch_u_diag(hipStream_t s, const Geo &g, const double *U, const double *det, double *part, double *out);
// dst (padded SoA, inner points) <- src (inner AoS 3N)
void launch_inner_aos_to_padded(hipStream_t s, const Geo &g, const double *src, double *dst);
// Final NaN-blind max of n partials into out[0].
void launch_max_final(hipStream_t s, const double *partials, int n, double *out);
// the static scratch of the allocation-free xNewStar entry point
long long noalloc_scratch_doubles();
double *noalloc_scratch(int device);
// Diagnostics over inner points from B, J (padded SoA) and det:
// out[0] = B2 (incl. dx dy dz), out[1] = JMax, out[2] = JxB_B2Max,
// out[3] = JxB_JBMax, out[4] = max|JxB|, out[5] = max |J||B|.  `scratch`
// holds 6*blocks doubles.
template <class R>
void launch_diag(hipStream_t s, const Geo &g, const R *B, const R *J, const double *det, double *scratch,
                 int max_blocks, double *out);

// Mesh and error diagnostics (src/diagnostics.cu:33-110): out[0] min
// convexity, out[1] min octant wedge, out[2] errB_1ez, out[3] errXb_XbAn,
// out[4] epsilon*, each only when its flag is set; per-point cell volume,
// convexity and wedge minimum (inner points, x fastest) when the arrays are
// given.  part holds extra_diag_parts() doubles.  Returns -1 when nz is too
// large for epsilon* (more than 3840 planes).
int extra_diag_parts();
template <class R>
int launch_extra_diag(hipStream_t s, const Geo &g, const double *x64, const float *xh, const float *xl, const R *B,
                      const R *J, double ar, double az, double phi1, int mesh, int errs, int eps, double *cellVol,
                      double *convex, double *wedge, double *part, double *out);

// ---- layout conversion between host AoS padded and device SoA --------
// dst[c*P + a] = src[3a + c] (src in double) with an optional hi/lo split.
void launch_aos_to_soa_f64(hipStream_t s, long long P, const double *src, double *dst);
void launch_soa_to_aos_f64(hipStream_t s, long long P, const double *src, double *dst);
// F32K positions: host AoS <-> (hi, lo) triples (x, y, z of padded point a
// at 3a); single FP32 arrays: component order, or triples (positions' lo
// words, B0)
void launch_aos_to_soa_f32k(hipStream_t s, long long P, const double *src, float *hi, float *lo);
void launch_soa_to_aos_f32k(hipStream_t s, long long P, const float *hi, const float *lo, double *dst);
void launch_soa_to_aos_f32(hipStream_t s, long long P, const float *src, double *dst, bool triples);
void launch_aos_to_soa_f32(hipStream_t s, long long P, const double *src, float *dst, bool triples);
// Inner AoS (3N) <- padded SoA (3P), as J/gradP/kk are laid out on the host;
// triples: the source holds (x, y, z) of padded point a at 3a (the F32K
// slopes)
template <class R>
void launch_padded_to_inner_aos(hipStream_t s, const Geo &g, const R *src, double *dst, bool triples = false);
// mode 0: copy src -> dst, 1: read src, 2: write dst; `bytes` per array
void launch_hbm_probe(hipStream_t s, int mode, const void *src, void *dst, long long bytes);
void launch_fill_f64(hipStream_t s, double *p, long long n, double v);
void launch_fill_f32(hipStream_t s, float *p, long long n, float v);

}  // namespace gmd
