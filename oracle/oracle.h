/*
 * GLEMuR CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A serial/OpenMP FP64 restatement of the reference relaxation step
 * (SimonCan/glemur, src/rungeKutta.cu, src/current.cu, src/gradP.cu,
 * src/boundaries.cu, src/maths.cu, src/init.cu, src/main.cu).  It is the
 * checker for the HIP library and the CPU baseline timed beside it; nothing
 * in the product links, loads or calls it.
 *
 * Layout is the reference's own: padded arrays of (nx+2)(ny+2)(nz+2) points,
 * vector components interleaved (AoS), x fastest:
 *     v[l + 3*(i + (nx+2)*(j + (ny+2)*k))]        (i,j,k padded indices)
 * Inner-only arrays (J, gradP, kk) use
 *     v[l + 3*(i + nx*(j + ny*k))]                 (i,j,k inner indices)
 * kk holds the six stage slopes back to back: kk[m*3N + ...].
 *
 * Arithmetic follows the reference operation by operation (same operand
 * order, no contraction: build with -ffp-contract=off) so the FP64 GPU path
 * can be compared bit for bit where it mirrors the same sequence.
 */
#ifndef GLEMUR_ORACLE_H
#define GLEMUR_ORACLE_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field-for-field the reference's struct Parameters with REAL = double
 * (src/global.h:44-108).  bool is one byte, as C++ bool and Fortran c_bool. */
typedef struct or_params {
    int    nx, ny, nz;
    double Lx, Ly, Lz;
    double Ox, Oy, Oz;
    double dx, dy, dz;
    double dx1, dy1, dz1;
    char   bInit[30];
    char   uInit[30];
    double ampl;
    double phi1, phi2;
    double L1, L2;
    double ar, az;
    int    nBlobs;
    double width;
    double minor, major;
    double stretch;
    double bGround;
    double blobXc[10], blobYc[10], blobZc[10], blobZl[10], blobTwist[10], blobScale[10];
    char   initDist[30];
    int    initDistCode;
    double initShearA;
    double initShearB;
    double initShearK;
    double pert;
    double twist;
    bool   fRestart;
    int    nCycle;
    double dt0;
    double dtMin;
    double maxError;
    bool   zUpdate;
    bool   inertia;
    bool   pressure;
    double nu;
    double beta;
    bool   jxbAver;
    double jxbAverWeight;
    bool   xPeri;
    bool   yPeri;
    bool   zPeri;
    bool   epsilonProf;
    char   jMethod[30];
    char   pMethod[30];
    int    nTs;
    double dtDump;
    double dtSave;
    bool   dumpJ, dumpDetJac, dumpCellVol, dumpConvexity, dumpWedgeMin;
    bool   redB2, redB2f, redBMax, redJMax, redJxB_B2Max, redEpsilonStar;
    bool   redErrB_1ez, redErrXb_XbAn, redConvex, redWedgeMin, redU2, redUMax;
} or_params;

/* Cash-Karp tableau, src/rungeKutta.cu:10-25 (b is 7x6 row-major). */
void or_assign_coefficients(double a[6], double b[42], double c[6], double cs[6]);

/* Host initial state: src/init.cu:30-226 (fields Pontin09, analytic, homZ,
 * blobs; undistorted grid), src/init.cu:388-419 (initDistortion),
 * src/boundaries.cu:492-634 (periodic host copies).  B0, xb: 3P each. */
int  or_init_state(const or_params *p, double *B0, double *xb);
/* src/init.cu:13-26 */
double or_init_residual(const or_params *p);

/* src/rungeKutta.cu:28-108: B and detJac at inner points. */
void or_b_jacb0(const or_params *p, const double *xb, const double *B0,
                double *B, double *detJac);
/* src/boundaries.cu:39-446: B and detJac on the six ghost faces. */
void or_update_bbound(const or_params *p, const double *xb, const double *B0,
                      double *B, double *detJac);
/* src/boundaries.cu:450-488, 532-582 (device twins of the host copies). */
void or_set_peri(const or_params *p, double *field);
void or_set_grid_peri(const or_params *p, double *xb);
/* src/current.cu:222-409: J (3N, inner AoS). */
void or_jstokes(const or_params *p, const double *xb, const double *B, double *J);
void or_jclassic(const or_params *p, const double *xb, const double *B0, const double *detJac, double *J);
void or_jstokes4th(const or_params *p, const double *xb, const double *B, double *J);
/* src/gradP.cu:10-156: gradP (3N, inner AoS). */
void or_gradp(const or_params *p, const double *xb, const double *detJac, double *gradP);
/* src/rungeKutta.cu:112-247 (magnetofrictional branch): writes kk[n] and
 * the inner points of xb_tmp. */
void or_kk(const or_params *p, const double b[42], int n, double dt,
           const double *xb, const double *B, const double *J,
           const double *gradP, double *kk, double *xb_tmp);
/* src/rungeKutta.cu:251-353 + src/maths.cu:58-153: inner points of xb_new;
 * returns the grid max of |x5 - x4| with the reference's NaN-blind max. */
double or_xnewstar(const or_params *p, const double c[6], const double cs[6],
                   const double *xb, const double *kk, double *xb_new, int *nan_seen);

/* src/rungeKutta.cu:112-247 (inertia branch): kk[n] holds six values per
 * point (position slope dt u, velocity slope); writes the inner points of
 * xb_tmp and uu_tmp (3N, inner AoS). */
void or_kk_inertia(const or_params *p, const double b[42], int n, double dt, const double *xb, const double *B,
                   const double *J, const double *gradP, const double *detJac, const double *uu, double *kk,
                   double *xb_tmp, double *uu_tmp);
/* src/rungeKutta.cu:251-353 (inertia): xb_new and uu_new; the error max is
 * taken over the positions only. */
double or_xnewstar_inertia(const or_params *p, const double c[6], const double cs[6], const double *xb,
                           const double *uu, const double *kk, double *xb_new, double *uu_new, int *nan_seen);
/* U2 (incl. dx dy dz) and UMax of the velocity (src/diagnostics.cu:114-130). */
double or_u2(const or_params *p, const double *uu, const double *detJac);
double or_umax(const or_params *p, const double *uu);

/* Reductions (src/maths.cu, src/diagnostics.cu:15-82). */
double or_b2(const or_params *p, const double *B, const double *detJac);   /* incl. dx dy dz */
double or_jmax(const or_params *p, const double *J);
double or_jxb_b2_max(const or_params *p, const double *B, const double *J);
/* max |JxB|/(|J||B|) with 0/0 := 0 (no reference counterpart). */
double or_jxb_jb_max(const or_params *p, const double *B, const double *J);
/* max |JxB| / max (|J||B|) (0 when J vanishes): the relaxation criterion
 * (no reference counterpart). */
double or_jxbmax_jbmax(const or_params *p, const double *B, const double *J);
void or_mesh(const or_params *p, const double *xb, double *cellVol, double *convex, double *wedge,
             double *convex_min, double *wedge_min);
double or_err_b_1ez(const or_params *p, const double *B);
double or_err_xb_xban(const or_params *p, const double *xb);
double or_epsilon_star(const or_params *p, const double *xb, const double *B, const double *J);

/* Whole-step driver state (src/main.cu:259-384). */
typedef struct or_state {
    double *B0, *xb, *B, *detJac, *J, *gradP, *kk, *xb_tmp, *xb_new;
    double t, dt;
    double a[6], b[42], c[6], cs[6];
    double *uu, *uu_tmp, *uu_new;   /* velocity, 3N inner AoS (inertia only; else NULL) */
} or_state;

/* Allocates every array and sets B = B0, xb_tmp = xb_new = xb, t = 0,
 * dt = dt0 (src/main.cu:181-256).  B0/xb are copied from the arguments;
 * with inertia the velocity starts at zero and kk holds 36N values. */
int  or_state_create(const or_params *p, const double *B0, const double *xb, or_state *s);
void or_state_free(or_state *s);

/* The six-stage loop + xNewStar of one attempted step at the state's dt
 * (src/main.cu:260-319).  Leaves xb = stage-6 position, xb_new = candidate. */
double or_attempt(const or_params *p, or_state *s, int *nan_seen);

/* One attempted step including accept/reject, the post-accept recompute of
 * B and J, and the dt update (src/main.cu:259-383).  Returns 1 on accept,
 * 0 on reject.  dt_used/maxDelta are reported for the time series. */
int  or_step(const or_params *p, or_state *s, double *dt_used, double *maxDelta,
             int *nan_seen);

/* B, detJac and J for the current xb (src/main.cu:343-353). */
void or_fields(const or_params *p, or_state *s);

int or_num_threads(void);
void or_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
