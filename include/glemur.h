/*
 * glemur.h -- C ABI of the MI355X-native GLEMuR relaxation library
 * (libglemur.so).  Plain C types only: no HIP or torch types cross it.
 *
 * The reference (SimonCan/glemur) is a monolithic CUDA executable; it has no
 * FFI.  Each entry point below replaces one step of its host call structure
 *   init -> boundaries -> current -> gradP -> rungeKutta -> diagnostics
 * and cites the reference function it stands in for.  Host arrays use the
 * reference's own layout (AoS, padded (nx+2)(ny+2)(nz+2) points, x fastest);
 * the library keeps its own device copies (FP64: one array per component;
 * FP32 with compensated positions: padded (x, y, z) triples).
 *
 * Every function returns GM_OK (0) or a negative GM_ERR_* code; the text of
 * the last error of the calling thread is available from gm_last_error().
 * One host thread per context; each context owns its own HIP stream.
 */
#ifndef GLEMUR_H
#define GLEMUR_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GM_OK               0
#define GM_ERR_ARG         -1
#define GM_ERR_HIP         -2
#define GM_ERR_IO          -3
#define GM_ERR_UNSUPPORTED -4
#define GM_ERR_NAN         -5
#define GM_ERR_PARSE       -6

/* Arithmetic of the stage kernels. */
#define GM_PREC_F64   0   /* FP64 throughout; follows the reference's operation order */
#define GM_PREC_F32K  1   /* FP32 stages, positions FP32 hi + compensated lo word;
                             determinant, error norm, energy and dt in FP64 */

/* Parameters: field for field the reference's struct Parameters with
 * REAL = double (src/global.h:44-108); bool is one byte. */
typedef struct gm_params {
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
} gm_params;

/* Reduction results (src/global.h:111-124) plus two measures of
 * force-freeness the reference does not compute: JxB_JBMax, the pointwise
 * max |JxB|/(|J||B|) (0/0 := 0), and JxBMax_JBMax = max|JxB| / max(|J||B|),
 * the relaxation criterion (the pointwise form stays near 1 wherever J is
 * at rounding level, e.g. in a uniform background field). */
typedef struct gm_reduction {
    double B2, B2res, BMax, JMax, JxB_B2Max, epsilonStar, errB_1ez, errXb_XbAn;
    double convex, wedgeMin, U2, UMax;
    double JxB_JBMax, JxBMax_JBMax;
} gm_reduction;

/* What one attempted step did (src/main.cu:309-383). */
typedef struct gm_step_info {
    int    accepted;    /* maxDelta < maxError */
    int    nan_seen;    /* a NaN error term was met (the reference's max skips it) */
    int    stopped;     /* the loop condition of src/main.cu:259 now fails */
    int    converged;   /* stop_ratio reached (gm_run only) */
    double dt_used;     /* dt of this attempt */
    double maxDelta;    /* grid max |x5 - x4| */
    double t;           /* time after the step */
    double dt;          /* dt for the next attempt */
    int64_t n_attempt, n_accept;
} gm_step_info;

/* Fields for gm_download / gm_upload_field. */
#define GM_FIELD_XB      0   /* 3P  positions (current)                    */
#define GM_FIELD_B0      1   /* 3P  initial field                          */
#define GM_FIELD_B       2   /* 3P  field B = J.B0/det (inner + ghost faces) */
#define GM_FIELD_DETJAC  3   /*  P  det of the grid Jacobian               */
#define GM_FIELD_J       4   /* 3N  current density, inner AoS             */
#define GM_FIELD_GRADP   5   /* 3N  pressure gradient, inner AoS           */
#define GM_FIELD_KK      6   /* 18N six stage slopes, inner AoS, back to back;
                                 36N with inertia: [stage][point][x0 x1 x2 u0 u1 u2]
                                 (src/rungeKutta.cu:236-244)                    */
#define GM_FIELD_XB_NEW  7   /* 3P  candidate positions of the last attempt */
#define GM_FIELD_XB_LO   8   /* 3P  compensation word of the positions (F32K; zeros in F64) */
#define GM_FIELD_CELLVOL 9   /*  N  signed cell volume, forward differences (src/maths.cu:443-494) */
#define GM_FIELD_CONVEXITY 10 /* N  +1 / -1: all 8 octant wedges positive (src/maths.cu:573-640) */
#define GM_FIELD_WEDGEMIN 11 /*  N  smallest oriented octant wedge (src/maths.cu:497-570) */
#define GM_FIELD_UU      12  /* 3N  velocity, inner AoS (inertia only; src/global.h:143) */

typedef struct gm_ctx gm_ctx;

/* ---- errors ---------------------------------------------------------- */
const char *gm_last_error(void);
int gm_version(void);

/* ---- parameters (src/readNamelist.f90) ------------------------------- */
/* Defaults of src/readNamelist.f90:92-150, literals rounded to single
 * precision as the Fortran default-real literals are, plus the derived
 * dx, dx1 (src/readNamelist.f90:185-186). */
int gm_params_default(gm_params *p);
/* readnamelist (src/readNamelist.f90:18-264): reads the groups &comp,
 * &start, &run, &io from a params.in file over the defaults. */
int gm_params_read(const char *path, gm_params *p);
/* Recompute dx, dy, dz, dx1, dy1, dz1 and initDistCode from the rest
 * (src/readNamelist.f90:185-186,257-262). */
int gm_params_derive(gm_params *p);

/* ---- host initial state (src/init.cu, src/boundaries.cu host twins) -- */
/* initState + initDistortion + setPeriHost + setGridPeriHost
 * (src/main.cu:166-169): B0 and xb, 3P doubles each, reference layout. */
int gm_init_state(const gm_params *p, double *B0, double *xb);
/* initResiduals (src/init.cu:13-26). */
double gm_init_residual(const gm_params *p);
/* Seeded braid: blobs from splitmix64(seed) as the workloads describe
 * (no reference counterpart; fills bInit="blobs" and the blob arrays). */
int gm_braid_seeded(gm_params *p, uint64_t seed, int nblobs_min, int nblobs_max,
                    int random_centres);

/* ---- context (src/memory.cu) ------------------------------------------ */
/* allocateMemory (src/memory.cu:10-73) + the constant uploads of
 * src/main.cu:227-256.  precision: GM_PREC_*. */
int gm_create(const gm_params *p, int precision, gm_ctx **ctx);
/* freeMemory (src/memory.cu:77-94). */
int gm_destroy(gm_ctx *ctx);
/* H2D of B0 and xb plus the device copies of src/main.cu:181-215. */
int gm_upload(gm_ctx *ctx, const double *B0, const double *xb, double t, double dt);
/* H2D of the velocity of inner points, inner AoS 3N (cudaMemcpy of h.uu,
 * src/main.cu:216-222).  Contexts with inertia only; gm_upload resets the
 * velocity to zero. */
int gm_upload_velocity(gm_ctx *ctx, const double *uu);
/* D2H of one field into reference layout (src/diagnostics.cu:145-152). */
int gm_download(gm_ctx *ctx, int field, double *host);
int gm_get_time(gm_ctx *ctx, double *t, double *dt);
int gm_set_time(gm_ctx *ctx, double t, double dt);
int gm_synchronize(gm_ctx *ctx);

/* ---- the reference call structure, one kernel per call (FP64) --------- */
int gm_b_jacb0(gm_ctx *ctx);                 /* B_JacB0,    src/rungeKutta.cu:28   */
int gm_set_bbound(gm_ctx *ctx);              /* setBbound,  src/boundaries.cu:637  */
int gm_set_peri(gm_ctx *ctx, int field);     /* setPeri,    src/boundaries.cu:645 (GM_FIELD_B) */
int gm_set_grid_peri(gm_ctx *ctx);           /* setGridPeri on xb, src/boundaries.cu:653 */
int gm_current(gm_ctx *ctx);                 /* current,    src/current.cu:1685    */
int gm_gradp(gm_ctx *ctx);                   /* gradP,      src/gradP.cu:160       */
/* kk<<<>>> (src/rungeKutta.cu:112) followed by xb <- xb_tmp (src/main.cu:288). */
int gm_rk_kk(gm_ctx *ctx, int n, double dt);
/* xNewStar<<<>>> (src/rungeKutta.cu:251) + findDevMax (src/maths.cu:101). */
int gm_rk_xnewstar(gm_ctx *ctx, double *maxDelta);
/* gm_rk_xnewstar without any allocation: no reference counterpart; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same results as
 * gm_rk_xnewstar; its scratch (per-block maxima, NaN flag, grid maximum) is
 * a static device array of 2^20 doubles (8 MiB) per device, so grids of
 * more than (2^20 - 8) * 256 inner points are rejected with GM_ERR_ARG.
 * Calls that overlap on one device race for that array. */
int gm_rk_xnewstar_noalloc(gm_ctx *ctx, double *maxDelta);
/* xb <- xb_new (src/main.cu:324), done by exchanging buffers. */
int gm_accept(gm_ctx *ctx);
/* reductions (src/diagnostics.cu:12-134) over the fields B, J, detJac the
 * last gm_fields / gm_b_jacb0+gm_set_bbound+gm_current left on the device:
 * B2, JMax, JxB_B2Max and JxB_JBMax. */
int gm_reductions(gm_ctx *ctx, gm_reduction *red);

/* ---- fused path ------------------------------------------------------- */
/* B, detJac, J (and gradP when pressure is on) for the current positions,
 * by the fused stage kernel (src/main.cu:343-353); any precision. */
int gm_fields(gm_ctx *ctx);
/* One attempted step (src/main.cu:259-383) with fused stage kernels and a
 * device-side step controller; blocks until the step is done. */
int gm_step(gm_ctx *ctx, gm_step_info *info);

/* Options of gm_run. */
typedef struct gm_run_opts {
    int64_t max_attempts;   /* upper bound on attempted steps               */
    int     n_cycle;        /* stop after this many accepted steps (<=0: params.nCycle) */
    double  stop_ratio;     /* stop when max|JxB| / max(|J||B|) <= stop_ratio (<=0: off) */
    int     batch;          /* steps enqueued between host checks (>=1)     */
    int     use_graph;      /* capture one batch in a hipGraph and replay it */
    double  dt_dump, dt_save;  /* stop after an accepted step that is due a dump
                               or a save by the reference's rule (src/main.cu:
                               362-376): (t - dt, t] with the adapted dt holds a
                               multiple of dt_dump or dt_save (<= 0: off)      */
} gm_run_opts;

typedef struct gm_run_stats {
    int64_t n_attempt, n_accept;
    double  t, dt;
    double  ratio;          /* last evaluated max|JxB| / max(|J||B|)         */
    double  ratio_pt;       /* last evaluated pointwise max |JxB|/(|J||B|)   */
    double  energy;         /* last evaluated B2                             */
    int     converged, stopped, nan_seen;
    double  seconds;        /* wall time of the run, host clock              */
    int     event;          /* the last accepted step is due a dump or save  */
} gm_run_stats;

/* Many attempted steps without host synchronisation inside a batch: the
 * device controller applies accept/reject and dt, and every kernel of a
 * finished run returns at once.  Relaxation criterion evaluated at every
 * attempt from the stage-0 fields. */
int gm_run(gm_ctx *ctx, const gm_run_opts *opts, gm_run_stats *stats);

/* Per-step record kept by the device controller (attempt order). */
typedef struct gm_step_record {
    double t, dt_used, maxDelta, ratio, ratio_pt, energy, jxb_b2, jmax;
    int32_t accepted, nan_seen;
} gm_step_record;
/* Copy the last n records (n <= capacity) into out; returns the count. */
int gm_step_log(gm_ctx *ctx, gm_step_record *out, int n);

/* ---- ensembles (no reference counterpart) ----------------------------- */
/* n independent relaxations (sizes and parameters may differ; pressure must
 * agree; line-tied faces) advanced together, in FP32 with compensated
 * positions (GM_PREC_F32K) or in FP64 (GM_PREC_F64, the fused FP64 stage
 * kernel: jMethod Stokes, no jxbAver, no inertia): each stage is one launch
 * over the workgroups of all members, one controller launch applies every
 * member's accept/reject, dt and stop rules, and finished members drop out
 * of every later launch.  A member's results equal its standalone
 * context's bit for bit with the same tiles and z chunks (FP64 positions
 * and decisions with any tiles; the energy sum's order follows the
 * chunks).  Members are ordinary contexts (gm_batch_context) for upload,
 * download, step logs and tiles; they belong to the ensemble and are
 * destroyed with it. */
typedef struct gm_batch gm_batch;
int gm_batch_create(const gm_params *params, int n, int precision, gm_batch **out);
int gm_batch_destroy(gm_batch *b);
int gm_batch_context(gm_batch *b, int i, gm_ctx **ctx);
/* Like gm_run for every member at once; stats has n entries (or NULL). */
int gm_batch_run(gm_batch *b, const gm_run_opts *opts, gm_run_stats *stats);
/* Ensemble tile autotuner: per stage kernel, times `reps` launches of the
 * whole ensemble for each candidate policy (every member's fewest-lines or
 * fewest-points rect within 512, 768 or 1024 points -- 512 for FP64
 * members -- shared z-chunk
 * length) on the current states and keeps the fastest (a later candidate
 * must be more than 2% faster to replace an earlier one); ms_best (6
 * entries, or NULL) receives the kept policies' mean times.  States, t and
 * dt unchanged. */
int gm_batch_autotune(gm_batch *b, int reps, double ms_best[6]);
/* Event timing of the ensemble's launches, as gm_timing_enable/read: index
 * 0..5 the stage launches (all members in one), 6 the controller launch. */
int gm_batch_timing_enable(gm_batch *b, int on);
int gm_batch_timing_read(gm_batch *b, double ms[7], int64_t count[7]);

/* ---- learned warm start (no reference counterpart) -------------------- */
/* A 3-level U-Net (3x3x3 convolutions with 32/64/128 channels, 3x3x3
 * average pooling with stride 2 down, trilinear x2 up, skip connections;
 * bfloat16 activations, FP32 accumulation on the bf16 matrix cores) maps
 * B0 of a braid on an n1 x n2 x n3 grid (each 4m+1) to the displacement of
 * its relaxed positions.  Input features per point: Bx, By, Bz - 1 and the
 * three coordinates scaled to [-1, 1].  Parameters are packed layer by
 * layer, W[co][t][ci] (tap t = (dx+1) + 3(dy+1) + 9(dz+1)) then the biases,
 * and last the output scale (prediction = scale * head output). */
typedef struct gm_unet gm_unet;
typedef struct gm_warm_info {
    int     used;          /* 1: warm state kept; 0: fell back to the cold state */
    int     mesh_ok;       /* min det > 0 and no NaN in the warm state */
    int64_t settle_attempts;           /* attempts run on the warm state before the check */
    double  ratio_cold, ratio_applied, ratio_warm;      /* max|JxB| / max(|J||B|): cold, right
                                          after the prediction, after the settle attempts */
    double  energy_cold, energy_applied, energy_warm;   /* B2 at the same three points */
    double  min_det;       /* of the warm state */
    double  ms;            /* wall time of the whole call */
} gm_warm_info;
#define GM_WARM_FORCE 1    /* keep the warm state whatever the checks say */
int64_t gm_unet_num_params(void);
/* capacity: training samples held in device memory (P * 28 bytes each). */
int gm_unet_create(int nx, int ny, int nz, int capacity, gm_unet **out);
int gm_unet_destroy(gm_unet *u);
/* He-normal weights from a splitmix64 stream; the output scale is set from
 * the data by the first gm_unet_train. */
int gm_unet_init_params(gm_unet *u, uint64_t seed);
int gm_unet_set_params(gm_unet *u, const double *params, int64_t n);   /* resets the optimiser */
int gm_unet_get_params(gm_unet *u, double *params, int64_t n);
/* Host features [P][8] -> prediction [P][3] (scale applied). */
int gm_unet_forward(gm_unet *u, const double *feat, double *y);
/* One sample: loss = mean over P and 3 of (head - target / scale)^2 and its
 * gradient in the packed parameters (n - 1 entries, no scale), or NULL. */
int gm_unet_loss_grad(gm_unet *u, const double *feat, const double *target, double *loss, double *grad);
/* Add the context's B0 features and its positions minus the uniform grid
 * as one training sample (the context is on the network's grid). */
int gm_unet_add_sample(gm_unet *u, gm_ctx *ctx);
int gm_unet_num_samples(gm_unet *u);
int gm_unet_clear_samples(gm_unet *u);
/* Adam (lr, 0.9, 0.999, 1e-8) over shuffled mini-batches of the resident
 * samples; loss = mean sample loss of the last epoch. */
int gm_unet_train(gm_unet *u, int epochs, int batch, double lr, uint64_t seed, double *loss);
/* Warm start of a context on the network's grid or its 2x refinement
 * (2n-1 points per axis): positions += trilinear x2 up-sampling of the
 * prediction, then `settle` attempted steps (gm_run) so that the
 * prediction's small-scale errors relax.  The warm state is kept when the
 * mesh stays valid and both max|JxB|/max(|J||B|) and the energy are then
 * below the cold state's (or flags has GM_WARM_FORCE); otherwise the cold
 * positions, time, dt and counters are restored (info->used = 0). */
int gm_warm_start(gm_ctx *ctx, gm_unet *u, int settle, int flags, gm_warm_info *info);

/* ---- kernel timing (HIP events on the context stream) ---------------- */
/* Enable/disable event timing of the fused stage kernels (resets the
 * totals); gm_timing_read returns the summed milliseconds and launch count
 * per kind: index 0..5 the stage kernels, 6 the step controller.  Graph
 * capture is off while timing. */
#define GM_TIMING_KINDS 7
int gm_timing_enable(gm_ctx *ctx, int on);
int gm_timing_read(gm_ctx *ctx, double ms[GM_TIMING_KINDS], int64_t count[GM_TIMING_KINDS]);
/* Bytes per grid point and stage the fused kernel must move (algorithmic). */
double gm_algorithmic_bytes_per_point_stage(int precision);
/* Plain device copy of `bytes`, `reps` times: mean kernel ms (the yardstick
 * for HBM byte counters; no reference counterpart).  Same as
 * gm_hbm_probe(bytes, 0, reps, ms_mean). */
int gm_copy_probe(int64_t bytes, int reps, double *ms_mean);
/* HBM yardsticks (no reference counterpart): mode 0 copies `bytes`, mode 1
 * reads `bytes`, mode 2 writes `bytes`; each lane keeps eight 16-byte words
 * in flight.  Mean kernel ms over `reps` launches. */
int gm_hbm_probe(int64_t bytes, int mode, int reps, double *ms_mean);
/* Fused-kernel tile configuration (x, y tile and z chunk); 0 = automatic. */
/* tx or ty <= 0 picks the tile automatically for about `cz` workgroups. */
int gm_set_tile(gm_ctx *ctx, int tx, int ty, int cz);
int gm_get_tile(gm_ctx *ctx, int *tx, int *ty, int *cz, int *nblocks);
/* Tile of one stage kernel (stage 0..5, -1 the fields-only pass); each stage
 * kernel has its own (gm_set_tile sets them all alike). */
int gm_get_stage_tile(gm_ctx *ctx, int stage, int *tx, int *ty, int *cz, int *nblocks);
/* Set the tile of one stage kernel (stage 0..5, -1 the fields-only pass):
 * tx x ty output columns per workgroup, z chunks of cz planes.  The rect
 * (tx+4)(ty+4) must fit the stage kernel's largest workgroup. */
int gm_set_stage_tile(gm_ctx *ctx, int stage, int tx, int ty, int cz);
/* Tile autotuner (no reference counterpart): for every stage kernel, times
 * `reps` launches of each candidate rect (the fewest-loads rect for each
 * workgroup size the kernel is built for, with one or two rounds of
 * workgroups over the CUs) on the current state and keeps the fastest (a
 * later candidate must be more than 2% faster to replace an earlier one);
 * ms_best (6 entries, or NULL) receives the kept tiles' mean times.  The
 * accepted state, t and dt are unchanged. */
int gm_autotune(gm_ctx *ctx, int reps, double ms_best[6]);

/* ---- files (src/io.cu) ------------------------------------------------ */
/* writeB0 (src/io.cu:167-205): data/B0.vtk in dir. */
int gm_write_b0(const char *dir, const gm_params *p, const double *B0, int precision);
/* dumpState (src/io.cu:209-386): n = -1 writes save.vtk, else dump<n>.vtk.
 * Arrays in reference layout; uu and J are inner AoS (uu written as
 * "ufield" when p->inertia); NULL skips optional fields. */
int gm_dump_state(const char *dir, const gm_params *p, double t, double dt, int n,
                  const double *xb, const double *B, const double *uu, const double *J,
                  const double *detJac, const double *cellVol, const double *convexity,
                  const double *wedgeMin, int precision);
/* readState (src/io.cu:390-499): xb, t, dt (and uu, inner AoS, when
 * p->inertia) from save.vtk, B0 from B0.vtk.  uu may be NULL otherwise. */
int gm_read_state(const char *dir, const gm_params *p, double *xb, double *B0, double *uu,
                  double *t, double *dt, int precision);
/* writeTs (src/io.cu:542-678); header when t == 0 (echoed to stdout when
 * echo != 0).  red may be NULL for the header.  jxb_jb_col != 0 appends the
 * columns JxB_JBMax and JxBMax_JBMax after the reference's. */
int gm_write_ts(const char *dir, const gm_params *p, const gm_reduction *red,
                double t, double dt, int it, double maxDelta, int echo, int jxb_jb_col);
/* writeCudaInfo (src/io.cu:10-82): device properties into <dir>/cuda.info. */
int gm_write_device_info(const char *dir, int active_device, int precision_bits);

#ifdef __cplusplus
}
#endif
#endif
