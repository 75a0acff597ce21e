// Arguments of the fused stage kernels and the step controller.
#pragma once

#include "gm_device.hpp"

namespace gmd {

// Tableau as the fused kernels use it: b, c, cs of src/rungeKutta.cu:10-25
// plus db[s][m] = b[s+1][m] - b[s][m] (the increments of the stage update,
// src/rungeKutta.cu:219) and ce = c - cs (the error weights).
struct TableauF {
    double b[42];
    double c[6];
    double cs[6];
    double db[36];
    double ce[6];
};

// Constant data of one line-tied ghost-face point: the two tangential
// Jacobian columns (from the fixed ghost-layer positions, with the
// reference's tile-edge rebuilding), the outer extrapolated ghost x2 and B0.
struct GhostPt {
    double tc0[3], tc1[3], x2[3], b0[3];
};

struct FusedArgs {
    Geo g;
    int stage;                        // 0..5, or -1 for fields only
    int TX, TY, CZ, tiles_x, tiles_y, nchunks, nblocks;
    int emit, diag, check_done;
    int face_table;                   // 1: every ghost face from the precomputed constants
                                      // (initial boundary distortion); 0: x/y faces direct (F32K)
    int rev;                          // F32K: march z downwards (plane q stored at nz + 1 - q)
    int rev_alt;                      // F32K: odd z chunks march the other way
    // F64 buffers
    double *X64[3];                   // positions: current, candidate, stage-6 (3P each)
    const double *B0_64;
    double *K64;                      // stage slopes, 6 x 3P
    double *eB64, *eJ64, *eG64;       // emitted B, J, gradP (3P)
    // F32K buffers
    float *Xh[2], *Xl[2];
    const float *B0_32;
    float *K32;                       // 5 x 3P: per slot, (x, y, z) of padded point a at 3a
    float *eB32, *eJ32, *eG32;
    double *edet;                     // emitted det (P)
    const double *epsx, *epsy, *epsz; // friction profile factors per axis
    const GhostPt *ghost;             // [face][side][v][u], offsets in ghost_off
    long long ghost_off[6];
    double *part_delta;               // [nblocks]
    double *part_diag;                // [6][nblocks]: energy, JMax, JxB_B2Max, JxB_JBMax, max|JxB|, max|J||B|
    Ctl *ctl;
};

struct CtlArgs {
    Ctl *ctl;
    const double *part_delta;
    const double *part_diag;
    int nb_delta, nb_diag;            // workgroups (partials) of the final stage and of stage 0
    int diag, nbuf;                   // nbuf: 3 (F64) or 2 (F32K) position buffers
    double maxError, dtMin, cell;     // cell = dx dy dz
    StepRecord *log;
    int log_cap;
};

// Largest rect (TX+4)(TY+4) a workgroup holds: one point per thread.
int fused_max_points(int f32k);
void set_tableau(const TableauF &t);
// Ghost-face constants from the (fixed) ghost positions and B0; X is the
// current position buffer (F64 SoA, or hi/lo for F32K).
void launch_ghost_setup(hipStream_t s, const FusedArgs &A, int f32k, GhostPt *out);
long long ghost_points(const Geo &g, long long off[6]);
void launch_fused(hipStream_t s, const FusedArgs &A, int f32k);
// FP32 stage kernel with compensated positions (kernels_f32.hip).
void set_tableau32(const TableauF &t);
void launch_fused32(hipStream_t s, const FusedArgs &A);
// Workgroup size the FP32 kernel uses for a rect of `rect` points, and the
// workgroups per CU its stage kernel is built for (stage -1: fields only).
int threads32(int rect);
int blocks_per_cu32(int rect, int stage, int press);
// Largest rect the FP32 kernel of this stage is built for (512, 768 or 1024).
int max_rect32(int stage, int press);
// Periodic ghost layers of the current F32K positions (buffer ctl->cur).
void launch_set_grid_peri32(hipStream_t s, const Geo &g, float *const Xh[2], float *const Xl[2], const Ctl *ctl);
// FP64 periodic ghost layers of the positions in buffer ctl->cur (read on the device)
void launch_set_grid_peri64(hipStream_t s, const FusedArgs &A);
void launch_controller(hipStream_t s, const CtlArgs &A);
// Ensemble launches: every workgroup of n instances in one launch of the
// FP32 stage kernel in `nt`-thread workgroups (512, or 768 / 1024 where the
// stage is built for them), and one controller workgroup per instance.
// batch_threads32: the workgroup size a rect needs there (at least 512).
int batch_threads32(int rect);
void launch_fused32_batch(hipStream_t s, int stage, int press, int nt, const FusedArgs *args, const int *start,
                          int n, int total);
// The FP64 stage kernel over n line-tied instances (512-thread workgroups).
void launch_fused64_batch(hipStream_t s, int stage, int press, const FusedArgs *args, const int *start, int n,
                          int total);
void launch_controller_batch(hipStream_t s, const CtlArgs *args, int n);

}  // namespace gmd
