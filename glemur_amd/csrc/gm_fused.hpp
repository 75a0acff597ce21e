// Arguments of the fused stage kernels and the step controller.
#pragma once

#include "gm_device.hpp"

namespace gmd {

// Compiled tile capacity: a workgroup of kFusedThreads threads holds a rect
// of (TX+4) x (TY+4) <= kFusedPts points per plane in LDS.
constexpr int kFusedThreads = 512;
constexpr int kFusedPts = 864;

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

struct FusedArgs {
    Geo g;
    TableauF tab;
    int stage;                        // 0..5, or -1 for fields only
    int TX, TY, CZ, tiles_x, tiles_y, nchunks, nblocks;
    int emit, diag, check_done;
    // F64 buffers
    double *X64[3];                   // positions: current, candidate, stage-6 (3P each)
    const double *B0_64;
    double *K64;                      // stage slopes, 6 x 3P
    double *eB64, *eJ64, *eG64;       // emitted B, J, gradP (3P)
    // F32K buffers
    float *Xh[2], *Xl[2];
    const float *B0_32;
    float *K32;                       // 5 x 3P
    float *eB32, *eJ32, *eG32;
    double *edet;                     // emitted det (P)
    const double *epsx, *epsy, *epsz; // friction profile factors per axis
    double *part_delta;               // [nblocks]
    double *part_diag;                // [4][nblocks]
    Ctl *ctl;
};

struct CtlArgs {
    Ctl *ctl;
    const double *part_delta;
    const double *part_diag;
    int nblocks, diag, nbuf;          // nbuf: 3 (F64) or 2 (F32K) position buffers
    double maxError, dtMin, cell;     // cell = dx dy dz
    StepRecord *log;
    int log_cap;
};

int fused_max_points();
size_t fused_lds_bytes(int f32k);
void launch_fused(hipStream_t s, const FusedArgs &A, int f32k);
void launch_controller(hipStream_t s, const CtlArgs &A);

}  // namespace gmd
