// Device state of one relaxation (a context), shared by the host units
// that launch work on it.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "gm_device.hpp"
#include "gm_fused.hpp"
#include "gm_internal.hpp"

// HIP-event timing of kernel launches by kind (0..5 the stage kernels, 6
// the step controller): a pair of events around each launch, folded into
// per-kind totals after a stream synchronisation.
struct EvTimer {
    bool on = false;
    std::vector<hipEvent_t> ev;
    std::vector<int> kind;
    int used = 0;
    double ms[GM_TIMING_KINDS] = {};
    long long count[GM_TIMING_KINDS] = {};
    void begin(hipStream_t s, int k)
    {
        if (!on) return;
        if (used + 2 > (int)ev.size()) {
            for (int i = 0; i < 64; ++i) {
                hipEvent_t e;
                (void)hipEventCreate(&e);
                ev.push_back(e);
                kind.push_back(0);
            }
        }
        kind[used] = k;
        (void)hipEventRecord(ev[used], s);
        used++;
    }
    void end(hipStream_t s)
    {
        if (!on) return;
        (void)hipEventRecord(ev[used], s);
        used++;
    }
    // after a synchronisation of the stream the events were recorded on
    void collect()
    {
        for (int i = 0; i + 1 < used; i += 2) {
            float t = 0;
            if (hipEventElapsedTime(&t, ev[i], ev[i + 1]) == hipSuccess) {
                ms[kind[i]] += t;
                count[kind[i]] += 1;
            }
        }
        used = 0;
    }
    void reset()
    {
        for (int i = 0; i < GM_TIMING_KINDS; ++i) { ms[i] = 0; count[i] = 0; }
    }
    void destroy()
    {
        for (auto e : ev) (void)hipEventDestroy(e);
        ev.clear();
        kind.clear();
        used = 0;
    }
};

struct gm_ctx {
    gm_params p{};
    int prec = GM_PREC_F64;
    int device = 0;
    hipStream_t stream = nullptr;
    gmd::Geo geo{};
    long long P = 0, N = 0;
    // F64
    double *X64[3] = {nullptr, nullptr, nullptr};
    double *B0_64 = nullptr, *K64 = nullptr, *B64 = nullptr, *J64 = nullptr, *G64 = nullptr, *Xtmp64 = nullptr;
    // F64 with inertia: velocity (padded SoA, inner points used; U64[ucur]
    // current, the other the candidate of the last attempt), its stage copy
    // and six velocity slopes
    double *U64[2] = {nullptr, nullptr}, *Utmp64 = nullptr, *KU64 = nullptr;
    int ucur = 0;
    // F32K
    float *Xh[2] = {nullptr, nullptr}, *Xl[2] = {nullptr, nullptr};
    float *B0_32 = nullptr, *K32 = nullptr, *B32 = nullptr, *J32 = nullptr, *G32 = nullptr;
    double *det = nullptr;
    double *epsx = nullptr, *epsy = nullptr, *epsz = nullptr;
    gmd::Ctl *ctl = nullptr;      // device
    gmd::Ctl *ctlh = nullptr;     // pinned host mirror
    double *part_delta = nullptr, *part_diag = nullptr;
    int part_cap = 0;
    gmd::Tableau *tab_dev = nullptr;
    gmd::TableauF tabf{};
    gmd::GhostPt *ghost = nullptr;
    long long ghost_off[6] = {0, 0, 0, 0, 0, 0};
    gmd::StepRecord *log = nullptr;
    int log_cap = 4096;
    double *red_out = nullptr, *red_host = nullptr;   // 16 doubles each
    double *extra_part = nullptr;                     // mesh / error diagnostics partials
    double *diag_scratch = nullptr;
    int diag_blocks = 2048;
    int *nanflag = nullptr;
    double *xns_part = nullptr;   // per-kernel path: per-block error maxima of k_xnewstar
    int cur = 0;                  // host copy of ctl->cur (valid after syncs)
    // F32K: odd stages march z downwards, so each stage starts on the planes
    // the previous one touched last, while they are in the 256 MB
    // memory-side cache; on up to 16 M padded points (a stage moves up to
    // about 1.3 GB there), off for larger grids and ensemble members
    bool zrev = false;
    bool zalt = false;
    // tiles
    // tiles per stage kernel (index 0..5 the stages, 6 the fields-only
    // pass): the stages exchange data through HBM only, so each takes the
    // rect and z chunking that suit its own workgroup size and occupancy
    int TX[7] = {}, TY[7] = {}, tiles_x[7] = {}, tiles_y[7] = {};
    int CZ[7] = {}, nchunks[7] = {}, nblocks[7] = {};
    int ncu = 256;
    EvTimer tm;                   // kernel timing (gm_timing_*)
    // graph of one batch
    hipGraphExec_t gexec = nullptr;
    int gbatch = 0;
};
