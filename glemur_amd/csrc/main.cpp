// glemur: command-line driver with the reference's run-directory contract
// (src/main.cu:44-394).  Run in a directory holding params.in and data/;
// writes cuda.info, data/B0.vtk, data/dump<k>.vtk, data/save.vtk and
// data/time_series.dat in the reference's formats.
//
//   glemur [-d device] [-p 64|32]
//     -p 64  FP64 stages (default; the reference's PRECISION = 64 build)
//     -p 32  FP32 stages with compensated positions
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unistd.h>
#include <vector>

#include "../../include/glemur.h"

// src/main.cu:362-376: a dump (save) is due after an accepted step when
// (t - dt, t] holds a multiple of the interval, dt being the adapted step
static bool gm_interval_due(double t, double dt, double every)
{
    return every > 0 && ceil((t - dt) / every) != ceil(t / every);
}

static int die(const char *what)
{
    printf("error: %s: %s\n", what, gm_last_error());
    return -1;
}

int main(int argc, char *argv[])
{
    int prec_bits = 64;
    long device = 0;
    int opt;
    printf("GLEMuR\n\n");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) {
        printf("error: could not locate HIP devices\n");
        return -1;
    }
    if (ndev == 0) {
        printf("error: there are no HIP devices on this machine\n");
        return -1;
    }
    while ((opt = getopt(argc, argv, "d:p:")) != -1) {
        if (opt == 'd') {
            char *end = nullptr;
            device = strtol(optarg, &end, 10);
            if (!end || *end) {
                printf("error: device number is not integer\n");
                return -1;
            }
            if (device > ndev - 1 || device < 0) {
                printf("error: device number is not valid\n");
                return -1;
            }
        } else if (opt == 'p') {
            prec_bits = atoi(optarg);
            if (prec_bits != 32 && prec_bits != 64) {
                printf("error: precision must be 32 or 64\n");
                return -1;
            }
        }
    }
    printf("PRECISION = %i\n", prec_bits);
    printf("HIP device count = %i\n", ndev);
    (void)hipSetDevice((int)device);
    printf("running on device %li\n", device);
    if (gm_write_device_info(".", (int)device, prec_bits) != GM_OK) {
        printf("error: could not write device properties into cuda.info\n");
        return -1;
    }

    gm_params p;
    gm_params_default(&p);
    if (gm_params_read("params.in", &p) != GM_OK) return die("params.in");
    printf("jMethod = %s\n", p.jMethod);
    printf("initial B: %s\n", p.bInit);
    if (p.pressure) printf("include pressure gradient with beta = %f\n", p.beta);

    const long long P3 = 3LL * (p.nx + 2) * (p.ny + 2) * (p.nz + 2);
    const long long N3 = 3LL * p.nx * p.ny * p.nz;
    std::vector<double> B0(P3), xb(P3), B(P3), J(N3), det(P3 / 3), cvol(N3 / 3), conv(N3 / 3), wmin(N3 / 3);
    std::vector<double> uu(p.inertia ? N3 : 0);   // inner AoS; zero unless restarted
    double t = 0, dt = p.dt0;
    gm_reduction red{};
    red.B2res = gm_init_residual(&p);
    if (!p.fRestart) {
        remove("data/save.vtk");
        if (gm_init_state(&p, B0.data(), xb.data()) != GM_OK) return die("initial state");
        if (gm_write_b0(".", &p, B0.data(), 64) != GM_OK) return die("B0");
        dt = p.dt0;
        t = 0;
        gm_write_ts(".", &p, &red, t, dt, 0, 0., 1, 1);
    } else {
        if (gm_read_state(".", &p, xb.data(), B0.data(), p.inertia ? uu.data() : nullptr, &t, &dt, 64) != GM_OK)
            return die("restart");
    }

    gm_ctx *ctx = nullptr;
    if (gm_create(&p, prec_bits == 64 ? GM_PREC_F64 : GM_PREC_F32K, &ctx) != GM_OK) return die("context");
    if (gm_upload(ctx, B0.data(), xb.data(), t, dt) != GM_OK) return die("upload");
    if (p.inertia && p.fRestart && gm_upload_velocity(ctx, uu.data()) != GM_OK) return die("upload velocity");

    // time stepping, src/main.cu:259-384.  The device controller runs the
    // attempts in batches without host round trips (gm_run); a batch ends at
    // the first accepted step that is due a dump or a save (the reference's
    // rule, the event stop of gm_run), after nCycle accepted steps, when dt
    // falls to dtMin, after 4096 attempts (the step log's capacity) and, when
    // a time-series column only the separate reductions give is requested,
    // at every nTs-th accepted step.  The other rows come from the step log:
    // the diagnostics recorded at the start of the attempt after an accepted
    // step are those of the state it produced.
    const bool extra = p.redEpsilonStar || p.redErrB_1ez || p.redErrXb_XbAn || p.redConvex || p.redWedgeMin ||
                       (p.inertia && (p.redU2 || p.redUMax));
    const int log_cap = 4096;
    std::vector<gm_step_record> recs(log_cap);
    int timeIndex = 0;
    // a due row whose diagnostics the next attempt's record will carry
    bool pending = false;
    int pend_it = 0;
    double pend_t = 0, pend_dt = 0, pend_md = 0;
    auto row_from = [&](const gm_step_record &r) {
        gm_reduction q{};
        q.B2 = r.energy;
        q.JMax = r.jmax;
        q.JxB_B2Max = r.jxb_b2;
        q.JxB_JBMax = r.ratio_pt;
        q.JxBMax_JBMax = r.ratio;
        q.B2res = red.B2res;
        gm_write_ts(".", &p, &q, pend_t, pend_dt, pend_it, pend_md, 1, 1);
        pending = false;
    };
    auto row_now = [&](int it, double tt, double dtu, double md) -> int {
        if (gm_fields(ctx) != GM_OK || gm_reductions(ctx, &red) != GM_OK) return -1;
        gm_write_ts(".", &p, &red, tt, dtu, it, md, 1, 1);
        return 0;
    };
    bool nan_stop = false;
    int64_t attempts_done = 0;
    const auto clock0 = std::chrono::steady_clock::now();
    double io_secs = 0;   // dumps and saves (downloads and file writes)
    while (timeIndex < p.nCycle && dt > p.dtMin && !nan_stop) {
        gm_run_opts o{};
        o.max_attempts = log_cap;
        int to_cycle = p.nCycle - timeIndex;
        if (extra) to_cycle = std::min(to_cycle, p.nTs - timeIndex % p.nTs);
        o.n_cycle = to_cycle;
        o.batch = 64;
        o.dt_dump = p.dtDump;
        o.dt_save = p.dtSave;
        gm_run_stats st{};
        const int r = gm_run(ctx, &o, &st);
        if (r != GM_OK && r != GM_ERR_NAN) return die("step");
        // this batch's attempts, oldest first
        const int n = (int)(st.n_attempt - attempts_done);
        attempts_done = st.n_attempt;
        const int got = gm_step_log(ctx, recs.data(), std::min(n, log_cap));
        if (got < 0) return die("step log");
        for (int k = 0; k < got; ++k) {
            const gm_step_record &rc = recs[k];
            if (pending && rc.ratio >= 0) row_from(rc);
            if (rc.nan_seen) {
                printf("error: NaN in the error estimate (the reference's max would skip it)\n");
                nan_stop = true;
                break;
            }
            if (rc.accepted == 1) {
                if (timeIndex % p.nTs == 0) {
                    pending = true;
                    pend_it = timeIndex;
                    pend_t = rc.t;
                    pend_dt = rc.dt_used;
                    pend_md = rc.maxDelta;
                }
                timeIndex++;
            }
        }
        t = st.t;
        dt = st.dt;
        if (nan_stop) break;
        const bool last = timeIndex >= p.nCycle;
        const bool dump = st.event && gm_interval_due(t, dt, p.dtDump);
        const bool save = (st.event && gm_interval_due(t, dt, p.dtSave)) || last;
        if (pending && (extra || dump || save || last || !(dt > p.dtMin))) {
            if (row_now(pend_it, pend_t, pend_dt, pend_md)) return die("reductions");
            pending = false;
        }
        if (dump || save) {
            const auto io0 = std::chrono::steady_clock::now();
            if (gm_fields(ctx) != GM_OK) return die("fields");
            if (gm_download(ctx, GM_FIELD_XB, xb.data()) || gm_download(ctx, GM_FIELD_B, B.data()) ||
                gm_download(ctx, GM_FIELD_J, J.data()) || gm_download(ctx, GM_FIELD_DETJAC, det.data()))
                return die("download");
            if ((p.dumpCellVol && gm_download(ctx, GM_FIELD_CELLVOL, cvol.data())) ||
                (p.dumpConvexity && gm_download(ctx, GM_FIELD_CONVEXITY, conv.data())) ||
                (p.dumpWedgeMin && gm_download(ctx, GM_FIELD_WEDGEMIN, wmin.data())) ||
                (p.inertia && gm_download(ctx, GM_FIELD_UU, uu.data())))
                return die("download");
            const double *u = p.inertia ? uu.data() : nullptr;
            if (dump)
                gm_dump_state(".", &p, t, dt, (int)ceil((t - dt) / p.dtDump), xb.data(), B.data(), u, J.data(),
                              det.data(), cvol.data(), conv.data(), wmin.data(), 64);
            if (save)
                gm_dump_state(".", &p, t, dt, -1, xb.data(), B.data(), u, J.data(), det.data(), cvol.data(),
                              conv.data(), wmin.data(), 64);
            io_secs += std::chrono::duration<double>(std::chrono::steady_clock::now() - io0).count();
        }
    }
    if (pending && row_now(pend_it, pend_t, pend_dt, pend_md)) return die("reductions");
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - clock0).count() - io_secs;
    printf("time stepping: %lld attempted steps (%d accepted) in %.3f s (dumps and saves: %.3f s more), "
           "%.4e grid-pt RK-stages/s\n",
           (long long)attempts_done, timeIndex, secs, io_secs,
           secs > 0 ? 6.0 * (double)p.nx * p.ny * p.nz * (double)attempts_done / secs : 0.0);
    if (dt < p.dtMin) printf("dt < dtMin = %e, stopping simulation\n", p.dtMin);
    gm_destroy(ctx);
    printf("glemur terminated successfully\n");
    return 0;
}
