// glemur: command-line driver with the reference's run-directory contract
// (src/main.cu:44-394).  Run in a directory holding params.in and data/;
// writes cuda.info, data/B0.vtk, data/dump<k>.vtk, data/save.vtk and
// data/time_series.dat in the reference's formats.
//
//   glemur [-d device] [-p 64|32]
//     -p 64  FP64 stages (default; the reference's PRECISION = 64 build)
//     -p 32  FP32 stages with compensated positions
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unistd.h>
#include <vector>

#include "../../include/glemur.h"

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

    // time stepping, src/main.cu:259-384
    int timeIndex;
    for (timeIndex = 0; timeIndex < p.nCycle && dt > p.dtMin; timeIndex++) {
        gm_step_info s{};
        int r = gm_step(ctx, &s);
        if (r == GM_ERR_NAN) {
            printf("error: %s\n", gm_last_error());
            break;
        }
        if (r != GM_OK) return die("step");
        double maxDelta = s.maxDelta;
        if (s.accepted) {
            t = s.t;
            bool fields = false;
            if (timeIndex % p.nTs == 0) {
                if (gm_fields(ctx) != GM_OK || gm_reductions(ctx, &red) != GM_OK) return die("reductions");
                fields = true;
                gm_write_ts(".", &p, &red, t, s.dt_used, timeIndex, maxDelta, 1, 1);
            }
            dt = s.dt;  // already adapted (src/main.cu:362-363)
            bool dump = ceil((t - dt) / p.dtDump) != ceil(t / p.dtDump);
            bool save = ceil((t - dt) / p.dtSave) != ceil(t / p.dtSave) || timeIndex == p.nCycle - 1;
            if (dump || save) {
                if (!fields && gm_fields(ctx) != GM_OK) return die("fields");
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
            }
        } else {
            timeIndex -= 1;
            dt = s.dt;
        }
    }
    if (dt < p.dtMin) printf("dt < dtMin = %e, stopping simulation\n", p.dtMin);
    gm_destroy(ctx);
    printf("glemur terminated successfully\n");
    return 0;
}
