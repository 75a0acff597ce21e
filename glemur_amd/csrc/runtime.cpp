// Host runtime of libglemur: device state of one relaxation (a context),
// the per-kernel entry points of the reference's call structure, and the
// fused step driver with its device-side controller.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gm_ctx.hpp"
#include "gm_device.hpp"
#include "gm_fused.hpp"
#include "gm_internal.hpp"

using gm::fail;

#define HIPCHK(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(GM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)


namespace {

template <class T>
int dalloc(T **p, size_t n)
{
    if (n == 0) { *p = nullptr; return GM_OK; }
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) return fail(GM_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    return GM_OK;
}
template <class T>
void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

gmd::Geo make_geo(const gm_params &p)
{
    gmd::Geo g{};
    g.nx = p.nx; g.ny = p.ny; g.nz = p.nz;
    g.sx = p.nx + 2;
    g.sxy = (long long)(p.nx + 2) * (p.ny + 2);
    g.P = g.sxy * (p.nz + 2);
    g.dx = p.dx; g.dy = p.dy; g.dz = p.dz;
    g.dx1 = p.dx1; g.dy1 = p.dy1; g.dz1 = p.dz1;
    g.Lx = p.Lx; g.Ly = p.Ly; g.Lz = p.Lz;
    g.Ox = p.Ox; g.Oy = p.Oy; g.Oz = p.Oz;
    g.xPeri = p.xPeri; g.yPeri = p.yPeri; g.zPeri = p.zPeri;
    g.epsilonProf = p.epsilonProf; g.zUpdate = p.zUpdate; g.pressure = p.pressure;
    g.beta = p.beta; g.maxError = p.maxError; g.dtMin = p.dtMin;
    g.jxbAver = p.jxbAver; g.jxbAverWeight = p.jxbAverWeight;
    g.inertia = p.inertia; g.nu = p.nu;
    g.dist.code = p.initDistCode;
    g.dist.shearA = p.initShearA; g.dist.shearB = p.initShearB; g.dist.shearK = p.initShearK;
    g.dist.dx = p.dx; g.dist.dy = p.dy; g.dist.Ox = p.Ox; g.dist.Oy = p.Oy;
    g.dist.Lx = p.Lx; g.dist.Ly = p.Ly;
    return g;
}

// Cash-Karp tableau exactly as src/rungeKutta.cu:10-25 assigns it.
void cash_karp(double b[42], double c[6], double cs[6])
{
    const double B[7][6] = {
        {0, 0, 0, 0, 0, 0},
        {0.2, 0, 0, 0, 0, 0},
        {3 / 40., 9 / 40., 0, 0, 0, 0},
        {0.3, -0.9, 1.2, 0, 0, 0},
        {-11 / 54., 2.5, -70 / 27., 35 / 27., 0, 0},
        {1631 / 55296., 175 / 512., 575 / 13824., 44275 / 110592., 253 / 4096., 0},
        {0, 0, 0, 0, 0, 0}};
    const double C[6] = {37 / 378., 0, 250 / 621., 125 / 594., 0, 512 / 1771.};
    const double CS[6] = {2825 / 27648., 0, 18575 / 48384., 13525 / 55296., 277 / 14336., 0.25};
    memcpy(b, B, sizeof B);
    memcpy(c, C, sizeof C);
    memcpy(cs, CS, sizeof CS);
}

bool any_peri(const gm_params &p) { return p.xPeri || p.yPeri || p.zPeri; }

bool distorted_ghosts(const gm_params &p)
{
    std::string d(p.initDist, strnlen(p.initDist, sizeof p.initDist));
    while (!d.empty() && (d.back() == ' ' || d.back() == '\0')) d.pop_back();
    return d == "initShearX" || d == "centerShift";
}

// Workgroups per CU a stage kernel is built for (k: 0..5 stages, 6 fields).
// (F32K grids with periodic faces run every stage in 512-thread workgroups.)
int stage_blocks_per_cu(const gm_ctx *c, int k)
{
    if (c->prec != GM_PREC_F32K) return 1;
    const int rect = any_peri(c->p) ? 512 : (c->TX[k] + 4) * (c->TY[k] + 4);
    return gmd::blocks_per_cu32(rect, k == 6 ? -1 : k, c->p.pressure);
}

// Largest rect a stage kernel holds (one point per thread).
int stage_max_rect(const gm_ctx *c, int k)
{
    if (c->prec != GM_PREC_F32K) return gmd::fused_max_points(0);
    if (any_peri(c->p)) return 512;
    return gmd::max_rect32(k == 6 ? -1 : k, c->p.pressure);
}

void set_tile_k(gm_ctx *c, int k, int tx, int ty, int cz)
{
    c->TX[k] = std::min(tx, c->p.nx);
    c->TY[k] = std::min(ty, c->p.ny);
    c->tiles_x[k] = (c->p.nx + c->TX[k] - 1) / c->TX[k];
    c->tiles_y[k] = (c->p.ny + c->TY[k] - 1) / c->TY[k];
    c->CZ[k] = std::max(1, std::min(cz, c->p.nz));
    c->nchunks[k] = (c->p.nz + c->CZ[k] - 1) / c->CZ[k];
    c->nblocks[k] = c->tiles_x[k] * c->tiles_y[k] * c->nchunks[k];
}

// z chunks per tile column: the workgroups run in rounds of `slots`
// (resident workgroups over all CUs) and each marches its chunk plus five
// planes of pipeline fill, so the launch takes about
// ceil(chunks x cols / slots) x (ceil(nz / chunks) + 5) plane steps; the
// fewest such steps, chunks never thinner than 8 planes
int best_chunks(int nz, int cols, int slots)
{
    int best = 1;
    long long best_cost = -1;
    for (int nch = 1; nch <= std::max(1, nz / 8); ++nch) {
        const long long rounds = ((long long)nch * cols + slots - 1) / slots;
        const long long cost = rounds * ((nz + nch - 1) / nch + 5);
        if (best_cost < 0 || cost < best_cost) { best_cost = cost; best = nch; }
    }
    return best;
}

// 128-byte lines a launch loads per plane of one array for tiles of
// (tlx x tly) rects of rw x rh points: a rect row of rw points at an
// arbitrary offset within a line touches (4 rw + 124) / 128 lines on
// average if a point is 4 bytes.  The F32K arrays hold 12-byte (x, y, z)
// triples, but the 4-byte count is kept on purpose as a ranking heuristic:
// counting 12 bytes picks the same tiles at 129 x 129 x 385 and moves the
// other workloads by under 1.5% either way
double rect_lines(int tlx, int tly, int rw, int rh)
{
    return (double)tlx * tly * rh * (4.0 * rw + 124.0) / 128.0;
}

// Tile choice per stage kernel: among the rects (TX+4)(TY+4) that fit the
// kernel's workgroup,
//  - F32K: the fewest 128-byte lines loaded per output point.  A load
//    instruction costs about one request per line its lanes touch, so wide
//    rects beat square ones with less halo: at 129 x 129 x 385, 65 x 10
//    tiles take stages 0-4 4-7% faster than 26 x 26;
//  - FP64: the fewest loaded points per output point (its register-queue
//    march is 20% slower on the wide rects);
// then the z chunks of best_chunks with `target` x the kernel's workgroups
// per CU resident at a time.  (Stage 0 took 512-point rects, two
// workgroups per CU, while its loads were issued one plane step ahead; with
// two steps it is faster on the wide rects: 98 -> 92 us at 129 x 129 x 385.)
void auto_tiles(gm_ctx *c, int target, bool member)
{
    const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    const bool lines = c->prec == GM_PREC_F32K;
    (void)member;   // ensemble members take the same rule
    for (int k = 0; k < 7; ++k) {
        const int cap = stage_max_rect(c, k);
        int bx = 4, by = 2;
        double best = 1e30;
        for (int ty = 2; ty <= 96; ++ty)
            for (int txm = 4; (txm + 4) * (ty + 4) <= cap; ++txm) {
                int tlx = (nx + txm - 1) / txm, tly = (ny + ty - 1) / ty;
                int tx = (nx + tlx - 1) / tlx, tyy = (ny + tly - 1) / tly;
                if ((tx + 4) * (tyy + 4) > cap) continue;
                const double cost = (lines ? rect_lines(tlx, tly, tx + 4, tyy + 4)
                                           : (double)tlx * tly * (tx + 4) * (tyy + 4)) / ((double)nx * ny);
                if (cost < best - 1e-12) { best = cost; bx = tx; by = tyy; }
            }
        set_tile_k(c, k, bx, by, nz);
        const int nch = best_chunks(nz, c->tiles_x[k] * c->tiles_y[k], target * stage_blocks_per_cu(c, k));
        set_tile_k(c, k, bx, by, (nz + nch - 1) / nch);
    }
}

// Workgroup rounds per launch: one round over the CUs (at each stage
// kernel's own occupancy).
int default_blocks(const gm_ctx *c)
{
    return c->ncu;
}

int ensure_partials(gm_ctx *c)
{
    int need = 0;
    for (int k = 0; k < 7; ++k) need = std::max(need, c->nblocks[k]);
    if (need <= c->part_cap) return GM_OK;
    dfree(c->part_delta);
    dfree(c->part_diag);
    int r;
    if ((r = dalloc(&c->part_delta, (size_t)need)) || (r = dalloc(&c->part_diag, (size_t)6 * need))) return r;
    c->part_cap = need;
    return GM_OK;
}

gmd::FusedArgs fused_args(gm_ctx *c, int stage, int emit, int diag, int check_done)
{
    gmd::FusedArgs A{};
    A.g = c->geo;
    A.stage = stage;
    const int k = stage < 0 ? 6 : stage;
    A.TX = c->TX[k]; A.TY = c->TY[k]; A.CZ = c->CZ[k];
    A.tiles_x = c->tiles_x[k]; A.tiles_y = c->tiles_y[k]; A.nchunks = c->nchunks[k]; A.nblocks = c->nblocks[k];
    A.emit = emit; A.diag = diag; A.check_done = check_done;
    // an initial distortion (src/init.cu:388-419) moves the x/y ghost
    // layers off the lattice, where the reference's tile-edge rebuilding and
    // the ghost layer itself differ: those runs take every face from the table
    A.face_table = distorted_ghosts(c->p);
    A.rev = c->zrev && stage >= 0 && (stage & 1);
    // odd chunks the other way when all of the launch's workgroups are
    // resident at once (neighbouring chunks then run side by side); not
    // stage 0, whose energy sum follows the march order
    A.rev_alt = c->zalt && stage >= 1 && c->nblocks[k] <= c->ncu * stage_blocks_per_cu(c, k);
    for (int i = 0; i < 3; ++i) A.X64[i] = c->X64[i];
    A.B0_64 = c->B0_64; A.K64 = c->K64;
    A.eB64 = c->B64; A.eJ64 = c->J64; A.eG64 = c->G64;
    for (int i = 0; i < 2; ++i) { A.Xh[i] = c->Xh[i]; A.Xl[i] = c->Xl[i]; }
    A.B0_32 = c->B0_32; A.K32 = c->K32;
    A.eB32 = c->B32; A.eJ32 = c->J32; A.eG32 = c->G32;
    A.edet = c->det;
    A.epsx = c->epsx; A.epsy = c->epsy; A.epsz = c->epsz;
    A.ghost = c->ghost;
    for (int i = 0; i < 6; ++i) A.ghost_off[i] = c->ghost_off[i];
    A.part_delta = c->part_delta; A.part_diag = c->part_diag;
    A.ctl = c->ctl;
    return A;
}

void ev_begin(gm_ctx *c, int kind) { c->tm.begin(c->stream, kind); }
void ev_end(gm_ctx *c) { c->tm.end(c->stream); }

// Enqueue one attempted step: stages 0..5 and the controller.
void enqueue_step(gm_ctx *c, int diag, int check_done)
{
    int f32k = c->prec == GM_PREC_F32K;
    for (int s = 0; s < 6; ++s) {
        gmd::FusedArgs A = fused_args(c, s, 0, s == 0 ? diag : 0, check_done);
        ev_begin(c, s);
        gmd::launch_fused(c->stream, A, f32k);
        ev_end(c);
    }
    gmd::CtlArgs C{};
    C.ctl = c->ctl;
    C.part_delta = c->part_delta;
    C.part_diag = c->part_diag;
    C.nb_delta = c->nblocks[5];
    C.nb_diag = c->nblocks[0];
    C.diag = diag;
    C.nbuf = f32k ? 2 : 3;
    C.maxError = c->p.maxError;
    C.dtMin = c->p.dtMin;
    C.cell = c->p.dx * c->p.dy * c->p.dz;
    C.log = c->log;
    C.log_cap = c->log_cap;
    ev_begin(c, 6);
    gmd::launch_controller(c->stream, C);
    ev_end(c);
    // periodic ghost layers of the positions the controller left current
    if (f32k && any_peri(c->p)) gmd::launch_set_grid_peri32(c->stream, c->geo, c->Xh, c->Xl, c->ctl);
    if (!f32k && any_peri(c->p)) gmd::launch_set_grid_peri64(c->stream, fused_args(c, 0, 0, 0, 0));
}

int sync_ctl(gm_ctx *c)
{
    HIPCHK(hipMemcpyAsync(c->ctlh, c->ctl, sizeof(gmd::Ctl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->tm.collect();
    c->cur = c->ctlh->cur;
    return GM_OK;
}
int push_ctl(gm_ctx *c)
{
    HIPCHK(hipMemcpyAsync(c->ctl, c->ctlh, sizeof(gmd::Ctl), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return GM_OK;
}

int need_f64(gm_ctx *c, const char *what)
{
    if (c->prec != GM_PREC_F64)
        return fail(GM_ERR_UNSUPPORTED, std::string(what) + ": the per-kernel path runs in FP64 contexts only");
    return GM_OK;
}

}  // namespace

// ===================================================================== API
extern "C" int gm_create(const gm_params *pp, int precision, gm_ctx **out)
{
    if (!pp || !out) return fail(GM_ERR_ARG, "gm_create: null pointer");
    if (precision != GM_PREC_F64 && precision != GM_PREC_F32K) return fail(GM_ERR_ARG, "unknown precision");
    if (pp->nx < 2 || pp->ny < 2 || pp->nz < 2) return fail(GM_ERR_ARG, "grid needs at least 2 inner points per axis");
    if (pp->inertia && precision != GM_PREC_F64)
        return fail(GM_ERR_UNSUPPORTED, "inertia = t runs on the FP64 per-kernel path only");
    if (pp->jxbAver && precision != GM_PREC_F64)
        return fail(GM_ERR_UNSUPPORTED, "jxbAver = t runs on the FP64 per-kernel path only");
    const bool classic = gm::starts(pp->jMethod, "Classic "), stokes4 = gm::starts(pp->jMethod, "Stokes4th ");
    if (!gm::starts(pp->jMethod, "Stokes ") && !classic && !stokes4)
        return fail(GM_ERR_UNSUPPORTED, std::string("jMethod '") + std::string(pp->jMethod, 29) +
                                            "': Stokes, Stokes4th and Classic are built");
    if ((classic || stokes4) && precision != GM_PREC_F64)
        return fail(GM_ERR_UNSUPPORTED, "jMethod = Classic / Stokes4th run on the FP64 per-kernel path only");
    if (classic && (pp->xPeri || pp->yPeri || pp->zPeri))
        return fail(GM_ERR_UNSUPPORTED, "jMethod = Classic with periodic faces: the reference leaves detJac unset "
                                        "on periodic ghost faces, which Classic reads");
    if (stokes4 && (pp->xPeri || pp->yPeri || pp->zPeri))
        return fail(GM_ERR_UNSUPPORTED, "jMethod = Stokes4th with periodic faces: the octilaterals read B on "
                                        "ghost edges, which the periodic copies do not define");
    if (pp->pressure && !gm::starts(pp->pMethod, "Classic "))
        return fail(GM_ERR_UNSUPPORTED, "pMethod: only Classic is built");
    if (precision == GM_PREC_F32K && 12LL * gm::npad(*pp) >= 0x7fffffffLL)
        return fail(GM_ERR_UNSUPPORTED, "F32K: the grid exceeds the 32-bit byte offsets of the stage kernels' buffer "
                                        "accesses (3 (nx+2)(ny+2)(nz+2) floats per array, at most 2 GiB)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(GM_ERR_HIP, "no HIP device");
    gm_ctx *c = new gm_ctx();
    c->p = *pp;
    c->prec = precision;
    (void)hipGetDevice(&c->device);
    c->geo = make_geo(c->p);
    c->P = c->geo.P;
    c->zrev = precision == GM_PREC_F32K && c->P <= (16LL << 20);
    c->zalt = precision == GM_PREC_F32K;
    c->N = (long long)pp->nx * pp->ny * pp->nz;
    const size_t P3 = 3 * (size_t)c->P;
    int r = GM_OK;
    auto bail = [&](int code) {
        gm_destroy(c);
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(fail(GM_ERR_HIP, "stream"));
    if (precision == GM_PREC_F64) {
        for (int i = 0; i < 3; ++i)
            if ((r = dalloc(&c->X64[i], P3))) return bail(r);
        if ((r = dalloc(&c->B0_64, P3)) || (r = dalloc(&c->K64, 6 * P3)) || (r = dalloc(&c->B64, P3)) ||
            (r = dalloc(&c->J64, P3)) || (r = dalloc(&c->G64, P3)))
            return bail(r);
        (void)hipMemset(c->K64, 0, 6 * P3 * sizeof(double));
        (void)hipMemset(c->J64, 0, P3 * sizeof(double));
        (void)hipMemset(c->G64, 0, P3 * sizeof(double));
        if (pp->inertia) {
            if ((r = dalloc(&c->U64[0], P3)) || (r = dalloc(&c->U64[1], P3)) || (r = dalloc(&c->Utmp64, P3)) ||
                (r = dalloc(&c->KU64, 6 * P3)))
                return bail(r);
            for (double *u : {c->U64[0], c->U64[1], c->Utmp64}) (void)hipMemset(u, 0, P3 * sizeof(double));
            (void)hipMemset(c->KU64, 0, 6 * P3 * sizeof(double));
        }
    } else {
        for (int i = 0; i < 2; ++i)
            if ((r = dalloc(&c->Xh[i], P3)) || (r = dalloc(&c->Xl[i], P3))) return bail(r);
        if ((r = dalloc(&c->B0_32, P3)) || (r = dalloc(&c->K32, 5 * P3)) || (r = dalloc(&c->B32, P3)) ||
            (r = dalloc(&c->J32, P3)) || (r = dalloc(&c->G32, P3)))
            return bail(r);
        (void)hipMemset(c->K32, 0, 5 * P3 * sizeof(float));
        (void)hipMemset(c->J32, 0, P3 * sizeof(float));
        (void)hipMemset(c->G32, 0, P3 * sizeof(float));
    }
    if ((r = dalloc(&c->det, (size_t)c->P))) return bail(r);
    (void)hipMemset(c->det, 0, c->P * sizeof(double));
    // friction profile factors (src/rungeKutta.cu:180-183), one table per axis
    std::vector<double> ex(pp->nx), ey(pp->ny), ez(pp->nz);
    for (int i = 0; i < pp->nx; ++i) ex[i] = gmd::eps_axis(i, pp->nx);
    for (int i = 0; i < pp->ny; ++i) ey[i] = gmd::eps_axis(i, pp->ny);
    for (int i = 0; i < pp->nz; ++i) ez[i] = gmd::eps_axis(i, pp->nz);
    if ((r = dalloc(&c->epsx, ex.size())) || (r = dalloc(&c->epsy, ey.size())) || (r = dalloc(&c->epsz, ez.size())))
        return bail(r);
    (void)hipMemcpy(c->epsx, ex.data(), ex.size() * 8, hipMemcpyHostToDevice);
    (void)hipMemcpy(c->epsy, ey.data(), ey.size() * 8, hipMemcpyHostToDevice);
    (void)hipMemcpy(c->epsz, ez.data(), ez.size() * 8, hipMemcpyHostToDevice);
    // tableaux
    gmd::Tableau tab{};
    cash_karp(tab.b, tab.c, tab.cs);
    memcpy(c->tabf.b, tab.b, sizeof tab.b);
    memcpy(c->tabf.c, tab.c, sizeof tab.c);
    memcpy(c->tabf.cs, tab.cs, sizeof tab.cs);
    for (int s = 0; s < 6; ++s)
        for (int m = 0; m < 6; ++m) c->tabf.db[s * 6 + m] = tab.b[m + (s + 1) * 6] - tab.b[m + s * 6];
    for (int m = 0; m < 6; ++m) c->tabf.ce[m] = tab.c[m] - tab.cs[m];
    if ((r = dalloc(&c->tab_dev, 1))) return bail(r);
    (void)hipMemcpy(c->tab_dev, &tab, sizeof tab, hipMemcpyHostToDevice);
    gmd::set_tableau(c->tabf);
    if ((r = dalloc(&c->ghost, (size_t)gmd::ghost_points(c->geo, c->ghost_off)))) return bail(r);
    // control block, log, reductions
    if ((r = dalloc(&c->ctl, 1))) return bail(r);
    if (hipHostMalloc((void **)&c->ctlh, sizeof(gmd::Ctl), hipHostMallocDefault) != hipSuccess)
        return bail(fail(GM_ERR_HIP, "hipHostMalloc"));
    memset(c->ctlh, 0, sizeof(gmd::Ctl));
    c->ctlh->dt = pp->dt0;
    c->ctlh->n_cycle = 1LL << 62;
    c->ctlh->max_attempt = 1LL << 62;
    if ((r = dalloc(&c->log, (size_t)c->log_cap))) return bail(r);
    if ((r = dalloc(&c->red_out, 16)) || (r = dalloc(&c->diag_scratch, (size_t)6 * c->diag_blocks)) ||
        (r = dalloc(&c->extra_part, (size_t)gmd::extra_diag_parts())) ||
        (r = dalloc(&c->nanflag, 1)))
        return bail(r);
    if (hipHostMalloc((void **)&c->red_host, 16 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return bail(fail(GM_ERR_HIP, "hipHostMalloc"));
    {
        hipDeviceProp_t pr;
        c->ncu = hipGetDeviceProperties(&pr, c->device) == hipSuccess ? pr.multiProcessorCount : 256;
    }
    auto_tiles(c, default_blocks(c), false);
    if ((r = ensure_partials(c))) return bail(r);
    if ((r = push_ctl(c))) return bail(r);
    *out = c;
    return GM_OK;
}

extern "C" int gm_destroy(gm_ctx *c)
{
    if (!c) return GM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    for (int i = 0; i < 3; ++i) dfree(c->X64[i]);
    dfree(c->B0_64); dfree(c->K64); dfree(c->B64); dfree(c->J64); dfree(c->G64); dfree(c->Xtmp64);
    dfree(c->U64[0]); dfree(c->U64[1]); dfree(c->Utmp64); dfree(c->KU64);
    for (int i = 0; i < 2; ++i) { dfree(c->Xh[i]); dfree(c->Xl[i]); }
    dfree(c->B0_32); dfree(c->K32); dfree(c->B32); dfree(c->J32); dfree(c->G32);
    dfree(c->det); dfree(c->epsx); dfree(c->epsy); dfree(c->epsz);
    dfree(c->ctl); dfree(c->part_delta); dfree(c->part_diag); dfree(c->tab_dev); dfree(c->log); dfree(c->ghost);
    dfree(c->red_out); dfree(c->diag_scratch); dfree(c->nanflag); dfree(c->extra_part); dfree(c->xns_part);
    if (c->ctlh) (void)hipHostFree(c->ctlh);
    if (c->red_host) (void)hipHostFree(c->red_host);
    c->tm.destroy();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GM_OK;
}

extern "C" int gm_upload(gm_ctx *c, const double *B0, const double *xb, double t, double dt)
{
    if (!c || !B0 || !xb) return fail(GM_ERR_ARG, "gm_upload: null pointer");
    const size_t P3 = 3 * (size_t)c->P;
    double *stage = nullptr;
    int r;
    if ((r = dalloc(&stage, P3))) return r;
    auto s = c->stream;
    HIPCHK(hipMemcpyAsync(stage, xb, P3 * 8, hipMemcpyHostToDevice, s));
    if (c->prec == GM_PREC_F64) {
        for (int i = 0; i < 3; ++i) gmd::launch_aos_to_soa_f64(s, c->P, stage, c->X64[i]);
    } else {
        for (int i = 0; i < 2; ++i) gmd::launch_aos_to_soa_f32k(s, c->P, stage, c->Xh[i], c->Xl[i]);
    }
    HIPCHK(hipMemcpyAsync(stage, B0, P3 * 8, hipMemcpyHostToDevice, s));
    if (c->prec == GM_PREC_F64) {
        gmd::launch_aos_to_soa_f64(s, c->P, stage, c->B0_64);
        // B starts as a copy of B0 (src/main.cu:195); ghost edges keep it
        HIPCHK(hipMemcpyAsync(c->B64, c->B0_64, P3 * 8, hipMemcpyDeviceToDevice, s));
    } else {
        // B0 as triples for the stage kernels; B (component order) starts as
        // a copy of it
        gmd::launch_aos_to_soa_f32(s, c->P, stage, c->B0_32, true);
        gmd::launch_aos_to_soa_f32(s, c->P, stage, c->B32, false);
    }
    if (c->p.inertia) {
        // the velocity starts at rest (uInit = 'nil'); gm_upload_velocity
        // sets another one
        for (int i = 0; i < 2; ++i) gmd::launch_fill_f64(s, c->U64[i], (long long)P3, 0.0);
        c->ucur = 0;
    }
    HIPCHK(hipStreamSynchronize(s));
    dfree(stage);
    gmd::Ctl &h = *c->ctlh;
    memset(&h, 0, sizeof h);
    h.t = t;
    h.dt = dt;
    h.cur = 0;
    h.n_cycle = 1LL << 62;
    h.max_attempt = 1LL << 62;
    c->cur = 0;
    int r2 = push_ctl(c);
    if (r2) return r2;
    // periodic ghost layers as the reference's setGridPeriHost leaves them
    if (c->prec == GM_PREC_F32K && any_peri(c->p)) gmd::launch_set_grid_peri32(s, c->geo, c->Xh, c->Xl, c->ctl);
    // constants of the line-tied ghost faces (the ghost layer never moves)
    gmd::FusedArgs A = fused_args(c, 0, 0, 0, 0);
    gmd::launch_ghost_setup(c->stream, A, c->prec == GM_PREC_F32K, c->ghost);
    HIPCHK(hipStreamSynchronize(c->stream));
    return GM_OK;
}

// cudaMemcpy(d.uu, h.uu), src/main.cu:216-222: the velocity of inner
// points, inner AoS (the reference's h.uu layout)
extern "C" int gm_upload_velocity(gm_ctx *c, const double *uu)
{
    if (!c || !uu) return fail(GM_ERR_ARG, "gm_upload_velocity: null pointer");
    if (!c->p.inertia) return fail(GM_ERR_ARG, "gm_upload_velocity: the context was created with inertia = f");
    const size_t N3 = 3 * (size_t)c->N;
    double *stage = nullptr;
    int r = dalloc(&stage, N3);
    if (r) return r;
    HIPCHK(hipMemcpyAsync(stage, uu, N3 * 8, hipMemcpyHostToDevice, c->stream));
    gmd::launch_inner_aos_to_padded(c->stream, c->geo, stage, c->U64[c->ucur]);
    HIPCHK(hipStreamSynchronize(c->stream));
    dfree(stage);
    return GM_OK;
}

extern "C" int gm_get_time(gm_ctx *c, double *t, double *dt)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    int r = sync_ctl(c);
    if (r) return r;
    if (t) *t = c->ctlh->t;
    if (dt) *dt = c->ctlh->dt;
    return GM_OK;
}

extern "C" int gm_set_time(gm_ctx *c, double t, double dt)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    int r = sync_ctl(c);
    if (r) return r;
    c->ctlh->t = t;
    c->ctlh->dt = dt;
    return push_ctl(c);
}

extern "C" int gm_synchronize(gm_ctx *c)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return GM_OK;
}

static int extra_diag(gm_ctx *c, int mesh, int errs, int eps, double *cellVol, double *convex, double *wedge);

extern "C" int gm_download(gm_ctx *c, int field, double *host)
{
    if (!c || !host) return fail(GM_ERR_ARG, "gm_download: null pointer");
    int r = sync_ctl(c);
    if (r) return r;
    const size_t P3 = 3 * (size_t)c->P, N3 = 3 * (size_t)c->N;
    auto s = c->stream;
    bool f64 = c->prec == GM_PREC_F64;
    double *tmp = nullptr;
    size_t n = 0;
    switch (field) {
    case GM_FIELD_XB:
    case GM_FIELD_XB_NEW:
    case GM_FIELD_B0:
    case GM_FIELD_B:
    case GM_FIELD_XB_LO: {
        n = P3;
        if ((r = dalloc(&tmp, n))) return r;
        int idx = c->cur;
        if (field == GM_FIELD_XB_NEW) idx = f64 ? (c->cur + 1) % 3 : c->cur ^ 1;
        if (field == GM_FIELD_XB || field == GM_FIELD_XB_NEW) {
            if (f64) gmd::launch_soa_to_aos_f64(s, c->P, c->X64[idx], tmp);
            else gmd::launch_soa_to_aos_f32k(s, c->P, c->Xh[idx], c->Xl[idx], tmp);
        } else if (field == GM_FIELD_XB_LO) {
            if (f64) (void)hipMemsetAsync(tmp, 0, n * 8, s);
            else gmd::launch_soa_to_aos_f32(s, c->P, c->Xl[idx], tmp, true);
        } else if (field == GM_FIELD_B0) {
            if (f64) gmd::launch_soa_to_aos_f64(s, c->P, c->B0_64, tmp);
            else gmd::launch_soa_to_aos_f32(s, c->P, c->B0_32, tmp, true);
        } else {
            if (f64) gmd::launch_soa_to_aos_f64(s, c->P, c->B64, tmp);
            else gmd::launch_soa_to_aos_f32(s, c->P, c->B32, tmp, false);
        }
        break;
    }
    case GM_FIELD_DETJAC:
        HIPCHK(hipMemcpyAsync(host, c->det, c->P * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return GM_OK;
    case GM_FIELD_J:
    case GM_FIELD_GRADP: {
        n = N3;
        if ((r = dalloc(&tmp, n))) return r;
        if (f64) gmd::launch_padded_to_inner_aos<double>(s, c->geo, field == GM_FIELD_J ? c->J64 : c->G64, tmp);
        else gmd::launch_padded_to_inner_aos<float>(s, c->geo, field == GM_FIELD_J ? c->J32 : c->G32, tmp);
        break;
    }
    case GM_FIELD_CELLVOL:
    case GM_FIELD_CONVEXITY:
    case GM_FIELD_WEDGEMIN: {
        n = (size_t)c->N;
        if ((r = dalloc(&tmp, n))) return r;
        r = extra_diag(c, 0, 0, 0, field == GM_FIELD_CELLVOL ? tmp : nullptr, field == GM_FIELD_CONVEXITY ? tmp : nullptr,
                       field == GM_FIELD_WEDGEMIN ? tmp : nullptr);
        if (r) { dfree(tmp); return r; }
        break;
    }
    case GM_FIELD_UU: {
        if (!c->p.inertia) return fail(GM_ERR_ARG, "gm_download: no velocity without inertia");
        n = N3;
        if ((r = dalloc(&tmp, n))) return r;
        gmd::launch_padded_to_inner_aos<double>(s, c->geo, c->U64[c->ucur], tmp);
        break;
    }
    case GM_FIELD_KK: {
        if (c->p.inertia) {
            // [stage][point][x0 x1 x2 u0 u1 u2], the reference's nVar = 6 layout
            n = 12 * N3;
            if ((r = dalloc(&tmp, n))) return r;
            for (int m = 0; m < 6; ++m) {
                gmd::launch_padded_to_inner_aos<double>(s, c->geo, c->K64 + m * P3, tmp + m * N3);
                gmd::launch_padded_to_inner_aos<double>(s, c->geo, c->KU64 + m * P3, tmp + (6 + m) * N3);
            }
            std::vector<double> h(n);
            hipError_t e = hipMemcpyAsync(h.data(), tmp, n * 8, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            dfree(tmp);
            if (e != hipSuccess) return fail(GM_ERR_HIP, std::string("download: ") + hipGetErrorString(e));
            const size_t N = (size_t)c->N;
            for (int m = 0; m < 6; ++m)
                for (size_t q = 0; q < N; ++q)
                    for (int l = 0; l < 3; ++l) {
                        host[6 * N * m + 6 * q + l] = h[N3 * m + 3 * q + l];
                        host[6 * N * m + 6 * q + 3 + l] = h[N3 * (6 + m) + 3 * q + l];
                    }
            return GM_OK;
        }
        n = 6 * N3;
        if ((r = dalloc(&tmp, n))) return r;
        (void)hipMemsetAsync(tmp, 0, n * 8, s);
        for (int m = 0; m < (f64 ? 6 : 5); ++m) {
            if (f64) gmd::launch_padded_to_inner_aos<double>(s, c->geo, c->K64 + m * P3, tmp + m * N3);
            else gmd::launch_padded_to_inner_aos<float>(s, c->geo, c->K32 + m * P3, tmp + m * N3, true);
        }
        break;
    }
    default:
        return fail(GM_ERR_ARG, "unknown field");
    }
    hipError_t e = hipMemcpyAsync(host, tmp, n * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dfree(tmp);
    if (e != hipSuccess) return fail(GM_ERR_HIP, std::string("download: ") + hipGetErrorString(e));
    return GM_OK;
}

// ------------------------------------------------------- per-kernel path
extern "C" int gm_b_jacb0(gm_ctx *c)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    if (int r = need_f64(c, "gm_b_jacb0")) return r;
    gmd::launch_b_jacb0(c->stream, c->geo, c->X64[c->cur], c->B0_64, c->B64, c->det);
    return GM_OK;
}
extern "C" int gm_set_bbound(gm_ctx *c)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    if (int r = need_f64(c, "gm_set_bbound")) return r;
    gmd::launch_bbound(c->stream, c->geo, c->X64[c->cur], c->B0_64, c->B64, c->det);
    return GM_OK;
}
extern "C" int gm_set_peri(gm_ctx *c, int field)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    if (int r = need_f64(c, "gm_set_peri")) return r;
    if (field != GM_FIELD_B) return fail(GM_ERR_ARG, "gm_set_peri: field must be GM_FIELD_B");
    gmd::launch_set_peri(c->stream, c->geo, c->B64);
    return GM_OK;
}
extern "C" int gm_set_grid_peri(gm_ctx *c)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    if (int r = need_f64(c, "gm_set_grid_peri")) return r;
    gmd::launch_set_grid_peri(c->stream, c->geo, c->X64[c->cur]);
    return GM_OK;
}
extern "C" int gm_current(gm_ctx *c)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    if (int r = need_f64(c, "gm_current")) return r;
    if (gm::starts(c->p.jMethod, "Classic "))
        gmd::launch_jclassic(c->stream, c->geo, c->X64[c->cur], c->B0_64, c->det, c->J64);
    else if (gm::starts(c->p.jMethod, "Stokes4th "))
        gmd::launch_jstokes4th(c->stream, c->geo, c->X64[c->cur], c->B64, c->J64);
    else
        gmd::launch_jstokes(c->stream, c->geo, c->X64[c->cur], c->B64, c->J64);
    return GM_OK;
}
extern "C" int gm_gradp(gm_ctx *c)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    if (int r = need_f64(c, "gm_gradp")) return r;
    gmd::launch_gradp(c->stream, c->geo, c->X64[c->cur], c->det, c->G64);
    return GM_OK;
}
extern "C" int gm_rk_kk(gm_ctx *c, int n, double dt)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    if (int r = need_f64(c, "gm_rk_kk")) return r;
    if (n < 0 || n > 5) return fail(GM_ERR_ARG, "stage index must be 0..5");
    const size_t P3 = 3 * (size_t)c->P;
    if (!c->Xtmp64) {
        int r = dalloc(&c->Xtmp64, P3);
        if (r) return r;
    }
    // xb_tmp starts as xb so its ghost layers match (src/main.cu:202)
    HIPCHK(hipMemcpyAsync(c->Xtmp64, c->X64[c->cur], P3 * 8, hipMemcpyDeviceToDevice, c->stream));
    const bool in = c->p.inertia;
    gmd::launch_kk(c->stream, c->geo, c->tab_dev, n, dt, c->X64[c->cur], c->B64, c->J64, c->G64, c->K64, c->Xtmp64,
                   in ? c->det : nullptr, in ? c->U64[c->ucur] : nullptr, in ? c->KU64 : nullptr,
                   in ? c->Utmp64 : nullptr);
    if (any_peri(c->p)) gmd::launch_set_grid_peri(c->stream, c->geo, c->Xtmp64);
    HIPCHK(hipMemcpyAsync(c->X64[c->cur], c->Xtmp64, P3 * 8, hipMemcpyDeviceToDevice, c->stream));
    // uu <- uu_tmp (src/main.cu:296-305)
    if (in) HIPCHK(hipMemcpyAsync(c->U64[c->ucur], c->Utmp64, P3 * 8, hipMemcpyDeviceToDevice, c->stream));
    return GM_OK;
}
extern "C" int gm_rk_xnewstar(gm_ctx *c, double *maxDelta)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    if (int r = need_f64(c, "gm_rk_xnewstar")) return r;
    if (!c->xns_part) {   // one per-block maximum per 256 inner points, kept for later attempts
        int r = dalloc(&c->xns_part, (size_t)((c->N + 255) / 256));
        if (r) return r;
    }
    double *parts = c->xns_part;
    (void)hipMemsetAsync(c->nanflag, 0, sizeof(int), c->stream);
    double *xnew = c->X64[(c->cur + 1) % 3];
    // xb_new starts as xb (src/main.cu:209) so its ghost layers match
    HIPCHK(hipMemcpyAsync(xnew, c->X64[c->cur], 3 * (size_t)c->P * 8, hipMemcpyDeviceToDevice, c->stream));
    const bool in = c->p.inertia;
    int n = gmd::launch_xnewstar(c->stream, c->geo, c->tab_dev, c->X64[c->cur], c->K64, xnew, parts, c->nanflag,
                                 in ? c->U64[c->ucur] : nullptr, in ? c->KU64 : nullptr,
                                 in ? c->U64[c->ucur ^ 1] : nullptr);
    if (any_peri(c->p)) gmd::launch_set_grid_peri(c->stream, c->geo, xnew);
    gmd::launch_max_final(c->stream, parts, n, c->red_out);
    HIPCHK(hipMemcpyAsync(c->red_host, c->red_out, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (maxDelta) *maxDelta = c->red_host[0];
    return GM_OK;
}
// xNewStar + findDevMax like gm_rk_xnewstar, with every piece of scratch
// (the per-block maxima, the NaN flag and the grid maximum) in the library's
// static device array instead of memory the library allocates: no
// hipMalloc, hipFree or allocating call.  At most noalloc_scratch_doubles()
// - 8 block maxima (256 inner points each); calls that overlap on one
// device race for the array.
extern "C" int gm_rk_xnewstar_noalloc(gm_ctx *c, double *maxDelta)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    if (int r = need_f64(c, "gm_rk_xnewstar_noalloc")) return r;
    const long long nb = (c->N + 255) / 256;
    if (nb + 8 > gmd::noalloc_scratch_doubles())
        return fail(GM_ERR_ARG, "gm_rk_xnewstar_noalloc: the grid needs " + std::to_string(nb) +
                                    " block maxima; the static scratch holds " +
                                    std::to_string(gmd::noalloc_scratch_doubles() - 8) +
                                    " (at most 256 of them per inner point count)");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != c->device) HIPCHK(hipSetDevice(c->device));
    double *S = gmd::noalloc_scratch(c->device);
    if (!S) return fail(GM_ERR_HIP, "gm_rk_xnewstar_noalloc: hipGetSymbolAddress of the scratch failed");
    double *out = S, *parts = S + 8;
    int *nan = reinterpret_cast<int *>(S + 1);
    gmd::launch_fill_f64(c->stream, S, 2, 0.0);   // grid maximum and NaN flag start at zero
    double *xnew = c->X64[(c->cur + 1) % 3];
    // xb_new starts as xb (src/main.cu:209) so its ghost layers match
    HIPCHK(hipMemcpyAsync(xnew, c->X64[c->cur], 3 * (size_t)c->P * 8, hipMemcpyDeviceToDevice, c->stream));
    const bool in = c->p.inertia;
    int n = gmd::launch_xnewstar(c->stream, c->geo, c->tab_dev, c->X64[c->cur], c->K64, xnew, parts, nan,
                                 in ? c->U64[c->ucur] : nullptr, in ? c->KU64 : nullptr,
                                 in ? c->U64[c->ucur ^ 1] : nullptr);
    if (any_peri(c->p)) gmd::launch_set_grid_peri(c->stream, c->geo, xnew);
    gmd::launch_max_final(c->stream, parts, n, out);
    HIPCHK(hipMemcpyAsync(c->red_host, out, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (maxDelta) *maxDelta = c->red_host[0];
    return GM_OK;
}
extern "C" int gm_accept(gm_ctx *c)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    int r = sync_ctl(c);
    if (r) return r;
    c->ctlh->cur = c->prec == GM_PREC_F64 ? (c->ctlh->cur + 1) % 3 : c->ctlh->cur ^ 1;
    c->cur = c->ctlh->cur;
    if (c->p.inertia) c->ucur ^= 1;   // uu <- uu_new (src/main.cu:331-338)
    return push_ctl(c);
}

// Mesh / error diagnostics of the current state into red_out[8..12] and,
// when given, per-point device arrays (inner points).
static int extra_diag(gm_ctx *c, int mesh, int errs, int eps, double *cellVol, double *convex, double *wedge)
{
    const bool f64 = c->prec == GM_PREC_F64;
    const int k = c->cur;
    const int rc = f64 ? gmd::launch_extra_diag<double>(c->stream, c->geo, c->X64[k], nullptr, nullptr, c->B64, c->J64,
                                                        c->p.ar, c->p.az, c->p.phi1, mesh, errs, eps, cellVol, convex,
                                                        wedge, c->extra_part, c->red_out + 8)
                       : gmd::launch_extra_diag<float>(c->stream, c->geo, nullptr, c->Xh[k], c->Xl[k], c->B32, c->J32,
                                                       c->p.ar, c->p.az, c->p.phi1, mesh, errs, eps, cellVol, convex,
                                                       wedge, c->extra_part, c->red_out + 8);
    if (rc) return fail(GM_ERR_UNSUPPORTED, "epsilonStar: more than 3840 planes along z");
    HIPCHK(hipGetLastError());
    return GM_OK;
}

extern "C" int gm_reductions(gm_ctx *c, gm_reduction *red)
{
    if (!c || !red) return fail(GM_ERR_ARG, "gm_reductions: null pointer");
    if (c->prec == GM_PREC_F64)
        gmd::launch_diag<double>(c->stream, c->geo, c->B64, c->J64, c->det, c->diag_scratch, c->diag_blocks, c->red_out);
    else
        gmd::launch_diag<float>(c->stream, c->geo, c->B32, c->J32, c->det, c->diag_scratch, c->diag_blocks, c->red_out);
    const gm_params &p = c->p;
    const int mesh = p.redConvex || p.redWedgeMin, errs = p.redErrB_1ez || p.redErrXb_XbAn, eps = p.redEpsilonStar;
    if (mesh || errs || eps) {
        if (int r = extra_diag(c, mesh, errs, eps, nullptr, nullptr, nullptr)) return r;
    }
    const bool udiag = p.inertia && (p.redU2 || p.redUMax);
    if (udiag) gmd::launch_u_diag(c->stream, c->geo, c->U64[c->ucur], c->det, c->diag_scratch, c->red_out + 13);
    HIPCHK(hipMemcpyAsync(c->red_host, c->red_out, 16 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    memset(red, 0, sizeof *red);
    if (udiag) {
        red->U2 = c->red_host[13];
        red->UMax = c->red_host[14];
    }
    red->B2 = c->red_host[0];
    red->JMax = c->red_host[1];
    red->JxB_B2Max = c->red_host[2];
    red->JxB_JBMax = c->red_host[3];
    red->JxBMax_JBMax = c->red_host[5] > 0 ? c->red_host[4] / c->red_host[5] : 0.0;
    if (mesh) {
        red->convex = c->red_host[8];
        red->wedgeMin = c->red_host[9];
    }
    if (errs) {
        red->errB_1ez = c->red_host[10];
        red->errXb_XbAn = c->red_host[11];
    }
    if (eps) red->epsilonStar = c->red_host[12];
    red->B2res = gm_init_residual(&c->p);
    return GM_OK;
}

// ------------------------------------------------------------ fused path
static int check_fused(gm_ctx *c)
{
    if (!gm::starts(c->p.jMethod, "Stokes "))
        return fail(GM_ERR_UNSUPPORTED, "jMethod Classic / Stokes4th run through the per-kernel path (gm_current ...)");
    if (c->p.jxbAver) return fail(GM_ERR_UNSUPPORTED, "jxbAver runs through the per-kernel path (gm_rk_kk ...)");
    if (c->p.inertia) return fail(GM_ERR_UNSUPPORTED, "inertia runs through the per-kernel path (gm_rk_kk ...)");
    return GM_OK;
}

// FP64 contexts the fused stage kernels do not cover (jMethod = Classic or
// Stokes4th, jxbAver, inertia) step through the per-kernel sequence
static bool per_kernel_path(gm_ctx *c)
{
    return c->prec == GM_PREC_F64 && (!gm::starts(c->p.jMethod, "Stokes ") || c->p.jxbAver || c->p.inertia);
}

// Fields of the current state through the per-kernel sequence
// (B_JacB0, setBbound, setPeri(B), current, gradP): the path of FP64
// contexts with periodic faces or jMethod = Classic.
static int fields_per_kernel(gm_ctx *c)
{
    int r;
    if ((r = gm_b_jacb0(c)) || (r = gm_set_bbound(c)) || (r = gm_set_peri(c, GM_FIELD_B)) || (r = gm_current(c)))
        return r;
    if (c->p.pressure && (r = gm_gradp(c))) return r;
    return GM_OK;
}

extern "C" int gm_fields(gm_ctx *c)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    if (per_kernel_path(c)) return fields_per_kernel(c);
    if (int r = check_fused(c)) return r;
    gmd::FusedArgs A = fused_args(c, -1, 1, 0, 0);
    gmd::launch_fused(c->stream, A, c->prec == GM_PREC_F32K);
    HIPCHK(hipGetLastError());
    if (c->prec == GM_PREC_F64 && any_peri(c->p)) {
        // B and detJac of every ghost face as the reference's updateBbound
        // leaves them (its periodic branch evaluates a periodic ghost at the
        // image plane, src/boundaries.cu:112-162), then B copied across the
        // periodic faces (setPeri): the stage kernels need neither, the
        // fields of a dump do
        gmd::launch_bbound(c->stream, c->geo, c->X64[c->cur], c->B0_64, c->B64, c->det);
        gmd::launch_set_peri(c->stream, c->geo, c->B64);
        HIPCHK(hipGetLastError());
    }
    return GM_OK;
}

static void fill_info(gm_ctx *c, gm_step_info *info)
{
    const gmd::Ctl &h = *c->ctlh;
    info->accepted = h.accepted;
    info->nan_seen = h.nan_seen;
    info->stopped = h.done && !h.converged;
    info->converged = h.converged;
    info->dt_used = h.dt_used;
    info->maxDelta = h.maxDelta;
    info->t = h.t;
    info->dt = h.dt;
    info->n_attempt = h.n_attempt;
    info->n_accept = h.n_accept;
}

// One attempted step of an FP64 context on the per-kernel path (periodic
// faces, jMethod = Classic): the
// reference's per-kernel sequence (src/main.cu:259-383) on the GPU, with
// the device controller's rules applied on the host (host pow, as the
// reference).  The stage positions update x in place (src/main.cu:288), so
// a rejected attempt continues from that x, as the reference does; the
// criterion of the starting state is evaluated before the stages.
// Expects c->ctlh synchronised; pushes it back.
static int step_per_kernel(gm_ctx *c)
{
    gmd::Ctl &h = *c->ctlh;
    if (h.done) return GM_OK;
    int r;
    if ((r = fields_per_kernel(c))) return r;
    gmd::launch_diag<double>(c->stream, c->geo, c->B64, c->J64, c->det, c->diag_scratch, c->diag_blocks, c->red_out);
    HIPCHK(hipMemcpyAsync(c->red_host, c->red_out, 6 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double *d = c->red_host;
    const double crit = d[5] > 0 ? d[4] / d[5] : 0.0;
    gmd::StepRecord rec{};
    rec.t = h.t;
    rec.dt_used = h.dt;
    h.energy = d[0];
    h.jmax = d[1];
    h.jxb_b2 = d[2];
    h.ratio = crit;
    h.ratio_pt = d[3];
    h.diag_valid = 1;
    rec.ratio = crit;
    rec.ratio_pt = d[3];
    rec.energy = d[0];
    rec.jxb_b2 = d[2];
    rec.jmax = d[1];
    auto log = [&](void) -> int {
        HIPCHK(hipMemcpy(c->log + (h.n_attempt % c->log_cap), &rec, sizeof rec, hipMemcpyHostToDevice));
        return GM_OK;
    };
    if (h.stop_ratio > 0 && crit <= h.stop_ratio) {
        h.converged = 1;
        h.done = 1;
        h.stop_reason = 3;
        h.accepted = 0;
        rec.accepted = -1;
        if ((r = log())) return r;
        return push_ctl(c);
    }
    const double dt = h.dt;
    for (int n = 0; n < 6; ++n) {
        if (n > 0 && (r = fields_per_kernel(c))) return r;
        if ((r = gm_rk_kk(c, n, dt))) return r;
    }
    double md = 0;
    if ((r = gm_rk_xnewstar(c, &md))) return r;
    int nan = 0;
    HIPCHK(hipMemcpy(&nan, c->nanflag, sizeof nan, hipMemcpyDeviceToHost));
    h.nan_seen |= nan;
    rec.nan_seen = nan;
    rec.maxDelta = md;
    h.dt_used = dt;
    h.maxDelta = md;
    const int acc = md < c->p.maxError;
    h.accepted = acc;
    if (acc) {
        h.t += dt;
        h.cur = (h.cur + 1) % 3;
        if (c->p.inertia) c->ucur ^= 1;
        h.n_accept += 1;
    }
    if (md > 0) h.dt = dt * std::pow(0.9 * c->p.maxError / md, 0.2);
    rec.accepted = acc;
    rec.t = h.t;
    if ((r = log())) return r;
    h.n_attempt += 1;
    h.event = acc && (gmd::interval_crossed(h.t, h.dt, h.ev_dump) || gmd::interval_crossed(h.t, h.dt, h.ev_save));
    if (h.n_accept >= h.n_cycle) { h.done = 1; h.stop_reason = 1; }
    else if (!(h.dt > c->p.dtMin)) { h.done = 1; h.stop_reason = 2; }
    else if (h.n_attempt >= h.max_attempt) { h.done = 1; h.stop_reason = 4; }
    else if (h.event) { h.done = 1; h.stop_reason = 5; }
    c->cur = h.cur;
    return push_ctl(c);
}


extern "C" int gm_step(gm_ctx *c, gm_step_info *info)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    if (!per_kernel_path(c))
        if (int r = check_fused(c)) return r;
    int r = sync_ctl(c);
    if (r) return r;
    gmd::Ctl &h = *c->ctlh;
    h.done = 0;
    h.converged = 0;
    h.stop_reason = 0;
    h.nan_seen = 0;
    h.stop_ratio = 0;
    h.n_cycle = 1LL << 62;
    h.max_attempt = 1LL << 62;
    h.event = 0;
    h.ev_dump = h.ev_save = 0;
    if ((r = push_ctl(c))) return r;
    if (per_kernel_path(c)) {
        if ((r = step_per_kernel(c))) return r;
    } else {
        enqueue_step(c, 1, 1);
    }
    HIPCHK(hipGetLastError());
    if ((r = sync_ctl(c))) return r;
    if (info) fill_info(c, info);
    if (h.nan_seen) return fail(GM_ERR_NAN, "NaN in the error estimate (the reference's max would skip it)");
    return GM_OK;
}

extern "C" int gm_run(gm_ctx *c, const gm_run_opts *o, gm_run_stats *st)
{
    if (!c || !o) return fail(GM_ERR_ARG, "gm_run: null pointer");
    if (!per_kernel_path(c))
        if (int r = check_fused(c)) return r;
    int r = sync_ctl(c);
    if (r) return r;
    gmd::Ctl &h = *c->ctlh;
    h.done = 0;
    h.converged = 0;
    h.stop_reason = 0;
    h.nan_seen = 0;
    h.diag_valid = 0;
    h.stop_ratio = o->stop_ratio;
    h.event = 0;
    h.ev_dump = o->dt_dump;
    h.ev_save = o->dt_save;
    long long ncyc = o->n_cycle > 0 ? o->n_cycle : c->p.nCycle;
    h.n_cycle = h.n_accept + ncyc;
    h.max_attempt = h.n_attempt + (o->max_attempts > 0 ? o->max_attempts : (1LL << 40));
    if ((r = push_ctl(c))) return r;
    int batch = std::max(1, o->batch);
    bool graph = o->use_graph && !c->tm.on && !per_kernel_path(c);
    auto t0 = std::chrono::steady_clock::now();
    if (graph && (!c->gexec || c->gbatch != batch)) {
        if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
        hipGraph_t gr;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < batch; ++i) enqueue_step(c, 1, 1);
        HIPCHK(hipStreamEndCapture(c->stream, &gr));
        HIPCHK(hipGraphInstantiate(&c->gexec, gr, nullptr, nullptr, 0));
        (void)hipGraphDestroy(gr);
        c->gbatch = batch;
    }
    while (per_kernel_path(c)) {
        if ((r = step_per_kernel(c))) return r;
        if (h.done) break;
    }
    while (!per_kernel_path(c)) {
        if (graph)
            HIPCHK(hipGraphLaunch(c->gexec, c->stream));
        else
            for (int i = 0; i < batch; ++i) enqueue_step(c, 1, 1);
        HIPCHK(hipGetLastError());
        if ((r = sync_ctl(c))) return r;
        if (h.done) break;
    }
    auto t1 = std::chrono::steady_clock::now();
    if (st) {
        st->n_attempt = h.n_attempt;
        st->n_accept = h.n_accept;
        st->t = h.t;
        st->dt = h.dt;
        st->ratio = h.diag_valid ? h.ratio : -1;
        st->ratio_pt = h.diag_valid ? h.ratio_pt : -1;
        st->energy = h.diag_valid ? h.energy : -1;
        st->converged = h.converged;
        st->stopped = h.done && !h.converged;
        st->nan_seen = h.nan_seen;
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
        st->event = h.event;
    }
    if (h.nan_seen) return fail(GM_ERR_NAN, "NaN in the error estimate");
    return GM_OK;
}

extern "C" int gm_step_log(gm_ctx *c, gm_step_record *out, int n)
{
    if (!c || !out) return fail(GM_ERR_ARG, "gm_step_log: null pointer");
    int r = sync_ctl(c);
    if (r) return r;
    long long total = c->ctlh->n_attempt + (c->ctlh->converged ? 1 : 0);
    long long cnt = std::min<long long>({(long long)n, total, (long long)c->log_cap});
    std::vector<gmd::StepRecord> all(c->log_cap);
    HIPCHK(hipMemcpy(all.data(), c->log, c->log_cap * sizeof(gmd::StepRecord), hipMemcpyDeviceToHost));
    for (long long i = 0; i < cnt; ++i) {
        long long idx = (total - cnt + i) % c->log_cap;
        const gmd::StepRecord &s = all[idx];
        out[i].t = s.t; out[i].dt_used = s.dt_used; out[i].maxDelta = s.maxDelta; out[i].ratio = s.ratio; out[i].ratio_pt = s.ratio_pt;
        out[i].energy = s.energy; out[i].jxb_b2 = s.jxb_b2; out[i].jmax = s.jmax;
        out[i].accepted = s.accepted; out[i].nan_seen = s.nan_seen;
    }
    return (int)cnt;
}

// ---------------------------------------------------------------- timing
extern "C" int gm_timing_enable(gm_ctx *c, int on)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    c->tm.collect();
    c->tm.on = on != 0;
    c->tm.reset();
    return GM_OK;
}
extern "C" int gm_timing_read(gm_ctx *c, double ms[GM_TIMING_KINDS], int64_t count[GM_TIMING_KINDS])
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    c->tm.collect();
    for (int i = 0; i < GM_TIMING_KINDS; ++i) {
        if (ms) ms[i] = c->tm.ms[i];
        if (count) count[i] = c->tm.count[i];
    }
    return GM_OK;
}
extern "C" double gm_algorithmic_bytes_per_point_stage(int precision)
{
    // per step: F64 792 B (stages 0-4: 72+24n, final 192), F32K 480 B
    // (stages 0-4: 48+12n, final 120); divided over six stages
    return precision == GM_PREC_F32K ? 80.0 : 132.0;
}
extern "C" int gm_set_tile(gm_ctx *c, int tx, int ty, int cz)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    if (tx <= 0 || ty <= 0) {
        auto_tiles(c, cz > 0 ? cz : default_blocks(c), false);
        return ensure_partials(c);
    }
    if ((tx + 4) * (ty + 4) > gmd::fused_max_points(c->prec == GM_PREC_F32K))
        return fail(GM_ERR_ARG, "tile rect (tx+4)(ty+4) exceeds the compiled capacity");
    for (int k = 0; k < 7; ++k) set_tile_k(c, k, tx, ty, cz > 0 ? cz : c->p.nz);
    return ensure_partials(c);
}
extern "C" int gm_get_tile(gm_ctx *c, int *tx, int *ty, int *cz, int *nblocks)
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    if (tx) *tx = c->TX[1];
    if (ty) *ty = c->TY[1];
    if (cz) *cz = c->CZ[1];           // stage 1 (the common configuration)
    if (nblocks) *nblocks = c->nblocks[1];
    return GM_OK;
}

// HBM yardsticks (no reference counterpart): a device copy of `bytes`
// (mode 0), a read of `bytes` (1) or a write of `bytes` (2), `reps` times;
// the mean kernel time in ms.  They calibrate the HBM byte counters and
// give the achievable bandwidth the stage kernels are compared with.
extern "C" int gm_hbm_probe(int64_t bytes, int mode, int reps, double *ms_mean)
{
    if (bytes < 16 || reps < 1 || mode < 0 || mode > 2) return fail(GM_ERR_ARG, "gm_hbm_probe: bad size or mode");
    if (bytes / 16 / 2048 >= (1LL << 31)) return fail(GM_ERR_ARG, "gm_hbm_probe: size too large");
    void *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, (size_t)bytes));
    if (hipMalloc(&b, (size_t)bytes) != hipSuccess) {
        (void)hipFree(a);
        return fail(GM_ERR_HIP, "gm_hbm_probe: out of device memory");
    }
    (void)hipMemset(a, 0, (size_t)bytes);
    (void)hipMemset(b, 0, (size_t)bytes);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    gmd::launch_hbm_probe(nullptr, mode, a, b, bytes);
    (void)hipEventRecord(e0, nullptr);
    for (int i = 0; i < reps; ++i) gmd::launch_hbm_probe(nullptr, mode, a, b, bytes);
    (void)hipEventRecord(e1, nullptr);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(GM_ERR_HIP, hipGetErrorString(e));
    if (ms_mean) *ms_mean = ms / reps;
    return GM_OK;
}
extern "C" int gm_copy_probe(int64_t bytes, int reps, double *ms_mean)
{
    return gm_hbm_probe(bytes, 0, reps, ms_mean);
}

// writeCudaInfo (src/io.cu:10-82): device properties, same key/value layout,
// file name kept as the reference's run helpers expect it.
extern "C" int gm_write_device_info(const char *dir, int active, int precision_bits)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) return fail(GM_ERR_HIP, "hipGetDeviceCount");
    std::string d = dir && *dir ? dir : ".";
    if (d.back() != '/') d += '/';
    std::string fn = d + "cuda.info";
    FILE *fd = fopen(fn.c_str(), "w");
    if (!fd) return fail(GM_ERR_IO, "could not open file '" + fn + "'");
    fprintf(fd, "number of CUDA devices: %i\n\n", ndev);
    fprintf(fd, "active device: %li\n\n", (long)active);
    fprintf(fd, "PRECISION = %i\n\n", precision_bits);
    for (int dev = 0; dev < ndev; ++dev) {
        hipDeviceProp_t pr;
        if (hipGetDeviceProperties(&pr, dev) != hipSuccess) continue;
        fprintf(fd, "device number                   %i\n", dev);
        fprintf(fd, "name                            %s\n", pr.name);
        fprintf(fd, "gcn arch                        %s\n", pr.gcnArchName);
        fprintf(fd, "global memory [bytes]           %zu\n", pr.totalGlobalMem);
        fprintf(fd, "shared memory per block [bytes] %zu\n", pr.sharedMemPerBlock);
        fprintf(fd, "registers per block             %d\n", pr.regsPerBlock);
        fprintf(fd, "threads per warp                %d\n", pr.warpSize);
        fprintf(fd, "max threads per block           %d\n", pr.maxThreadsPerBlock);
        fprintf(fd, "clock rate [kHz]                %d\n", pr.clockRate);
        fprintf(fd, "multiprocessors                 %d\n", pr.multiProcessorCount);
        fprintf(fd, "memory clock rate [kHz]         %d\n", pr.memoryClockRate);
        fprintf(fd, "memory bus width [bits]         %d\n", pr.memoryBusWidth);
        fprintf(fd, "l2 cache size [bytes]           %d\n", pr.l2CacheSize);
        fprintf(fd, "maximum threads per multiproc.  %d\n", pr.maxThreadsPerMultiProcessor);
        fprintf(fd, "\n");
    }
    fclose(fd);
    return GM_OK;
}



// ================================================================ autotune
namespace {

// The rect in (lo, hi] points with the fewest loaded points (lines = false)
// or 128-byte lines (lines = true) per output point.
bool best_rect(const gm_ctx *c, int lo, int hi, bool lines, int &bx, int &by)
{
    const int nx = c->p.nx, ny = c->p.ny;
    double best = 1e30;
    for (int ty = 2; ty <= 96; ++ty)
        for (int txm = 4; (txm + 4) * (ty + 4) <= hi; ++txm) {
            int tlx = (nx + txm - 1) / txm, tly = (ny + ty - 1) / ty;
            int tx = (nx + tlx - 1) / tlx, tyy = (ny + tly - 1) / tly;
            int rect = (tx + 4) * (tyy + 4);
            if (rect > hi || rect <= lo) continue;
            double cost = (lines ? rect_lines(tlx, tly, tx + 4, tyy + 4) : (double)tlx * tly * rect) / ((double)nx * ny);
            if (cost < best - 1e-12) { best = cost; bx = tx; by = tyy; }
        }
    return best < 1e30;
}

// A later candidate replaces the one kept so far only when it is more than
// 2% faster: launches of one tile vary by about 1% from run to run, and
// near-equal tiles (the 64-member ensemble's two stage-0 policies differ by
// about that) would otherwise be picked by noise.
constexpr double kTuneKeep = 0.98;

}  // namespace

// Per stage kernel: time `reps` launches of each candidate tile on the
// current state and keep the fastest (by the keep margin).  The stage
// kernels alone write only the slope buffers and the candidate position
// buffer, so the accepted state, t and dt are unchanged.
extern "C" int gm_autotune(gm_ctx *c, int reps, double ms_best[6])
{
    if (!c) return fail(GM_ERR_ARG, "null context");
    if (int r = check_fused(c)) return r;
    if (c->tm.on) return fail(GM_ERR_ARG, "gm_autotune: disable kernel timing first");
    reps = std::max(1, reps);
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    const int f32k = c->prec == GM_PREC_F32K;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const int bounds[6] = {0, 256, 384, 512, 768, 1024};
    int rc = GM_OK;
    for (int k = 0; k < 7 && rc == GM_OK; ++k) {
        const int stage = k == 6 ? -1 : k;
        const int cap = stage_max_rect(c, k);
        int bTX = c->TX[k], bTY = c->TY[k], bCZ = c->CZ[k];
        double best = 1e30;
        // per workgroup size: the fewest-points rect and, for F32K, the
        // fewest-lines one
        std::vector<std::pair<int, int>> rects;
        for (int w = 0; w < 5; ++w) {
            if (bounds[w] >= cap) break;
            for (int m = 0; m < (f32k ? 2 : 1); ++m) {
                int tx, ty;
                if (!best_rect(c, bounds[w], std::min(bounds[w + 1], cap), m == 1, tx, ty)) continue;
                if (std::find(rects.begin(), rects.end(), std::make_pair(tx, ty)) == rects.end()) rects.emplace_back(tx, ty);
            }
        }
        for (const auto &rt : rects) {
            if (rc != GM_OK) break;
            const int tx = rt.first, ty = rt.second;
            // z chunks: the plane-step model's choice, and one or two full
            // rounds over the CUs
            set_tile_k(c, k, tx, ty, c->p.nz);
            const int cols = c->tiles_x[k] * c->tiles_y[k], slots = c->ncu * stage_blocks_per_cu(c, k);
            int cand[3] = {best_chunks(c->p.nz, cols, slots), slots / cols, 2 * slots / cols};
            for (int q = 0; q < 3; ++q) {
                const int nch = cand[q] = std::max(1, std::min(cand[q], std::max(1, c->p.nz / 8)));
                bool timed = false;
                for (int r = 0; r < q; ++r) timed |= cand[r] == nch;
                if (timed) continue;
                set_tile_k(c, k, tx, ty, (c->p.nz + nch - 1) / nch);
                if ((rc = ensure_partials(c))) break;
                gmd::FusedArgs A = fused_args(c, stage, 0, stage == 0 ? 1 : 0, 0);
                gmd::launch_fused(c->stream, A, f32k);   // warm (and first-launch setup)
                if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = fail(GM_ERR_HIP, "event"); break; }
                for (int i = 0; i < reps; ++i) gmd::launch_fused(c->stream, A, f32k);
                if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                    rc = fail(GM_ERR_HIP, "event");
                    break;
                }
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (ms / reps < best * kTuneKeep) { best = ms / reps; bTX = tx; bTY = ty; bCZ = c->CZ[k]; }
            }
        }
        set_tile_k(c, k, bTX, bTY, bCZ);
        if (ms_best && k < 6) ms_best[k] = best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    return ensure_partials(c);
}

extern "C" int gm_set_stage_tile(gm_ctx *c, int stage, int tx, int ty, int cz)
{
    if (!c || stage < -1 || stage > 5 || tx <= 0 || ty <= 0 || cz <= 0)
        return fail(GM_ERR_ARG, "gm_set_stage_tile: bad arguments");
    const int k = stage < 0 ? 6 : stage;
    if ((std::min(tx, c->p.nx) + 4) * (std::min(ty, c->p.ny) + 4) > stage_max_rect(c, k))
        return fail(GM_ERR_ARG, "gm_set_stage_tile: rect (tx+4)(ty+4) exceeds the stage kernel's workgroup");
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    set_tile_k(c, k, tx, ty, cz);
    return ensure_partials(c);
}

extern "C" int gm_get_stage_tile(gm_ctx *c, int stage, int *tx, int *ty, int *cz, int *nblocks)
{
    if (!c || stage < -1 || stage > 5) return fail(GM_ERR_ARG, "gm_get_stage_tile: bad arguments");
    const int k = stage < 0 ? 6 : stage;
    if (tx) *tx = c->TX[k];
    if (ty) *ty = c->TY[k];
    if (cz) *cz = c->CZ[k];
    if (nblocks) *nblocks = c->nblocks[k];
    return GM_OK;
}

// ================================================================ ensemble
// Many independent relaxations advanced together: every stage is one launch
// over the workgroups of all members (each member keeps its own tiles,
// buffers, dt and control block), and one controller launch applies each
// member's accept/reject, dt and stop rules.  Members that are done return
// at once from every later kernel of the batch (their convergence mask).
struct gm_batch {
    std::vector<gm_ctx *> inst;
    hipStream_t stream = nullptr;
    int press = 0;
    bool f64 = false;               // FP64 members (the FP64 stage kernel), else F32K
    gmd::FusedArgs *args[6] = {};   // per stage: one argument block per member
    int *start[6] = {};             // per stage: first workgroup of each member, then the total
    int total[6] = {};
    int nt[6] = {};                 // per stage: workgroup size of the launch
    gmd::CtlArgs *cargs = nullptr;
    std::vector<gmd::Ctl *> ctlh;   // pinned host mirrors (the members' own)
    hipGraphExec_t gexec = nullptr;
    int gbatch = 0;
    EvTimer tm;                     // kernel timing (gm_batch_timing_*)
};

namespace {

// Workgroup size of a batch stage launch (k = 6: the fields pass, which
// members run on their own): the largest any member's rect needs (the FP64
// stage kernel always runs 512-thread workgroups)
int batch_nt(const gm_batch *b, int k)
{
    if (b->f64) return gmd::fused_max_points(0);
    int nt = 512;
    for (const gm_ctx *c : b->inst) nt = std::max(nt, gmd::batch_threads32((c->TX[k] + 4) * (c->TY[k] + 4)));
    return nt;
}

// Largest member rect a batch stage launch holds.
int batch_max_rect(const gm_batch *b, int k)
{
    return b->f64 ? gmd::fused_max_points(0) : gmd::max_rect32(k, b->press);
}

// Resident workgroups per CU of a batch stage launch in `nt`-thread workgroups.
int batch_blocks_per_cu(const gm_batch *b, int nt, int k)
{
    return b->f64 ? 1 : gmd::blocks_per_cu32(nt, k == 6 ? -1 : k, b->press);
}

void batch_launch(const gm_batch *b, int s, int nt, const gmd::FusedArgs *args, const int *start, int n, int total)
{
    if (b->f64) gmd::launch_fused64_batch(b->stream, s, b->press, args, start, n, total);
    else gmd::launch_fused32_batch(b->stream, s, b->press, nt, args, start, n, total);
}

void batch_free_dev(gm_batch *b)
{
    for (int s = 0; s < 6; ++s) { dfree(b->args[s]); dfree(b->start[s]); }
    dfree(b->cargs);
}

// Launch arguments from the members' current tiles and buffers; each
// stage launches in workgroups as large as its largest member rect needs.
int batch_prepare(gm_batch *b)
{
    const int n = (int)b->inst.size();
    batch_free_dev(b);
    for (int s = 0; s < 6; ++s) {
        b->nt[s] = batch_nt(b, s);
        bool fits = b->nt[s] <= batch_max_rect(b, s);
        for (const gm_ctx *c : b->inst) fits = fits && (c->TX[s] + 4) * (c->TY[s] + 4) <= b->nt[s];
        if (!fits) return fail(GM_ERR_ARG, "ensemble member tile exceeds the stage kernel's workgroup");
    }
    std::vector<gmd::FusedArgs> A(n);
    std::vector<int> st(n + 1);
    std::vector<gmd::CtlArgs> C(n);
    int r;
    for (int s = 0; s < 6; ++s) {
        st[0] = 0;
        for (int i = 0; i < n; ++i) {
            A[i] = fused_args(b->inst[i], s, 0, s == 0 ? 1 : 0, 1);
            st[i + 1] = st[i] + A[i].nblocks;
        }
        b->total[s] = st[n];
        if ((r = dalloc(&b->args[s], (size_t)n)) || (r = dalloc(&b->start[s], (size_t)n + 1))) return r;
        HIPCHK(hipMemcpy(b->args[s], A.data(), n * sizeof(gmd::FusedArgs), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b->start[s], st.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice));
    }
    for (int i = 0; i < n; ++i) {
        gm_ctx *c = b->inst[i];
        gmd::CtlArgs &K = C[i];
        K = gmd::CtlArgs{};
        K.ctl = c->ctl;
        K.part_delta = c->part_delta;
        K.part_diag = c->part_diag;
        K.nb_delta = c->nblocks[5];
        K.nb_diag = c->nblocks[0];
        K.diag = 1;
        K.nbuf = b->f64 ? 3 : 2;
        K.maxError = c->p.maxError;
        K.dtMin = c->p.dtMin;
        K.cell = c->p.dx * c->p.dy * c->p.dz;
        K.log = c->log;
        K.log_cap = c->log_cap;
    }
    if ((r = dalloc(&b->cargs, (size_t)n))) return r;
    HIPCHK(hipMemcpy(b->cargs, C.data(), n * sizeof(gmd::CtlArgs), hipMemcpyHostToDevice));
    return GM_OK;
}

void batch_enqueue_step(gm_batch *b)
{
    const int n = (int)b->inst.size();
    for (int s = 0; s < 6; ++s) {
        b->tm.begin(b->stream, s);
        batch_launch(b, s, b->nt[s], b->args[s], b->start[s], n, b->total[s]);
        b->tm.end(b->stream);
    }
    b->tm.begin(b->stream, 6);
    gmd::launch_controller_batch(b->stream, b->cargs, n);
    b->tm.end(b->stream);
}

}  // namespace

// Ensemble z chunks: one launch runs every member's workgroups, so it lasts
// as long as the longest march.  Per stage kernel the members share one
// chunk length L (a member shorter than L takes one chunk): the fewest
// ceil(workgroups / resident slots) x (L + 5) plane steps.
static void batch_chunks_k(gm_batch *b, int k)
{
    gm_ctx *c0 = b->inst[0];
    {
        const long long slots = (long long)c0->ncu * batch_blocks_per_cu(b, batch_nt(b, k), k);
        int maxnz = 1;
        for (gm_ctx *c : b->inst) maxnz = std::max(maxnz, c->p.nz);
        int bestL = maxnz;
        long long best = -1;
        for (int L = std::min(8, maxnz); L <= maxnz; ++L) {
            long long wg = 0;
            for (gm_ctx *c : b->inst)
                wg += (long long)c->tiles_x[k] * c->tiles_y[k] * ((c->p.nz + L - 1) / L);
            const long long cost = ((wg + slots - 1) / slots) * (L + 5);
            if (best < 0 || cost < best) { best = cost; bestL = L; }
        }
        for (gm_ctx *c : b->inst) set_tile_k(c, k, c->TX[k], c->TY[k], bestL);
    }
}
static void batch_chunks(gm_batch *b)
{
    for (int k = 0; k < 7; ++k) batch_chunks_k(b, k);
}

extern "C" int gm_batch_create(const gm_params *params, int n, int precision, gm_batch **out)
{
    if (!params || !out || n < 1) return fail(GM_ERR_ARG, "gm_batch_create: bad arguments");
    if (precision != GM_PREC_F32K && precision != GM_PREC_F64) return fail(GM_ERR_ARG, "unknown precision");
    for (int i = 0; i < n; ++i) {
        if (params[i].pressure != params[0].pressure)
            return fail(GM_ERR_ARG, "ensemble members must agree on pressure (one kernel variant per launch)");
        if (any_peri(params[i])) return fail(GM_ERR_UNSUPPORTED, "ensemble members: periodic boundaries run the per-kernel path");
    }
    gm_batch *b = new gm_batch();
    auto bail = [&](int code) {
        gm_batch_destroy(b);
        return code;
    };
    int r;
    for (int i = 0; i < n; ++i) {
        gm_ctx *c = nullptr;
        if ((r = gm_create(&params[i], precision, &c))) return bail(r);
        b->inst.push_back(c);
        if (per_kernel_path(c))
            return bail(fail(GM_ERR_UNSUPPORTED, "ensemble members: jMethod Classic / Stokes4th, jxbAver and inertia "
                                                 "run the per-kernel path"));
        c->zrev = false;   // the members together move several GB per stage
        c->zalt = false;
        auto_tiles(c, std::max(1, (c->ncu + n - 1) / n), true);
        b->ctlh.push_back(c->ctlh);
    }
    b->press = params[0].pressure ? 1 : 0;
    b->f64 = precision == GM_PREC_F64;
    batch_chunks(b);
    for (gm_ctx *c : b->inst)
        if ((r = ensure_partials(c))) return bail(r);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bail(fail(GM_ERR_HIP, "stream"));
    *out = b;
    return GM_OK;
}

extern "C" int gm_batch_destroy(gm_batch *b)
{
    if (!b) return GM_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    batch_free_dev(b);
    b->tm.destroy();
    for (gm_ctx *c : b->inst) gm_destroy(c);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return GM_OK;
}

extern "C" int gm_batch_context(gm_batch *b, int i, gm_ctx **ctx)
{
    if (!b || !ctx || i < 0 || i >= (int)b->inst.size()) return fail(GM_ERR_ARG, "gm_batch_context: bad member index");
    *ctx = b->inst[i];
    return GM_OK;
}

extern "C" int gm_batch_run(gm_batch *b, const gm_run_opts *o, gm_run_stats *st)
{
    if (!b || !o) return fail(GM_ERR_ARG, "gm_batch_run: null pointer");
    const int n = (int)b->inst.size();
    int r;
    for (gm_ctx *c : b->inst) {
        HIPCHK(hipStreamSynchronize(c->stream));
        if ((r = sync_ctl(c))) return r;
        gmd::Ctl &h = *c->ctlh;
        h.done = 0;
        h.converged = 0;
        h.stop_reason = 0;
        h.nan_seen = 0;
        h.diag_valid = 0;
        h.stop_ratio = o->stop_ratio;
        h.event = 0;
        h.ev_dump = o->dt_dump;
        h.ev_save = o->dt_save;
        long long ncyc = o->n_cycle > 0 ? o->n_cycle : c->p.nCycle;
        h.n_cycle = h.n_accept + ncyc;
        h.max_attempt = h.n_attempt + (o->max_attempts > 0 ? o->max_attempts : (1LL << 40));
        if ((r = push_ctl(c))) return r;
    }
    if ((r = batch_prepare(b))) return r;
    const int batch = std::max(1, o->batch);
    auto t0 = std::chrono::steady_clock::now();
    if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (o->use_graph && !b->tm.on) {
        hipGraph_t gr;
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < batch; ++i) batch_enqueue_step(b);
        HIPCHK(hipStreamEndCapture(b->stream, &gr));
        HIPCHK(hipGraphInstantiate(&b->gexec, gr, nullptr, nullptr, 0));
        (void)hipGraphDestroy(gr);
        b->gbatch = batch;
    }
    while (true) {
        if (b->gexec)
            HIPCHK(hipGraphLaunch(b->gexec, b->stream));
        else
            for (int i = 0; i < batch; ++i) batch_enqueue_step(b);
        HIPCHK(hipGetLastError());
        for (int i = 0; i < n; ++i)
            HIPCHK(hipMemcpyAsync(b->ctlh[i], b->inst[i]->ctl, sizeof(gmd::Ctl), hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        b->tm.collect();
        bool all = true;
        for (int i = 0; i < n; ++i) all = all && b->ctlh[i]->done;
        if (all) break;
    }
    auto t1 = std::chrono::steady_clock::now();
    bool nan = false;
    for (int i = 0; i < n; ++i) {
        gm_ctx *c = b->inst[i];
        const gmd::Ctl &h = *c->ctlh;
        c->cur = h.cur;
        nan = nan || h.nan_seen;
        if (st) {
            gm_run_stats &s = st[i];
            s.n_attempt = h.n_attempt;
            s.n_accept = h.n_accept;
            s.t = h.t;
            s.dt = h.dt;
            s.ratio = h.diag_valid ? h.ratio : -1;
            s.ratio_pt = h.diag_valid ? h.ratio_pt : -1;
            s.energy = h.diag_valid ? h.energy : -1;
            s.converged = h.converged;
            s.stopped = h.done && !h.converged;
            s.nan_seen = h.nan_seen;
            s.seconds = std::chrono::duration<double>(t1 - t0).count();
            s.event = h.event;
        }
    }
    if (nan) return fail(GM_ERR_NAN, "NaN in the error estimate of an ensemble member");
    return GM_OK;
}

// Ensemble tile autotuner (no reference counterpart).  Per stage kernel the
// candidates are tile policies for all members at once: each member takes
// its fewest-lines (or fewest-points) rect within a workgroup bound of 512,
// 768 or 1024 points (as far as the stage kernel is built for it), the
// members then share the plane-step model's z-chunk length, and `reps`
// launches of the whole ensemble's stage are timed on the current states.
// The fastest policy is kept, by the same 2% keep margin as gm_autotune.
// Stage launches alone write only slopes and candidate positions, so the
// members' accepted states, t and dt stay put.
extern "C" int gm_batch_autotune(gm_batch *b, int reps, double ms_best[6])
{
    if (!b) return fail(GM_ERR_ARG, "null ensemble");
    reps = std::max(1, reps);
    const int n = (int)b->inst.size();
    int r;
    for (gm_ctx *c : b->inst) HIPCHK(hipStreamSynchronize(c->stream));
    if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    gmd::FusedArgs *dargs = nullptr;
    int *dstart = nullptr;
    if ((r = dalloc(&dargs, (size_t)n)) || (r = dalloc(&dstart, (size_t)n + 1))) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return r;
    }
    std::vector<gmd::FusedArgs> A(n);
    std::vector<int> st(n + 1);
    std::vector<int> bTX(n), bTY(n), bCZ(n);
    const int bounds[3] = {512, 768, 1024};
    int rc = GM_OK;
    for (int s = 0; s < 6 && rc == GM_OK; ++s) {
        double best = 1e30;
        for (int i = 0; i < n; ++i) { bTX[i] = b->inst[i]->TX[s]; bTY[i] = b->inst[i]->TY[s]; bCZ[i] = b->inst[i]->CZ[s]; }
        for (int w = 0; w < 3 && rc == GM_OK; ++w) {
            if (bounds[w] > batch_max_rect(b, s)) break;
            for (int m = 0; m < 2 && rc == GM_OK; ++m) {
                for (gm_ctx *c : b->inst) {
                    int tx = c->TX[s], ty = c->TY[s];
                    if (!best_rect(c, 0, bounds[w], m == 0, tx, ty)) { tx = c->TX[s]; ty = c->TY[s]; }
                    set_tile_k(c, s, tx, ty, c->p.nz);
                }
                batch_chunks_k(b, s);
                st[0] = 0;
                for (int i = 0; i < n; ++i) {
                    if ((rc = ensure_partials(b->inst[i]))) break;
                    A[i] = fused_args(b->inst[i], s, 0, s == 0 ? 1 : 0, 0);
                    st[i + 1] = st[i] + A[i].nblocks;
                }
                if (rc) break;
                if (hipMemcpy(dargs, A.data(), n * sizeof(gmd::FusedArgs), hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemcpy(dstart, st.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
                    rc = fail(GM_ERR_HIP, "gm_batch_autotune: argument upload");
                    break;
                }
                const int nt = batch_nt(b, s);
                batch_launch(b, s, nt, dargs, dstart, n, st[n]);   // warm
                if (hipEventRecord(e0, b->stream) != hipSuccess) { rc = fail(GM_ERR_HIP, "event"); break; }
                for (int q = 0; q < reps; ++q) batch_launch(b, s, nt, dargs, dstart, n, st[n]);
                if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                    rc = fail(GM_ERR_HIP, "gm_batch_autotune: launch");
                    break;
                }
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (ms / reps < best * kTuneKeep) {
                    best = ms / reps;
                    for (int i = 0; i < n; ++i) { bTX[i] = b->inst[i]->TX[s]; bTY[i] = b->inst[i]->TY[s]; bCZ[i] = b->inst[i]->CZ[s]; }
                }
            }
        }
        for (int i = 0; i < n; ++i) set_tile_k(b->inst[i], s, bTX[i], bTY[i], bCZ[i]);
        if (ms_best) ms_best[s] = best;
    }
    dfree(dargs);
    dfree(dstart);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    for (gm_ctx *c : b->inst)
        if ((r = ensure_partials(c))) return r;
    return GM_OK;
}

// Event timing of the ensemble's launches (no reference counterpart): like
// gm_timing_enable / gm_timing_read, per kind 0..5 the stage launches over
// all members and 6 the controller launch.  Graph capture is off while on.
extern "C" int gm_batch_timing_enable(gm_batch *b, int on)
{
    if (!b) return fail(GM_ERR_ARG, "null ensemble");
    HIPCHK(hipStreamSynchronize(b->stream));
    b->tm.collect();
    b->tm.on = on != 0;
    b->tm.reset();
    return GM_OK;
}
extern "C" int gm_batch_timing_read(gm_batch *b, double ms[GM_TIMING_KINDS], int64_t count[GM_TIMING_KINDS])
{
    if (!b) return fail(GM_ERR_ARG, "null ensemble");
    HIPCHK(hipStreamSynchronize(b->stream));
    b->tm.collect();
    for (int i = 0; i < GM_TIMING_KINDS; ++i) {
        if (ms) ms[i] = b->tm.ms[i];
        if (count) count[i] = b->tm.count[i];
    }
    return GM_OK;
}
