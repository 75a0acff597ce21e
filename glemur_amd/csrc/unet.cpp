// Learned warm start: a 3-level U-Net that maps B0 of a braid (sampled on a
// coarse grid) to the displacement of its relaxed state, trained on
// relaxations held in HBM, and applied to a context by trilinear
// up-sampling, with a fallback to the cold state when the residual grows.
// No reference counterpart.
//
// Network (level 1 = the coarse grid n, level 2 = (n+1)/2, level 3 =
// (n2+1)/2; every convolution 3x3x3 with zero padding 1, ReLU except the
// head):
//   a1 = c(8->32, x0)  a2 = c(32->32, a1)
//   a3 = c(32->64, pool a2)  a4 = c(64->64, a3)
//   a5 = c(64->128, pool a4)  a6 = c(128->128, a5)
//   a7 = c(192->64, [up a6, a4])  a8 = c(64->64, a7)
//   a9 = c(96->32, [up a8, a2])  a10 = c(32->32, a9)  h = c(32->3, a10)
// Prediction y = scale * h.  Skip tensors are written straight into their
// concatenation buffers, so the concatenation costs nothing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gm_ctx.hpp"
#include "gm_internal.hpp"
#include "unet_kernels.hpp"

using gm::fail;
using gmu::bf16_t;

#define HIPCHK(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(GM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

namespace {

constexpr int NL = 11;
constexpr int LCIN[NL] = {8, 32, 32, 64, 64, 128, 192, 64, 96, 32, 32};
constexpr int LCOUT[NL] = {32, 32, 64, 64, 128, 128, 64, 64, 32, 32, 3};
constexpr int LEVEL[NL] = {1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 1};
constexpr int WCHUNK = 2048;   // points per weight-gradient chunk

int pad32(int k) { return (k + 31) / 32 * 32; }

struct ULayer {
    int cin = 0, cout = 0, coutp = 0, kpad = 0, kpadT = 0;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;   // [coutp][kpad]
    float *b = nullptr, *bm = nullptr, *bv = nullptr, *bg = nullptr; // [coutp]
    bf16_t *w16 = nullptr;   // [coutp][kpad]
    bf16_t *wT = nullptr;    // [cin][kpadT]: flipped transpose (input gradient)
};

}  // namespace

struct gm_unet {
    int n[4][3] = {};            // grid of levels 1..3
    long long P[4] = {};
    ULayer L[NL];
    double scale = 1.0;
    int step = 0;
    hipStream_t stream = nullptr;
    bf16_t *x0 = nullptr, *a1 = nullptr, *cat1 = nullptr, *p2 = nullptr, *a3 = nullptr, *cat2 = nullptr;
    bf16_t *p3 = nullptr, *a5 = nullptr, *a6 = nullptr, *a7 = nullptr, *a8 = nullptr, *a9 = nullptr, *a10 = nullptr;
    float *y = nullptr;          // [P1][16]
    bf16_t *g11 = nullptr, *g10 = nullptr, *g9 = nullptr, *dcat1 = nullptr, *g8 = nullptr, *g7 = nullptr;
    bf16_t *dcat2 = nullptr, *g6 = nullptr, *g5 = nullptr, *dp3 = nullptr, *g4 = nullptr, *g3 = nullptr;
    bf16_t *dp2 = nullptr, *g2 = nullptr, *g1 = nullptr;
    float *part = nullptr, *bpart = nullptr, *tgt = nullptr;
    double *dpart = nullptr, *dacc = nullptr;   // 1024 + 4 doubles
    bf16_t *ds_feat = nullptr;   // [cap][P1][8]
    float *ds_tgt = nullptr;     // [cap][P1][3]
    int cap = 0, nsamples = 0;
    std::vector<void *> allocs;
};

namespace {

template <class T>
int ualloc(gm_unet *u, T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) return GM_OK;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) return fail(GM_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    u->allocs.push_back(*p);
    (void)hipMemset(*p, 0, n * sizeof(T));
    return GM_OK;
}

long long nchunks(long long P) { return (P + WCHUNK - 1) / WCHUNK; }

void conv(gm_unet *u, int l, const bf16_t *in, long long ldi, bf16_t *out, long long ldo, float *outf = nullptr,
          long long ldf = 0)
{
    const ULayer &L = u->L[l];
    const int *n = u->n[LEVEL[l]];
    gmu::ConvArgs a;
    a.in = in; a.ldi = ldi;
    a.w = L.w16; a.kpad = L.kpad;
    a.bias = L.b;
    a.out = out; a.ldo = ldo;
    a.outf = outf; a.ldf = ldf;
    a.nx = n[0]; a.ny = n[1]; a.nz = n[2];
    a.cin = L.cin; a.coutp = L.coutp; a.cout = L.coutp;
    a.relu = l != NL - 1;
    gmu::launch_conv(u->stream, a);
}

// input gradient of layer l: out = W_l^T g (masked by `mask` > 0 when given)
void conv_bwd(gm_unet *u, int l, const bf16_t *g, const bf16_t *mask, long long ldm, bf16_t *out, long long ldo)
{
    const ULayer &L = u->L[l];
    const int *n = u->n[LEVEL[l]];
    gmu::ConvArgs a;
    a.in = g; a.ldi = L.coutp;
    a.w = L.wT; a.kpad = L.kpadT;
    a.out = out; a.ldo = ldo;
    a.mask = mask; a.ldm = ldm;
    a.nx = n[0]; a.ny = n[1]; a.nz = n[2];
    a.cin = L.coutp; a.coutp = L.cin; a.cout = L.cin;
    gmu::launch_conv(u->stream, a);
}

void wgrad(gm_unet *u, int l, const bf16_t *in, long long ldi, const bf16_t *g)
{
    const ULayer &L = u->L[l];
    const int *n = u->n[LEVEL[l]];
    const long long P = u->P[LEVEL[l]];
    gmu::WgradArgs a;
    a.in = in; a.ldi = ldi;
    a.g = g; a.ldg = L.coutp;
    a.nx = n[0]; a.ny = n[1]; a.nz = n[2];
    a.cin = L.cin; a.cout = L.coutp;
    a.chunk = WCHUNK;
    a.nchunks = (int)nchunks(P);
    a.part = u->part;
    gmu::launch_wgrad(u->stream, a);
    gmu::launch_wgrad_reduce(u->stream, u->part, a.nchunks, L.coutp, L.cin, L.kpad, L.g, 1);
    gmu::launch_bias_grad(u->stream, g, L.coutp, P, L.coutp, u->bpart, L.bg, 1);
}

void forward(gm_unet *u, const bf16_t *x0)
{
    const int *n1 = u->n[1], *n2 = u->n[2], *n3 = u->n[3];
    hipStream_t s = u->stream;
    conv(u, 0, x0, 8, u->a1, 32);
    conv(u, 1, u->a1, 32, u->cat1 + 64, 96);                     // a2
    gmu::launch_pool(s, u->cat1 + 64, 96, n1[0], n1[1], n1[2], 32, u->p2, 32);
    conv(u, 2, u->p2, 32, u->a3, 64);
    conv(u, 3, u->a3, 64, u->cat2 + 128, 192);                   // a4
    gmu::launch_pool(s, u->cat2 + 128, 192, n2[0], n2[1], n2[2], 64, u->p3, 64);
    conv(u, 4, u->p3, 64, u->a5, 128);
    conv(u, 5, u->a5, 128, u->a6, 128);
    gmu::launch_up(s, u->a6, 128, n3[0], n3[1], n3[2], 128, u->cat2, 192);
    conv(u, 6, u->cat2, 192, u->a7, 64);
    conv(u, 7, u->a7, 64, u->a8, 64);
    gmu::launch_up(s, u->a8, 64, n2[0], n2[1], n2[2], 64, u->cat1, 96);
    conv(u, 8, u->cat1, 96, u->a9, 32);
    conv(u, 9, u->a9, 32, u->a10, 32);
    conv(u, 10, u->a10, 32, nullptr, 0, u->y, 16);
}

// Gradients of the loss of one sample are added to every layer's g / bg;
// the loss is added to dacc[0].
void backward(gm_unet *u, const bf16_t *x0, const float *target)
{
    const int *n1 = u->n[1], *n2 = u->n[2], *n3 = u->n[3];
    hipStream_t s = u->stream;
    gmu::launch_loss(s, u->y, target, (float)(1.0 / u->scale), u->P[1], u->g11, u->dpart, u->dacc);
    wgrad(u, 10, u->a10, 32, u->g11);
    conv_bwd(u, 10, u->g11, u->a10, 32, u->g10, 32);
    wgrad(u, 9, u->a9, 32, u->g10);
    conv_bwd(u, 9, u->g10, u->a9, 32, u->g9, 32);
    wgrad(u, 8, u->cat1, 96, u->g9);
    conv_bwd(u, 8, u->g9, nullptr, 0, u->dcat1, 96);
    gmu::launch_up_bwd(s, u->dcat1, 96, n2[0], n2[1], n2[2], 64, u->a8, 64, u->g8, 64);
    wgrad(u, 7, u->a7, 64, u->g8);
    conv_bwd(u, 7, u->g8, u->a7, 64, u->g7, 64);
    wgrad(u, 6, u->cat2, 192, u->g7);
    conv_bwd(u, 6, u->g7, nullptr, 0, u->dcat2, 192);
    gmu::launch_up_bwd(s, u->dcat2, 192, n3[0], n3[1], n3[2], 128, u->a6, 128, u->g6, 128);
    wgrad(u, 5, u->a5, 128, u->g6);
    conv_bwd(u, 5, u->g6, u->a5, 128, u->g5, 128);
    wgrad(u, 4, u->p3, 64, u->g5);
    conv_bwd(u, 4, u->g5, nullptr, 0, u->dp3, 64);
    gmu::launch_pool_bwd(s, u->dp3, 64, n2[0], n2[1], n2[2], 64, u->dcat2 + 128, 192, u->cat2 + 128, 192, u->g4, 64);
    wgrad(u, 3, u->a3, 64, u->g4);
    conv_bwd(u, 3, u->g4, u->a3, 64, u->g3, 64);
    wgrad(u, 2, u->p2, 32, u->g3);
    conv_bwd(u, 2, u->g3, nullptr, 0, u->dp2, 32);
    gmu::launch_pool_bwd(s, u->dp2, 32, n1[0], n1[1], n1[2], 32, u->dcat1 + 64, 96, u->cat1 + 64, 96, u->g2, 32);
    wgrad(u, 1, u->a1, 32, u->g2);
    conv_bwd(u, 1, u->g2, u->a1, 32, u->g1, 32);
    wgrad(u, 0, x0, 8, u->g1);
}

void zero_grads(gm_unet *u)
{
    for (auto &L : u->L) {
        gmu::launch_zero(u->stream, L.g, (long long)L.coutp * L.kpad * 4);
        gmu::launch_zero(u->stream, L.bg, (long long)L.coutp * 4);
    }
}

void refresh_bf16(gm_unet *u)
{
    for (int l = 0; l < NL; ++l) {
        ULayer &L = u->L[l];
        gmu::launch_f32_to_bf16(u->stream, L.w, (long long)L.coutp * L.kpad, L.w16);
        if (L.wT) gmu::launch_flip(u->stream, L.w16, L.coutp, L.cin, L.kpad, L.wT, L.kpadT);
    }
}

void adam(gm_unet *u, float lr, float gscale)
{
    ++u->step;
    for (auto &L : u->L) {
        gmu::launch_adam(u->stream, L.w, L.g, L.m, L.v, (long long)L.coutp * L.kpad, gscale, lr, 0.9f, 0.999f, 1e-8f,
                         u->step, L.w16);
        gmu::launch_adam(u->stream, L.b, L.bg, L.bm, L.bv, L.coutp, gscale, lr, 0.9f, 0.999f, 1e-8f, u->step, nullptr);
        if (L.wT) gmu::launch_flip(u->stream, L.w16, L.coutp, L.cin, L.kpad, L.wT, L.kpadT);
    }
}

int upload_feat(gm_unet *u, const double *feat)
{
    const long long n = u->P[1] * 8;
    std::vector<float> f(feat, feat + n);
    HIPCHK(hipMemcpyAsync(u->y, f.data(), n * 4, hipMemcpyHostToDevice, u->stream));   // y as fp32 staging
    gmu::launch_f32_to_bf16(u->stream, u->y, n, u->x0);
    HIPCHK(hipStreamSynchronize(u->stream));
    return GM_OK;
}

}  // namespace

extern "C" int64_t gm_unet_num_params(void)
{
    int64_t n = 0;
    for (int l = 0; l < NL; ++l) n += (int64_t)LCOUT[l] * 27 * LCIN[l] + LCOUT[l];
    return n + 1;
}

extern "C" int gm_unet_destroy(gm_unet *u)
{
    if (!u) return GM_OK;
    if (u->stream) (void)hipStreamSynchronize(u->stream);
    for (void *p : u->allocs) (void)hipFree(p);
    if (u->stream) (void)hipStreamDestroy(u->stream);
    delete u;
    return GM_OK;
}

extern "C" int gm_unet_create(int nx, int ny, int nz, int capacity, gm_unet **out)
{
    if (!out) return fail(GM_ERR_ARG, "gm_unet_create: null output");
    *out = nullptr;
    const int dims[3] = {nx, ny, nz};
    for (int d = 0; d < 3; ++d)
        if (dims[d] < 5 || dims[d] % 4 != 1)
            return fail(GM_ERR_ARG, "gm_unet_create: every grid size must be 4m+1 >= 5 (two exact coarsenings)");
    if (capacity < 0) return fail(GM_ERR_ARG, "gm_unet_create: negative capacity");
    gm_unet *u = new gm_unet();
    for (int d = 0; d < 3; ++d) {
        u->n[1][d] = dims[d];
        u->n[2][d] = (dims[d] + 1) / 2;
        u->n[3][d] = (u->n[2][d] + 1) / 2;
    }
    for (int v = 1; v <= 3; ++v) u->P[v] = (long long)u->n[v][0] * u->n[v][1] * u->n[v][2];
    auto bail = [&](int r) { gm_unet_destroy(u); return r; };
    if (hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(GM_ERR_HIP, "hipStreamCreate"));
    long long part_cap = 0;
    for (int l = 0; l < NL; ++l) {
        ULayer &L = u->L[l];
        L.cin = LCIN[l];
        L.cout = LCOUT[l];
        L.coutp = (L.cout + 15) / 16 * 16;
        L.kpad = pad32(27 * L.cin);
        L.kpadT = pad32(27 * L.coutp);
        const size_t nw = (size_t)L.coutp * L.kpad;
        int r;
        if ((r = ualloc(u, &L.w, nw)) || (r = ualloc(u, &L.m, nw)) || (r = ualloc(u, &L.v, nw)) ||
            (r = ualloc(u, &L.g, nw)) || (r = ualloc(u, &L.b, L.coutp)) || (r = ualloc(u, &L.bm, L.coutp)) ||
            (r = ualloc(u, &L.bv, L.coutp)) || (r = ualloc(u, &L.bg, L.coutp)) || (r = ualloc(u, &L.w16, nw)))
            return bail(r);
        if (l > 0 && (r = ualloc(u, &L.wT, (size_t)L.cin * L.kpadT))) return bail(r);
        part_cap = std::max(part_cap, nchunks(u->P[LEVEL[l]]) * L.coutp * 27 * L.cin);
    }
    const long long P1 = u->P[1], P2 = u->P[2], P3 = u->P[3];
    int r;
    if ((r = ualloc(u, &u->x0, P1 * 8)) || (r = ualloc(u, &u->a1, P1 * 32)) || (r = ualloc(u, &u->cat1, P1 * 96)) ||
        (r = ualloc(u, &u->p2, P2 * 32)) || (r = ualloc(u, &u->a3, P2 * 64)) || (r = ualloc(u, &u->cat2, P2 * 192)) ||
        (r = ualloc(u, &u->p3, P3 * 64)) || (r = ualloc(u, &u->a5, P3 * 128)) || (r = ualloc(u, &u->a6, P3 * 128)) ||
        (r = ualloc(u, &u->a7, P2 * 64)) || (r = ualloc(u, &u->a8, P2 * 64)) || (r = ualloc(u, &u->a9, P1 * 32)) ||
        (r = ualloc(u, &u->a10, P1 * 32)) || (r = ualloc(u, &u->y, P1 * 16)) || (r = ualloc(u, &u->g11, P1 * 16)) ||
        (r = ualloc(u, &u->g10, P1 * 32)) || (r = ualloc(u, &u->g9, P1 * 32)) || (r = ualloc(u, &u->dcat1, P1 * 96)) ||
        (r = ualloc(u, &u->g8, P2 * 64)) || (r = ualloc(u, &u->g7, P2 * 64)) || (r = ualloc(u, &u->dcat2, P2 * 192)) ||
        (r = ualloc(u, &u->g6, P3 * 128)) || (r = ualloc(u, &u->g5, P3 * 128)) || (r = ualloc(u, &u->dp3, P3 * 64)) ||
        (r = ualloc(u, &u->g4, P2 * 64)) || (r = ualloc(u, &u->g3, P2 * 64)) || (r = ualloc(u, &u->dp2, P2 * 32)) ||
        (r = ualloc(u, &u->g2, P1 * 32)) || (r = ualloc(u, &u->g1, P1 * 32)) || (r = ualloc(u, &u->part, part_cap)) ||
        (r = ualloc(u, &u->bpart, (size_t)gmu::bias_grad_chunks(P1) * 128)) || (r = ualloc(u, &u->tgt, P1 * 3)) ||
        (r = ualloc(u, &u->dpart, 1024 + (size_t)gmu::loss_blocks(P1))) || (r = ualloc(u, &u->dacc, 4)))
        return bail(r);
    if (capacity > 0) {
        if ((r = ualloc(u, &u->ds_feat, (size_t)capacity * P1 * 8)) || (r = ualloc(u, &u->ds_tgt, (size_t)capacity * P1 * 3)))
            return bail(r);
    }
    u->cap = capacity;
    HIPCHK(hipDeviceSynchronize());
    *out = u;
    return GM_OK;
}

extern "C" int gm_unet_set_params(gm_unet *u, const double *params, int64_t n)
{
    if (!u || !params) return fail(GM_ERR_ARG, "gm_unet_set_params: null pointer");
    if (n != gm_unet_num_params()) return fail(GM_ERR_ARG, "gm_unet_set_params: wrong parameter count");
    if (!(params[n - 1] > 0)) return fail(GM_ERR_ARG, "gm_unet_set_params: the output scale must be > 0");
    const double *q = params;
    for (auto &L : u->L) {
        std::vector<float> w((size_t)L.coutp * L.kpad, 0.f), b(L.coutp, 0.f);
        for (int co = 0; co < L.cout; ++co)
            for (int k = 0; k < 27 * L.cin; ++k) w[(size_t)co * L.kpad + k] = (float)q[(size_t)co * 27 * L.cin + k];
        q += (size_t)L.cout * 27 * L.cin;
        for (int co = 0; co < L.cout; ++co) b[co] = (float)q[co];
        q += L.cout;
        HIPCHK(hipMemcpy(L.w, w.data(), w.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(L.b, b.data(), b.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemset(L.m, 0, w.size() * 4));
        HIPCHK(hipMemset(L.v, 0, w.size() * 4));
        HIPCHK(hipMemset(L.bm, 0, b.size() * 4));
        HIPCHK(hipMemset(L.bv, 0, b.size() * 4));
    }
    u->scale = *q;
    u->step = 0;
    refresh_bf16(u);
    HIPCHK(hipStreamSynchronize(u->stream));
    return GM_OK;
}

extern "C" int gm_unet_get_params(gm_unet *u, double *params, int64_t n)
{
    if (!u || !params) return fail(GM_ERR_ARG, "gm_unet_get_params: null pointer");
    if (n != gm_unet_num_params()) return fail(GM_ERR_ARG, "gm_unet_get_params: wrong parameter count");
    HIPCHK(hipStreamSynchronize(u->stream));
    double *q = params;
    for (auto &L : u->L) {
        std::vector<float> w((size_t)L.coutp * L.kpad), b(L.coutp);
        HIPCHK(hipMemcpy(w.data(), L.w, w.size() * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(b.data(), L.b, b.size() * 4, hipMemcpyDeviceToHost));
        for (int co = 0; co < L.cout; ++co)
            for (int k = 0; k < 27 * L.cin; ++k) q[(size_t)co * 27 * L.cin + k] = w[(size_t)co * L.kpad + k];
        q += (size_t)L.cout * 27 * L.cin;
        for (int co = 0; co < L.cout; ++co) q[co] = b[co];
        q += L.cout;
    }
    *q = u->scale;
    return GM_OK;
}

// He-normal weights (std sqrt(2 / (27 cin)); the head's scaled by 0.1),
// zero biases, output scale 1 (gm_unet_train sets it from the data).
extern "C" int gm_unet_init_params(gm_unet *u, uint64_t seed)
{
    if (!u) return fail(GM_ERR_ARG, "null network");
    std::vector<double> p((size_t)gm_unet_num_params());
    gm::SplitMix64 rng(seed);
    size_t o = 0;
    for (int l = 0; l < NL; ++l) {
        const double sd = std::sqrt(2.0 / (27.0 * LCIN[l])) * (l == NL - 1 ? 0.1 : 1.0);
        const size_t nw = (size_t)LCOUT[l] * 27 * LCIN[l];
        for (size_t i = 0; i < nw; ++i) {
            // Box-Muller
            const double u1 = 1.0 - rng.uniform(), u2 = rng.uniform();
            p[o + i] = sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        }
        o += nw + LCOUT[l];
    }
    p[o] = 1.0;
    int r = gm_unet_set_params(u, p.data(), (int64_t)p.size());
    if (r) return r;
    u->scale = 0.0;   // set from the data on the first training call
    return GM_OK;
}

extern "C" int gm_unet_forward(gm_unet *u, const double *feat, double *y)
{
    if (!u || !feat || !y) return fail(GM_ERR_ARG, "gm_unet_forward: null pointer");
    if (!(u->scale > 0)) return fail(GM_ERR_ARG, "gm_unet_forward: no output scale (set parameters or train first)");
    if (int r = upload_feat(u, feat)) return r;
    forward(u, u->x0);
    const long long P1 = u->P[1];
    std::vector<float> h((size_t)P1 * 16);
    HIPCHK(hipMemcpyAsync(h.data(), u->y, h.size() * 4, hipMemcpyDeviceToHost, u->stream));
    HIPCHK(hipStreamSynchronize(u->stream));
    for (long long p = 0; p < P1; ++p)
        for (int c = 0; c < 3; ++c) y[p * 3 + c] = u->scale * h[p * 16 + c];
    return GM_OK;
}

extern "C" int gm_unet_loss_grad(gm_unet *u, const double *feat, const double *target, double *loss, double *grad)
{
    if (!u || !feat || !target) return fail(GM_ERR_ARG, "gm_unet_loss_grad: null pointer");
    if (!(u->scale > 0)) return fail(GM_ERR_ARG, "gm_unet_loss_grad: no output scale");
    if (int r = upload_feat(u, feat)) return r;
    const long long P1 = u->P[1];
    std::vector<float> t(target, target + P1 * 3);
    HIPCHK(hipMemcpyAsync(u->tgt, t.data(), t.size() * 4, hipMemcpyHostToDevice, u->stream));
    zero_grads(u);
    gmu::launch_zero(u->stream, u->dacc, 8);
    forward(u, u->x0);
    backward(u, u->x0, u->tgt);
    double l = 0;
    HIPCHK(hipMemcpyAsync(&l, u->dacc, 8, hipMemcpyDeviceToHost, u->stream));
    HIPCHK(hipStreamSynchronize(u->stream));
    if (loss) *loss = l;
    if (grad) {
        double *q = grad;
        for (auto &L : u->L) {
            std::vector<float> w((size_t)L.coutp * L.kpad), b(L.coutp);
            HIPCHK(hipMemcpy(w.data(), L.g, w.size() * 4, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(b.data(), L.bg, b.size() * 4, hipMemcpyDeviceToHost));
            for (int co = 0; co < L.cout; ++co)
                for (int k = 0; k < 27 * L.cin; ++k) q[(size_t)co * 27 * L.cin + k] = w[(size_t)co * L.kpad + k];
            q += (size_t)L.cout * 27 * L.cin;
            for (int co = 0; co < L.cout; ++co) q[co] = b[co];
            q += L.cout;
        }
    }
    return GM_OK;
}

static int ctx_check(gm_ctx *c, gm_unet *u, int *stride)
{
    int s = 0;
    const int cn[3] = {c->p.nx, c->p.ny, c->p.nz};
    for (int d = 0; d < 3; ++d) {
        const int sd = cn[d] == u->n[1][d] ? 1 : cn[d] == 2 * u->n[1][d] - 1 ? 2 : 0;
        if (!sd || (s && sd != s))
            return fail(GM_ERR_ARG, "context grid is neither the network's grid nor its 2x refinement");
        s = sd;
    }
    if (c->p.xPeri || c->p.yPeri || c->p.zPeri)
        return fail(GM_ERR_UNSUPPORTED, "warm start: line-tied boundaries only");
    *stride = s;
    return GM_OK;
}

static int ctx_cur(gm_ctx *c)
{
    gmd::Ctl h;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&h, c->ctl, sizeof h, hipMemcpyDeviceToHost));
    c->cur = h.cur;
    return GM_OK;
}

extern "C" int gm_unet_add_sample(gm_unet *u, gm_ctx *c)
{
    if (!u || !c) return fail(GM_ERR_ARG, "gm_unet_add_sample: null pointer");
    int stride;
    if (int r = ctx_check(c, u, &stride)) return r;
    if (stride != 1) return fail(GM_ERR_ARG, "gm_unet_add_sample: the context must be on the network's grid");
    if (u->nsamples >= u->cap) return fail(GM_ERR_ARG, "gm_unet_add_sample: dataset capacity reached");
    if (int r = ctx_cur(c)) return r;
    const long long P1 = u->P[1];
    const bool f64 = c->prec == GM_PREC_F64;
    const int sx = c->p.nx + 2;
    const long long sxy = (long long)(c->p.nx + 2) * (c->p.ny + 2);
    gmu::launch_features(c->stream, f64 ? c->B0_64 : nullptr, f64 ? nullptr : c->B0_32, c->P, sx, sxy, 1, u->n[1][0],
                         u->n[1][1], u->n[1][2], u->ds_feat + (size_t)u->nsamples * P1 * 8);
    gmu::launch_target(c->stream, f64 ? c->X64[c->cur] : nullptr, f64 ? nullptr : c->Xh[c->cur],
                       f64 ? nullptr : c->Xl[c->cur], c->P, sx, sxy, u->n[1][0], u->n[1][1], u->n[1][2], c->p.Ox,
                       c->p.Oy, c->p.Oz, c->p.dx, c->p.dy, c->p.dz, u->ds_tgt + (size_t)u->nsamples * P1 * 3);
    HIPCHK(hipStreamSynchronize(c->stream));
    ++u->nsamples;
    return GM_OK;
}

extern "C" int gm_unet_num_samples(gm_unet *u) { return u ? u->nsamples : 0; }
extern "C" int gm_unet_clear_samples(gm_unet *u)
{
    if (!u) return fail(GM_ERR_ARG, "null network");
    u->nsamples = 0;
    return GM_OK;
}

extern "C" int gm_unet_train(gm_unet *u, int epochs, int batch, double lr, uint64_t seed, double *loss)
{
    if (!u) return fail(GM_ERR_ARG, "null network");
    if (u->nsamples < 1 || epochs < 1 || batch < 1) return fail(GM_ERR_ARG, "gm_unet_train: nothing to train");
    const long long P1 = u->P[1];
    if (!(u->scale > 0)) {
        // output scale = RMS of the targets
        gmu::launch_sumsq(u->stream, u->ds_tgt, (long long)u->nsamples * P1 * 3, u->dpart, u->dacc + 1);
        double ss = 0;
        HIPCHK(hipMemcpyAsync(&ss, u->dacc + 1, 8, hipMemcpyDeviceToHost, u->stream));
        HIPCHK(hipStreamSynchronize(u->stream));
        u->scale = ss > 0 ? std::sqrt(ss / ((double)u->nsamples * P1 * 3)) : 1.0;
    }
    gm::SplitMix64 rng(seed);
    std::vector<int> order(u->nsamples);
    double last = 0;
    for (int e = 0; e < epochs; ++e) {
        for (int i = 0; i < u->nsamples; ++i) order[i] = i;
        for (int i = u->nsamples - 1; i > 0; --i) std::swap(order[i], order[(int)(rng.next() % (uint64_t)(i + 1))]);
        gmu::launch_zero(u->stream, u->dacc, 8);
        for (int b0 = 0; b0 < u->nsamples; b0 += batch) {
            const int nb = std::min(batch, u->nsamples - b0);
            zero_grads(u);
            for (int j = 0; j < nb; ++j) {
                const size_t sidx = order[b0 + j];
                const bf16_t *x0 = u->ds_feat + sidx * P1 * 8;
                forward(u, x0);
                backward(u, x0, u->ds_tgt + sidx * P1 * 3);
            }
            adam(u, (float)lr, 1.f / nb);
        }
        HIPCHK(hipMemcpyAsync(&last, u->dacc, 8, hipMemcpyDeviceToHost, u->stream));
        HIPCHK(hipStreamSynchronize(u->stream));
        last /= u->nsamples;
    }
    if (loss) *loss = last;
    return GM_OK;
}

// mesh check of the context's current state (after gm_fields): min det
// over inner points and the NaN flag
static int mesh_check(gm_ctx *c, gm_unet *u, double *min_det, int *ok, const gm_reduction &red)
{
    const int sx = c->p.nx + 2;
    const long long sxy = (long long)(c->p.nx + 2) * (c->p.ny + 2);
    gmu::launch_min_det(c->stream, c->det, c->p.nx, c->p.ny, c->p.nz, sx, sxy, u->dpart, u->dacc + 2);
    double md[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(md, u->dacc + 2, 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *min_det = md[0];
    *ok = md[1] == 0 && md[0] > 0 && std::isfinite(red.JxBMax_JBMax) && std::isfinite(red.B2);
    return GM_OK;
}

extern "C" int gm_warm_start(gm_ctx *c, gm_unet *u, int settle, int flags, gm_warm_info *info)
{
    if (!c || !u) return fail(GM_ERR_ARG, "gm_warm_start: null pointer");
    if (!(u->scale > 0)) return fail(GM_ERR_ARG, "gm_warm_start: the network has no output scale");
    if (settle < 0) return fail(GM_ERR_ARG, "gm_warm_start: negative settle count");
    int stride;
    if (int r = ctx_check(c, u, &stride)) return r;
    const auto t0 = std::chrono::steady_clock::now();
    if (int r = ctx_cur(c)) return r;
    gm_warm_info w{};
    gm_reduction red{};
    int r;
    if ((r = gm_fields(c)) || (r = gm_reductions(c, &red))) return r;
    w.ratio_cold = red.JxBMax_JBMax;
    w.energy_cold = red.B2;
    // cold state: controller (time, dt, counters, buffer index) and positions
    gmd::Ctl ctl0;
    HIPCHK(hipMemcpy(&ctl0, c->ctl, sizeof ctl0, hipMemcpyDeviceToHost));
    const bool f64 = c->prec == GM_PREC_F64;
    const size_t half = f64 ? 3 * (size_t)c->P * 8 : 3 * (size_t)c->P * 4;
    const size_t nsave = f64 ? half : 2 * half;
    char *save = nullptr;
    HIPCHK(hipMalloc((void **)&save, nsave));
    auto done = [&](int rc) { (void)hipFree(save); return rc; };
    auto copy_pos = [&](bool restore) {
        const int k = ctl0.cur;
        void *a = f64 ? (void *)c->X64[k] : (void *)c->Xh[k];
        bool ok = hipMemcpyAsync(restore ? a : save, restore ? save : a, half, hipMemcpyDeviceToDevice, c->stream) ==
                  hipSuccess;
        if (!f64)
            ok = ok && hipMemcpyAsync(restore ? (void *)c->Xl[k] : save + half, restore ? save + half : (void *)c->Xl[k],
                                      half, hipMemcpyDeviceToDevice, c->stream) == hipSuccess;
        return ok && hipStreamSynchronize(c->stream) == hipSuccess;
    };
    if (!copy_pos(false)) return done(fail(GM_ERR_HIP, "warm start: save positions"));
    // prediction on the context's B0, added to the positions
    const int sx = c->p.nx + 2;
    const long long sxy = (long long)(c->p.nx + 2) * (c->p.ny + 2);
    gmu::launch_features(u->stream, f64 ? c->B0_64 : nullptr, f64 ? nullptr : c->B0_32, c->P, sx, sxy, stride,
                         u->n[1][0], u->n[1][1], u->n[1][2], u->x0);
    forward(u, u->x0);
    gmu::launch_apply(u->stream, u->y, u->n[1][0], u->n[1][1], u->n[1][2], stride, u->scale,
                      f64 ? c->X64[c->cur] : nullptr, f64 ? nullptr : c->Xh[c->cur], f64 ? nullptr : c->Xl[c->cur],
                      c->P, sx, sxy);
    if (hipStreamSynchronize(u->stream) != hipSuccess) return done(fail(GM_ERR_HIP, "warm start: network"));
    if ((r = gm_fields(c)) || (r = gm_reductions(c, &red))) return done(r);
    w.ratio_applied = w.ratio_warm = red.JxBMax_JBMax;
    w.energy_applied = w.energy_warm = red.B2;
    if ((r = mesh_check(c, u, &w.min_det, &w.mesh_ok, red))) return done(r);
    if (w.mesh_ok && settle > 0) {
        gm_run_opts o{};
        o.max_attempts = settle;
        o.n_cycle = 0x7fffffff;
        o.batch = settle;
        gm_run_stats st{};
        if ((r = gm_run(c, &o, &st))) return done(r);
        w.settle_attempts = st.n_attempt - ctl0.n_attempt;
        if ((r = ctx_cur(c)) || (r = gm_fields(c)) || (r = gm_reductions(c, &red))) return done(r);
        w.ratio_warm = red.JxBMax_JBMax;
        w.energy_warm = red.B2;
        if ((r = mesh_check(c, u, &w.min_det, &w.mesh_ok, red))) return done(r);
    }
    w.used = (flags & GM_WARM_FORCE) ||
             (w.mesh_ok && w.ratio_warm < w.ratio_cold && w.energy_warm <= w.energy_cold);
    if (!w.used) {
        // fallback: the cold controller state and positions
        HIPCHK(hipMemcpy(c->ctl, &ctl0, sizeof ctl0, hipMemcpyHostToDevice));
        c->cur = ctl0.cur;
        if (!copy_pos(true)) return done(fail(GM_ERR_HIP, "warm start: restore"));
        if ((r = gm_fields(c))) return done(r);
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) return done(fail(GM_ERR_HIP, "warm start: sync"));
    w.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (info) *info = w;
    return done(GM_OK);
}
