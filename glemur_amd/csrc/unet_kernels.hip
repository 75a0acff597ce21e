// Device kernels of the warm-start U-Net.  Convolutions are implicit GEMMs
// on the gfx950 bf16 matrix instruction v_mfma_f32_16x16x32_bf16 (16 x 16
// outputs, K = 32 per instruction, FP32 accumulation): rows are output
// points, columns output channels, K the (tap, input channel) pairs.  A lane
// supplies row (lane & 15) of A and column (lane & 15) of B over
// k = 8 (lane >> 4) .. +7, and receives column (lane & 15), rows
// 4 (lane >> 4) .. +3 of the result.
#include "unet_kernels.hpp"

#include <cmath>

namespace gmu {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline float bf2f(bf16_t h) { return __uint_as_float((uint32_t)h << 16); }

// float -> bfloat16, round to nearest even (NaN stays NaN)
__device__ inline bf16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}

__device__ inline u32x4 ld16(const bf16_t *p) { return *reinterpret_cast<const u32x4 *>(p); }
__device__ inline void st16(bf16_t *p, u32x4 v) { *reinterpret_cast<u32x4 *>(p) = v; }
__device__ inline bf16x8 as_bf(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }
__device__ inline float lo16(uint32_t w) { return __uint_as_float(w << 16); }
__device__ inline float hi16(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
__device__ inline void unpack8(u32x4 v, float *f)
{
    for (int e = 0; e < 4; ++e) {
        f[2 * e] = lo16(v[e]);
        f[2 * e + 1] = hi16(v[e]);
    }
}
__device__ inline u32x4 pack8(const float *f)
{
    u32x4 v;
    for (int e = 0; e < 4; ++e) v[e] = (uint32_t)f2bf(f[2 * e]) | ((uint32_t)f2bf(f[2 * e + 1]) << 16);
    return v;
}

__device__ inline void decomp(long long p, int nx, int ny, int &x, int &y, int &z)
{
    x = (int)(p % nx);
    long long q = p / nx;
    y = (int)(q % ny);
    z = (int)(q / ny);
}

// ------------------------------------------------------------------ conv
// Workgroup: 4 waves x 32 output points; each wave keeps 2 x NF tiles of
// 16 x 16 FP32 accumulators.  A comes straight from global memory (16 bytes
// per lane: 8 channels of one neighbour point), B (weights) likewise; the
// 27-fold reuse of activations and the reuse of weights across workgroups
// are served by the caches.
template <int NF>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const long long P = (long long)a.nx * a.ny * a.nz;
    const long long m0 = ((long long)blockIdx.x * 4 + wave) * 32;
    int px[2], py[2], pz[2];
    bool pv[2];
    for (int r = 0; r < 2; ++r) {
        const long long p = m0 + 16 * r + r16;
        pv[r] = p < P;
        decomp(pv[r] ? p : 0, a.nx, a.ny, px[r], py[r], pz[r]);
    }
    f32x4 acc[2][NF];
    for (int r = 0; r < 2; ++r)
        for (int f = 0; f < NF; ++f) acc[r][f] = f32x4{0.f, 0.f, 0.f, 0.f};
    const u32x4 zero = {0u, 0u, 0u, 0u};
    const bf16_t *wl = a.w + (long long)r16 * a.kpad + 8 * g;
    for (int k0 = 0; k0 < a.kpad; k0 += 32) {
        const int k = k0 + 8 * g;
        const int t = k / a.cin, ci = k - t * a.cin;
        u32x4 A[2] = {zero, zero};
        if (t < 27) {
            const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
            for (int r = 0; r < 2; ++r) {
                const int qx = px[r] + dx, qy = py[r] + dy, qz = pz[r] + dz;
                if (pv[r] && qx >= 0 && qx < a.nx && qy >= 0 && qy < a.ny && qz >= 0 && qz < a.nz)
                    A[r] = ld16(a.in + (qx + (long long)a.nx * (qy + (long long)a.ny * qz)) * a.ldi + ci);
            }
        }
        const bf16x8 A0 = as_bf(A[0]), A1 = as_bf(A[1]);
        for (int f = 0; f < NF; ++f) {
            const bf16x8 B = as_bf(ld16(wl + (long long)16 * f * a.kpad + k0));
            acc[0][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, B, acc[0][f], 0, 0, 0);
            acc[1][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, B, acc[1][f], 0, 0, 0);
        }
    }
    for (int f = 0; f < NF; ++f) {
        const int co = 16 * f + r16;
        if (co >= a.cout) continue;
        const float b = a.bias ? a.bias[co] : 0.f;
        for (int r = 0; r < 2; ++r)
            for (int i = 0; i < 4; ++i) {
                const long long p = m0 + 16 * r + 4 * g + i;
                if (p >= P) continue;
                float v = acc[r][f][i] + b;
                if (a.relu && !(v > 0.f)) v = 0.f;
                if (a.mask && !(bf2f(a.mask[p * a.ldm + co]) > 0.f)) v = 0.f;
                if (a.out) a.out[p * a.ldo + co] = f2bf(v);
                if (a.outf) a.outf[p * a.ldf + co] = v;
            }
    }
}

// ------------------------------------------------------------- weight grad
// Workgroup: one chunk of points, one tap, a 64 x 64 block of (co, ci).
// Tiles of 64 points of g and of the tap-shifted input are transposed into
// LDS ([channel][point]) so that K (points) runs along each lane's 8
// operand elements; wave w owns the 16 rows co0 + 16w .. and all four
// 16-column blocks.
constexpr int WP = 64;        // points per tile
constexpr int LP = WP + 8;    // LDS row pitch (bf16 elements)

__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a)
{
    __shared__ __attribute__((aligned(16))) bf16_t Gs[64 * LP];
    __shared__ __attribute__((aligned(16))) bf16_t Xs[64 * LP];
    const int chunk = blockIdx.x, t = blockIdx.y;
    const int nciB = (a.cin + 63) / 64;
    const int co0 = (blockIdx.z / nciB) * 64, ci0 = (blockIdx.z % nciB) * 64;
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    const long long P = (long long)a.nx * a.ny * a.nz;
    const long long pbeg = (long long)chunk * a.chunk;
    const long long pend = pbeg + a.chunk < P ? pbeg + a.chunk : P;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, g = lane >> 4, r16 = lane & 15;
    const u32x4 zero = {0u, 0u, 0u, 0u};
    f32x4 acc[4];
    for (int f = 0; f < 4; ++f) acc[f] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long long pb = pbeg; pb < pend; pb += WP) {
        __syncthreads();
        for (int j = 0; j < 2; ++j) {
            const int s = tid + 256 * j, pt = s >> 3, cg = s & 7;
            const long long p = pb + pt;
            const int co = co0 + 8 * cg, ci = ci0 + 8 * cg;
            u32x4 gv = zero, xv = zero;
            if (p < pend) {
                if (co < a.cout) gv = ld16(a.g + p * a.ldg + co);
                if (ci < a.cin) {
                    int x, y, z;
                    decomp(p, a.nx, a.ny, x, y, z);
                    const int qx = x + dx, qy = y + dy, qz = z + dz;
                    if (qx >= 0 && qx < a.nx && qy >= 0 && qy < a.ny && qz >= 0 && qz < a.nz)
                        xv = ld16(a.in + (qx + (long long)a.nx * (qy + (long long)a.ny * qz)) * a.ldi + ci);
                }
            }
            for (int e = 0; e < 4; ++e) {
                Gs[(8 * cg + 2 * e) * LP + pt] = (bf16_t)(gv[e] & 0xffffu);
                Gs[(8 * cg + 2 * e + 1) * LP + pt] = (bf16_t)(gv[e] >> 16);
                Xs[(8 * cg + 2 * e) * LP + pt] = (bf16_t)(xv[e] & 0xffffu);
                Xs[(8 * cg + 2 * e + 1) * LP + pt] = (bf16_t)(xv[e] >> 16);
            }
        }
        __syncthreads();
        for (int ks = 0; ks < WP / 32; ++ks) {
            const bf16x8 A = as_bf(*reinterpret_cast<const u32x4 *>(&Gs[(16 * wave + r16) * LP + 32 * ks + 8 * g]));
            for (int f = 0; f < 4; ++f) {
                const bf16x8 B = as_bf(*reinterpret_cast<const u32x4 *>(&Xs[(16 * f + r16) * LP + 32 * ks + 8 * g]));
                acc[f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[f], 0, 0, 0);
            }
        }
    }
    for (int f = 0; f < 4; ++f) {
        const int ci = ci0 + 16 * f + r16;
        if (ci >= a.cin) continue;
        for (int i = 0; i < 4; ++i) {
            const int co = co0 + 16 * wave + 4 * g + i;
            if (co >= a.cout) continue;
            a.part[(((long long)chunk * a.cout + co) * 27 + t) * a.cin + ci] = acc[f][i];
        }
    }
}

__global__ void k_wgrad_reduce(const float *part, int nchunks, int cout, int cin, int kpad, float *grad, int accumulate)
{
    const long long n = (long long)cout * 27 * cin;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int c = 0; c < nchunks; ++c) s += part[c * n + i];
    const int co = (int)(i / (27LL * cin));
    const long long k = i - (long long)co * 27 * cin;
    float *d = grad + (long long)co * kpad + k;
    *d = accumulate ? *d + s : s;
}

constexpr int BG_CHUNK = 4096;

__global__ __launch_bounds__(256) void k_bias_part(const bf16_t *g, long long ldg, long long P, int c, float *part)
{
    __shared__ float red[256];
    const int ph = 256 / c, ch = threadIdx.x % c, k = threadIdx.x / c;
    const long long p0 = (long long)blockIdx.x * BG_CHUNK;
    const long long p1 = p0 + BG_CHUNK < P ? p0 + BG_CHUNK : P;
    float s = 0.f;
    for (long long p = p0 + k; p < p1; p += ph) s += bf2f(g[p * ldg + ch]);
    red[threadIdx.x] = s;
    __syncthreads();
    if (k == 0) {
        for (int j = 1; j < ph; ++j) s += red[j * c + ch];
        part[(long long)blockIdx.x * c + ch] = s;
    }
}

__global__ void k_bias_final(const float *part, int nparts, int c, float *bgrad, int accumulate)
{
    const int ch = threadIdx.x;
    if (ch >= c) return;
    float s = 0.f;
    for (int j = 0; j < nparts; ++j) s += part[(long long)j * c + ch];
    bgrad[ch] = accumulate ? bgrad[ch] + s : s;
}

// ------------------------------------------------------------ pool / up
__global__ void k_pool(const bf16_t *in, long long ldi, int nx, int ny, int nz, int c, bf16_t *out, long long ldo)
{
    const int cx = (nx + 1) / 2, cy = (ny + 1) / 2, cz = (nz + 1) / 2, ng = c / 8;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)cx * cy * cz * ng) return;
    const long long q = idx / ng;
    const int cg = (int)(idx % ng);
    int X, Y, Z;
    decomp(q, cx, cy, X, Y, Z);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int t = 0; t < 27; ++t) {
        const int x = 2 * X + t % 3 - 1, y = 2 * Y + (t / 3) % 3 - 1, z = 2 * Z + t / 9 - 1;
        if (x < 0 || x >= nx || y < 0 || y >= ny || z < 0 || z >= nz) continue;
        float v[8];
        unpack8(ld16(in + (x + (long long)nx * (y + (long long)ny * z)) * ldi + 8 * cg), v);
        for (int e = 0; e < 8; ++e) s[e] += v[e];
    }
    for (int e = 0; e < 8; ++e) s[e] /= 27.f;
    st16(out + q * ldo + 8 * cg, pack8(s));
}

__global__ void k_pool_bwd(const bf16_t *dP, long long ldp, int nx, int ny, int nz, int c, const bf16_t *skip,
                           long long lds, const bf16_t *mask, long long ldm, bf16_t *out, long long ldo)
{
    const int cx = (nx + 1) / 2, cy = (ny + 1) / 2, ng = c / 8;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nx * ny * nz * ng) return;
    const long long p = idx / ng;
    const int cg = (int)(idx % ng);
    int x, y, z;
    decomp(p, nx, ny, x, y, z);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // coarse Q with |2Q - f| <= 1: f/2 (f even) or (f-1)/2, (f+1)/2 (f odd)
    const int x0 = x >> 1, x1 = (x + 1) >> 1, y0 = y >> 1, y1 = (y + 1) >> 1, z0 = z >> 1, z1 = (z + 1) >> 1;
    for (int Z = z0; Z <= z1; ++Z)
        for (int Y = y0; Y <= y1; ++Y)
            for (int X = x0; X <= x1; ++X) {
                float v[8];
                unpack8(ld16(dP + (X + (long long)cx * (Y + (long long)cy * Z)) * ldp + 8 * cg), v);
                for (int e = 0; e < 8; ++e) s[e] += v[e];
            }
    float k[8], m[8];
    unpack8(ld16(skip + p * lds + 8 * cg), k);
    unpack8(ld16(mask + p * ldm + 8 * cg), m);
    for (int e = 0; e < 8; ++e) s[e] = m[e] > 0.f ? s[e] / 27.f + k[e] : 0.f;
    st16(out + p * ldo + 8 * cg, pack8(s));
}

__device__ inline void lin1(int f, int &i0, int &i1, float &w0, float &w1)
{
    if ((f & 1) == 0) { i0 = i1 = f >> 1; w0 = 1.f; w1 = 0.f; }
    else { i0 = (f - 1) >> 1; i1 = i0 + 1; w0 = 0.5f; w1 = 0.5f; }
}

__global__ void k_up(const bf16_t *in, long long ldi, int nx, int ny, int nz, int c, bf16_t *out, long long ldo)
{
    const int fx = 2 * nx - 1, fy = 2 * ny - 1, fz = 2 * nz - 1, ng = c / 8;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)fx * fy * fz * ng) return;
    const long long p = idx / ng;
    const int cg = (int)(idx % ng);
    int x, y, z;
    decomp(p, fx, fy, x, y, z);
    int ix[2], iy[2], iz[2];
    float wx[2], wy[2], wz[2];
    lin1(x, ix[0], ix[1], wx[0], wx[1]);
    lin1(y, iy[0], iy[1], wy[0], wy[1]);
    lin1(z, iz[0], iz[1], wz[0], wz[1]);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int e = 0; e < 2; ++e) {
                const float w = wx[e] * wy[b] * wz[a];
                if (w == 0.f) continue;
                float v[8];
                unpack8(ld16(in + (ix[e] + (long long)nx * (iy[b] + (long long)ny * iz[a])) * ldi + 8 * cg), v);
                for (int j = 0; j < 8; ++j) s[j] += w * v[j];
            }
    st16(out + p * ldo + 8 * cg, pack8(s));
}

__global__ void k_up_bwd(const bf16_t *dF, long long ldf, int nx, int ny, int nz, int c, const bf16_t *mask,
                         long long ldm, bf16_t *out, long long ldo)
{
    const int fx = 2 * nx - 1, fy = 2 * ny - 1, fz = 2 * nz - 1, ng = c / 8;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nx * ny * nz * ng) return;
    const long long q = idx / ng;
    const int cg = (int)(idx % ng);
    int X, Y, Z;
    decomp(q, nx, ny, X, Y, Z);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int dz = -1; dz <= 1; ++dz) {
        const int z = 2 * Z + dz;
        if (z < 0 || z >= fz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = 2 * Y + dy;
            if (y < 0 || y >= fy) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = 2 * X + dx;
                if (x < 0 || x >= fx) continue;
                const float w = (dx ? 0.5f : 1.f) * (dy ? 0.5f : 1.f) * (dz ? 0.5f : 1.f);
                float v[8];
                unpack8(ld16(dF + (x + (long long)fx * (y + (long long)fy * z)) * ldf + 8 * cg), v);
                for (int j = 0; j < 8; ++j) s[j] += w * v[j];
            }
        }
    }
    float m[8];
    unpack8(ld16(mask + q * ldm + 8 * cg), m);
    for (int j = 0; j < 8; ++j) s[j] = m[j] > 0.f ? s[j] : 0.f;
    st16(out + q * ldo + 8 * cg, pack8(s));
}

// ------------------------------------------------------------ loss, adam
constexpr int LOSS_PTS = 1024;   // points per workgroup (256 threads x 4)

__global__ __launch_bounds__(256) void k_loss(const float *y, const float *target, float inv_scale, long long P,
                                              bf16_t *g, double *part)
{
    __shared__ double red[256];
    double s = 0;
    const float gs = 2.f / (3.f * (float)P);
    for (int j = 0; j < 4; ++j) {
        const long long p = (long long)blockIdx.x * LOSS_PTS + j * 256 + threadIdx.x;
        if (p >= P) break;
        float v[16];
        for (int c = 0; c < 16; ++c) v[c] = 0.f;
        for (int c = 0; c < 3; ++c) {
            const float d = y[p * 16 + c] - target[p * 3 + c] * inv_scale;
            s += (double)d * d;
            v[c] = gs * d;
        }
        st16(g + p * 16, pack8(v));
        st16(g + p * 16 + 8, pack8(v + 8));
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_loss_final(const double *part, int n, long long P, double *acc)
{
    if (threadIdx.x != 0) return;
    double s = 0;
    for (int i = 0; i < n; ++i) s += part[i];
    acc[0] += s / (3.0 * (double)P);
}

__global__ void k_adam(float *w, const float *g, float *m, float *v, long long n, float gscale, float lr, float b1,
                       float b2, float eps, float c1, float c2, bf16_t *w16)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] * gscale;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    if (w16) w16[i] = f2bf(wi);
}

__global__ void k_flip(const bf16_t *w16, int coutp, int cin, int kpad, bf16_t *wT, int kpadT)
{
    const long long n = (long long)cin * 27 * coutp;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int co = (int)(i % coutp);
    const int tp = (int)((i / coutp) % 27);
    const int ci = (int)(i / (27LL * coutp));
    wT[(long long)ci * kpadT + tp * coutp + co] = w16[(long long)co * kpad + (26 - tp) * cin + ci];
}

__global__ void k_f2b(const float *in, long long n, bf16_t *out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

// ------------------------------------------------- context <-> network
__global__ void k_features(const double *B64, const float *B32, long long Pf, int fsx, long long fsxy, int stride,
                           int nx, int ny, int nz, bf16_t *feat)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)nx * ny * nz) return;
    int i, j, k;
    decomp(q, nx, ny, i, j, k);
    const long long gi = (stride * i + 1) + (long long)fsx * (stride * j + 1) + fsxy * (stride * k + 1);
    float f[8];
    for (int c = 0; c < 3; ++c) {
        const double b = B64 ? B64[c * Pf + gi] : (double)B32[3 * gi + c];   // FP32 B0: triples
        f[c] = (float)(c == 2 ? b - 1.0 : b);
    }
    f[3] = (float)(2.0 * i / (nx - 1) - 1.0);
    f[4] = (float)(2.0 * j / (ny - 1) - 1.0);
    f[5] = (float)(2.0 * k / (nz - 1) - 1.0);
    f[6] = f[7] = 0.f;
    st16(feat + q * 8, pack8(f));
}

__global__ void k_target(const double *X64, const float *Xh, const float *Xl, long long Pf, int fsx, long long fsxy,
                         int nx, int ny, int nz, double Ox, double Oy, double Oz, double dx, double dy, double dz,
                         float *target)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)nx * ny * nz) return;
    int i, j, k;
    decomp(q, nx, ny, i, j, k);
    const long long gi = (i + 1) + (long long)fsx * (j + 1) + fsxy * (k + 1);
    const double o[3] = {Ox + i * dx, Oy + j * dy, Oz + k * dz};
    for (int c = 0; c < 3; ++c) {
        const double x = X64 ? X64[c * Pf + gi] : (double)Xh[3 * gi + c] + (double)Xl[3 * gi + c];
        target[q * 3 + c] = (float)(x - o[c]);
    }
}

__global__ void k_apply(const float *y, int cnx, int cny, int cnz, int stride, double scale, double *X64, float *Xh,
                        float *Xl, long long Pf, int fsx, long long fsxy, int nx, int ny, int nz)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)nx * ny * nz) return;
    int i, j, k;
    decomp(q, nx, ny, i, j, k);
    int ix[2], iy[2], iz[2];
    float wx[2], wy[2], wz[2];
    if (stride == 2) {
        lin1(i, ix[0], ix[1], wx[0], wx[1]);
        lin1(j, iy[0], iy[1], wy[0], wy[1]);
        lin1(k, iz[0], iz[1], wz[0], wz[1]);
    } else {
        ix[0] = ix[1] = i; iy[0] = iy[1] = j; iz[0] = iz[1] = k;
        wx[0] = wy[0] = wz[0] = 1.f;
        wx[1] = wy[1] = wz[1] = 0.f;
    }
    double d[3] = {0, 0, 0};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int e = 0; e < 2; ++e) {
                const double w = (double)wx[e] * wy[b] * wz[a];
                if (w == 0.0) continue;
                const float *s = y + (ix[e] + (long long)cnx * (iy[b] + (long long)cny * iz[a])) * 16;
                for (int c = 0; c < 3; ++c) d[c] += w * s[c];
            }
    const long long gi = (i + 1) + (long long)fsx * (j + 1) + fsxy * (k + 1);
    for (int c = 0; c < 3; ++c) {
        const double dc = scale * d[c];
        if (X64) {
            X64[c * Pf + gi] += dc;
        } else {
            const double x = (double)Xh[3 * gi + c] + (double)Xl[3 * gi + c] + dc;   // F32K: triples
            const float h = (float)x;
            Xh[3 * gi + c] = h;
            Xl[3 * gi + c] = (float)(x - (double)h);
        }
    }
    (void)cnz;
}

__global__ __launch_bounds__(256) void k_min_det(const double *det, int nx, int ny, int nz, int sx, long long sxy,
                                                 double *part)
{
    __shared__ double rmin[256];
    __shared__ int rnan[256];
    const long long N = (long long)nx * ny * nz;
    double mn = 1e300;
    int nan = 0;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < N; q += (long long)gridDim.x * 256) {
        int i, j, k;
        decomp(q, nx, ny, i, j, k);
        const double d = det[(i + 1) + (long long)sx * (j + 1) + sxy * (k + 1)];
        if (d != d) nan = 1;
        else if (d < mn) mn = d;
    }
    rmin[threadIdx.x] = mn;
    rnan[threadIdx.x] = nan;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            rmin[threadIdx.x] = fmin(rmin[threadIdx.x], rmin[threadIdx.x + w]);
            rnan[threadIdx.x] |= rnan[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = rmin[0];
        part[2 * blockIdx.x + 1] = rnan[0];
    }
}

__global__ void k_min_det_final(const double *part, int n, double *out)
{
    if (threadIdx.x != 0) return;
    double mn = 1e300, nan = 0;
    for (int i = 0; i < n; ++i) {
        mn = fmin(mn, part[2 * i]);
        if (part[2 * i + 1] != 0) nan = 1;
    }
    out[0] = mn;
    out[1] = nan;
}

__global__ __launch_bounds__(256) void k_sumsq(const float *v, long long n, double *part)
{
    __shared__ double red[256];
    double s = 0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        s += (double)v[i] * v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_sum_final(const double *part, int n, double *out)
{
    if (threadIdx.x != 0) return;
    double s = 0;
    for (int i = 0; i < n; ++i) s += part[i];
    out[0] = s;
}

inline unsigned nblk(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void launch_conv(hipStream_t s, const ConvArgs &a)
{
    const long long P = (long long)a.nx * a.ny * a.nz;
    const unsigned grid = (unsigned)((P + 127) / 128);
    switch (a.coutp) {
    case 16: k_conv<1><<<grid, 256, 0, s>>>(a); break;
    case 32: k_conv<2><<<grid, 256, 0, s>>>(a); break;
    case 64: k_conv<4><<<grid, 256, 0, s>>>(a); break;
    case 96: k_conv<6><<<grid, 256, 0, s>>>(a); break;
    case 128: k_conv<8><<<grid, 256, 0, s>>>(a); break;
    case 192: k_conv<12><<<grid, 256, 0, s>>>(a); break;
    default: break;   // the host validates coutp
    }
}

void launch_wgrad(hipStream_t s, const WgradArgs &a)
{
    const int ncoB = (a.cout + 63) / 64, nciB = (a.cin + 63) / 64;
    k_wgrad<<<dim3(a.nchunks, 27, ncoB * nciB), 256, 0, s>>>(a);
}

void launch_wgrad_reduce(hipStream_t s, const float *part, int nchunks, int cout, int cin, int kpad, float *grad,
                         int accumulate)
{
    const long long n = (long long)cout * 27 * cin;
    k_wgrad_reduce<<<nblk(n), 256, 0, s>>>(part, nchunks, cout, cin, kpad, grad, accumulate);
}

int bias_grad_chunks(long long P) { return (int)((P + BG_CHUNK - 1) / BG_CHUNK); }

void launch_bias_grad(hipStream_t s, const bf16_t *g, long long ldg, long long P, int c, float *part, float *bgrad,
                      int accumulate)
{
    const int n = bias_grad_chunks(P);
    k_bias_part<<<n, 256, 0, s>>>(g, ldg, P, c, part);
    k_bias_final<<<1, 256, 0, s>>>(part, n, c, bgrad, accumulate);
}

void launch_pool(hipStream_t s, const bf16_t *in, long long ldi, int nx, int ny, int nz, int c, bf16_t *out,
                 long long ldo)
{
    const long long n = (long long)((nx + 1) / 2) * ((ny + 1) / 2) * ((nz + 1) / 2) * (c / 8);
    k_pool<<<nblk(n), 256, 0, s>>>(in, ldi, nx, ny, nz, c, out, ldo);
}

void launch_pool_bwd(hipStream_t s, const bf16_t *dP, long long ldp, int nx, int ny, int nz, int c,
                     const bf16_t *skip, long long lds, const bf16_t *mask, long long ldm, bf16_t *out, long long ldo)
{
    const long long n = (long long)nx * ny * nz * (c / 8);
    k_pool_bwd<<<nblk(n), 256, 0, s>>>(dP, ldp, nx, ny, nz, c, skip, lds, mask, ldm, out, ldo);
}

void launch_up(hipStream_t s, const bf16_t *in, long long ldi, int nx, int ny, int nz, int c, bf16_t *out,
               long long ldo)
{
    const long long n = (long long)(2 * nx - 1) * (2 * ny - 1) * (2 * nz - 1) * (c / 8);
    k_up<<<nblk(n), 256, 0, s>>>(in, ldi, nx, ny, nz, c, out, ldo);
}

void launch_up_bwd(hipStream_t s, const bf16_t *dF, long long ldf, int nx, int ny, int nz, int c,
                   const bf16_t *mask, long long ldm, bf16_t *out, long long ldo)
{
    const long long n = (long long)nx * ny * nz * (c / 8);
    k_up_bwd<<<nblk(n), 256, 0, s>>>(dF, ldf, nx, ny, nz, c, mask, ldm, out, ldo);
}

int loss_blocks(long long P) { return (int)((P + LOSS_PTS - 1) / LOSS_PTS); }

void launch_loss(hipStream_t s, const float *y, const float *target, float inv_scale, long long P, bf16_t *g,
                 double *part, double *loss_acc)
{
    const int n = loss_blocks(P);
    k_loss<<<n, 256, 0, s>>>(y, target, inv_scale, P, g, part);
    k_loss_final<<<1, 64, 0, s>>>(part, n, P, loss_acc);
}

void launch_adam(hipStream_t s, float *w, const float *g, float *m, float *v, long long n, float gscale, float lr,
                 float b1, float b2, float eps, int step, bf16_t *w16)
{
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, step)));
    k_adam<<<nblk(n), 256, 0, s>>>(w, g, m, v, n, gscale, lr, b1, b2, eps, c1, c2, w16);
}

void launch_flip(hipStream_t s, const bf16_t *w16, int coutp, int cin, int kpad, bf16_t *wT, int kpadT)
{
    k_flip<<<nblk((long long)cin * 27 * coutp), 256, 0, s>>>(w16, coutp, cin, kpad, wT, kpadT);
}

void launch_f32_to_bf16(hipStream_t s, const float *in, long long n, bf16_t *out)
{
    k_f2b<<<nblk(n), 256, 0, s>>>(in, n, out);
}

void launch_zero(hipStream_t s, void *p, long long bytes) { (void)hipMemsetAsync(p, 0, (size_t)bytes, s); }

void launch_features(hipStream_t s, const double *B0_64, const float *B0_32, long long Pf, int fsx, long long fsxy,
                     int stride, int nx, int ny, int nz, bf16_t *feat)
{
    k_features<<<nblk((long long)nx * ny * nz), 256, 0, s>>>(B0_64, B0_32, Pf, fsx, fsxy, stride, nx, ny, nz, feat);
}

void launch_target(hipStream_t s, const double *X64, const float *Xh, const float *Xl, long long Pf, int fsx,
                   long long fsxy, int nx, int ny, int nz, double Ox, double Oy, double Oz, double dx, double dy,
                   double dz, float *target)
{
    k_target<<<nblk((long long)nx * ny * nz), 256, 0, s>>>(X64, Xh, Xl, Pf, fsx, fsxy, nx, ny, nz, Ox, Oy, Oz, dx,
                                                           dy, dz, target);
}

void launch_apply(hipStream_t s, const float *y, int cnx, int cny, int cnz, int stride, double scale, double *X64,
                  float *Xh, float *Xl, long long Pf, int fsx, long long fsxy)
{
    const int nx = stride == 2 ? 2 * cnx - 1 : cnx, ny = stride == 2 ? 2 * cny - 1 : cny;
    const int nz = stride == 2 ? 2 * cnz - 1 : cnz;
    k_apply<<<nblk((long long)nx * ny * nz), 256, 0, s>>>(y, cnx, cny, cnz, stride, scale, X64, Xh, Xl, Pf, fsx, fsxy,
                                                          nx, ny, nz);
}

void launch_sumsq(hipStream_t s, const float *v, long long n, double *part, double *out)
{
    const int nb = 512;
    k_sumsq<<<nb, 256, 0, s>>>(v, n, part);
    k_sum_final<<<1, 64, 0, s>>>(part, nb, out);
}

void launch_min_det(hipStream_t s, const double *det, int nx, int ny, int nz, int sx, long long sxy, double *part,
                    double *out)
{
    const int nb = 512;
    k_min_det<<<nb, 256, 0, s>>>(det, nx, ny, nz, sx, sxy, part);
    k_min_det_final<<<1, 64, 0, s>>>(part, nb, out);
}

}  // namespace gmu
