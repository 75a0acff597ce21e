// Fused Cash-Karp stage kernel, FP32 with compensated positions (F32K).
//
// Same 2.5D march as the FP64 kernel (kernels_fused.hip): a workgroup owns
// an x-y tile plus a 2-point halo (one thread per column) and a z chunk,
// and per plane p it forms the stage position of plane p+1, issues the HBM
// loads of a plane further ahead, evaluates B at plane p and J, force and
// slope at plane p-1, with one barrier per plane.  What differs, for FP32:
//   * a point's position (hi x, y, z + small x | small y, z) is one 16-byte
//     and one 8-byte LDS word, B one 16-byte word: a neighbour costs two
//     LDS reads instead of six;
//   * a column's own z-neighbours are read back from the LDS rings (4
//     position planes, 3 B planes) instead of being queued in registers;
//   * the loads of a plane are issued two plane steps before its use (one
//     for the kernels whose registers hold only one input slot), into two
//     input slots that alternate with the march unrolled by two.  Every
//     lane issues every load and store of a step (planes and points clamped
//     into range, stores of lanes without an output dropped at offset
//     2^31), so the compiler's wait before a step consumes its slot covers
//     that slot's loads only;
//   * workgroups: up to 1024 threads for stages 0-4 and the fields pass,
//     768 for the final stage, 768 for every kernel with pressure (the host
//     picks the smallest that holds a rect);
//   * positions (hi, lo), B0 and the slopes are stored as (x, y, z)
//     triples, one 12-byte load or store per point and array: a rect row
//     of an array is one run of 12 (TX + 4) bytes, so its loads touch about
//     a fifth fewer 128-byte lines than three component rows of 4 (TX + 4)
//     bytes;
//   * on grids up to 16 M padded points odd stages march z downwards, so
//     the planes the previous stage wrote last (still in the memory-side
//     cache) are the first the next stage reads; reversal maps virtual
//     plane q to nz + 1 - q and negates z differences, which is exact;
//   * line-tied faces x = const and y = const go through the interior
//     formula: the column beyond the face holds the outer ghost x_g -+ h n
//     of its face point, the tangential neighbours are the ghost layer, and
//     B . n = B0 . n is set afterwards; the table of precomputed face
//     constants is used for the z faces (first and last plane only) and for
//     every face of a grid with an initial boundary distortion;
//   * the final stage carries, per plane, sum_m (c_m - b5_m) k_m and
//     sum_m (c_m - c*_m) k_m over the five stored slopes and adds
//     c_5 k_5 at the point: x_new = x_5 + sum_m (c_m - b5_m) k_m by a
//     two-sum into (hi, lo);
//   * gradP from the cofactor vectors of the current, with the grid
//     spacings cancelled, and its mixed second differences from a two-slot
//     LDS ring of each point's y and z differences;
//   * variants: odd z chunks marching the other way (stages 1-5 of a
//     launch of one round, stages 1-4 of an ensemble), and periodic faces
//     (512 threads), whose ghost images the kernel forms from the opposite
//     side's inputs at every stage.
#include "fused_common.hpp"
#include "gm_device.hpp"
#include "gm_fused.hpp"

namespace gmd {

__constant__ TableauF c_tab32;

namespace {

// Plane steps between a plane's loads and its use: two for the stage
// kernels whose registers hold a second input slot without spilling (bit s
// of the mask: stage s, 6 the fields pass; pressure variants take one)
#ifndef GM_PF2_STAGES
#define GM_PF2_STAGES 0x4f
#endif
#ifndef GM_PF2_STAGES_P
#define GM_PF2_STAGES_P 0x1f
#endif
constexpr int pf_depth(int stage, bool press) { return ((press ? GM_PF2_STAGES_P : GM_PF2_STAGES) >> stage) & 1 ? 2 : 1; }
constexpr int kNXL = 4;      // position planes in LDS
constexpr int kNBL = 3;      // B planes in LDS

struct Pt3 {
    float h[3], s[3];        // high words and small parts (lo + stage increment)
};

__device__ __forceinline__ float dif(const Pt3 &a, const Pt3 &b, int l)
{
    return (a.h[l] - b.h[l]) + (a.s[l] - b.s[l]);
}

// NaN-blind running maximum (a NaN candidate leaves m unchanged)
__device__ __forceinline__ float nb_maxf(float m, float v) { return v > m ? v : m; }

template <int NT>
struct Lds {
    float4 *xa;   // [kNXL][NT] hi x, y, z, small x
    float2 *xb;   // [kNXL][NT] small y, z
    float4 *bl;   // [kNBL][NT] B x, y, z
    __device__ __forceinline__ static int xs(int q) { return (unsigned)(q + 4 * kNXL) % (unsigned)kNXL; }
    __device__ __forceinline__ static int bs(int q) { return (unsigned)(q + 4 * kNBL) % (unsigned)kNBL; }
    __device__ __forceinline__ void ldx(int q, int pt, Pt3 &o) const
    {
        const int i = xs(q) * NT + pt;
        float4 a = xa[i];
        float2 b = xb[i];
        o.h[0] = a.x; o.h[1] = a.y; o.h[2] = a.z;
        o.s[0] = a.w; o.s[1] = b.x; o.s[2] = b.y;
    }
    __device__ __forceinline__ void stx(int q, int pt, const Pt3 &v) const
    {
        const int i = xs(q) * NT + pt;
        xa[i] = make_float4(v.h[0], v.h[1], v.h[2], v.s[0]);
        xb[i] = make_float2(v.s[1], v.s[2]);
    }
    __device__ __forceinline__ void ldb(int q, int pt, float o[3]) const
    {
        float4 a = bl[bs(q) * NT + pt];
        o[0] = a.x; o[1] = a.y; o[2] = a.z;
    }
    __device__ __forceinline__ void stb(int q, int pt, const float v[3]) const
    {
        bl[bs(q) * NT + pt] = make_float4(v[0], v[1], v[2], 0.f);
    }
};

// Raw buffer view of one float array: loads and stores take a 32-bit byte
// offset per lane (and one held in a scalar register), so no per-lane
// 64-bit address arithmetic runs (4% off every stage at 129 x 129 x 385);
// offsets past the array read 0 and drop writes.
struct BufF {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ float ld(int voff, int soff) const
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    }
    __device__ __forceinline__ void st(float v, int voff, int soff) const
    {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
    }
    // three consecutive floats (x, y, z of one point) in one 12-byte access.
    // The loaded vector is bit-cast whole: a bit_cast of one of its
    // elements compiles to a single-dword load whose value lands in every
    // component
    __device__ __forceinline__ void ld3(int voff, float o[3], int soff = 0) const
    {
        typedef float v3f __attribute__((ext_vector_type(3)));
        const v3f f = __builtin_bit_cast(v3f, __builtin_amdgcn_raw_buffer_load_b96(r, voff, soff, 0));
        o[0] = f[0];
        o[1] = f[1];
        o[2] = f[2];
    }
    __device__ __forceinline__ void st3(const float o[3], int voff) const
    {
        typedef unsigned v3u __attribute__((ext_vector_type(3)));
        v3u v;
        v.x = __builtin_bit_cast(unsigned, o[0]);
        v.y = __builtin_bit_cast(unsigned, o[1]);
        v.z = __builtin_bit_cast(unsigned, o[2]);
        __builtin_amdgcn_raw_buffer_store_b96(v, r, voff, 0, 0);
    }
};
// Store offset of lanes without an output point: 2^31 bytes lies past
// every array (an F32K array holds fewer than 2^31 bytes), so the buffer
// range check drops the store
constexpr int kNoStore = (int)0x80000000u;
// 0x00020000: the resource word-3 format of the gfx9 family
__device__ __forceinline__ BufF buf_of(const float *p, long long nfloats)
{
    const long long bytes = nfloats * 4;
    return BufF{__builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(p), 0,
                                                  (int)(bytes < 0x7fffffffLL ? bytes : 0x7fffffffLL), 0x00020000)};
}
// The pressure kernels put the y and z central differences of each B-ring
// point in LDS for one plane: gradP's mixed second differences at a tile
// point are differences of them between its in-plane neighbours,
// (x(+a+b) - x(+a-b)) - (x(-a+b) - x(-a-b)) = dr_b(+a) - dr_b(-a), the same
// floating-point operations on the same values as from the positions, with
// 8 LDS reads instead of 24 (257 x 257 x 769 with pressure: stages 0-4 5%
// faster, the final stage 7%).  The final stage's per-plane sums
// (c - b5) k and (c - c*) k stay in registers: in LDS instead they cost the
// final stage without pressure 5 us at 129 x 129 x 385, and take the LDS the
// ring needs with pressure
constexpr bool dr_lds(bool press) { return press; }
template <int NT, int STAGE = 0, bool PRESS = false>
constexpr size_t lds_bytes32()
{
    // positions, B and, with pressure, the two-slot ring of differences
    return (size_t)kNXL * NT * (sizeof(float4) + sizeof(float2)) + (size_t)kNBL * NT * sizeof(float4) +
           (dr_lds(PRESS) ? (size_t)2 * NT * (sizeof(float4) + sizeof(float2)) : 0);
}

// det of an FP32 Jacobian in FP64 by cofactors (products of two floats are
// exact in FP64, so each cofactor carries a single rounding)
__device__ __forceinline__ double det32(const float a[3][3])
{
    double c0 = fma((double)a[1][1], (double)a[2][2], -((double)a[1][2] * (double)a[2][1]));
    double c1 = fma((double)a[1][2], (double)a[2][0], -((double)a[1][0] * (double)a[2][2]));
    double c2 = fma((double)a[1][0], (double)a[2][1], -((double)a[1][1] * (double)a[2][0]));
    return fma((double)a[0][0], c0, fma((double)a[0][1], c1, (double)a[0][2] * c2));
}

// FP32 Jacobian -> B = jac.B0 / det with det in FP64 (normal < 0: none held)
__device__ __forceinline__ double cauchy32(const float jac[3][3], const float b0[3], int normal, float B[3])
{
    double det = det32(jac);
    const float rd = __builtin_amdgcn_rcpf((float)det);   // 1 ulp: below the FP32 rounding of B
#pragma unroll
    for (int l = 0; l < 3; ++l)
        B[l] = l == normal ? b0[l] : (jac[l][0] * b0[0] + jac[l][1] * b0[1] + jac[l][2] * b0[2]) * rd;
    return det;
}

}  // namespace

// Waves per SIMD a stage kernel is register-allocated for.  At 512 threads
// the stages without pressure other than the final one fit 128 VGPRs without
// spilling, so two workgroups share a CU (4 waves per SIMD); the final stage
// (five slopes in flight plus its per-plane sums) and the pressure variants
// keep the compiler's allocation and one workgroup per CU.  768 threads:
// three waves per SIMD (168 VGPRs), one workgroup per CU.  Smaller
// workgroups: two per CU.
constexpr int waves32(int nt, int stage, bool press)
{
    return nt <= 256 ? 4 : nt <= 384 ? 3 : nt == 768 ? 3 : nt >= 1024 ? 4 : (!press && stage != 5) ? 4 : 2;
}
// 1024-thread workgroups (16 waves, 128 VGPRs) only for the kernels that
// fit 128 VGPRs without spilling
constexpr bool fits1024(int stage, bool press) { return !press && stage != 5; }

// One workgroup's march: `lb` is its logical index in the instance (after
// the XCD-aware remap), which picks the tile and z chunk and the slot of
// its partial results.
template <int NT, int STAGE, bool PRESS, bool ALT, bool PERI>
__device__ __forceinline__ void stage32_body(const FusedArgs &A, const int lb)
{
#pragma clang fp contract(fast)
    constexpr bool FINAL = STAGE == 5;
    constexpr int stage = STAGE == 6 ? -1 : STAGE;
    constexpr int n = STAGE == 6 ? 0 : STAGE;   // slopes entering the stage position
    constexpr bool DIAG = STAGE == 0;            // diagnostics ride on stage 0 only
    // only the fields pass writes B, J, det, gradP; the pressure kernels keep
    // the (runtime-off) branch: without it their stage 0 takes 30 more VGPRs
    constexpr bool EMIT = STAGE == 6 || PRESS;
    constexpr int kPF = pf_depth(STAGE, PRESS);
    constexpr bool kDrLds = dr_lds(PRESS);   // y, z differences in LDS

    extern __shared__ __align__(16) unsigned char smem[];
    Lds<NT> L;
    L.xa = reinterpret_cast<float4 *>(smem);
    L.xb = reinterpret_cast<float2 *>(L.xa + kNXL * NT);
    L.bl = reinterpret_cast<float4 *>(L.xb + kNXL * NT);
    // two-slot ring [2][NT] of dr_y | dr_z of a plane (kDrLds)
    float4 *fqa = L.bl + kNBL * NT;
    float2 *fqb = reinterpret_cast<float2 *>(fqa + 2 * NT);
    (void)fqa; (void)fqb;
    __shared__ double red[7][NT / 64];

    const Geo &g = A.g;
    Ctl *ctl = A.ctl;
    if (A.check_done && ctl->done) return;
    const int cur = ctl->cur;
    const float dt = (float)ctl->dt;
    const long long P = g.P, P3 = 3 * P;

    const int Lb = lb;
    const int tx = Lb % A.tiles_x, ty = (Lb / A.tiles_x) % A.tiles_y, ch = Lb / (A.tiles_x * A.tiles_y);
    const int ix0 = tx * A.TX, iy0 = ty * A.TY;
    // the march runs over virtual planes q; downwards (A.rev), q lives at
    // physical plane nz + 1 - q, and z differences of the march order change
    // sign (zs).  With ALT odd chunks march the other way than A.rev says:
    // neighbouring chunks then read the planes they share (two beyond each
    // end) at about the same time, both first or both last, so the second
    // read finds them in cache.  (A direction that depends on the workgroup
    // changes the compiler's schedule of the whole kernel, by up to 5% on
    // the pressure kernels, so it is a variant of its own.)
    int Ks, Ke;
    bool rev;
    if constexpr (ALT) {
        const int P0 = ch * A.CZ + 1, P1 = min(ch * A.CZ + A.CZ, g.nz);   // physical planes of the chunk
        rev = (A.rev != 0) != ((ch & 1) != 0);
        Ks = rev ? g.nz + 1 - P1 : P0;
        Ke = rev ? g.nz + 1 - P0 : P1;
    } else {
        Ks = ch * A.CZ + 1;
        Ke = min(ch * A.CZ + A.CZ, g.nz);
        rev = A.rev != 0;
    }
    auto zp = [&](int q) { return rev ? g.nz + 1 - q : q; };
    const float zs = rev ? -1.f : 1.f;
    const int RW = A.TX + 4, RH = A.TY + 4;
    const int ixe = min(ix0 + A.TX, g.nx), iye = min(iy0 + A.TY, g.ny);

    // periodic axes (PERI variants): columns, rows and planes beyond the
    // inner range hold the opposite side's points shifted by -+(L + h) along
    // the axis, the reference's ghost copies (src/boundaries.cu:532-582)
    // formed at every stage from the opposite side's inputs; the ghost
    // layer's B then takes the interior formula, which equals the
    // reference's copy of the opposite layer (setPeri) in exact arithmetic
    const bool px = PERI && g.xPeri, py = PERI && g.yPeri, pz = PERI && g.zPeri;

    // ---- this thread's column
    const int t = threadIdx.x;
    const int c = t % RW, r = t / RW;
    const int gi = ix0 - 1 + c, gj = iy0 - 1 + r;
    const bool col = PERI ? t < RW * RH && gi >= (px ? -1 : 0) && gi <= g.nx + (px ? 2 : 1) && gj >= (py ? -1 : 0) &&
                                gj <= g.ny + (py ? 2 : 1)
                          : t < RW * RH && gi >= 0 && gi <= g.nx + 1 && gj >= 0 && gj <= g.ny + 1;
    const bool gx = gi == 0 || gi == g.nx + 1, gy = gj == 0 || gj == g.ny + 1;
    const bool lx = gx && !px, ly = gy && !py;   // ghost columns / rows of line-tied faces
    // virtual columns one step beyond a line-tied x or y face hold the outer
    // ghost x_g -+ h n of the face point next to them (src/boundaries.cu:
    // 164-186, 291-313), so the face point's Jacobian is the interior
    // central-difference formula and no wave runs a separate face branch
    const int vx = (!PERI || !px) && t < RW * RH && gj >= 0 && gj <= g.ny + 1 ? (gi == -1 ? -1 : gi == g.nx + 2 ? 1 : 0) : 0;
    const int vy = (!PERI || !py) && t < RW * RH && gi >= 0 && gi <= g.nx + 1 ? (gj == -1 ? -1 : gj == g.ny + 2 ? 1 : 0) : 0;
    const bool vcol = !A.face_table && (vx != 0 || vy != 0);
    const bool ldc = col || vcol;       // columns whose stage positions go to LDS
    const bool inB = col && c >= 1 && c <= RW - 2 && r >= 1 && r <= RH - 2 &&
                     !((c == 1 || c == RW - 2) && (r == 1 || r == RH - 2)) && !(gx && gy);
    const bool inJ = col && c >= 2 && c <= RW - 3 && r >= 2 && r <= RH - 3 && gi >= 1 && gi <= ixe && gj >= 1 &&
                     gj <= iye;
    // the padded point a lane loads: a virtual column's face point; lanes
    // past the rect repeat its last point and lanes outside the grid the
    // nearest padded point, so every lane's loads are in bounds and a step's
    // loads are issued by every lane (unconditionally: the compiler then
    // counts them and waits only for the slot a step consumes)
    const int tl = min(t, RW * RH - 1);
    const int gil = ix0 - 1 + tl % RW, gjl = iy0 - 1 + tl / RW;
    // abase: the lane's own padded point (stores); aload: the point it
    // loads, on a periodic axis the opposite side's, shifted by sgx, sgy
    // periods when its stage position is formed
    const long long abase = (long long)min(max(gil, 0), g.nx + 1) + g.sx * (long long)min(max(gjl, 0), g.ny + 1);
    const int sgx = px ? (gil <= 0 ? -1 : gil >= g.nx + 1 ? 1 : 0) : 0;
    const int sgy = py ? (gjl <= 0 ? -1 : gjl >= g.ny + 1 ? 1 : 0) : 0;
    const long long aload = PERI ? (long long)(px ? gil - sgx * g.nx : min(max(gil, 0), g.nx + 1)) +
                                       g.sx * (long long)(py ? gjl - sgy * g.ny : min(max(gjl, 0), g.ny + 1))
                                 : abase;
    // byte offsets of 12-byte points: the lane's point within a plane (for
    // its loads and its stores) and, per plane, the plane's first byte, a
    // workgroup-uniform value that the loads take in their scalar offset
    const int vl12 = 12 * (int)aload, vs12 = 12 * (int)abase;
    const int pl12 = 12 * (int)g.sxy;
    const int ii = gi - 1, jj = gj - 1;
    const float eps_xy = (g.epsilonProf && inJ) ? (float)(A.epsx[ii] * A.epsy[jj]) : 1.0f;
    // buffer views: positions (hi, lo), B0, one per slope slot; each holds
    // (x, y, z) of the point at padded index a at bytes 12a .. 12a + 11
    const BufF bXh = buf_of(A.Xh[cur], P3), bXl = buf_of(A.Xl[cur], P3), bB0 = buf_of(A.B0_32, P3);
    BufF bK[6];
    const BufF bXhd = buf_of(A.Xh[cur ^ 1], P3), bXld = buf_of(A.Xl[cur ^ 1], P3);   // final stage: x_new
    (void)bXhd; (void)bXld;
#pragma unroll
    for (int m = 0; m < 6; ++m) bK[m] = buf_of(A.K32 + (long long)(m < 5 ? m : 4) * P3, P3);

    // ---- registers carried between planes.  A new plane's differences and
    // det are locals of its plane step, set only in the branch that uses
    // them: carried through the branches that set them, the old plane's
    // values would be copied between registers on every plane
    float b0q[2][3];      // B0 at planes p, p+1
    double dq0;           // det at plane p-1
    float drq0[3][3];     // central differences x(+a) - x(-a) at plane p-1
    struct In {
        float x[3], lo[3], k[5][3], b0[3];
    } pf0, pf1;           // inputs of one plane, loaded kPF plane steps ahead
    float fg[3][3], fe[3][3];   // final stage: per plane sum (c - b5) k and sum (c - c*) k
    (void)fg; (void)fe;

    double dmax = 0, e_sum = 0;
    float jmax2 = 0.f, r1q = 0.f, r2q = 0.f, r3q = 0.f, r4q = 0.f;   // stage-0 maxima, squared
    int nan_seen = 0;

    // planes a chunk touches: two beyond its ends, within the padded grid
    // (one more on each side of a periodic z axis)
    auto qlo = [&]() { return PERI ? max(Ks - 2, pz ? -1 : 0) : max(Ks - 2, 0); };
    auto qhi = [&]() { return PERI ? min(Ke + 2, g.nz + (pz ? 2 : 1)) : min(Ke + 2, g.nz + 1); };
    auto plane_ok = [&](int q) { return q >= qlo() && q <= qhi(); };
    // periodic z: the physical plane a virtual plane loads, and its period shift
    auto zsrc = [&](int q) { const int z = zp(q); return pz ? (z <= 0 ? z + g.nz : z >= g.nz + 1 ? z - g.nz : z) : z; };
    auto zshift = [&](int q) { const int z = zp(q); return pz ? (z <= 0 ? -1 : z >= g.nz + 1 ? 1 : 0) : 0; };
    (void)zshift;

    // Every lane loads every step; planes outside the chunk's range repeat
    // its first or last plane (cache hits, no extra HBM traffic).  Slopes of
    // ghost points are zero in memory (never written), so a ghost point's
    // stage position needs no test
    auto load_plane = [&](int q, In &in) {
        const int qc = min(max(q, qlo()), qhi());
        const int so = pl12 * (PERI ? zsrc(qc) : zp(qc));
        bXh.ld3(vl12, in.x, so);
        bXl.ld3(vl12, in.lo, so);
        bB0.ld3(vl12, in.b0, so);
#pragma unroll
        for (int m = 0; m < 5; ++m)
            if (m < n) bK[m].ld3(vl12, in.k[m], so);   // xyz triples at 12 a
    };

    // ---- march.  Plane p+1+kPF is loaded at step p into one of kPF input
    // slots and turned into a stage position kPF steps later; with two
    // slots the loop is unrolled by two so that each step names its slot
    // statically (no register indexing)
    auto plane_step = [&](const int p, In &pf) {
        float drn[3][3];   // this plane's differences and det, set by (2)
        double dqn;
        // (1) stage position of plane p+1 from the inputs loaded kPF
        // iterations ago.  Every lane loads every plane, so B0 and the final
        // stage's sums of plane p+1 are taken without a branch (lanes and
        // planes outside the range use none of them)
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            b0q[1][l] = pf.b0[l];
            if constexpr (FINAL) {
                float a1 = 0, a2 = 0;
#pragma unroll
                for (int m = 0; m < 5; ++m) {
                    a1 += (float)(c_tab32.c[m] - c_tab32.b[5 * 6 + m]) * pf.k[m][l];
                    a2 += (float)c_tab32.ce[m] * pf.k[m][l];
                }
                fg[2][l] = a1;
                fe[2][l] = a2;
            }
        }
        Pt3 xn;
        if (ldc && plane_ok(p + 1)) {
#pragma unroll
            for (int l = 0; l < 3; ++l) {
                float s = pf.lo[l];
#pragma unroll
                for (int m = 0; m < 5; ++m)
                    if (m < n) s += (float)c_tab32.b[n * 6 + m] * pf.k[m][l];
                xn.h[l] = pf.x[l];
                xn.s[l] = s;
            }
            if (vcol) {
                // outer ghost: the face point shifted by h along the face
                // normal, the shift in the small part so that the face
                // point's difference (h_in - h_face) + (s_in - s_face -+ h)
                // keeps the high words' exact difference
                if (vx) xn.s[0] += (float)vx * (float)g.dx;
                if (vy) xn.s[1] += (float)vy * (float)g.dy;
            }
            if constexpr (PERI) {
                // periodic images: x -+ (L + h) on the axis, the period split
                // into two floats and added by an exact two-sum to the high
                // word, its rounding error and the period's low part to the
                // small word
                const int sg[3] = {sgx, sgy, zshift(p + 1)};
                const double per[3] = {g.Lx + g.dx, g.Ly + g.dy, g.Lz + g.dz};
#pragma unroll
                for (int l = 0; l < 3; ++l) {
                    if (sg[l]) {
                        const float ph = (float)per[l], pl = (float)(per[l] - (double)ph);
                        float hs, e;
                        two_sum(xn.h[l], (float)sg[l] * ph, hs, e);
                        xn.h[l] = hs;
                        xn.s[l] += e + (float)sg[l] * pl;
                    }
                }
            }
            L.stx(p + 1, t, xn);
        }
        // (0) loads of plane p+1+kPF into the slot just read, in flight
        // while B and J are computed (and, with kPF = 2, while the next
        // plane step runs on the other slot)
        load_plane(p + 1 + kPF, pf);
        __builtin_amdgcn_sched_barrier(0);

        // (2) B at (column, p); ghost edges keep B0 and are never read by J
        if (inB && p >= Ks - 1 && p <= Ke + 1 && p >= 0 && p <= g.nz + 1 &&
            !((gx || gy) && (p == 0 || p == g.nz + 1))) {
            const bool zg = p == 0 || p == g.nz + 1;
            const bool zgl = zg && !pz;   // ghost plane of a line-tied z face
            float Bv[3];
            double detv;
            Pt3 xc, xzp, xzm;
            if (!zgl) {
                L.ldx(p, t, xc);
                xzp = xn;                 // own plane p+1, formed above
                L.ldx(p - 1, t, xzm);
            }
            if (!zgl && !(A.face_table && (lx || ly))) {
                // central differences of x along each axis, one neighbour
                // pair at a time; they give the Jacobian here and, kept for
                // one plane, the current at this point in the next iteration.
                // On a line-tied x or y face the neighbour beyond the face
                // is the outer ghost x_g -+ h n (a virtual column), the
                // tangential ones the ghost layer itself, and B . n = B0 . n
                const float hx = 0.5f * (float)g.dx1, hy = 0.5f * (float)g.dy1, hz = 0.5f * (float)g.dz1;
                {
                    Pt3 a, b;
                    L.ldx(p, t + 1, a);
                    L.ldx(p, t - 1, b);
#pragma unroll
                    for (int l = 0; l < 3; ++l) drn[0][l] = dif(a, b, l);
                }
                __builtin_amdgcn_sched_barrier(0);
                {
                    Pt3 a, b;
                    L.ldx(p, t + RW, a);
                    L.ldx(p, t - RW, b);
#pragma unroll
                    for (int l = 0; l < 3; ++l) drn[1][l] = dif(a, b, l);
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int l = 0; l < 3; ++l) drn[2][l] = zs * dif(xzp, xzm, l);
                // jac = dr diag(hx, hy, hz): det(jac) = det(dr) hx hy hz
                // (det(dr) from its rows, the transpose of jac's columns),
                // and jac B0 = dr (h * B0), so the spacing scales the three
                // components of B0 instead of the nine of dr
                detv = det32(drn) * ((double)hx * (double)hy * (double)hz);
                const float rd = __builtin_amdgcn_rcpf((float)detv);
                const float bs[3] = {b0q[0][0] * hx, b0q[0][1] * hy, b0q[0][2] * hz};
#pragma unroll
                for (int l = 0; l < 3; ++l)
                    Bv[l] = (drn[0][l] * bs[0] + drn[1][l] * bs[1] + drn[2][l] * bs[2]) * rd;
                if (lx) Bv[0] = b0q[0][0];
                if (ly) Bv[1] = b0q[0][1];
            } else {
                // z faces (and every face of a grid with an initial boundary
                // distortion): precomputed face constants
                int face, side, u, v;
                double xin[3];
                if (zgl) {
                    face = 2;
                    side = zp(p) != 0;
                    u = ii;
                    v = jj;
                    Pt3 pin;
                    if (p != 0) L.ldx(p - 1, t, pin);   // the inner neighbour in march order
                    else pin = xn;
#pragma unroll
                    for (int l = 0; l < 3; ++l) xin[l] = (double)pin.h[l] + (double)pin.s[l];
                } else {
                    face = lx ? 0 : 1;
                    side = lx ? (gi != 0) : (gj != 0);
                    u = lx ? jj : ii;
                    v = zp(p) - 1;
                    const int st = face == 0 ? 1 : RW;
                    Pt3 pin;
                    L.ldx(p, side ? t - st : t + st, pin);
#pragma unroll
                    for (int l = 0; l < 3; ++l) xin[l] = (double)pin.h[l] + (double)pin.s[l];
                }
                const GhostPt &G = A.ghost[A.ghost_off[face * 2 + side] + u + (long long)nface(g, face) * v];
                double Bd[3];
                ghost_b(g, G, face, side, xin, Bd, detv);
#pragma unroll
                for (int l = 0; l < 3; ++l) Bv[l] = (float)Bd[l];
                if constexpr (kDrLds) {
                    // x and y faces from the table: their y and z differences
                    // still feed the mixed terms of the tile points beside them
                    if (!zgl) {
                        Pt3 a, b;
                        L.ldx(p, t + RW, a);
                        L.ldx(p, t - RW, b);
#pragma unroll
                        for (int l = 0; l < 3; ++l) {
                            drn[1][l] = dif(a, b, l);
                            drn[2][l] = zs * dif(xzp, xzm, l);
                        }
                    }
                }
            }
            if constexpr (kDrLds) {
                if (!zgl) {
                    fqa[(p & 1) * NT + t] = make_float4(drn[1][0], drn[1][1], drn[1][2], drn[2][0]);
                    fqb[(p & 1) * NT + t] = make_float2(drn[2][1], drn[2][2]);
                }
            }
            if (EMIT && A.emit && (gx || gy || zg)) {
                int ci = min(max(gi, 1), g.nx) - 1, cj = min(max(gj, 1), g.ny) - 1, ck = min(max(p, 1), g.nz);
                if (ci >= ix0 && ci < ixe && cj >= iy0 && cj < iye && ck >= Ks && ck <= Ke) {
                    long long a = abase + g.sxy * (long long)zp(p);
#pragma unroll
                    for (int l = 0; l < 3; ++l) A.eB32[l * P + a] = Bv[l];
                    A.edet[a] = detv;
                }
            }
            L.stb(p, t, Bv);
            dqn = detv;
        }

        __builtin_amdgcn_sched_barrier(0);
        // (3) J, force and slope at (column, p-1): in-plane data of plane p-1
        // was published before the previous barrier, B(p) is this thread's own
        const int pj = p - 1;
        // this step's stores are issued by every lane: lanes without an
        // output point store at byte offset 2^31, past every array, where
        // the buffer store is dropped
        int st_off = kNoStore;
        float kst[3] = {0.f, 0.f, 0.f}, hn[3], ln[3];
        (void)kst; (void)hn; (void)ln;
        if (inJ && pj >= Ks && pj <= Ke) {
            const long long a = abase + g.sxy * (long long)zp(pj);
            st_off = vs12 + pl12 * zp(pj);
            // Stokes current from central differences only.  For the
            // diamond through the face neighbours v0..v3 of the reference
            // (src/current.cu:291-382) the trapezoid circulation is
            // (1/2)(db02 . dr13 - db13 . dr02) and the mean area vector
            // (1/2) dr02 x dr13, with d.. = v(+) - v(-) along an axis.  With
            // the diamonds xy, yz, zx the 3x3 system has the cofactor
            // vectors of [dx dy dz] as rows, so its solution is
            //   J = (C_x x db_x + C_y x db_y + C_z x db_z) / det[dx dy dz],
            // C_x = dy x dz, C_y = dz x dx, C_z = dx x dy -- identical to the
            // reference's circulation over area in exact arithmetic, without
            // reading the six neighbour positions again.
            float Bc[3], db[3][3];
            L.ldb(pj, t, Bc);
            {
                float bp[3], bm[3];
                L.ldb(pj, t + 1, bp);
                L.ldb(pj, t - 1, bm);
#pragma unroll
                for (int l = 0; l < 3; ++l) db[0][l] = bp[l] - bm[l];
                L.ldb(pj, t + RW, bp);
                L.ldb(pj, t - RW, bm);
#pragma unroll
                for (int l = 0; l < 3; ++l) db[1][l] = bp[l] - bm[l];
                L.ldb(pj + 1, t, bp);
                L.ldb(pj - 1, t, bm);
#pragma unroll
                for (int l = 0; l < 3; ++l) db[2][l] = zs * (bp[l] - bm[l]);
            }
            const float (&dr)[3][3] = drq0;
            // cofactor vectors C_a = dr[a+1] x dr[a+2]: the rows of
            // det(dr) dr^-1 (dr's columns are the three differences)
            float C[3][3];
#pragma unroll
            for (int a0 = 0; a0 < 3; ++a0) cross3(dr[(a0 + 1) % 3], dr[(a0 + 2) % 3], C[a0]);
            float Jv[3] = {0.f, 0.f, 0.f}, kv[3];
#pragma unroll
            for (int a0 = 0; a0 < 3; ++a0) {
                // Jv += C x db[a0], two fused products per component
                Jv[0] = fmaf(C[a0][1], db[a0][2], fmaf(-C[a0][2], db[a0][1], Jv[0]));
                Jv[1] = fmaf(C[a0][2], db[a0][0], fmaf(-C[a0][0], db[a0][2], Jv[1]));
                Jv[2] = fmaf(C[a0][0], db[a0][1], fmaf(-C[a0][1], db[a0][0], Jv[2]));
            }
            // det[dx dy dz] = 8 hx hy hz det(jac), det(jac) from FP64
            const float rD = __builtin_amdgcn_rcpf((float)(dq0 * (8.0 * g.dx * g.dy * g.dz)));
#pragma unroll
            for (int l = 0; l < 3; ++l) Jv[l] *= rD;
            kv[0] = Jv[1] * Bc[2] - Jv[2] * Bc[1];
            kv[1] = Jv[2] * Bc[0] - Jv[0] * Bc[2];
            kv[2] = Jv[0] * Bc[1] - Jv[1] * Bc[0];
            const int kk = zp(pj) - 1;
            float eps = 1.f;
            if (g.epsilonProf) eps = eps_xy * (float)A.epsz[kk];
            float gv[3] = {0, 0, 0};
            const double detv = dq0;
            if constexpr (PRESS) {
                // gradP of src/gradP.cu:93-154,
                //   g_c = -beta det1 sum_{a,b,l} d2[a][b][l] jac1[b][a] jac1[l][c]
                // (a a component, b and l axes, jac1 = jac^-1, det1 = 1/det(jac)),
                // with the spacings cancelled: jac = dr diag(1/2h), so
                // jac1 = diag(2h) C / det(dr), and the second derivatives are
                // the raw second differences over h_b^2 (diagonal) or 4 h_b h_l
                // (mixed).  Hence
                //   g = -beta det1 / det(dr)^2 sum_l N_l C_l,
                //   N_l = 4 (dd_l . C_l) + sum_{b != l} (dm_bl . C_b),
                // dd_l = x(+l) - 2x + x(-l) and dm_bl = dr_b(+l) - dr_b(-l):
                // 45 operations on top of the differences instead of about 190
                // for the inverse Jacobian and the 27-term sum
                float dd[3][3], mxy[3], mxz[3], myz[3];
                // diagonal: x(+a) - 2x + x(-a) = 2 (x(+a) - x) - dr_a along x and
                // y, with the point's own central differences dr: one
                // compensated difference per axis instead of two, and two
                // position reads fewer (257 x 257 x 769 with pressure: about
                // 3% faster).  Along z, whose march direction may reverse,
                // (x(+) - x) - (x - x(-)), which a reversal maps to itself bit
                // for bit (the form above would round differently)
                Pt3 xc;
                L.ldx(pj, t, xc);
#pragma unroll
                for (int a2 = 0; a2 < 2; ++a2) {
                    Pt3 xp;
                    L.ldx(pj, a2 == 0 ? t + 1 : t + RW, xp);
#pragma unroll
                    for (int l = 0; l < 3; ++l) dd[a2][l] = fmaf(2.f, dif(xp, xc, l), -dr[a2][l]);
                }
                {
                    Pt3 xm, xp;
                    L.ldx(pj - 1, t, xm);
                    L.ldx(pj + 1, t, xp);
#pragma unroll
                    for (int l = 0; l < 3; ++l) dd[2][l] = dif(xp, xc, l) - dif(xc, xm, l);
                }
                // mixed: (x(+a+b) - x(+a-b)) - (x(-a+b) - x(-a-b)) = dr_b(+a) -
                // dr_b(-a), from the ring; dr_z carries the march sign zs
                // already (exact: zs = -+1)
                {
                    const int sl = (pj & 1) * NT;
                    const float4 ea = fqa[sl + t + 1], wa = fqa[sl + t - 1], na = fqa[sl + t + RW], sa = fqa[sl + t - RW];
                    const float2 eb = fqb[sl + t + 1], wb = fqb[sl + t - 1], nb = fqb[sl + t + RW], sb = fqb[sl + t - RW];
                    const float dyE[3] = {ea.x, ea.y, ea.z}, dyW[3] = {wa.x, wa.y, wa.z};
                    const float dzE[3] = {ea.w, eb.x, eb.y}, dzW[3] = {wa.w, wb.x, wb.y};
                    const float dzN[3] = {na.w, nb.x, nb.y}, dzS[3] = {sa.w, sb.x, sb.y};
#pragma unroll
                    for (int l = 0; l < 3; ++l) {
                        mxy[l] = dyE[l] - dyW[l];
                        mxz[l] = dzE[l] - dzW[l];
                        myz[l] = dzN[l] - dzS[l];
                    }
                }
                auto dot = [](const float u[3], const float v[3]) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
                const float N0 = 4.f * dot(dd[0], C[0]) + dot(mxy, C[1]) + dot(mxz, C[2]);
                const float N1 = 4.f * dot(dd[1], C[1]) + dot(mxy, C[0]) + dot(myz, C[2]);
                const float N2 = 4.f * dot(dd[2], C[2]) + dot(mxz, C[0]) + dot(myz, C[1]);
                const float f = -(float)g.beta * __builtin_amdgcn_rcpf((float)detv) * (rD * rD);
#pragma unroll
                for (int l = 0; l < 3; ++l) gv[l] = (N0 * C[0][l] + N1 * C[1][l] + N2 * C[2][l]) * f;
            }
#pragma unroll
            for (int l = 0; l < 3; ++l) {
                kv[l] *= dt * eps;
                if (PRESS) kv[l] += -dt * gv[l];
            }
            if (EMIT && A.emit) {
#pragma unroll
                for (int l = 0; l < 3; ++l) {
                    A.eB32[l * P + a] = Bc[l];
                    A.eJ32[l * P + a] = Jv[l];
                    if (PRESS) A.eG32[l * P + a] = gv[l];
                }
                A.edet[a] = detv;
            }
            if (DIAG && A.diag) {
                // squared magnitudes in FP32, maxima kept squared and rooted
                // once per thread after the march, per-point quotients by
                // the hardware reciprocal; only the energy sum runs in FP64:
                // |B|^2 / 2^bound (an exact power-of-two scaling) times det
                float jxb[3];
                cross3(Jv, Bc, jxb);
                const float b2 = fmaf(Bc[0], Bc[0], fmaf(Bc[1], Bc[1], Bc[2] * Bc[2]));
                const float nc2 = fmaf(jxb[0], jxb[0], fmaf(jxb[1], jxb[1], jxb[2] * jxb[2]));
                const float nj2 = fmaf(Jv[0], Jv[0], fmaf(Jv[1], Jv[1], Jv[2] * Jv[2]));
                const float den2 = nj2 * b2;
                const int bound = (ii == 0) + (jj == 0) + (kk == 0) + (ii == g.nx - 1) + (jj == g.ny - 1) +
                                  (kk == g.nz - 1);
                const float wb = __builtin_bit_cast(float, (127 - bound) << 23);   // 2^-bound
                e_sum = fma((double)(b2 * wb), detv, e_sum);
                jmax2 = nb_maxf(jmax2, nj2);
                r1q = nb_maxf(r1q, nc2 * __builtin_amdgcn_rcpf(b2 * b2));
                r2q = nb_maxf(r2q, den2 > 0.f ? nc2 * __builtin_amdgcn_rcpf(den2) : 0.f);
                r3q = nb_maxf(r3q, nc2);
                r4q = nb_maxf(r4q, den2);
            }
            if constexpr (!FINAL) {
#pragma unroll
                for (int l = 0; l < 3; ++l) kst[l] = kv[l];
            } else {
                // x_new = x_5 + sum_m (c_m - b5_m) k_m (k_5 included), into
                // (hi, lo) by a two-sum; error |sum_m (c_m - c*_m) k_m| in FP64
                const int va = st_off;   // the point's byte offset, 12 a
                Pt3 xc;
                L.ldx(pj, t, xc);
                const float (&fg0)[3] = fg[0], (&fe0)[3] = fe[0];   // the sums of plane pj
#pragma unroll
                for (int l = 0; l < 3; ++l) {
                    float hi = xc.h[l], lo;
                    if (l < 2 || g.zUpdate) {
                        float d = xc.s[l] + (fg0[l] + (float)c_tab32.c[5] * kv[l]);
                        float s, e;
                        two_sum(hi, d, s, e);
                        float hi2 = s + e;
                        lo = e - (hi2 - s);
                        hi = hi2;
                        double dd = fabs((double)fe0[l] + c_tab32.ce[5] * (double)kv[l]);
                        if (dd != dd) nan_seen = 1;
                        dmax = nb_max(dmax, dd);
                    } else {
                        // z held: x_5's small part is the old lo word (no slope terms on z)
                        lo = bXl.ld(va + 4 * l, 0);
                    }
                    hn[l] = hi;
                    ln[l] = lo;
                }
            }
        }
        if constexpr (!FINAL) {
            if constexpr (stage >= 0) bK[stage].st3(kst, st_off);
        } else {
            bXhd.st3(hn, st_off);
            bXld.st3(ln, st_off);
        }

        // one barrier per plane
        __syncthreads();

#pragma unroll
        for (int l = 0; l < 3; ++l) b0q[0][l] = b0q[1][l];
        dq0 = dqn;
#pragma unroll
        for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
            for (int l = 0; l < 3; ++l) drq0[a2][l] = drn[a2][l];
        if constexpr (FINAL) {
#pragma unroll
            for (int l = 0; l < 3; ++l) {
                fg[0][l] = fg[1][l]; fg[1][l] = fg[2][l];
                fe[0][l] = fe[1][l]; fe[1][l] = fe[2][l];
            }
        }
    };
    const int p_first = Ks - 3 - kPF, p_last = Ke + 1;
    if constexpr (kPF == 1) {
        for (int p = p_first; p <= p_last; ++p) plane_step(p, pf0);
    } else {
        int p = p_first - ((p_first + 1) & 1);   // plane p+1 even: slot 0 first
        for (; p + 1 <= p_last; p += 2) {
            plane_step(p, pf0);
            plane_step(p + 1, pf1);
        }
        if (p <= p_last) plane_step(p, pf0);
    }

    double jmax = 0, r1 = 0, r2 = 0, r3 = 0, r4 = 0;
    if (DIAG && A.diag) {   // the maxima were kept squared
        jmax = sqrt((double)jmax2);
        r1 = sqrt((double)r1q);
        r2 = sqrt((double)r2q);
        r3 = sqrt((double)r3q);
        r4 = sqrt((double)r4q);
    }
    // ---- block partials (fixed order: deterministic)
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (FINAL) {
        dmax = wave_max(dmax);
        if (lane == 0) red[0][w] = dmax;
        if (nan_seen) atomicOr(&ctl->nan_seen, 1);
    }
    if (DIAG && A.diag) {
        e_sum = wave_sum(e_sum);
        jmax = wave_max(jmax);
        r1 = wave_max(r1);
        r2 = wave_max(r2);
        r3 = wave_max(r3);
        r4 = wave_max(r4);
        if (lane == 0) {
            red[1][w] = e_sum;
            red[2][w] = jmax;
            red[3][w] = r1;
            red[4][w] = r2;
            red[5][w] = r3;
            red[6][w] = r4;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (FINAL) {
            double m = 0;
            for (int q = 0; q < NT / 64; ++q) m = nb_max(m, red[0][q]);
            A.part_delta[lb] = m;
        }
        if (DIAG && A.diag) {
            double s = 0, am[5] = {0, 0, 0, 0, 0};
            for (int q = 0; q < NT / 64; ++q) {
                s += red[1][q];
                for (int m = 0; m < 5; ++m) am[m] = nb_max(am[m], red[m + 2][q]);
            }
            A.part_diag[lb] = s;
            for (int m = 0; m < 5; ++m) A.part_diag[(m + 1) * A.nblocks + lb] = am[m];
        }
    }
}

template <int NT, int STAGE, bool PRESS, bool ALT = false, bool PERI = false>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(waves32(NT, STAGE, PRESS), 8)))
k_stage32(FusedArgs A)
{
    stage32_body<NT, STAGE, PRESS, ALT, PERI>(A, xcd_remap(blockIdx.x, A.nblocks));
}

// Ensemble: one launch covers the workgroups of n instances; start[i] is
// the first (remapped) workgroup of instance i, start[n] the total.
template <int NT, int STAGE, bool PRESS, bool ALT = false>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(waves32(NT, STAGE, PRESS), 8)))
k_stage32_batch(const FusedArgs *__restrict__ args, const int *__restrict__ start, int n)
{
    const int L = xcd_remap(blockIdx.x, start[n]);
    int lo = 0, hi = n - 1;
    while (lo < hi) {   // last instance whose first workgroup is <= L
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= L) lo = mid;
        else hi = mid - 1;
    }
    stage32_body<NT, STAGE, PRESS, ALT, false>(args[lo], L - start[lo]);
}

namespace {
// chunk-alternating variants exist for stages 1-5 in 768- and 1024-thread
// workgroups; other launches ignore A.rev_alt
constexpr bool alt_built(int nt, int stage) { return nt >= 768 && stage >= 1 && stage <= 5; }
template <int NT, int STAGE, bool PRESS, bool ALT, bool PERI = false>
void launch_k32(hipStream_t s, const FusedArgs &A)
{
    static bool attr = false;
    constexpr size_t lds = lds_bytes32<NT, STAGE, PRESS>();
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)k_stage32<NT, STAGE, PRESS, ALT, PERI>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        attr = true;
    }
    hipLaunchKernelGGL((k_stage32<NT, STAGE, PRESS, ALT, PERI>), dim3(A.nblocks), dim3(NT), lds, s, A);
}
template <int NT, int STAGE, bool PRESS>
void launch_one32(hipStream_t s, const FusedArgs &A)
{
    if constexpr (alt_built(NT, STAGE)) {
        if (A.rev_alt) return launch_k32<NT, STAGE, PRESS, true>(s, A);
    }
    launch_k32<NT, STAGE, PRESS, false>(s, A);
}
template <int NT, bool PRESS>
void launch_stage32(hipStream_t s, const FusedArgs &A)
{
    switch (A.stage) {
    case 0: launch_one32<NT, 0, PRESS>(s, A); break;
    case 1: launch_one32<NT, 1, PRESS>(s, A); break;
    case 2: launch_one32<NT, 2, PRESS>(s, A); break;
    case 3: launch_one32<NT, 3, PRESS>(s, A); break;
    case 4: launch_one32<NT, 4, PRESS>(s, A); break;
    case 5: launch_one32<NT, 5, PRESS>(s, A); break;
    default: launch_one32<NT, 6, PRESS>(s, A); break;
    }
}
template <int NT>
void launch_nt32(hipStream_t s, const FusedArgs &A)
{
    if (A.g.pressure) launch_stage32<NT, true>(s, A);
    else launch_stage32<NT, false>(s, A);
}
// grids with periodic faces: 512-thread workgroups for every stage (the
// host keeps their rects within 512 points)
template <bool PRESS>
void launch_peri32(hipStream_t s, const FusedArgs &A)
{
    switch (A.stage) {
    case 0: launch_k32<512, 0, PRESS, false, true>(s, A); break;
    case 1: launch_k32<512, 1, PRESS, false, true>(s, A); break;
    case 2: launch_k32<512, 2, PRESS, false, true>(s, A); break;
    case 3: launch_k32<512, 3, PRESS, false, true>(s, A); break;
    case 4: launch_k32<512, 4, PRESS, false, true>(s, A); break;
    case 5: launch_k32<512, 5, PRESS, false, true>(s, A); break;
    default: launch_k32<512, 6, PRESS, false, true>(s, A); break;
    }
}
// 1024 threads: stages 0-4 and the fields pass without pressure
void launch_1024(hipStream_t s, const FusedArgs &A)
{
    switch (A.stage) {
    case 0: launch_one32<1024, 0, false>(s, A); break;
    case 1: launch_one32<1024, 1, false>(s, A); break;
    case 2: launch_one32<1024, 2, false>(s, A); break;
    case 3: launch_one32<1024, 3, false>(s, A); break;
    case 4: launch_one32<1024, 4, false>(s, A); break;
    default: launch_one32<1024, 6, false>(s, A); break;
    }
}
}  // namespace

int threads32(int rect) { return rect <= 256 ? 256 : rect <= 384 ? 384 : rect <= 512 ? 512 : rect <= 768 ? 768 : 1024; }

// 768-thread workgroups (12 waves, 168 VGPRs, three waves per SIMD): every
// stage kernel fits that budget without spilling.  The final stage without
// pressure: 266 instead of 272 us at 129 x 129 x 385 when it was set; the
// pressure stages 0-4 at 257 x 257 x 769: 4-11% faster than in 512-thread
// workgroups, the pressure final stage 2466 -> 1952 us
constexpr bool fits768(int, bool) { return true; }

int max_rect32(int stage, int press)
{
    const int s = stage < 0 ? 6 : stage;
    return fits1024(s, press != 0) ? 1024 : fits768(s, press != 0) ? 768 : 512;
}

int blocks_per_cu32(int rect, int stage, int press)
{
    const int nt = threads32(rect);
    if (nt >= 768) return 1;
    return nt < 512 || waves32(512, stage < 0 ? 6 : stage, press != 0) == 4 ? 2 : 1;
}

void set_tableau32(const TableauF &t) { (void)hipMemcpyToSymbol(HIP_SYMBOL(c_tab32), &t, sizeof t); }

// Periodic ghost layers of the current F32K positions (setGridPeriFace,
// src/boundaries.cu:532-582): the ghost layer takes the opposite inner
// layer, shifted by -+(L + h) along the axis, with (hi + lo) -+ (L + h)
// formed in FP64 and split again.  The stage kernels form these images
// themselves; the layers serve downloads and diagnostics.
__global__ void k_set_grid_peri32(Geo g, float *Xh0, float *Xl0, float *Xh1, float *Xl1, const Ctl *ctl, int face)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int na = face == 0 ? g.ny + 2 : g.nx + 2, nb = face == 2 ? g.ny + 2 : g.nz + 2;
    if (t >= (long long)na * nb) return;
    const int cur = ctl->cur;
    float *Xh = cur ? Xh1 : Xh0, *Xl = cur ? Xl1 : Xl0;
    const int u = (int)(t % na), v = (int)(t / na);
    long long lo, hi, srclo, srchi;
    double per;
    if (face == 2) {
        lo = gidx(g, u, v, 0); hi = gidx(g, u, v, g.nz + 1);
        srclo = gidx(g, u, v, g.nz); srchi = gidx(g, u, v, 1);
        per = g.Lz + g.dz;
    } else if (face == 1) {
        lo = gidx(g, u, 0, v); hi = gidx(g, u, g.ny + 1, v);
        srclo = gidx(g, u, g.ny, v); srchi = gidx(g, u, 1, v);
        per = g.Ly + g.dy;
    } else {
        lo = gidx(g, 0, u, v); hi = gidx(g, g.nx + 1, u, v);
        srclo = gidx(g, g.nx, u, v); srchi = gidx(g, 1, u, v);
        per = g.Lx + g.dx;
    }
    for (int l = 0; l < 3; ++l) {
        double vlo = (double)Xh[3 * srclo + l] + (double)Xl[3 * srclo + l];
        double vhi = (double)Xh[3 * srchi + l] + (double)Xl[3 * srchi + l];
        if (l == face) {
            vlo -= per;
            vhi += per;
        }
        const float hlo = (float)vlo, hhi = (float)vhi;
        Xh[3 * lo + l] = hlo;
        Xl[3 * lo + l] = (float)(vlo - (double)hlo);
        Xh[3 * hi + l] = hhi;
        Xl[3 * hi + l] = (float)(vhi - (double)hhi);
    }
}

void launch_set_grid_peri32(hipStream_t s, const Geo &g, float *const Xh[2], float *const Xl[2], const Ctl *ctl)
{
    // z, then y, then x, as the reference: edges and corners take the
    // composed images
    const int face_on[3] = {g.xPeri, g.yPeri, g.zPeri};
    for (int face = 2; face >= 0; --face) {
        if (!face_on[face]) continue;
        const long long n = face == 2 ? (long long)(g.nx + 2) * (g.ny + 2)
                          : face == 1 ? (long long)(g.nx + 2) * (g.nz + 2) : (long long)(g.ny + 2) * (g.nz + 2);
        hipLaunchKernelGGL(k_set_grid_peri32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, Xh[0], Xl[0],
                           Xh[1], Xl[1], ctl, face);
    }
}

namespace {
template <int NT, int STAGE, bool PRESS>
void launch_batch_one(hipStream_t s, const FusedArgs *args, const int *start, int n, int total)
{
    // stages 1-4 of an ensemble with odd z chunks marching the other way: a
    // member's workgroups are dispatched together, so its neighbouring
    // chunks mostly run side by side (64 braids: stages 2-4 13-20 us faster
    // per launch; the final stage 5 us slower, so it keeps the plain march)
    constexpr bool alt = alt_built(NT, STAGE) && STAGE <= 4;
    static bool attr = false;
    constexpr size_t lds = lds_bytes32<NT, STAGE, PRESS>();
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)k_stage32_batch<NT, STAGE, PRESS, alt>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        attr = true;
    }
    hipLaunchKernelGGL((k_stage32_batch<NT, STAGE, PRESS, alt>), dim3(total), dim3(NT), lds, s, args, start, n);
}
// The workgroup sizes a stage kernel is built for: 512 always, 768 and
// 1024 where its registers fit (fits768, fits1024)
template <int STAGE, bool PRESS>
void launch_batch_nt(hipStream_t s, int nt, const FusedArgs *args, const int *start, int n, int total)
{
    if constexpr (fits1024(STAGE, PRESS))
        if (nt == 1024) return launch_batch_one<1024, STAGE, PRESS>(s, args, start, n, total);
    if constexpr (fits768(STAGE, PRESS))
        if (nt == 768) return launch_batch_one<768, STAGE, PRESS>(s, args, start, n, total);
    launch_batch_one<512, STAGE, PRESS>(s, args, start, n, total);
}
template <bool PRESS>
void launch_batch_stage(hipStream_t s, int stage, int nt, const FusedArgs *args, const int *start, int n,
                        int total)
{
    switch (stage) {
    case 0: launch_batch_nt<0, PRESS>(s, nt, args, start, n, total); break;
    case 1: launch_batch_nt<1, PRESS>(s, nt, args, start, n, total); break;
    case 2: launch_batch_nt<2, PRESS>(s, nt, args, start, n, total); break;
    case 3: launch_batch_nt<3, PRESS>(s, nt, args, start, n, total); break;
    case 4: launch_batch_nt<4, PRESS>(s, nt, args, start, n, total); break;
    default: launch_batch_nt<5, PRESS>(s, nt, args, start, n, total); break;
    }
}
}  // namespace

int batch_threads32(int rect) { return std::max(512, threads32(rect)); }

void launch_fused32_batch(hipStream_t s, int stage, int press, int nt, const FusedArgs *args, const int *start,
                          int n, int total)
{
    if (press) launch_batch_stage<true>(s, stage, nt, args, start, n, total);
    else launch_batch_stage<false>(s, stage, nt, args, start, n, total);
}

void launch_fused32(hipStream_t s, const FusedArgs &A)
{
    if (A.g.xPeri || A.g.yPeri || A.g.zPeri) {
        if (A.g.pressure) launch_peri32<true>(s, A);
        else launch_peri32<false>(s, A);
        return;
    }
    // the smallest workgroup that holds the rect (the runtime checked the
    // stage kernel's limit)
    switch (threads32((A.TX + 4) * (A.TY + 4))) {
    case 256: launch_nt32<256>(s, A); break;
    case 384: launch_nt32<384>(s, A); break;
    case 512: launch_nt32<512>(s, A); break;
    case 768: launch_nt32<768>(s, A); break;
    default: launch_1024(s, A); break;   // the runtime offers 1024 only where fits1024
    }
}

}  // namespace gmd
