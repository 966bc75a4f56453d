// pps_nn.hip -- BF16 MFMA convolutions of the learned preconditioner, their
// backward passes, the training loss, Adam, and the network driver.
//
// Convolution = implicit GEMM on v_mfma_f32_16x16x32_bf16:
//   A (16 x 32)  weights: row = output channel, k = (tap pair, input channel)
//   B (32 x 16)  inputs:  column = one voxel of a 16-voxel x-row
//   C (16 x 16)  lane l, register r -> channel 4*(l/16)+r of voxel l%16
// A lane holds row l%16 and the 8 consecutive k of block l/16, B likewise,
// so one 16-byte LDS read per lane feeds one MFMA: tap 2s + (l/16)/2,
// channels 8*((l/16)&1) .. +7 of voxel l%16.  14 steps cover 28 taps
// (the 28th has zero weights).  All 14 weight operands stay in registers.
//
// A block stages 18 x 18 voxel input planes (16 x 16 output tile + halo) in
// LDS and marches along z with a ring of 4 planes, one barrier per plane.
// Each of the 4 wavefronts computes 4 x-rows of 16 voxels per plane.
#include "pps_nn.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>

namespace pps {
namespace nn {

const LayerDef kNet[kLayers] = {
    {1, 16, 0, IN_FIELD, -1, -1, 1},   // 0  lift     r -> act0 (skip of level 0)
    {16, 16, 1, IN_POOL, 0, -1, 1},    // 1  b0       pool(act0)
    {16, 16, 1, IN_PLAIN, 1, -1, 1},   // 2  b1       -> act2 (skip of level 1)
    {16, 16, 2, IN_POOL, 2, -1, 1},    // 3  c0       pool(act2)
    {16, 16, 2, IN_PLAIN, 3, -1, 1},   // 4  c1
    {16, 16, 2, IN_PLAIN, 4, -1, 1},   // 5  c2
    {16, 16, 2, IN_PLAIN, 5, -1, 1},   // 6  c3
    {16, 16, 2, IN_PLAIN, 6, -1, 1},   // 7  c4
    {16, 16, 2, IN_PLAIN, 7, -1, 1},   // 8  c5
    {16, 16, 2, IN_PLAIN, 8, -1, 1},   // 9  c6
    {16, 16, 2, IN_PLAIN, 9, -1, 1},   // 10 c7
    {16, 16, 1, IN_UPADD, 10, 2, 1},   // 11 e0       up(act10) + act2
    {16, 16, 1, IN_PLAIN, 11, -1, 1},  // 12 e1
    {16, 1, 0, IN_UPADD, 12, 0, 0},    // 13 out      up(act12) + act0 -> z (linear)
};

// activation i is read through a 2x2x2 average pool by a later layer: its
// producer then also writes the pooled copy, which that layer reads as it is
static bool pooled_act(int i)
{
    for (int l = 0; l < kLayers; ++l)
        if (kNet[l].mode_in == IN_POOL && kNet[l].in_act == i) return true;
    return false;
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct DwArgs {
    int mode_in;      // the layer's input, as in the forward pass
    ActRef in, in2;
    FieldRef fin;
    int dz_mode;      // 0: dZ = dY masked by Y > 0; 1: dZ = dY; 2: dZ = one-channel FP32 field
    ActRef dy, y;
    FieldRef fdy;
    int B, D, H, W;   // extent of the layer's level
    float* part;      // per-block partial sums, kLayerParams each
};

constexpr int kTW = 16, kTHt = 16;          // output tile per plane
constexpr int kLW = kTW + 2, kLH = kTHt + 2;  // staged input plane with halo
// Staged plane of 16-byte chunks (8 channels): [channel half][y][x], the
// second half at kHalf.  A 16-byte LDS read is served 16 lanes per cycle, lanes
// 0-7 with 24-31 and 8-15 with 16-23 (and likewise in the upper half-wave); it
// is free of bank conflicts when those 16 lanes hit 16 different chunk
// addresses mod 16.  The two channel halves of one operand are read by lane
// blocks j and j+1, so kHalf is a multiple of 16 chunks.
constexpr int kHalf = (kLH * kLW + 15) / 16 * 16;
constexpr int kPlane = 2 * kHalf;
static_assert(kHalf % 16 == 0, "channel halves 16 chunks apart mod 16");
constexpr int kZPlane = kTHt * 2 * kTW;      // chunks of one staged gradient plane

// ------------------------------------------------------------ BF16 helpers

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// FP32 -> BF16, round to nearest even (v_cvt_pk_bf16_f32): one value in the low
// half, or two values packed (a low, b high)
__device__ __forceinline__ unsigned f2bf(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }
__device__ __forceinline__ unsigned pack2bf(float a, float b)
{
    const f32x2 v = {a, b};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
}
// value of the neighbouring lane (lane ^ 1) by a DPP quad permute
__device__ __forceinline__ float xor1(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float bf2f(unsigned h) { return __uint_as_float(h << 16); }
__device__ __forceinline__ void unpack8(const u32x4 c, float* f)
{
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        f[2 * i] = __uint_as_float(c[i] << 16);
        f[2 * i + 1] = __uint_as_float(c[i] & 0xffff0000u);
    }
}
__device__ __forceinline__ u32x4 pack8(const float* f)
{
    u32x4 c;
#pragma unroll
    for (int i = 0; i < 4; ++i) c[i] = pack2bf(f[2 * i], f[2 * i + 1]);
    return c;
}
__device__ __forceinline__ bool bf_pos(unsigned h) { return (h & 0x8000u) == 0 && (h & 0x7fffu) != 0; }

__device__ __forceinline__ long long vox(const ActRef& a, int b, int x, int y, int z)
{
    return ((((long long)b * a.D + z) * a.H + y) * a.W + x) * kC;
}

__device__ __forceinline__ float field_scale(const FieldRef& f)
{
    if (!f.scale_sum) return f.scale;
    const double ms = *f.scale_sum / f.scale_count;
    return ms > 0.0 ? (float)(1.0 / sqrt(ms)) : 1.0f;
}
// Raw buffer access.  A buffer load whose offset lies at or beyond the
// resource's num_records returns 0 and such a store is dropped, so voxels
// outside the domain (zero padding, lanes without an output) need no branch
// and no mask: their offset is kOOB, and a plane outside the domain gets
// num_records = 0.
constexpr unsigned kOOB = 0x80000000u;     // beyond every num_records used below
constexpr int kInRange = 0x7fffffff;        // num_records of a resource that is in range
// p and nrec must be uniform; they are read from the first lane so that the
// resource lives in scalar registers (a per-lane resource costs a loop per access)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, int nrec)
{
    const unsigned long long a = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    void* q = (void*)(((unsigned long long)hi << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(q, (short)0, __builtin_amdgcn_readfirstlane(nrec), 0x00020000);
}

__device__ __forceinline__ long long fidx(const FieldRef& f, int b, int x, int y, int z)
{
    return f.off + x * f.sx + y * f.sy + z * f.sz + b * f.sb;
}

// one 16-byte chunk (8 channels) of the convolution input at voxel (x,y,z)
// of a level with extent D x H x W; zero outside (Dirichlet padding)
template <int MI>
__device__ __forceinline__ u32x4 load_chunk(const ActRef& in, const ActRef& in2, const FieldRef& f, float fscale,
                                            int D, int H, int W, int b, int x, int y, int z, int ch)
{
    u32x4 r = {0u, 0u, 0u, 0u};
    if (x < 0 || y < 0 || z < 0 || x >= W || y >= H || z >= D) return r;
    if constexpr (MI == IN_PLAIN) {
        return *(const u32x4*)(in.p + vox(in, b, x, y, z) + ch * 8);
    } else if constexpr (MI == IN_MASK) {
        const long long o = vox(in, b, x, y, z) + ch * 8;
        const u32x4 v = *(const u32x4*)(in.p + o);
        const u32x4 m = *(const u32x4*)(in2.p + o);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            r[i] = (bf_pos(m[i] & 0xffffu) ? (v[i] & 0xffffu) : 0u) | (bf_pos(m[i] >> 16) ? (v[i] & 0xffff0000u) : 0u);
        return r;
    } else if constexpr (MI == IN_POOL) {
        float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int fx = 2 * x + (c & 1), fy = 2 * y + ((c >> 1) & 1), fz = 2 * z + (c >> 2);
            if (fx < in.W && fy < in.H && fz < in.D) {
                float t[8];
                unpack8(*(const u32x4*)(in.p + vox(in, b, fx, fy, fz) + ch * 8), t);
#pragma unroll
                for (int e = 0; e < 8; ++e) acc[e] += t[e];
            }
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[e] *= 0.125f;
        return pack8(acc);
    } else if constexpr (MI == IN_UPADD) {
        float a[8], s[8];
        unpack8(*(const u32x4*)(in.p + vox(in, b, x >> 1, y >> 1, z >> 1) + ch * 8), a);
        unpack8(*(const u32x4*)(in2.p + vox(in2, b, x, y, z) + ch * 8), s);
#pragma unroll
        for (int e = 0; e < 8; ++e) a[e] += s[e];
        return pack8(a);
    } else {  // IN_FIELD: one channel
        if (ch) return r;
        const long long i = fidx(f, b, x, y, z);
        const float v = f.is64 ? (float)(((const double*)f.p)[i] * (double)fscale) : ((const float*)f.p)[i] * fscale;
        r[0] = f2bf(v);
        return r;
    }
}

// ------------------------------------------------------------ convolution
//
// Tap pairs.  The 9 (dz, dx) combinations c = 3*dz + dx are paired as
// (0,1) (2,3) (4,5) (6,7) (8,-): one MFMA of pair p at dy covers taps
// (dz, dy, dx) of both members, its B operand is the input row y + dy at the
// members' (dz, dx) shifts.  That operand is the same for output row y at dy
// and row y - 1 at dy + 1, so each LDS read feeds up to three MFMAs:
// a wavefront computing 4 output rows reads 6 input rows x 5 pairs = 30
// operands for 60 MFMAs.  Weight operand o = 3*p + dy.

__host__ __device__ constexpr int pair_combo(int p, int m) { return 2 * p + m; }  // < 9 valid

// Block -> tile.  The hardware deals consecutive workgroups round-robin to the
// 8 XCDs, each with its own L2; renumbered, each XCD owns a contiguous run of
// tiles (x fastest, then y, then z-chunk), so neighbouring tiles' halo reads
// and partial output lines meet in one L2 instead of going to HBM once per XCD.
struct TileIdx {
    int x, y, z;
};
__device__ __forceinline__ TileIdx tile_of_block()
{
    const int nx = gridDim.x, ny = gridDim.y;
    const int nb = nx * ny * gridDim.z;
    const int b = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
    const int full = nb & ~7;
    const int t = b >= full ? b : (b & 7) * (full >> 3) + (b >> 3);
    return {t % nx, (t / nx) % ny, t / (nx * ny)};
}

// raw staging values: global loads are issued one plane ahead into these and
// converted to BF16 chunks only when the plane is written to LDS
template <int MI> struct Raw { u32x4 a; };
template <> struct Raw<IN_MASK> { u32x4 a, m; };
template <> struct Raw<IN_UPADD> { u32x4 a, s; };
template <> struct Raw<IN_FIELD> { u32x2 a; };  // raw bits: FP64, or FP32 in .x

template <int MI>
__device__ __forceinline__ Raw<MI> fetch(const ActRef& in, const ActRef& in2, const FieldRef& f, float fscale, int D,
                                         int H, int W, int b, int x, int y, int z, int ch)
{
    Raw<MI> r{};
    if (x < 0 || y < 0 || z < 0 || x >= W || y >= H || z >= D) return r;
    if constexpr (MI == IN_PLAIN) {
        r.a = *(const u32x4*)(in.p + vox(in, b, x, y, z) + ch * 8);
    } else if constexpr (MI == IN_MASK) {
        const long long o = vox(in, b, x, y, z) + ch * 8;
        r.a = *(const u32x4*)(in.p + o);
        r.m = *(const u32x4*)(in2.p + o);
    } else if constexpr (MI == IN_UPADD) {
        r.a = *(const u32x4*)(in.p + vox(in, b, x >> 1, y >> 1, z >> 1) + ch * 8);
        r.s = *(const u32x4*)(in2.p + vox(in2, b, x, y, z) + ch * 8);
    } else if constexpr (MI == IN_POOL) {
        r.a = load_chunk<IN_POOL>(in, in2, f, fscale, D, H, W, b, x, y, z, ch);
    } else {
        if (ch == 0) {
            const long long i = fidx(f, b, x, y, z);
            if (f.is64) r.a = ((const u32x2*)f.p)[i];
            else r.a.x = ((const unsigned*)f.p)[i];
        }
    }
    return r;
}

// fetch without bounds checks: the caller passes a voxel inside the level
// (clamped) and masks the result when it is outside
template <int MI>
__device__ __forceinline__ Raw<MI> fetch_raw(const ActRef& in, const ActRef& in2, const FieldRef& f, float fscale, int D,
                                             int H, int W, int b, int x, int y, int z, int ch)
{
    Raw<MI> r{};
    if constexpr (MI == IN_PLAIN) {
        r.a = *(const u32x4*)(in.p + vox(in, b, x, y, z) + ch * 8);
    } else if constexpr (MI == IN_MASK) {
        const long long o = vox(in, b, x, y, z) + ch * 8;
        r.a = *(const u32x4*)(in.p + o);
        r.m = *(const u32x4*)(in2.p + o);
    } else if constexpr (MI == IN_UPADD) {
        r.a = *(const u32x4*)(in.p + vox(in, b, x >> 1, y >> 1, z >> 1) + ch * 8);
        r.s = *(const u32x4*)(in2.p + vox(in2, b, x, y, z) + ch * 8);
    } else if constexpr (MI == IN_POOL) {
        r.a = load_chunk<IN_POOL>(in, in2, f, fscale, D, H, W, b, x, y, z, ch);
    } else {
        const long long i = fidx(f, b, x, y, z);
        if (f.is64) r.a = ((const u32x2*)f.p)[i];
        else r.a.x = ((const unsigned*)f.p)[i];
    }
    return r;
}

template <int MI>
__device__ __forceinline__ u32x4 finish(const Raw<MI>& r, const FieldRef& f, float fscale)
{
    if constexpr (MI == IN_PLAIN || MI == IN_POOL) {
        return r.a;
    } else if constexpr (MI == IN_MASK) {
        u32x4 o;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            o[i] = (bf_pos(r.m[i] & 0xffffu) ? (r.a[i] & 0xffffu) : 0u) | (bf_pos(r.m[i] >> 16) ? (r.a[i] & 0xffff0000u) : 0u);
        return o;
    } else if constexpr (MI == IN_UPADD) {
        float a[8], s[8];
        unpack8(r.a, a);
        unpack8(r.s, s);
#pragma unroll
        for (int e = 0; e < 8; ++e) a[e] += s[e];
        return pack8(a);
    } else {
        u32x4 o = {0u, 0u, 0u, 0u};
        const float v = f.is64 ? (float)(__builtin_bit_cast(double, r.a) * (double)fscale) : __uint_as_float(r.a.x) * fscale;
        o[0] = f2bf(v);
        return o;
    }
}

template <int MO>
__device__ __forceinline__ void store_out(const ConvArgs& a, const f32x4& acc, const float* bias, int b, int x, int y,
                                          int z, int j)
{
    if constexpr (MO == OUT_FIELD) {
        if (j == 0) {
            const long long o = fidx(a.fout, b, x, y, z);
            const float v = acc[0] + bias[0];
            if (a.fout.is64) ((double*)a.fout.p)[o] = (double)v;
            else ((float*)a.fout.p)[o] = v;
        }
    } else {
        float v[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            v[r] = acc[r] + bias[r];
            if (MO == OUT_RELU) v[r] = fmaxf(v[r], 0.0f);
        }
        u32x2 o;
        o[0] = pack2bf(v[0], v[1]);
        o[1] = pack2bf(v[2], v[3]);
        *(u32x2*)(a.out.p + vox(a.out, b, x, y, z) + 4 * j) = o;
    }
}

constexpr int kChunksPerThread = (2 * kLH * kLW + 255) / 256;  // staged chunks per thread and plane

// PO: also write the 2x2x2-pooled copy of the (RELU) output to a.pout
template <int MI, int MO, bool PO = false>
__global__ __launch_bounds__(256, 3) void k_conv(ConvArgs a, int zc)
{
    if ((a.gate && *a.gate) || (a.gate2 && *a.gate2)) return;
    __shared__ u32x4 lds[4 * kPlane + 1];  // + a dummy slot for threads without a chunk
    const int D = a.out.D, H = a.out.H, W = a.out.W;
    const int nzc = (D + zc - 1) / zc;
    const TileIdx ti = tile_of_block();
    const int b = ti.z / nzc;
    const int z0 = (ti.z - b * nzc) * zc;
    const int z1 = min(z0 + zc, D);
    const int x0 = ti.x * kTW, y0 = ti.y * kTHt;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = lane >> 4, n = lane & 15;
    float fscale = 1.0f;
    if constexpr (MI == IN_FIELD) fscale = field_scale(a.fin);

    bf16x8 wr[kOps];
    const u32x4* wl = (const u32x4*)a.wl;
#pragma unroll
    for (int o = 0; o < kOps; ++o) wr[o] = __builtin_bit_cast(bf16x8, wl[o * 64 + lane]);
    float bias[4] = {0.f, 0.f, 0.f, 0.f};
    if (MO != OUT_LINEAR && a.bias) {
#pragma unroll
        for (int r = 0; r < 4; ++r) bias[r] = a.bias[4 * j + r];
    }

    // this thread's staging chunks: source voxel clamped into the level, LDS
    // slot (the dummy slot for threads without one) and whether the voxel is
    // inside the level in x and y (else the chunk is zero padding).  The loads
    // run without bounds branches, so they stay in flight until the plane is
    // written.  Consecutive threads load the two halves of one voxel record,
    // so a wavefront's loads are contiguous.
    constexpr int kStaged = 2 * kLH * kLW;  // chunks staged per plane
    int sx[kChunksPerThread], sy[kChunksPerThread], sc[kChunksPerThread], sl[kChunksPerThread];
    unsigned sval = 0u;
#pragma unroll
    for (int k = 0; k < kChunksPerThread; ++k) {
        const int t = tid + 256 * k;
        const int yy = t / (2 * kLW);
        const int rem = t - yy * (2 * kLW);
        const int xs = x0 - 1 + (rem >> 1), ys = y0 - 1 + yy;
        sx[k] = min(max(xs, 0), W - 1);
        sy[k] = min(max(ys, 0), H - 1);
        sc[k] = rem & 1;
        sl[k] = t < kStaged ? (rem & 1) * kHalf + yy * kLW + (rem >> 1) : 4 * kPlane;
        if (t < kStaged && xs >= 0 && ys >= 0 && xs < W && ys < H && (MI != IN_FIELD || sc[k] == 0)) sval |= 1u << k;
    }
    // the inference input stages (plain, up + skip) load through buffer
    // offsets: a chunk outside the level, or a plane outside it, reads zeros
    constexpr bool kBuf = MI == IN_PLAIN || MI == IN_UPADD;
    unsigned boff[kChunksPerThread], bcoff[kChunksPerThread];
    if constexpr (kBuf) {
#pragma unroll
        for (int k = 0; k < kChunksPerThread; ++k) {
            const bool v = (sval >> k) & 1u;
            boff[k] = v ? (unsigned)(((sy[k] * W + sx[k]) * kC + sc[k] * 8) * 2) : kOOB;
            // UPADD: the coarse voxel under (x, y) in the level below
            bcoff[k] = v ? (unsigned)((((sy[k] >> 1) * a.in.W + (sx[k] >> 1)) * kC + sc[k] * 8) * 2) : kOOB;
        }
    }
    // planes in flight: preA / preB hold raw loads issued one or two steps earlier
    Raw<MI> preA[kChunksPerThread], preB[kChunksPerThread];
    auto fetch_plane = [&](Raw<MI>* pre, int zi) {
        if constexpr (kBuf) {
            const bool zin = zi >= 0 && zi < D;  // uniform
            const int zz = zin ? zi : 0, nrec = zin ? kInRange : 0;
            if constexpr (MI == IN_PLAIN) {
                const __amdgpu_buffer_rsrc_t r = buf_rsrc(a.in.p + vox(a.in, b, 0, 0, zz), nrec);
#pragma unroll
                for (int k = 0; k < kChunksPerThread; ++k) pre[k].a = __builtin_amdgcn_raw_buffer_load_b128(r, boff[k], 0, 0);
            } else {
                const __amdgpu_buffer_rsrc_t rc = buf_rsrc(a.in.p + vox(a.in, b, 0, 0, zz >> 1), nrec);
                const __amdgpu_buffer_rsrc_t rs = buf_rsrc(a.in2.p + vox(a.in2, b, 0, 0, zz), nrec);
#pragma unroll
                for (int k = 0; k < kChunksPerThread; ++k) {
                    pre[k].a = __builtin_amdgcn_raw_buffer_load_b128(rc, bcoff[k], 0, 0);
                    pre[k].s = __builtin_amdgcn_raw_buffer_load_b128(rs, boff[k], 0, 0);
                }
            }
        } else {
            const int zc_ = min(max(zi, 0), D - 1);
#pragma unroll
            for (int k = 0; k < kChunksPerThread; ++k)
                pre[k] = fetch_raw<MI>(a.in, a.in2, a.fin, fscale, D, H, W, b, sx[k], sy[k], zc_, sc[k]);
        }
    };
    auto store_plane = [&](const Raw<MI>* pre, int zi) {
        u32x4* dst = lds + (zi & 3) * kPlane;
        const bool zin = zi >= 0 && zi < D;
#pragma unroll
        for (int k = 0; k < kChunksPerThread; ++k) {
            const u32x4 v = finish<MI>(pre[k], a.fin, fscale);
            u32x4* const d = sl[k] < 4 * kPlane ? dst + sl[k] : lds + 4 * kPlane;
            if constexpr (kBuf) *d = v;  // zeros already where the padding is
            else *d = (zin && ((sval >> k) & 1u)) ? v : u32x4{0u, 0u, 0u, 0u};
        }
    };
    // lane geometry of the two pair members: member j>>1, channel half j&1
    const int m = j >> 1;
    fetch_plane(preA, z0 - 1);
    store_plane(preA, z0 - 1);
    fetch_plane(preA, z0);
    store_plane(preA, z0);
    // prefetch depth: two planes where the raw loads fit the register budget of
    // three blocks per CU, one plane for the two-source and pooled input stages
    constexpr int kDepth = (MI == IN_UPADD || MI == IN_MASK || MI == IN_POOL || PO) ? 1 : 2;
    fetch_plane(preA, z0 + 1);
    if constexpr (kDepth == 2) fetch_plane(preB, z0 + 2);
    const int x = x0 + n;
    float ps[PO ? 2 : 1][4] = {};  // pooled-copy sums over two planes
    // ReLU output stores through buffer offsets: this lane's 4 channels of
    // row i inside an output plane, and of its pooled voxel for row pair ip
    // (kOOB outside the level and for odd lanes: those stores are dropped)
    unsigned ooff[4], poff[2];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int y = y0 + wave * 4 + i;
        ooff[i] = (x < W && y < H) ? (unsigned)(((y * W + x) * kC + 4 * j) * 2) : kOOB;
    }
#pragma unroll
    for (int ip = 0; ip < 2; ++ip) {
        const int yf = y0 + wave * 4 + 2 * ip;
        poff[ip] = (PO && !(n & 1) && x < W && yf < H) ? (unsigned)((((yf >> 1) * a.pout.W + (x >> 1)) * kC + 4 * j) * 2)
                                                        : kOOB;
    }
    // one plane; the buffer just written is refilled kDepth planes ahead
    // (clamped beyond the chunk, unused there)
    auto step = [&](int z, Raw<MI>* buf) {
        store_plane(buf, z + 1);
        fetch_plane(buf, z + 1 + kDepth);
        __syncthreads();
        f32x4 acc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
        // B operands in (pair, input row) order, read kPf steps ahead of their MFMAs
        int pbase[5];
#pragma unroll
        for (int p = 0; p < 5; ++p) {
            // member geometry (the empty member of pair 4 reuses member 0's, its weights are zero)
            const int c = pair_combo(p, m) < 9 ? pair_combo(p, m) : pair_combo(p, 0);
            const int dz = c / 3, dx = c % 3;
            pbase[p] = ((z + dz - 1) & 3) * kPlane + (j & 1) * kHalf + n + dx + wave * 4 * kLW;
        }
        // two steps ahead where the registers allow (the pooled and up+skip
        // input stages hold more raw staging data; with the one-channel field
        // output they are left to the compiler's schedule)
        constexpr bool kPipe = !((MI == IN_POOL || MI == IN_UPADD) && MO == OUT_FIELD);
        constexpr int kSteps = 30, kPf = !kPipe ? 0 : (MI == IN_POOL || MI == IN_UPADD) ? 1 : 2;
        u32x4 bq[kPf + 1];
#pragma unroll
        for (int q = 0; q < kPf; ++q) bq[q] = lds[pbase[q / 6] + (q % 6) * kLW];
#pragma unroll
        for (int q = 0; q < kSteps; ++q) {
            if (q + kPf < kSteps)
                bq[(q + kPf) % (kPf + 1)] = lds[pbase[(q + kPf) / 6] + ((q + kPf) % 6) * kLW];
            const bf16x8 B = __builtin_bit_cast(bf16x8, bq[q % (kPf + 1)]);
            const int p = q / 6, t = q % 6;
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
                const int i = t - dy;
                if (i >= 0 && i < 4) acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[3 * p + dy], B, acc[i], 0, 0, 0);
            }
            if constexpr (kPipe) __builtin_amdgcn_sched_barrier(0);  // keep the reads kPf steps ahead
        }
        if constexpr (MO == OUT_RELU) {
            const __amdgpu_buffer_rsrc_t ro = buf_rsrc(a.out.p + vox(a.out, b, 0, 0, z), kInRange);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                u32x2 o;
                o[0] = pack2bf(fmaxf(acc[i][0] + bias[0], 0.0f), fmaxf(acc[i][1] + bias[1], 0.0f));
                o[1] = pack2bf(fmaxf(acc[i][2] + bias[2], 0.0f), fmaxf(acc[i][3] + bias[3], 0.0f));
                __builtin_amdgcn_raw_buffer_store_b64(o, ro, ooff[i], 0, 0);
                if constexpr (PO) {
                    // a voxel outside the level counts as zero in the average
                    // (with an odd extent its pair partner is inside)
                    const u32x2 om = ooff[i] != kOOB ? o : u32x2{0u, 0u};
                    ps[i >> 1][0] += bf2f(om[0] & 0xffffu);
                    ps[i >> 1][1] += __uint_as_float(om[0] & 0xffff0000u);
                    ps[i >> 1][2] += bf2f(om[1] & 0xffffu);
                    ps[i >> 1][3] += __uint_as_float(om[1] & 0xffff0000u);
                }
            }
            // pooled copy for a layer that reads this output through a 2x2x2
            // average: x pairs are neighbouring lanes, y pairs rows of this
            // wavefront, z pairs consecutive planes (chunks start on even planes)
            if (PO && ((z & 1) || z == D - 1)) {  // uniform
                const __amdgpu_buffer_rsrc_t rp = buf_rsrc(a.pout.p + vox(a.pout, b, 0, 0, z >> 1), kInRange);
#pragma unroll
                for (int ip = 0; ip < 2; ++ip) {
                    float v[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) v[r] = 0.125f * (ps[ip][r] + xor1(ps[ip][r]));
                    u32x2 o;
                    o[0] = pack2bf(v[0], v[1]);
                    o[1] = pack2bf(v[2], v[3]);
                    __builtin_amdgcn_raw_buffer_store_b64(o, rp, poff[ip], 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; ++r) ps[ip][r] = 0.0f;
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int y = y0 + wave * 4 + i;
                if (x < W && y < H) store_out<MO>(a, acc[i], bias, b, x, y, z, j);
            }
        }
    };
    if constexpr (kDepth == 1) {
        for (int z = z0; z < z1; ++z) step(z, preA);
    } else {
        for (int z = z0; z < z1; z += 2) {
            step(z, preA);
            if (z + 1 < z1) step(z + 1, preB);
        }
    }
}

// Lift operand from the quad ring (z-packed: quad = input planes s-1, s, s+1
// and a zero, at one x-y position): one K = 16 MFMA per dy.  Lane block
// j < 3 holds the quad at dx = j of its voxel, block 3 carries zero weights
// and reads block 2's quad (a broadcast).  The operand of dy for output row y
// is quad row y + dy, so each 8-byte row read serves up to three output rows.
// The two blocks of a half-wave read 17 consecutive quads between them (the
// common ones as broadcasts), so the reads are free of LDS bank conflicts.
// The chain of one row's three MFMAs uses this one instruction only: a
// 16x16x32 MFMA accumulated straight into a 16x16x16 one gave wrong results.
typedef short s16x4 __attribute__((ext_vector_type(4)));  // BF16 bits, K = 16 MFMA operand
__device__ __forceinline__ int lift_dx(int j) { return j < 3 ? j : 2; }

// The fused output kernel keeps two K = 32 MFMAs o = 0, 1 per row (three
// K = 16 ones cost it more MFMA time than the shared reads save): the
// operand of lane block j is the quad pair at (dy, dx) = kQuad[o][j] and
// (dy, dx + 1), one 2 x 8-byte read.  Together they cover the 9 (dy, dx)
// positions; the other slots carry zero weights (q_real).  With the ring row
// of 18 quads, the pairs of blocks 0/1 and of 2/3 lie 16 quads apart (mod 32)
// or coincide, which keeps the 8-byte reads of each half-wave on distinct LDS
// banks.
__constant__ signed char kQuad[2][4][2] = {{{0, 2}, {1, 0}, {1, 2}, {2, 0}}, {{0, 0}, {0, 0}, {2, 2}, {2, 2}}};
// whether quad h (0, 1) of MFMA o's pair for block j is a tap position
__host__ __device__ constexpr bool q_real(int o, int j, int h)
{
    return o == 0 ? !(h == 1 && (j == 0 || j == 2)) : (j == 0 || (j == 2 && h == 0));
}
// where the fused output kernel's operands sit in layer 0's packed weights
constexpr int kLiftPairOff = 1024;
static_assert(kLiftPairOff >= 768 && kLiftPairOff + 1024 <= kPackedHalves, "both lift layouts fit one layer's slot");

// packed BF16 helpers of the lift kernels
typedef short s16x2 __attribute__((ext_vector_type(2)));
// relu of two packed BF16 values: as int16 a BF16 with the sign bit set is negative
__device__ __forceinline__ unsigned relu_bf2(unsigned h)
{
    return __builtin_bit_cast(unsigned, __builtin_elementwise_max(__builtin_bit_cast(s16x2, h), s16x2{0, 0}));
}

// c + h / 8 for the low (lo8) or high (hi8) BF16 h of a packed pair, in one
// BF16 dot product: the products by 1/8 and 0 are exact, so the sum rounds
// once, as c + h / 8 in FP32 does
__device__ __forceinline__ float dot_lo8(unsigned h, float c)
{
    return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, h), __builtin_bit_cast(bf16x2, 0x00003e00u), c, false);
}
__device__ __forceinline__ float dot_hi8(unsigned h, float c)
{
    return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, h), __builtin_bit_cast(bf16x2, 0x3e000000u), c, false);
}

// Staging of the one-channel network input (an FP64 or FP32 field) into a
// BF16 LDS plane: a thread owns kPer slots of the staged region.  Slots
// outside the domain (and beyond the region) load 0 from an out-of-range
// buffer offset; the loads have no control flow around them and stay in
// flight until the plane is converted.
template <bool F64, int kPer, int kThreads = 256>
struct FieldStage {
    using Raw = std::conditional_t<F64, u32x2, unsigned>;
    const char* base;  // element (0, 0, 0) of sample b
    long long szb;     // plane stride in bytes
    unsigned soff[kPer];  // byte offset of slot k inside its plane, or kOOB
    float scale;
    // region origin (rx, ry), row width rw, slots per plane np
    __device__ void init(const FieldRef& f, int b, int rx, int ry, int rw, int np, int W, int H, int tid)
    {
        constexpr int esz = F64 ? 8 : 4;
        base = (const char*)f.p + (f.off + b * f.sb) * esz;
        szb = f.sz * esz;
        scale = field_scale(f);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int t = tid + kThreads * k;
            const int yy = ry + t / rw, xx = rx + t % rw;
            const bool in = t < np && xx >= 0 && yy >= 0 && xx < W && yy < H;
            soff[k] = in ? (unsigned)((xx * f.sx + yy * f.sy) * esz) : kOOB;
        }
    }
    // planes at or beyond zend (the domain's or the chunk's end) read zeros
    __device__ void fetch(Raw* pre, int zi, int zend) const
    {
        const bool zin = zi >= 0 && zi < zend;  // uniform
        const __amdgpu_buffer_rsrc_t r = buf_rsrc(base + (zin ? zi : 0) * szb, zin ? kInRange : 0);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if constexpr (F64) pre[k] = __builtin_amdgcn_raw_buffer_load_b64(r, soff[k], 0, 0);
            else pre[k] = __builtin_amdgcn_raw_buffer_load_b32(r, soff[k], 0, 0);
        }
    }
    // the plane at byte offset off from element (0, 0, 0), or zeros (!in)
    __device__ void fetch_at(Raw* pre, long long off, bool in) const
    {
        const __amdgpu_buffer_rsrc_t r = buf_rsrc(base + (in ? off : 0), in ? kInRange : 0);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if constexpr (F64) pre[k] = __builtin_amdgcn_raw_buffer_load_b64(r, soff[k], 0, 0);
            else pre[k] = __builtin_amdgcn_raw_buffer_load_b32(r, soff[k], 0, 0);
        }
    }
    // BF16 bits of slot k (0 outside the domain: the load returned 0)
    __device__ uint16_t bf(const Raw* pre, int k) const
    {
        float v;
        if constexpr (F64) v = (float)(__builtin_bit_cast(double, pre[k]) * (double)scale);
        else v = __uint_as_float(pre[k]) * scale;
        return (uint16_t)f2bf(v);
    }
};

// Lift (1 -> 16 channels), training form: writes the full-resolution output
// and its pooled copy (inference, which keeps only the pooled copy, runs the
// wide-tile k_liftw below).  The 27 taps of one channel as three K = 16 MFMAs
// per 16 voxels over the z-packed quad ring; lane (voxel n, block j)
// receives channels 4j .. 4j+3.  The accumulator starts at the bias, or at
// -1e30 outside the domain, so relu leaves zeros there.  Ring entry q mod 2
// holds the quads (plane q-1, q, q+1, 0) of output plane q.  Each staging
// thread keeps its slots' last two planes (BF16) in a register and writes the
// whole quad of entry p-1 once plane p arrives: one 8-byte store per slot and
// plane.  With one barrier per plane a wavefront ahead of the others writes
// entry z+1 while the others still read entry z: two entries keep them apart.
// Input stage as IN_FIELD (scaled to unit RMS).
// UPD: the input is the residual after the Krylov step's update, formed here
// from the previous residual (a.fin) and A p (a.fap): r = fin - alpha fap
// (fin while *a.skip), and written to a.fcur on this block's own tile.
template <bool F64, bool UPD>
__global__ __launch_bounds__(256, 4) void k_lift(ConvArgs a, int zc)
{
    if ((a.gate && *a.gate) || (a.gate2 && *a.gate2)) return;
    constexpr int kP = kLH * kLW;
    // quads per entry: the 324 positions and the slots of the staging threads
    // beyond the region (2 x 256), which store their zeros there instead of
    // branching
    constexpr int kE = 512;
    static_assert(kE >= kP, "ring entry holds the region and every staging slot");
    __shared__ __align__(16) u32x2 lq[2 * kE];
    const int D = a.out.D, H = a.out.H, W = a.out.W;
    const int nzc = (D + zc - 1) / zc;
    const TileIdx ti = tile_of_block();
    const int b = ti.z / nzc;
    const int z0 = (ti.z - b * nzc) * zc;
    const int z1 = min(z0 + zc, D);
    const int x0 = ti.x * kTW, y0 = ti.y * kTHt;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = lane >> 4, n = lane & 15;
    s16x4 wq[3];  // per dy
#pragma unroll
    for (int dy = 0; dy < 3; ++dy) wq[dy] = __builtin_bit_cast(s16x4, ((const u32x2*)a.wl)[64 * dy + lane]);
    const int x = x0 + n;
    f32x4 cinit[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int y = y0 + wave * 4 + i;
#pragma unroll
        for (int r = 0; r < 4; ++r) cinit[i][r] = (x < W && y < H) ? (a.bias ? a.bias[4 * j + r] : 0.0f) : -1e30f;
    }
    // this lane's quad in ring row 4 wave (the first of the 6 rows it reads)
    const int pq = wave * 4 * kLW + n + lift_dx(j);
    constexpr int kPer = (kP + 255) / 256;
    unsigned hist[kPer] = {};  // BF16 of the slot's planes p-2 (low) and p-1 (high)
    using T = std::conditional_t<F64, double, float>;
    using Raw = typename FieldStage<F64, kPer>::Raw;
    FieldStage<F64, kPer> fs, fa;
    fs.init(a.fin, b, x0 - 1, y0 - 1, kLW, kP, W, H, tid);
    // residual update: A p staged like the input, and the byte offsets of the
    // slots on this block's own tile (the others are kOOB: their store is dropped)
    T alpha = (T)0;
    bool skip = true;
    unsigned coff[kPer];
    const char* cbase = nullptr;
    if constexpr (UPD) {
        FieldRef fr = a.fin;
        fr.p = const_cast<void*>(a.fap);
        fa.init(fr, b, x0 - 1, y0 - 1, kLW, kP, W, H, tid);
        alpha = (T)*a.alpha;
        skip = *a.skip != 0;
        cbase = (const char*)a.fcur + (a.fin.off + b * a.fin.sb) * (F64 ? 8 : 4);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int t = tid + 256 * k;
            const int yy = t / kLW, xx = t - (t / kLW) * kLW;
            const bool own = t < kP && xx >= 1 && xx <= kTW && yy >= 1 && yy <= kTHt;
            coff[k] = own ? fs.soff[k] : kOOB;
        }
    }
    // planes in flight: buffer i holds the raw loads of a plane issued NB - i
    // steps ahead of its use (FP64 loads take twice the registers)
    constexpr int NB = F64 ? 2 : 3;
    Raw pre[NB][kPer], ap[NB][kPer];
    // the chunk reads planes z0-1 .. z1 (z1 as the halo of its last plane)
    const int zend = min(z1 + 1, D);
    auto fetch = [&](Raw* p, Raw* q, int zi) {
        fs.fetch(p, zi, zend);
        if constexpr (UPD) fa.fetch(q, zi, zend);
    };
    // plane zi completes the quads of entry zi-1
    // (rc: where its residual goes, with no records for a plane outside the chunk)
    auto store_plane = [&](const Raw* p, const Raw* q, int zi, __amdgpu_buffer_rsrc_t rc) {
        u32x2* const dq = lq + ((zi - 1) & 1) * kE;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            // slots beyond the region hold zeros (their offset is kOOB) and
            // land in the entry's tail
            const int t = tid + 256 * k;
            uint16_t h;
            if constexpr (UPD) {
                T r, v;
                if constexpr (F64) {
                    r = __builtin_bit_cast(double, p[k]);
                    v = __builtin_bit_cast(double, q[k]);
                } else {
                    r = __uint_as_float(p[k]);
                    v = __uint_as_float(q[k]);
                }
                const T rn = skip ? r : r - alpha * v;
                if constexpr (F64) {
                    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, rn), rc, coff[k], 0, 0);
                    h = (uint16_t)f2bf((float)(rn * (double)fs.scale));
                } else {
                    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(rn), rc, coff[k], 0, 0);
                    h = (uint16_t)f2bf(rn * fs.scale);
                }
            } else {
                h = fs.bf(p, k);
            }
            dq[t] = u32x2{hist[k], (unsigned)h};
            hist[k] = (hist[k] >> 16) | ((unsigned)h << 16);
        }
    };
    // the two planes before the chunk fill the slots' history (their quads
    // land in entries that are rewritten before they are read); all loads of
    // the prologue are issued before the first is used
    {
        Raw h0[kPer], h1[kPer], g0[kPer], g1[kPer];
        fetch(h0, g0, z0 - 1);
        fetch(h1, g1, z0);
#pragma unroll
        for (int i = 0; i < NB; ++i) fetch(pre[i], ap[i], z0 + 1 + i);
        store_plane(h0, g0, z0 - 1, buf_rsrc(cbase, 0));
        store_plane(h1, g1, z0, buf_rsrc(cbase + z0 * fs.szb, kInRange));
    }
    // pooled copy (2x2x2 average of the BF16 outputs) for the next level:
    // y pairs are neighbouring rows of this wave and z pairs consecutive planes
    // (chunks start on even planes), summed per lane (each output times 1/8);
    // x pairs are neighbouring lanes, added when the pooled voxel is written
    float ps[2][4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
    // byte offset of this lane's pooled voxel in a pooled plane (row pair ip),
    // kOOB for odd lanes and voxels outside the level: their store is dropped
    unsigned poff[2];
#pragma unroll
    for (int ip = 0; ip < 2; ++ip) {
        const int yf = y0 + wave * 4 + 2 * ip;
        poff[ip] = (!(n & 1) && x < W && yf < H) ? (unsigned)(2 * (((yf >> 1) * a.pout.W + (x >> 1)) * kC + 4 * j))
                                                  : kOOB;
    }
    // one plane: NB planes of loads in flight, each buffer refilled once
    // written.  The four rows have no branch between them, so their MFMA
    // chains overlap.
    // byte offset of the plane the next step fetches (plane z + 1 + NB)
    long long foff = (long long)(z0 + 1 + NB) * fs.szb;
    const long long back = (long long)NB * fs.szb;
    auto step = [&](int z, Raw* buf, Raw* abuf) {
        const bool own = z + 1 < z1;  // uniform: plane z + 1 is one of the chunk's
        store_plane(buf, abuf, z + 1, buf_rsrc(cbase + (foff - back), own ? kInRange : 0));
        const bool in = z + 1 + NB < zend;  // zeros beyond the chunk's last halo plane
        fs.fetch_at(buf, foff, in);
        if constexpr (UPD) fa.fetch_at(abuf, foff, in);
        foff += fs.szb;
        __syncthreads();
        const u32x2* const q = lq + (z & 1) * kE;  // entry of output plane z
        u32x2 qr[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) qr[r] = q[pq + r * kLW];
        f32x4 acc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = cinit[i];
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wq[dy], __builtin_bit_cast(s16x4, qr[i + dy]), acc[i], 0, 0, 0);
        u32x2 o[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            o[i][0] = relu_bf2(pack2bf(acc[i][0], acc[i][1]));
            o[i][1] = relu_bf2(pack2bf(acc[i][2], acc[i][3]));
            ps[i >> 1][0] = dot_lo8(o[i][0], ps[i >> 1][0]);
            ps[i >> 1][1] = dot_hi8(o[i][0], ps[i >> 1][1]);
            ps[i >> 1][2] = dot_lo8(o[i][1], ps[i >> 1][2]);
            ps[i >> 1][3] = dot_hi8(o[i][1], ps[i >> 1][3]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int y = y0 + wave * 4 + i;
            if (x < W && y < H) *(u32x2*)(a.out.p + vox(a.out, b, x, y, z) + 4 * j) = o[i];
        }
        if (a.pout.p && ((z & 1) || z == D - 1)) {  // uniform
            const __amdgpu_buffer_rsrc_t rp = buf_rsrc(a.pout.p + vox(a.pout, b, 0, 0, z >> 1), kInRange);
#pragma unroll
            for (int ip = 0; ip < 2; ++ip) {
                float v[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) v[r] = ps[ip][r] + xor1(ps[ip][r]);
                u32x2 ov;
                ov[0] = pack2bf(v[0], v[1]);
                ov[1] = pack2bf(v[2], v[3]);
                __builtin_amdgcn_raw_buffer_store_b64(ov, rp, poff[ip], 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r) ps[ip][r] = 0.0f;
            }
        }
    };
    // the loop body runs all NB steps: a conditional step inside it would
    // leave fewer loads behind each wait on the path that skips it, and the
    // waits of every step would drain the prefetch.  The rest of the chunk
    // follows the loop.
    int z = z0;
    for (; z + NB <= z1; z += NB) {
#pragma unroll
        for (int i = 0; i < NB; ++i) step(z + i, pre[i], ap[i]);
    }
#pragma unroll
    for (int i = 0; i < NB - 1; ++i)
        if (z + i < z1) step(z + i, pre[i], ap[i]);
}

// Wide-tile lift for inference, which keeps only the pooled copy: a block
// owns 64 x 16 output voxels per plane, its 8 wavefronts 4 rows x 2 segments
// of 16 voxels each (wavefront w: rows 4 (w & 3) .., segments at x offsets
// 32 (w >> 2) and 32 (w >> 2) + 16).  The staged rows are 66 values wide, so
// the input loads and the residual stores cover whole cache lines, and the
// halo adds 16 % to the loads (27 % with 16 x 16 tiles).  Operands, ring and
// residual update as in k_lift; the ring entry holds 18 x 66 quads.
constexpr int kWX = 64, kWY = 16, kWLW = kWX + 2, kWLH = kWY + 2, kWThreads = 512;

template <bool F64, bool UPD>
__global__ __launch_bounds__(kWThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_liftw(ConvArgs a, int zc)
{
    if ((a.gate && *a.gate) || (a.gate2 && *a.gate2)) return;
    constexpr int kP = kWLH * kWLW;
    constexpr int kPer = (kP + kWThreads - 1) / kWThreads;
    // ring entry: every slot of the staging threads (those beyond the region
    // store their zeros in the tail instead of branching)
    constexpr int kE = kPer * kWThreads;
    __shared__ __align__(16) u32x2 lq[2 * kE];
    const int D = a.out.D, H = a.out.H, W = a.out.W;
    const int nzc = (D + zc - 1) / zc;
    const TileIdx ti = tile_of_block();
    const int b = ti.z / nzc;
    const int z0 = (ti.z - b * nzc) * zc;
    const int z1 = min(z0 + zc, D);
    const int x0 = ti.x * kWX, y0 = ti.y * kWY;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = lane >> 4, n = lane & 15;
    const int wr = 4 * (wave & 3), sx = 32 * (wave >> 2);
    s16x4 wq[3];  // per dy
#pragma unroll
    for (int dy = 0; dy < 3; ++dy) wq[dy] = __builtin_bit_cast(s16x4, ((const u32x2*)a.wl)[64 * dy + lane]);
    // the accumulators start at the bias, or at -1e30 for voxels outside the
    // domain (relu leaves zeros there); vin bit 4 s + i: voxel of segment s,
    // row i inside
    f32x4 cb;
#pragma unroll
    for (int r = 0; r < 4; ++r) cb[r] = a.bias ? a.bias[4 * j + r] : 0.0f;
    unsigned vin = 0u;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (x0 + sx + 16 * s + n < W && y0 + wr + i < H) vin |= 1u << (4 * s + i);
    // this lane's quad in ring row wr for segment 0 (segment 1: + 16)
    const int pq = wr * kWLW + sx + n + lift_dx(j);
    unsigned hist[kPer] = {};  // BF16 of the slot's planes p-2 (low) and p-1 (high)
    using T = std::conditional_t<F64, double, float>;
    using Raw = typename FieldStage<F64, kPer, kWThreads>::Raw;
    FieldStage<F64, kPer, kWThreads> fs;
    fs.init(a.fin, b, x0 - 1, y0 - 1, kWLW, kP, W, H, tid);
    // residual update: A p staged like the input (same layout, so through the
    // same slot offsets), and the slots on this block's own tile (bit k of
    // ownm; the other slots' stores go to kOOB and are dropped)
    T alpha = (T)0;
    bool skip = true;
    unsigned ownm = 0u;
    const char* cbase = nullptr;
    const char* abase = nullptr;
    if constexpr (UPD) {
        abase = (const char*)a.fap + (a.fin.off + b * a.fin.sb) * (F64 ? 8 : 4);
        alpha = (T)*a.alpha;
        skip = *a.skip != 0;
        cbase = (const char*)a.fcur + (a.fin.off + b * a.fin.sb) * (F64 ? 8 : 4);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int t = tid + kWThreads * k;
            const int yy = t / kWLW, xx = t - (t / kWLW) * kWLW;
            if (t < kP && xx >= 1 && xx <= kWX && yy >= 1 && yy <= kWY) ownm |= 1u << k;
        }
    }
    auto fetch_ap = [&](Raw* q, long long off, bool in) {
        const __amdgpu_buffer_rsrc_t r = buf_rsrc(abase + (in ? off : 0), in ? kInRange : 0);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if constexpr (F64) q[k] = __builtin_amdgcn_raw_buffer_load_b64(r, fs.soff[k], 0, 0);
            else q[k] = __builtin_amdgcn_raw_buffer_load_b32(r, fs.soff[k], 0, 0);
        }
    };
    // planes of loads in flight: buffer i holds plane z + 1 + i at step z
    // (FP64 with the update: one, as two do not fit the registers)
    constexpr int NB = (F64 && UPD) ? 1 : 2;
    Raw pre[NB][kPer], ap[NB][kPer];
    // the chunk reads planes z0-1 .. z1 (z1 as the halo of its last plane)
    const int zend = min(z1 + 1, D);
    // plane zi completes the quads of entry zi-1 (rc: where its residual goes,
    // with no records for a plane outside the chunk)
    auto store_plane = [&](const Raw* p, const Raw* q, int zi, __amdgpu_buffer_rsrc_t rc) {
        u32x2* const dq = lq + ((zi - 1) & 1) * kE;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int t = tid + kWThreads * k;
            uint16_t h;
            if constexpr (UPD) {
                T r, v;
                if constexpr (F64) {
                    r = __builtin_bit_cast(double, p[k]);
                    v = __builtin_bit_cast(double, q[k]);
                } else {
                    r = __uint_as_float(p[k]);
                    v = __uint_as_float(q[k]);
                }
                const T rn = skip ? r : r - alpha * v;
                if constexpr (F64) {
                    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, rn), rc, ((ownm >> k) & 1u) ? fs.soff[k] : kOOB, 0, 0);
                    h = (uint16_t)f2bf((float)(rn * (double)fs.scale));
                } else {
                    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(rn), rc, ((ownm >> k) & 1u) ? fs.soff[k] : kOOB, 0, 0);
                    h = (uint16_t)f2bf(rn * fs.scale);
                }
            } else {
                h = fs.bf(p, k);
            }
            dq[t] = u32x2{hist[k], (unsigned)h};
            hist[k] = (hist[k] >> 16) | ((unsigned)h << 16);
        }
    };
    // prologue: the two planes before the chunk fill the slots' history; all
    // its loads are issued before the first is used
    {
        Raw h0[kPer], h1[kPer], g0[kPer], g1[kPer];
        fs.fetch(h0, z0 - 1, zend);
        fs.fetch(h1, z0, zend);
#pragma unroll
        for (int i = 0; i < NB; ++i) fs.fetch(pre[i], z0 + 1 + i, zend);
        if constexpr (UPD) {
            fetch_ap(g0, (long long)(z0 - 1) * fs.szb, z0 >= 1);
            fetch_ap(g1, (long long)z0 * fs.szb, true);
#pragma unroll
            for (int i = 0; i < NB; ++i) fetch_ap(ap[i], (long long)(z0 + 1 + i) * fs.szb, z0 + 1 + i < zend);
        }
        store_plane(h0, g0, z0 - 1, buf_rsrc(cbase, 0));
        store_plane(h1, g1, z0, buf_rsrc(cbase + z0 * fs.szb, kInRange));
    }
    // pooled copy (2x2x2 average of the BF16 outputs): per segment s and row
    // pair ip, this lane's sum of the outputs times 1/8 over the pair's rows
    // and the plane pair (chunks start on even planes); x pairs are
    // neighbouring lanes, added when the pooled voxel is written (kOOB: odd
    // lanes and voxels outside the level, whose stores are dropped)
    float ps[2][2][4];
    // pooled voxel of segment s, row pair ip: pbase + 256 s + prow ip (bytes);
    // it exists where the segment's voxel of row 2 ip does (vin)
    const unsigned pbase = (unsigned)(2 * ((((y0 + wr) >> 1) * a.pout.W + ((x0 + sx + n) >> 1)) * kC + 4 * j));
    const unsigned prow = (unsigned)(2 * a.pout.W * kC);
    long long foff = (long long)(z0 + 1 + NB) * fs.szb;  // byte offset of the plane the next step fetches
    const long long back = NB * fs.szb;
    // one plane.  FIRST: the first plane of a pair (the pooled sums start
    // over); EDGE: the block's tile reaches beyond the domain
    auto step = [&](int z, Raw* buf, Raw* abuf, auto first_c, auto edge_c) {
        constexpr bool FIRST = decltype(first_c)::value, EDGE = decltype(edge_c)::value;
        const bool own = z + 1 < z1;  // uniform: plane z + 1 is one of the chunk's
        store_plane(buf, abuf, z + 1, buf_rsrc(cbase + (foff - back), own ? kInRange : 0));
        const bool in = z + 1 + NB < zend;  // zeros beyond the chunk's last halo plane
        fs.fetch_at(buf, foff, in);
        if constexpr (UPD) fetch_ap(abuf, foff, in);
        foff += fs.szb;
        __syncthreads();
        const u32x2* const q = lq + (z & 1) * kE;  // entry of output plane z
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            u32x2 qr[6];
#pragma unroll
            for (int r = 0; r < 6; ++r) qr[r] = q[pq + 16 * s + r * kWLW];
            f32x4 acc[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if constexpr (EDGE) acc[i] = ((vin >> (4 * s + i)) & 1u) ? cb : f32x4{-1e30f, -1e30f, -1e30f, -1e30f};
                else acc[i] = cb;
            }
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wq[dy], __builtin_bit_cast(s16x4, qr[i + dy]), acc[i], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const unsigned o0 = relu_bf2(pack2bf(acc[i][0], acc[i][1]));
                const unsigned o1 = relu_bf2(pack2bf(acc[i][2], acc[i][3]));
                float* const p = ps[s][i >> 1];
                const bool fresh = FIRST && !(i & 1);  // compile-time
                p[0] = dot_lo8(o0, fresh ? 0.0f : p[0]);
                p[1] = dot_hi8(o0, fresh ? 0.0f : p[1]);
                p[2] = dot_lo8(o1, fresh ? 0.0f : p[2]);
                p[3] = dot_hi8(o1, fresh ? 0.0f : p[3]);
            }
            // one segment's operands and accumulators live at a time: its
            // sums are formed here (not sunk into the pooled store's branch)
#pragma unroll
            for (int ip = 0; ip < 2; ++ip)
#pragma unroll
                for (int c = 0; c < 4; ++c) asm volatile("" : "+v"(ps[s][ip][c]));
            __builtin_amdgcn_sched_barrier(0);
        }
        if (a.pout.p && ((z & 1) || z == D - 1)) {  // uniform
            const __amdgpu_buffer_rsrc_t rp = buf_rsrc(a.pout.p + vox(a.pout, b, 0, 0, z >> 1), kInRange);
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int ip = 0; ip < 2; ++ip) {
                    float v[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) v[r] = ps[s][ip][r] + xor1(ps[s][ip][r]);
                    u32x2 ov;
                    ov[0] = pack2bf(v[0], v[1]);
                    ov[1] = pack2bf(v[2], v[3]);
                    const bool pin = !(n & 1) && ((vin >> (4 * s + 2 * ip)) & 1u);
                    __builtin_amdgcn_raw_buffer_store_b64(ov, rp, pin ? pbase + 256u * s + prow * ip : kOOB, 0, 0);
                }
        }
    };
    // the loop body runs both steps of a plane pair (a conditional step inside
    // it would leave fewer loads behind the waits of the path that skips it,
    // and every wait would drain the prefetch); an odd last plane follows it
    auto run = [&](auto edge_c) {
        int z = z0;
        for (; z + 2 <= z1; z += 2) {
            step(z, pre[0], ap[0], std::true_type{}, edge_c);
            step(z + 1, pre[NB - 1], ap[NB - 1], std::false_type{}, edge_c);
        }
        if (z < z1) step(z, pre[0], ap[0], std::true_type{}, edge_c);
    };
    if (x0 + kWX > W || y0 + kWY > H) run(std::true_type{});
    else run(std::false_type{});
}

// z-chunk of the wide-tile lift: at least 2048 blocks (4 per CU), chunks of
// at least 16 planes, starting on even planes
static int pick_zc_w(int D, int H, int W, int B)
{
    const long long tiles = (long long)((W + kWX - 1) / kWX) * ((H + kWY - 1) / kWY) * B;
    int zc = D;
    while (zc > 16 && tiles * ((D + zc - 1) / zc) < 2048) zc = (zc + 1) / 2;
    if ((zc & 1) && zc < D) ++zc;
    return std::max(zc, 1);
}

// z-chunk: enough blocks to fill the chip, at most one chunk per 16 planes
static int pick_zc(int D, int H, int W, int B)
{
    const long long tiles = (long long)((W + kTW - 1) / kTW) * ((H + kTHt - 1) / kTHt) * B;
    int zc = D;
    while (zc > 16 && tiles * ((D + zc - 1) / zc) < 8192) zc = (zc + 1) / 2;
    if ((zc & 1) && zc < D) ++zc;  // chunks start on even planes (z pairs of the pooled lift output)
    return std::max(zc, 1);
}
// z-chunk of the forward convolutions.  Three blocks per CU are resident;
// when a level's tiles fit one such round but pick_zc's chunks make more
// blocks than that, the fewest chunks that fill one round run instead (fewer
// halo planes, no partly filled last round: 512^3's level-1 layers 5-8 %
// faster with 86-plane chunks than with 16).  Smaller launches keep pick_zc's
// chunks (at level 2 of 512^3, 512 blocks of 16 planes beat 704 of 12).
static int cu_count()
{
    static int n = 0;
    if (n == 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            n = v;
        else
            n = 256;
    }
    return n;
}
static int pick_zc_conv(int D, int H, int W, int B)
{
    int zc = pick_zc(D, H, W, B);
    const long long tiles = (long long)((W + kTW - 1) / kTW) * ((H + kTHt - 1) / kTHt) * B;
    const long long slots = 3LL * cu_count();
    if (tiles <= slots && tiles * ((D + zc - 1) / zc) > slots) {
        const long long nz = slots / tiles;
        zc = (int)((D + nz - 1) / nz);
        if ((zc & 1) && zc < D) ++zc;  // chunks start on even planes
    }
    return zc;
}
static dim3 conv_grid(int D, int H, int W, int B, int zc)
{
    return dim3((unsigned)((W + kTW - 1) / kTW), (unsigned)((H + kTHt - 1) / kTHt), (unsigned)(B * ((D + zc - 1) / zc)));
}

template <int MI>
static void conv_mo(const ConvArgs& a, hipStream_t s, dim3 grid, int zc)
{
    switch (a.mode_out) {
    case OUT_RELU:
        // pooled copies come from plain-input layers only (checked by Net::init)
        if constexpr (MI == IN_PLAIN) {
            if (a.pout.p) {
                hipLaunchKernelGGL((k_conv<MI, OUT_RELU, true>), grid, dim3(256), 0, s, a, zc);
                break;
            }
        }
        hipLaunchKernelGGL((k_conv<MI, OUT_RELU>), grid, dim3(256), 0, s, a, zc);
        break;
    case OUT_LINEAR: hipLaunchKernelGGL((k_conv<MI, OUT_LINEAR>), grid, dim3(256), 0, s, a, zc); break;
    default: hipLaunchKernelGGL((k_conv<MI, OUT_FIELD>), grid, dim3(256), 0, s, a, zc); break;
    }
}

void launch_conv(const ConvArgs& a, hipStream_t s)
{
    const int zc = pick_zc_conv(a.out.D, a.out.H, a.out.W, a.out.B);
    const dim3 grid = conv_grid(a.out.D, a.out.H, a.out.W, a.out.B, zc);
    switch (a.mode_in) {
    case IN_PLAIN: conv_mo<IN_PLAIN>(a, s, grid, zc); break;
    case IN_POOL: conv_mo<IN_POOL>(a, s, grid, zc); break;
    case IN_UPADD: conv_mo<IN_UPADD>(a, s, grid, zc); break;
    case IN_MASK: conv_mo<IN_MASK>(a, s, grid, zc); break;
    default: conv_mo<IN_FIELD>(a, s, grid, zc); break;
    }
}

void launch_lift(const ConvArgs& a, hipStream_t s)
{
    if (a.out.p) {  // training: the full-resolution output and its pooled copy
        const int zc = pick_zc(a.out.D, a.out.H, a.out.W, a.out.B);
        const dim3 grid = conv_grid(a.out.D, a.out.H, a.out.W, a.out.B, zc);
        if (a.fcur) {
            if (a.fin.is64) hipLaunchKernelGGL((k_lift<true, true>), grid, dim3(256), 0, s, a, zc);
            else hipLaunchKernelGGL((k_lift<false, true>), grid, dim3(256), 0, s, a, zc);
        } else {
            if (a.fin.is64) hipLaunchKernelGGL((k_lift<true, false>), grid, dim3(256), 0, s, a, zc);
            else hipLaunchKernelGGL((k_lift<false, false>), grid, dim3(256), 0, s, a, zc);
        }
        return;
    }
    // inference keeps only the pooled copy: the wide-tile kernel
    const int zw = pick_zc_w(a.out.D, a.out.H, a.out.W, a.out.B);
    const dim3 gw((unsigned)((a.out.W + kWX - 1) / kWX), (unsigned)((a.out.H + kWY - 1) / kWY),
                  (unsigned)(a.out.B * ((a.out.D + zw - 1) / zw)));
    if (a.fcur) {
        if (a.fin.is64) hipLaunchKernelGGL((k_liftw<true, true>), gw, dim3(kWThreads), 0, s, a, zw);
        else hipLaunchKernelGGL((k_liftw<false, true>), gw, dim3(kWThreads), 0, s, a, zw);
    } else {
        if (a.fin.is64) hipLaunchKernelGGL((k_liftw<true, false>), gw, dim3(kWThreads), 0, s, a, zw);
        else hipLaunchKernelGGL((k_liftw<false, false>), gw, dim3(kWThreads), 0, s, a, zw);
    }
}

// Inference form of the output layer with the lift fused in:
//   z = out(up(E1) + relu(lift(r)))
// so the 16-channel level-0 activation never goes to HBM.  A block owns a
// 14 x 18 output tile; its X region (tile + 1-voxel halo) is 16 x 20 voxels,
// one 16-voxel x-row per MFMA group, 5 rows per wavefront.  Per plane s and row:
//   1. lift: two K = 32 MFMAs over quad pairs of the input ring (64 slots
//      for the 27 taps, the rest with zero weights; accumulator started at
//      the bias, or at -1e30 outside the domain so that relu gives the zero
//      padding); lane (voxel n, block j) then holds channels
//      4j .. 4j+3; relu + BF16 in packed 16-bit form.  With e = up(E1) of the
//      same channels, (a, e) is lane (n, j)'s A operand X;
//   2. row partials P[dz,dy][n] = sum_dx sum_ci W[dz,dy,dx][ci] (a + e)[n+dx-1][ci]:
//      X shifted one voxel left and right inside the 16-lane row (DPP), three
//      MFMAs (one per dx) into columns m = 3 dz + dy; the skip sum a + e
//      happens inside the MFMA (the weight operand repeats W[.][4j .. 4j+3]
//      for both halves).  The 9 rows go to LDS as 16-byte writes of 4 voxels;
// in the next step (after its one barrier) every thread adds the 9 partials of
// its output voxel (rows y-1 .. y+1) into three running output planes; the
// gather of plane s completes output plane s-1.
// The output columns are 1 .. 14 of each row, so a shift never reaches past
// the row for a value that is used.  With a.dot set the kernel also sums
// z . dot over its outputs (FP64) and writes one partial per block to a.dpart
// (the Krylov step's gamma = z . Ap_old without another pass over z).  The unfused path rounds a + e to BF16
// before its MFMA; here the sum is exact, so outputs differ from it by that
// rounding (relative ~1e-3 at most) and FP32 order.
constexpr int kFW = 14, kFH = 18;                 // output tile
constexpr int kXW = kFW + 2, kXH = kFH + 2;       // X region: 16 x 20
constexpr int kFRW = kXW + 2, kFRH = kXH + 2;     // input ring plane: 18 x 22
constexpr int kFRP = kFRW * kFRH;
// quads per ring entry: the 396 positions, junk reads past the last row, and
// the slots of the staging threads beyond the region (2 x 256), which store
// their zeros there instead of branching
constexpr int kQE = 512;
static_assert(kQE >= kFRP + 4 && kQE >= 512, "ring entry holds the region, the junk reads and every staging slot");
constexpr int kXQ = kXW * kXH;
// P row stride: 4 mod 64 banks (measured: a stride of 8 mod 64 doubles the
// bank conflicts of the 16-byte partial stores)
constexpr int kXQp = kXQ + 4;
constexpr int kFG = kXH / 4;                      // x-rows (groups) per wavefront
static_assert(kXW == 16 && kXH % 4 == 0, "one 16-voxel x-row per group, equal rows per wavefront");

// (FP64 input: three blocks per CU, the raw FP64 prefetch does not fit the
// register budget of four).  O64: the output field (and the dot operand) is
// FP64; a template argument, so the plane loop holds one store and one
// conversion.
template <bool F64, bool O64>
__global__ __launch_bounds__(256, 4) void k_outf(ConvArgs a, int zc)
{
    if ((a.gate && *a.gate) || (a.gate2 && *a.gate2)) return;
    __shared__ __align__(16) u32x2 lq[2 * kQE];  // input quad ring: 2 entries
    // P[3 dz + dy][voxel of the X region], double-buffered by X plane parity
    // so that the partials of plane s and the gather of plane s - 1 share one
    // barrier; the tail pads the gathers of the four threads without an
    // output voxel
    constexpr int kLPS = 9 * kXQp + 32;
    __shared__ __align__(16) float lp[2 * kLPS];
    const int D = a.out.D, H = a.out.H, W = a.out.W;
    const TileIdx ti = tile_of_block();
    const int z0 = ti.z * zc;  // one sample
    const int z1 = min(z0 + zc, D);
    const int x0 = ti.x * kFW, y0 = ti.y * kFH;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = lane >> 4, n = lane & 15;
    // lift operands of the two quad-pair MFMAs
    const bf16x8 wqa = __builtin_bit_cast(bf16x8, ((const u32x4*)a.wl2)[lane]);
    const bf16x8 wqb = __builtin_bit_cast(bf16x8, ((const u32x4*)a.wl2)[64 + lane]);
    bf16x8 wp[3];  // per dx: column m = 3 dz + dy
#pragma unroll
    for (int dx = 0; dx < 3; ++dx) wp[dx] = __builtin_bit_cast(bf16x8, ((const u32x4*)a.wl)[64 * dx + lane]);
    const float bout = a.bias ? a.bias[0] : 0.0f;
    // X voxel of lane n: column gx; rows y0 - 1 + wave + 4k (k < kFG)
    const int gx = x0 - 1 + n;
    const bool xin = gx >= 0 && gx < W;
    unsigned yin = 0u;  // rows of this wavefront inside the domain
#pragma unroll
    for (int k = 0; k < kFG; ++k) {
        const int gy = y0 - 1 + wave + 4 * k;
        if (gy >= 0 && gy < H) yin |= 1u << k;
    }
    yin = __builtin_amdgcn_readfirstlane(yin);
    // lift accumulators start at the bias, or at -1e30 outside the domain
    // (relu then gives the zero padding)
    f32x4 cinit[kFG];
#pragma unroll
    for (int k = 0; k < kFG; ++k)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            cinit[k][r] = (xin && ((yin >> k) & 1u)) ? (a.bias2 ? a.bias2[4 * j + r] : 0.0f) : -1e30f;
    // coarse activation E1: byte offset of this lane's 4 channels in row k
    // (kOOB outside the domain: the load returns zeros)
    const int cW = a.in.W, cH = a.in.H;
    const long long cplane = (long long)cH * cW * kC;
    const int cy0 = (y0 - 1 + wave) >> 1;  // coarse row of k = 0 (rows k step by 2)
    unsigned eoff[kFG];
#pragma unroll
    for (int k = 0; k < kFG; ++k)
        eoff[k] = (xin && ((yin >> k) & 1u)) ? (unsigned)(2 * (((cy0 + 2 * k) * cW + (gx >> 1)) * kC + 4 * j)) : kOOB;
    // lift operand quads of this lane for X voxel (row wave, column n)
    const int pa = (kQuad[0][j][0] + wave) * kFRW + kQuad[0][j][1] + n;
    const int pb = (kQuad[1][j][0] + wave) * kFRW + kQuad[1][j][1] + n;
    // input ring over the region (x0 - 2 .., y0 - 2 ..), 18 x 22 positions:
    // entry q mod 2 holds the quads (plane q-1, q, q+1, 0) of X plane q.  Each
    // staging thread keeps its slots' last two planes (BF16) in a register and
    // writes the whole quad of entry p-1 once plane p arrives.
    constexpr int kPer = (kFRP + 255) / 256;
    unsigned hist[kPer] = {};  // BF16 of the slot's planes p-2 (low) and p-1 (high)
    FieldStage<F64, kPer> fs;
    fs.init(a.fin, 0, x0 - 2, y0 - 2, kFRW, kFRP, W, H, tid);
    using Raw = typename FieldStage<F64, kPer>::Raw;
    Raw preA[kPer], preB[kPer];
    // plane zi completes the quads of entry zi-1
    auto store_plane = [&](const Raw* pre, int zi) {
        u32x2* const dq = lq + ((zi - 1) & 1) * kQE;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int t = tid + 256 * k;
            // slots beyond the region hold zeros and land in the entry's tail
            const unsigned h = fs.bf(pre, k);
            dq[t] = u32x2{hist[k], h};
            hist[k] = (hist[k] >> 16) | (h << 16);
        }
    };
    // up(E1) of this wavefront's rows, loaded one plane ahead (zeros outside
    // the domain)
    u32x2 ev[kFG];
#pragma unroll
    for (int k = 0; k < kFG; ++k) ev[k] = u32x2{0u, 0u};
    auto fetch_e = [&](u32x2* e, int s) {
        const __amdgpu_buffer_rsrc_t r = buf_rsrc(a.in.p + (s >> 1) * cplane, kInRange);
#pragma unroll
        for (int k = 0; k < kFG; ++k) e[k] = __builtin_amdgcn_raw_buffer_load_b64(r, eoff[k], 0, 0);
    };
    // row k of X from the quad ring: lift (two MFMAs, relu to BF16) beside
    // up(E1) as the partials' operand
    auto lift_row = [&](const u32x2* q, const u32x2* e, int k) {
        const u32x2 qa0 = q[pa + 4 * k * kFRW], qa1 = q[pa + 4 * k * kFRW + 1];
        const u32x2 qb0 = q[pb + 4 * k * kFRW], qb1 = q[pb + 4 * k * kFRW + 1];
        const u32x4 ba = {qa0[0], qa0[1], qa1[0], qa1[1]}, bb = {qb0[0], qb0[1], qb1[0], qb1[1]};
        f32x4 c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wqa, __builtin_bit_cast(bf16x8, ba), cinit[k], 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wqb, __builtin_bit_cast(bf16x8, bb), c, 0, 0, 0);
        u32x4 bx;
        bx[0] = relu_bf2(pack2bf(c[0], c[1]));
        bx[1] = relu_bf2(pack2bf(c[2], c[3]));
        bx[2] = e[k][0];
        bx[3] = e[k][1];
        return bx;
    };
    // one plane of X -> tap partials, this wavefront's rows.  On FP32 fields
    // the lift of row k + 1 is issued before the partials of row k so that
    // the two chains overlap; the FP64 variant has no registers to spare for
    // that at 4 blocks per CU.
    auto partials = [&](const u32x2* e, int es) {
        const u32x2* const q = lq + es * kQE;  // entry of X plane s (es = s mod 2)
        u32x4 bx = lift_row(q, e, 0);
#pragma unroll
        for (int k = 0; k < kFG; ++k) {
            u32x4 bn;
            if (!F64 && k + 1 < kFG) bn = lift_row(q, e, k + 1);
            // the row's neighbours: voxel n-1 (row_shr:1) and n+1 (row_shl:1)
            u32x4 bm, bp;
#pragma unroll
            for (int e8 = 0; e8 < 4; ++e8) {
                bm[e8] = (unsigned)__builtin_amdgcn_mov_dpp((int)bx[e8], 0x111, 0xF, 0xF, true);
                bp[e8] = (unsigned)__builtin_amdgcn_mov_dpp((int)bx[e8], 0x101, 0xF, 0xF, true);
            }
            f32x4 pr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, bm), wp[0],
                                                               f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
            pr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, bx), wp[1], pr, 0, 0, 0);
            pr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, bp), wp[2], pr, 0, 0, 0);
            // lane (column m, j): voxels 4j .. 4j+3 of row wave + 4k; the MFMA
            // columns m >= 9 carry no partial
            if (n < 9) *(f32x4*)(lp + es * kLPS + n * kXQp + (wave + 4 * k) * kXW + 4 * j) = pr;
            if (k + 1 < kFG) bx = F64 ? lift_row(q, e, k + 1) : bn;
        }
    };
    // this thread's output voxel: byte offset inside an output plane, kOOB
    // for threads without one (their stores are dropped, dot loads give 0)
    const int oy = tid / kFW, ox = tid - oy * kFW;
    const int x = x0 + ox, y = y0 + oy;
    const bool mine = tid < kFW * kFH && x < W && y < H;
    const int q0 = oy * kXW + ox + 1;  // X region voxel of the output's (x, y - 1) neighbour
    float am1 = 0.0f, a0 = 0.0f, ap1 = 0.0f;
    constexpr bool o64 = O64;
    constexpr int oesz = o64 ? 8 : 4;
    const unsigned ooff = mine ? (unsigned)((x * a.fout.sx + y * a.fout.sy) * oesz) : kOOB;
    const char* obase = (const char*)a.fout.p + a.fout.off * oesz;
    const char* dbase = a.dot ? (const char*)a.dot + a.fout.off * oesz : nullptr;
    const long long oszb = a.fout.sz * oesz;
    // dot operands: loaded with the input prefetch two steps ahead of their
    // output plane's step (one buffer per step parity, like the prefetch:
    // step s loads the operand of output plane s) and converted where the step
    // waits for that prefetch anyway.  The load is one 8-byte load in every
    // case (an FP32 operand is its low word; a plane outside the chunk, or no
    // dot at all, gets num_records 0): a load under a branch would make the
    // compiler wait for the prefetch right after it.
    u32x2 dA = {0u, 0u}, dB = {0u, 0u};
    double dcur = 0.0, dacc = 0.0;

    // planes s = z0-1 .. z1 of X; the ring holds input planes s-1 .. s+1
    // planes z0 - 2 and z0 - 1 fill the slots' history (their quads land in
    // entries that are rewritten before they are read)
    fs.fetch(preA, z0 - 2, D);
    store_plane(preA, z0 - 2);
    fs.fetch(preA, z0 - 1, D);
    store_plane(preA, z0 - 1);
    fs.fetch(preA, z0, D);
    fs.fetch(preB, z0 + 1, D);
    if (z0 >= 1) fetch_e(ev, z0 - 1);  // uniform
    // step s: stage input plane s+1, partials of X plane s, gather of X plane
    // s - 1 and output plane s - 2.  Two planes of input loads in flight,
    // each buffer refilled once written; es = s mod 2: the ring entry and the
    // partials buffer of X plane s.  One barrier per step: the ring entry of
    // plane s is staged; the partials of plane s - 1 are complete; the gather
    // of plane s - 2, which read the buffer that plane s fills, is done.
    auto step = [&](int s, int es, Raw* buf, u32x2& dbuf, auto full_c, auto keep_c) {
        // FULL: a step of the chunk's interior (z0 + 2 <= s <= z1 - 2), which
        // has every part below; KEEP: s is even, so ev already holds the
        // coarse plane (s + 1) / 2 (loaded by step s - 1)
        constexpr bool FULL = decltype(full_c)::value, KEEP = decltype(keep_c)::value;
        store_plane(buf, s + 1);
        {
            const double d64 = __builtin_bit_cast(double, dbuf), d32 = (double)__uint_as_float(dbuf[0]);
            if constexpr (O64) dcur = d64;  // the operand of output plane s-2
            else dcur = d32;
        }
        fs.fetch(buf, s + 3, D);  // zeros beyond the domain; unused beyond the chunk
        {
            const bool din = dbase != nullptr && s >= z0 && s < z1;  // uniform: output plane s
            const __amdgpu_buffer_rsrc_t r = buf_rsrc(din ? dbase + s * oszb : obase, din ? kInRange : 0);
            dbuf = __builtin_amdgcn_raw_buffer_load_b64(r, ooff, 0, 0);
        }
        __syncthreads();
        if (FULL || (s >= 0 && s < D && s <= z1)) partials(ev, es);  // uniform
        if constexpr (!KEEP)
            if (FULL || (s + 1 <= z1 && s + 1 < D)) fetch_e(ev, s + 1);
        const int sg = s - 1;                        // X plane gathered
        if (FULL || (sg >= 0 && sg < D && sg >= z0 - 1)) {  // uniform; every thread gathers,
            const float* const lg = lp + (es ^ 1) * kLPS;  // those without an output read the padding
            float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
                const float* pq = lg + dy * kXQp + q0 + dy * kXW;
                s0 += pq[0];
                s1 += pq[3 * kXQp];
                s2 += pq[6 * kXQp];
            }
            am1 += s2;
            a0 += s1;
            ap1 += s0;
        }
        const int zo = s - 2;
        if (FULL || (zo >= z0 && zo < z1)) {  // uniform
            const float v = am1 + bout;
            const __amdgpu_buffer_rsrc_t r = buf_rsrc(obase + zo * oszb, kInRange);
            if constexpr (O64) __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, (double)v), r, ooff, 0, 0);
            else __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, ooff, 0, 0);
            // threads without an output hold a dot operand of 0 but may hold a
            // non-finite v (gathered from the padding)
            const double pr = (double)v * dcur;
            dacc += mine ? pr : 0.0;
        }
        am1 = a0;
        a0 = ap1;
        ap1 = 0.0f;
    };
    // steps s = z0 - 1 .. z1 + 1 alternate the buffers A, B from A.  The
    // interior pairs run without a conditional step or part (a skipped step
    // would leave fewer loads behind the waits of the path that skips it, and
    // every wait would drain the prefetch); the steps at the chunk's ends
    // check which parts apply.
    // (chunks start on even planes, so the interior pairs start on even s)
    using T_ = std::true_type;
    using F_ = std::false_type;
    step(z0 - 1, (z0 - 1) & 1, preA, dA, F_{}, F_{});
    step(z0, z0 & 1, preB, dB, F_{}, F_{});
    step(z0 + 1, (z0 + 1) & 1, preA, dA, F_{}, F_{});
    int s = z0 + 2;
    for (; s + 1 <= z1 - 2; s += 2) {
        step(s, s & 1, preB, dB, T_{}, T_{});
        step(s + 1, (s + 1) & 1, preA, dA, T_{}, F_{});
    }
    // the steps left up to z1 + 1: 3 or 4, or 1 or 2 for a chunk of 1 or 2 planes
    if (s <= z1 + 1) step(s, s & 1, preB, dB, F_{}, F_{});
    if (s + 1 <= z1 + 1) step(s + 1, (s + 1) & 1, preA, dA, F_{}, F_{});
    if (s + 2 <= z1 + 1) step(s + 2, (s + 2) & 1, preB, dB, F_{}, F_{});
    if (s + 3 <= z1 + 1) step(s + 3, (s + 3) & 1, preA, dA, F_{}, F_{});
    if (a.dot) {
        __shared__ double red[4];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) dacc += __shfl_xor(dacc, o, 64);
        if (lane == 0) red[wave] = dacc;
        __syncthreads();
        if (tid == 0)
            a.dpart[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = (red[0] + red[1]) + (red[2] + red[3]);
    }
}

// grid of the fused output kernel: 14 x 18 tiles, z-chunks so that the launch
// has at least 8192 blocks (chunks of 16 planes or more)
static dim3 outf_grid(int D, int H, int W, int* zc)
{
    const int tiles = ((W + kFW - 1) / kFW) * ((H + kFH - 1) / kFH);
    int c = D;
    while (c > 16 && (long long)tiles * ((D + c - 1) / c) < 8192) c = (c + 1) / 2;
    if ((c & 1) && c < D) ++c;  // chunks start on even planes
    *zc = c;
    return dim3((unsigned)((W + kFW - 1) / kFW), (unsigned)((H + kFH - 1) / kFH), (unsigned)((D + c - 1) / c));
}

static long long outf_blocks(int D, int H, int W)
{
    int zc;
    const dim3 g = outf_grid(D, H, W, &zc);
    return (long long)g.x * g.y * g.z;
}

void launch_outf(const ConvArgs& a, hipStream_t s)
{
    int zc;
    const dim3 grid = outf_grid(a.out.D, a.out.H, a.out.W, &zc);
    if (a.fin.is64) {
        if (a.fout.is64) hipLaunchKernelGGL((k_outf<true, true>), grid, dim3(256), 0, s, a, zc);
        else hipLaunchKernelGGL((k_outf<true, false>), grid, dim3(256), 0, s, a, zc);
    } else {
        if (a.fout.is64) hipLaunchKernelGGL((k_outf<false, true>), grid, dim3(256), 0, s, a, zc);
        else hipLaunchKernelGGL((k_outf<false, false>), grid, dim3(256), 0, s, a, zc);
    }
}

// ------------------------------------------------------------ weight gradient
//
// dW[tap][ci][co] = sum_v X[v + off(tap)][ci] * dZ[v][co]: per tap one MFMA
// per 32 voxels (two x-rows of the tile) with A = X^T (row = ci, k = voxel)
// and B = dZ (k = voxel, column = co).  Wavefront w owns taps 7w .. 7w+6;
// wavefront 3's spare slot sums the bias gradient.  Each block writes its
// partial sums; a second kernel adds the partials in a fixed order.

// staged input plane of the weight-gradient kernel: [y][channel half][x]
constexpr int kPlaneDw = kLH * 2 * kLW;

template <int MI, int MZ>
__global__ __launch_bounds__(256) void k_dw(DwArgs a, int zc)
{
    __shared__ u32x4 lx[4 * kPlaneDw];
    __shared__ u32x4 lz[2 * kZPlane];
    const int D = a.D, H = a.H, W = a.W;
    const int nzc = (D + zc - 1) / zc;
    const int b = blockIdx.z / nzc;
    const int z0 = (blockIdx.z - b * nzc) * zc;
    const int z1 = min(z0 + zc, D);
    const int x0 = blockIdx.x * kTW, y0 = blockIdx.y * kTHt;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = lane >> 4, n = lane & 15;
    float fscale = 1.0f;
    if constexpr (MI == IN_FIELD) fscale = field_scale(a.fin);

    auto stage_x = [&](int zi) {
        u32x4* dst = lx + (zi & 3) * kPlaneDw;
        for (int t = tid; t < kPlaneDw; t += 256) {
            const int yy = t / (2 * kLW);
            const int rem = t - yy * (2 * kLW);
            const int xx = rem >> 1, ch = rem & 1;
            dst[(yy * 2 + ch) * kLW + xx] =
                load_chunk<MI>(a.in, a.in2, a.fin, fscale, D, H, W, b, x0 - 1 + xx, y0 - 1 + yy, zi, ch);
        }
    };
    constexpr int MZI = MZ == 0 ? IN_MASK : (MZ == 1 ? IN_PLAIN : IN_FIELD);
    auto stage_z = [&](int zi) {
        u32x4* dst = lz + (zi & 1) * kZPlane;
        for (int t = tid; t < kZPlane; t += 256) {
            const int yy = t >> 5, rem = t & 31;
            const int xx = rem >> 1, ch = rem & 1;
            dst[(yy * 2 + ch) * kTW + xx] = load_chunk<MZI>(a.dy, a.y, a.fdy, 1.0f, D, H, W, b, x0 + xx, y0 + yy, zi, ch);
        }
    };

    f32x4 acc[7];
#pragma unroll
    for (int t = 0; t < 7; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    float db = 0.0f;
    const int ntap = wave == 3 ? 6 : 7;

    stage_x(z0 - 1);
    stage_x(z0);
    for (int z = z0; z < z1; ++z) {
        stage_x(z + 1);
        stage_z(z);
        __syncthreads();
        const uint16_t* X = (const uint16_t*)lx;
        const uint16_t* Zp = (const uint16_t*)(lz + (z & 1) * kZPlane);
        for (int rp = 0; rp < kTHt / 2; ++rp) {
            const int row = 2 * rp + (j >> 1), xb = 8 * (j & 1);
            const uint16_t* zb = Zp + ((row * 2 + (n >> 3)) * kTW + xb) * 8 + (n & 7);
            u32x4 bw;
#pragma unroll
            for (int e = 0; e < 4; ++e) bw[e] = (unsigned)zb[16 * e] | ((unsigned)zb[16 * e + 8] << 16);
            if (wave == 3) {
#pragma unroll
                for (int e = 0; e < 4; ++e) db += bf2f(bw[e] & 0xffffu) + __uint_as_float(bw[e] & 0xffff0000u);
            }
            const bf16x8 Bv = __builtin_bit_cast(bf16x8, bw);
#pragma unroll
            for (int t = 0; t < 7; ++t) {
                if (t >= ntap) break;
                const int tap = 7 * wave + t;
                const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
                const uint16_t* xp =
                    X + ((z + dz - 1) & 3) * kPlaneDw * 8 + (((row + dy) * 2 + (n >> 3)) * kLW + xb + dx) * 8 + (n & 7);
                u32x4 aw;
#pragma unroll
                for (int e = 0; e < 4; ++e) aw[e] = (unsigned)xp[16 * e] | ((unsigned)xp[16 * e + 8] << 16);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, aw), Bv, acc[t], 0, 0, 0);
            }
        }
    }
    const long long blk = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    float* part = a.part + blk * kLayerParams;
#pragma unroll
    for (int t = 0; t < 7; ++t) {
        if (t >= ntap) break;
        const int tap = 7 * wave + t;
#pragma unroll
        for (int r = 0; r < 4; ++r) part[tap * 256 + (4 * j + r) * 16 + n] = acc[t][r];
    }
    if (wave == 3) {
        db += __shfl_xor(db, 16);
        db += __shfl_xor(db, 32);
        if (j == 0) part[kTaps * 256 + n] = db;
    }
}

int dw_blocks(const DwArgs& a)
{
    const int zc = pick_zc(a.D, a.H, a.W, a.B);
    const dim3 g = conv_grid(a.D, a.H, a.W, a.B, zc);
    return (int)(g.x * g.y * g.z);
}

template <int MI>
static void dw_mz(const DwArgs& a, hipStream_t s, dim3 grid, int zc)
{
    switch (a.dz_mode) {
    case 0: hipLaunchKernelGGL((k_dw<MI, 0>), grid, dim3(256), 0, s, a, zc); break;
    case 1: hipLaunchKernelGGL((k_dw<MI, 1>), grid, dim3(256), 0, s, a, zc); break;
    default: hipLaunchKernelGGL((k_dw<MI, 2>), grid, dim3(256), 0, s, a, zc); break;
    }
}

void launch_dw(const DwArgs& a, hipStream_t s)
{
    const int zc = pick_zc(a.D, a.H, a.W, a.B);
    const dim3 grid = conv_grid(a.D, a.H, a.W, a.B, zc);
    switch (a.mode_in) {
    case IN_PLAIN: dw_mz<IN_PLAIN>(a, s, grid, zc); break;
    case IN_POOL: dw_mz<IN_POOL>(a, s, grid, zc); break;
    case IN_UPADD: dw_mz<IN_UPADD>(a, s, grid, zc); break;
    default: dw_mz<IN_FIELD>(a, s, grid, zc); break;
    }
}

__global__ void k_dw_reduce(const float* __restrict__ part, int nblocks, float* __restrict__ g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kLayerParams) return;
    float s = 0.0f;
    for (int b = 0; b < nblocks; ++b) s += part[(long long)b * kLayerParams + i];
    g[i] = s;
}

void launch_dw_reduce(const float* part, int nblocks, float* g, hipStream_t s)
{
    hipLaunchKernelGGL(k_dw_reduce, dim3((kLayerParams + 255) / 256), dim3(256), 0, s, part, nblocks, g);
}

// ------------------------------------------------ gradient routing of pool / up

// dfine[v] (+)= dcoarse[parent(v)] / 8
__global__ void k_pool_back(ActRef dc, ActRef df, int accumulate)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nch = (long long)df.B * df.D * df.H * df.W * 2;
    if (t >= nch) return;
    const int ch = (int)(t & 1);
    long long v = t >> 1;
    const int x = (int)(v % df.W);
    v /= df.W;
    const int y = (int)(v % df.H);
    v /= df.H;
    const int z = (int)(v % df.D);
    const int b = (int)(v / df.D);
    float c[8];
    unpack8(*(const u32x4*)(dc.p + vox(dc, b, x >> 1, y >> 1, z >> 1) + ch * 8), c);
    u32x4* dst = (u32x4*)(df.p + vox(df, b, x, y, z) + ch * 8);
    float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (accumulate) unpack8(*dst, f);
#pragma unroll
    for (int e = 0; e < 8; ++e) f[e] += 0.125f * c[e];
    *dst = pack8(f);
}

void launch_pool_back(const ActRef& dcoarse, const ActRef& dfine, int accumulate, hipStream_t s)
{
    const long long nch = (long long)dfine.B * dfine.D * dfine.H * dfine.W * 2;
    hipLaunchKernelGGL(k_pool_back, dim3((unsigned)((nch + 255) / 256)), dim3(256), 0, s, dcoarse, dfine, accumulate);
}

// dskip[v] (+)= dx[v] for the 8 children v of a coarse voxel; dcoarse = their sum
__global__ void k_upadd_back(ActRef dx, ActRef dskip, int accumulate_skip, ActRef dco)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nch = (long long)dco.B * dco.D * dco.H * dco.W * 2;
    if (t >= nch) return;
    const int ch = (int)(t & 1);
    long long v = t >> 1;
    const int x = (int)(v % dco.W);
    v /= dco.W;
    const int y = (int)(v % dco.H);
    v /= dco.H;
    const int z = (int)(v % dco.D);
    const int b = (int)(v / dco.D);
    float sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int fx = 2 * x + (c & 1), fy = 2 * y + ((c >> 1) & 1), fz = 2 * z + (c >> 2);
        if (fx >= dx.W || fy >= dx.H || fz >= dx.D) continue;
        const long long o = vox(dx, b, fx, fy, fz) + ch * 8;
        float g[8];
        unpack8(*(const u32x4*)(dx.p + o), g);
#pragma unroll
        for (int e = 0; e < 8; ++e) sum[e] += g[e];
        if (accumulate_skip) {
            float s[8];
            unpack8(*(const u32x4*)(dskip.p + o), s);
#pragma unroll
            for (int e = 0; e < 8; ++e) s[e] += g[e];
            *(u32x4*)(dskip.p + o) = pack8(s);
        } else {
            *(u32x4*)(dskip.p + o) = *(const u32x4*)(dx.p + o);
        }
    }
    *(u32x4*)(dco.p + vox(dco, b, x, y, z) + ch * 8) = pack8(sum);
}

void launch_upadd_back(const ActRef& dx, const ActRef& dskip, int accumulate_skip, const ActRef& dcoarse, hipStream_t s)
{
    const long long nch = (long long)dcoarse.B * dcoarse.D * dcoarse.H * dcoarse.W * 2;
    hipLaunchKernelGGL(k_upadd_back, dim3((unsigned)((nch + 255) / 256)), dim3(256), 0, s, dx, dskip, accumulate_skip,
                       dcoarse);
}

// ------------------------------------------------------------ weights, Adam

// lane-arranged A operands, wl[o][lane][e] for operand o = 3*pair + dy:
// W'[tap][k][m], m = lane%16 (output channel), k = 8*((lane/16)&1) + e (input
// channel), tap = 9*dz + 3*dy + dx of pair member (lane/16)/2 (zero for the
// empty member).
// forward:        W'[tap][k][m] = W[tap][ci = k][co = m]
// backward data:  W'[tap][k][m] = W[26 - tap][ci = m][co = k]   (flipped, transposed)
// lift (mode 4, 1 input channel): three K = 16 operands, one per dy, for the
// lift kernel: k = 4*(lane/16) + d holds plane d of the quad at dx = lane/16;
// and from kLiftPairOff two K = 32 operands o for the fused output kernel:
// k = 8*(lane/16) + e holds plane e%4 of quad e/4 of the pair kQuad[o][lane/16]
// row partials (mode 3, 1 output channel): three operands o = dx, column
// m = 3 dz + dy, k = 8*(lane/16) + e -> input channel 4*(lane/16) + e%4
__global__ void k_pack(const float* __restrict__ W, int cin, int cout, int mode, uint16_t* __restrict__ wl)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int e = t & 7, lane = (t >> 3) & 63, o = t >> 9;
    const int j = lane >> 4, m = lane & 15;
    float v = 0.0f;
    if (mode == 4) {
        if (t < 768) {
            // the lift kernel's operands: K = 16 MFMA dy (one per 256
            // halves), lane (channel m, block jj), plane d; plane slot 3 and
            // block 3 carry zeros
            const int dy = t >> 8, ln = (t >> 2) & 63, d = t & 3;
            const int jj = ln >> 4, mm = ln & 15;
            if (d < 3 && jj < 3 && mm < cout) v = W[(9 * d + 3 * dy + jj) * 256 + mm];
        } else if (t >= kLiftPairOff && t < kLiftPairOff + 1024) {
            // the fused output kernel's operands: MFMA o, lane, slot e = quad
            // h (e / 4) and plane d (e % 4); plane slot 3 and the quads
            // without a tap position carry zeros
            const int u = t - kLiftPairOff, oo = u >> 9, ln = (u >> 3) & 63, d = u & 3, hh = (u >> 2) & 1;
            const int jj = ln >> 4, mm = ln & 15;
            const int dy = kQuad[oo][jj][0], dx = kQuad[oo][jj][1] + hh;
            if (d < 3 && q_real(oo, jj, hh) && mm < cout) v = W[(9 * d + 3 * dy + dx) * 256 + mm];
        } else {
            return;
        }
    } else if (mode == 3) {
        // the fused output layer's row-partial weights: operand o = dx, column
        // m = 3 dz + dy (m < 9), k = 8*(lane/16) + e -> input channel
        // 4*(lane/16) + e%4 (the lift half and the skip half of the A operand
        // share the weights)
        if (t >= 3 * 512) return;
        const int ci = 4 * j + (e & 3);
        if (m < 9 && ci < cin && cout > 0) v = W[(9 * (m / 3) + 3 * (m % 3) + o) * 256 + ci * 16];
    } else {
        if (t >= kPackedHalves) return;
        const int p = o / 3, dy = o % 3, c = pair_combo(p, j >> 1), k = 8 * (j & 1) + e;
        if (c < 9) {
            const int tap = (c / 3) * 9 + dy * 3 + c % 3;
            if (mode == 0) {
                if (k < cin && m < cout) v = W[tap * 256 + k * 16 + m];
            } else {
                if (m < cin && k < cout) v = W[(kTaps - 1 - tap) * 256 + m * 16 + k];
            }
        }
    }
    wl[t] = (uint16_t)f2bf(v);
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int n, float lr, float b1, float b2, float eps, float bc1, float bc2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// ------------------------------------------------------------ training loss
//
// Per sample s (residual r normalised to unit RMS, network output y):
//   loss_s = 1 - (y.r)^2 / ((y.Ky) G_s),   K = -(index-space 7-point Laplacian)
// which is 1 - cos^2 of the angle between y and the error e in the energy
// norm of K, with G_s = e.Ke (both sides in the units of the normalised r).
// It is invariant to the scale of y, as FCG is.  Its gradient:
//   dloss_s/dy = 2 a^2/(c^2 G) Ky - 2 a/(c G) r,   a = y.r, c = y.Ky
// Batch loss = mean over samples.

__device__ __forceinline__ double block_sum_d(double v, double* sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void k_loss1(const float* __restrict__ y, const float* __restrict__ r,
                                               float* __restrict__ ky, int n0, int n1, int n2, int per_block,
                                               double* __restrict__ part)
{
    __shared__ double sh[256];
    const int b = blockIdx.y;
    const long long nv = (long long)n0 * n1 * n2;
    const float* yb = y + b * nv;
    const float* rb = r + b * nv;
    float* kb = ky + b * nv;
    double sa = 0.0, sc = 0.0;
    const long long v0 = (long long)blockIdx.x * per_block;
    const long long v1 = min(v0 + per_block, nv);
    for (long long v = v0 + threadIdx.x; v < v1; v += blockDim.x) {
        const int x = (int)(v % n0);
        const int yy = (int)((v / n0) % n1);
        const int z = (int)(v / ((long long)n0 * n1));
        const float c = yb[v];
        float s = 6.0f * c;
        if (x > 0) s -= yb[v - 1];
        if (x < n0 - 1) s -= yb[v + 1];
        if (yy > 0) s -= yb[v - n0];
        if (yy < n1 - 1) s -= yb[v + n0];
        if (z > 0) s -= yb[v - (long long)n0 * n1];
        if (z < n2 - 1) s -= yb[v + (long long)n0 * n1];
        kb[v] = s;
        sa += (double)c * (double)rb[v];
        sc += (double)c * (double)s;
    }
    sa = block_sum_d(sa, sh);
    sc = block_sum_d(sc, sh);
    if (threadIdx.x == 0) {
        part[((long long)b * gridDim.x + blockIdx.x) * 2] = sa;
        part[((long long)b * gridDim.x + blockIdx.x) * 2 + 1] = sc;
    }
}

__global__ void k_loss2(const double* __restrict__ part, int nblk, const double* __restrict__ G, int B,
                        double* __restrict__ coef, double* __restrict__ lossv)
{
    const int b = blockIdx.x;
    if (threadIdx.x != 0) return;
    double a = 0.0, c = 0.0;
    for (int i = 0; i < nblk; ++i) {
        a += part[((long long)b * nblk + i) * 2];
        c += part[((long long)b * nblk + i) * 2 + 1];
    }
    c = fmax(c, 1e-300);
    const double g = G[b];
    lossv[b] = 1.0 - a * a / (c * g);
    coef[2 * b] = 2.0 * a * a / (c * c * g * B);
    coef[2 * b + 1] = -2.0 * a / (c * g * B);
}

__global__ void k_loss3(const float* __restrict__ ky, const float* __restrict__ r, const double* __restrict__ coef,
                        long long nv, int B, float* __restrict__ dy)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv * B) return;
    const int b = (int)(t / nv);
    dy[t] = (float)(coef[2 * b] * (double)ky[t] + coef[2 * b + 1] * (double)r[t]);
}

// ================================================================ Net

int Net::fail(const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err_ = buf;
    return -1;
}

int Net::hip(hipError_t e, const char* what)
{
    if (e == hipSuccess) return 0;
    fail("%s: %s", what, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? -3 : -2;
}

Net::~Net()
{
    release(inf_);
    release(trn_);
    for (void* p : {(void*)params_, (void*)grads_, (void*)adam_m_, (void*)adam_v_, (void*)wl_, (void*)wlt_, (void*)wlo_,
                    (void*)data_r_})
        if (p) (void)hipFree(p);
    if (stream_) (void)hipStreamDestroy(stream_);
}

long long Net::effective_params()
{
    long long s = 0;
    for (const LayerDef& d : kNet) s += (long long)kTaps * d.cin * d.cout + d.cout;
    return s;
}

int Net::init(unsigned long long seed)
{
    // a pooled activation's producer writes its pooled copy: the lift, or a
    // plain-input convolution with a ReLU output
    for (int i = 0; i < kActs; ++i)
        if (pooled_act(i) && !(i == 0 || (kNet[i].mode_in == IN_PLAIN && i < kLayers - 1)))
            return fail("network table: activation %d is pooled but its producer cannot write the pooled copy", i);
    int rc = hip(hipStreamCreate(&stream_), "hipStreamCreate");
    if (!rc) rc = hip(hipMalloc(&params_, kParams * sizeof(float)), "hipMalloc params");
    if (!rc) rc = hip(hipMalloc(&grads_, kParams * sizeof(float)), "hipMalloc grads");
    if (!rc) rc = hip(hipMalloc(&adam_m_, kParams * sizeof(float)), "hipMalloc adam");
    if (!rc) rc = hip(hipMalloc(&adam_v_, kParams * sizeof(float)), "hipMalloc adam");
    if (!rc) rc = hip(hipMalloc(&wl_, (size_t)kLayers * kPackedHalves * 2), "hipMalloc weights");
    if (!rc) rc = hip(hipMalloc(&wlt_, (size_t)kLayers * kPackedHalves * 2), "hipMalloc weights");
    if (!rc) rc = hip(hipMalloc(&wlo_, 3 * 512 * 2), "hipMalloc weights");
    if (rc) return rc;
    // He initialisation (std = sqrt(2 / fan_in)) of the ReLU layers, sqrt(1 / fan_in) for the output
    std::vector<float> h(kParams, 0.0f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (int l = 0; l < kLayers; ++l) {
        const LayerDef& d = kNet[l];
        const double sd = std::sqrt((d.relu ? 2.0 : 1.0) / (kTaps * d.cin));
        for (int t = 0; t < kTaps; ++t)
            for (int ci = 0; ci < d.cin; ++ci)
                for (int co = 0; co < d.cout; ++co) h[(size_t)l * kLayerParams + t * 256 + ci * 16 + co] = (float)(sd * nd(rng));
    }
    return set_params(h.data(), kParams);
}

void Net::pack_all()
{
    for (int l = 0; l < kLayers; ++l) {
        const float* W = params_ + (size_t)l * kLayerParams;
        // the first layer (one input channel) runs on the lift kernels' quad-ring operands
        hipLaunchKernelGGL(k_pack, dim3(kPackedHalves / 256), dim3(256), 0, stream_, W, kNet[l].cin, kNet[l].cout,
                           l == 0 ? 4 : 0, wl_ + (size_t)l * kPackedHalves);
        hipLaunchKernelGGL(k_pack, dim3(kPackedHalves / 256), dim3(256), 0, stream_, W, kNet[l].cin, kNet[l].cout, 1,
                           wlt_ + (size_t)l * kPackedHalves);
    }
    // the output layer again as row-partial operands, for the fused inference kernel
    hipLaunchKernelGGL(k_pack, dim3(6), dim3(256), 0, stream_, params_ + (size_t)(kLayers - 1) * kLayerParams,
                       kNet[kLayers - 1].cin, kNet[kLayers - 1].cout, 3, wlo_);
}

int Net::set_params(const float* host, long long n)
{
    if (n != kParams) return fail("parameter vector must have %d entries (got %lld)", kParams, n);
    // entries outside a layer's channel counts must stay zero
    std::vector<float> h(host, host + n);
    for (int l = 0; l < kLayers; ++l) {
        const LayerDef& d = kNet[l];
        for (int t = 0; t < kTaps; ++t)
            for (int ci = 0; ci < 16; ++ci)
                for (int co = 0; co < 16; ++co)
                    if (ci >= d.cin || co >= d.cout) h[(size_t)l * kLayerParams + t * 256 + ci * 16 + co] = 0.0f;
        for (int co = d.cout; co < 16; ++co) h[(size_t)l * kLayerParams + kTaps * 256 + co] = 0.0f;
    }
    int rc = hip(hipMemcpyAsync(params_, h.data(), n * sizeof(float), hipMemcpyHostToDevice, stream_), "upload params");
    if (!rc) rc = hip(hipMemsetAsync(adam_m_, 0, kParams * sizeof(float), stream_), "memset");
    if (!rc) rc = hip(hipMemsetAsync(adam_v_, 0, kParams * sizeof(float), stream_), "memset");
    if (rc) return rc;
    adam_t_ = 0;
    pack_all();
    return hip(hipStreamSynchronize(stream_), "sync");
}

int Net::get_params(float* host, long long n)
{
    if (n != kParams) return fail("parameter vector must have %d entries (got %lld)", kParams, n);
    int rc = hip(hipMemcpyAsync(host, params_, n * sizeof(float), hipMemcpyDeviceToHost, stream_), "download params");
    if (rc) return rc;
    return hip(hipStreamSynchronize(stream_), "sync");
}

void Net::release(BufSet& bs)
{
    for (void* p : bs.owned) (void)hipFree(p);
    bs = BufSet{};
}

ActRef Net::level(const BufSet& bs, int lv, uint16_t* p) const
{
    ActRef a;
    a.p = p;
    a.B = bs.B;
    a.D = bs.dims[lv][0];
    a.H = bs.dims[lv][1];
    a.W = bs.dims[lv][2];
    return a;
}
ActRef Net::act(const BufSet& bs, int i) const { return level(bs, kNet[i].level, bs.act[i]); }
ActRef Net::dact(const BufSet& bs, int i) const { return level(bs, kNet[i].level, bs.dact[i]); }

int Net::ensure(BufSet& bs, int B, const int n[3], bool train)
{
    if (bs.B == B && bs.n[0] == n[0] && bs.n[1] == n[1] && bs.n[2] == n[2] && bs.train == train) return 0;
    release(bs);
    bs.B = B;
    bs.train = train;
    for (int d = 0; d < 3; ++d) bs.n[d] = n[d];
    bs.dims[0][0] = n[2];
    bs.dims[0][1] = n[1];
    bs.dims[0][2] = n[0];
    for (int l = 1; l < 3; ++l)
        for (int d = 0; d < 3; ++d) bs.dims[l][d] = (bs.dims[l - 1][d] + 1) / 2;
    auto vox_count = [&](int lv) { return (size_t)B * bs.dims[lv][0] * bs.dims[lv][1] * bs.dims[lv][2]; };
    auto alloc = [&](size_t bytes, void** p) -> int {
        int rc = hip(hipMalloc(p, std::max<size_t>(bytes, 256)), "hipMalloc network buffers");
        if (!rc) bs.owned.push_back(*p);
        return rc;
    };
    bs.act.assign(kActs, nullptr);
    bs.dact.assign(kActs, nullptr);
    bs.pooled.assign(kActs, nullptr);
    int rc = 0;
    for (int i = 0; i < kActs && !rc; ++i)
        if (pooled_act(i)) rc = alloc(vox_count(kNet[i].level + 1) * kC * 2, (void**)&bs.pooled[i]);
    if (!train) {
        // inference: an activation takes a buffer of its level that holds no live
        // activation (live = read by a later layer), so each level keeps only
        // as many buffers as it has activations alive at once
        int last_use[kActs];
        for (int i = 0; i < kActs; ++i) last_use[i] = i;
        for (int l = 0; l < kLayers; ++l) {
            if (kNet[l].in_act >= 0) last_use[kNet[l].in_act] = std::max(last_use[kNet[l].in_act], l);
            if (kNet[l].skip_act >= 0) last_use[kNet[l].skip_act] = std::max(last_use[kNet[l].skip_act], l);
        }
        std::vector<std::pair<int, uint16_t*>> slots[3];  // (holder, buffer) per level
        for (int i = 0; i < kActs && !rc; ++i) {
            if (i == 0) continue;  // the lift's output lives only in the fused output kernel (and its pooled copy)
            auto& pool = slots[kNet[i].level];
            int pick = -1;
            for (size_t k = 0; k < pool.size(); ++k)
                if (last_use[pool[k].first] < i) pick = (int)k;
            if (pick < 0) {
                uint16_t* p = nullptr;
                rc = alloc(vox_count(kNet[i].level) * kC * 2, (void**)&p);
                pool.push_back({i, p});
                pick = (int)pool.size() - 1;
            }
            pool[pick].first = i;
            bs.act[i] = pool[pick].second;
        }
        return rc;
    }
    for (int i = 0; i < kActs && !rc; ++i) rc = alloc(vox_count(kNet[i].level) * kC * 2, (void**)&bs.act[i]);
    for (int i = 0; i < kActs && !rc; ++i) rc = alloc(vox_count(kNet[i].level) * kC * 2, (void**)&bs.dact[i]);
    for (int l = 0; l < 3 && !rc; ++l) rc = alloc(vox_count(l) * kC * 2, (void**)&bs.dx[l]);
    const size_t nv = vox_count(0);
    if (!rc) rc = alloc(nv * 4, (void**)&bs.in);
    if (!rc) rc = alloc(nv * 4, (void**)&bs.y);
    if (!rc) rc = alloc(nv * 4, (void**)&bs.dy);
    if (!rc) rc = alloc(nv * 4, (void**)&bs.ky);
    if (!rc) rc = alloc(B * 8, (void**)&bs.g);
    if (!rc) rc = alloc(B * 16, (void**)&bs.coef);
    if (!rc) rc = alloc(B * 8, (void**)&bs.lossv);
    const long long per_sample = (long long)n[0] * n[1] * n[2];
    bs.lblocks = (int)std::min<long long>(256, (per_sample + 1023) / 1024);
    if (!rc) rc = alloc((size_t)B * bs.lblocks * 16, (void**)&bs.lpart);
    long long pb = 0;
    for (int l = 0; l < kLayers; ++l) {
        DwArgs w{};
        w.B = B;
        w.D = bs.dims[kNet[l].level][0];
        w.H = bs.dims[kNet[l].level][1];
        w.W = bs.dims[kNet[l].level][2];
        pb = std::max<long long>(pb, dw_blocks(w));
    }
    bs.part_blocks = pb;
    if (!rc) rc = alloc((size_t)pb * kLayerParams * 4, (void**)&bs.part);
    return rc;
}

ConvArgs Net::layer_args(const BufSet& bs, int i, const FieldRef& fin, const FieldRef& fout, const int* gate,
                         const int* gate2) const
{
    const LayerDef& d = kNet[i];
    ConvArgs a{};
    a.mode_in = d.mode_in;
    a.mode_out = i == kLayers - 1 ? OUT_FIELD : OUT_RELU;
    a.out = i < kActs ? act(bs, i) : level(bs, 0, nullptr);
    if (d.mode_in == IN_FIELD) a.fin = fin;
    else a.in = act(bs, d.in_act);
    if (d.mode_in == IN_UPADD) a.in2 = act(bs, d.skip_act);
    if (i < kActs && bs.pooled[i]) a.pout = level(bs, d.level + 1, bs.pooled[i]);
    if (d.mode_in == IN_POOL) {
        // the producer has written pool(act): read it as it is
        a.mode_in = IN_PLAIN;
        a.in = level(bs, d.level, bs.pooled[d.in_act]);
    }
    a.fout = fout;
    a.wl = wl_ + (size_t)i * kPackedHalves;
    a.bias = params_ + (size_t)i * kLayerParams + kTaps * 256;
    a.gate = gate;
    a.gate2 = gate2;
    if (!bs.train && bs.B == 1 && i == kLayers - 1 && d.mode_in == IN_UPADD && d.skip_act == 0 &&
        kNet[0].mode_in == IN_FIELD) {
        // inference: the output layer recomputes the lift (act0 is never stored)
        a.fused = 1;
        a.fin = fin;
        a.wl = wlo_;
        a.wl2 = wl_ + kLiftPairOff;  // layer 0's lift operands of this kernel
        a.bias2 = params_ + kTaps * 256;
    }
    return a;
}

void Net::launch_layer(const ConvArgs& a, int i)
{
    if (kNet[i].mode_in == IN_FIELD) launch_lift(a, stream_);
    else if (a.fused) launch_outf(a, stream_);
    else launch_conv(a, stream_);
}

void Net::forward(BufSet& bs, const FieldRef& fin, const FieldRef& fout, const int* gate, const int* gate2,
                  const void* dot, double* dpart, const NNResUpdate* upd)
{
    for (int i = 0; i < kLayers; ++i) {
        ConvArgs a = layer_args(bs, i, fin, fout, gate, gate2);
        if (a.fused && dot) {
            a.dot = dot;
            a.dpart = dpart;
        }
        if (upd && kNet[i].mode_in == IN_FIELD) {
            // the lift forms the input from the previous residual and writes it
            a.fin.p = const_cast<void*>(upd->r_prev);
            a.fap = upd->ap_prev;
            a.fcur = upd->r_cur;
            a.alpha = upd->alpha;
            a.skip = upd->skip;
        }
        launch_layer(a, i);
    }
}

void Net::backward(BufSet& bs, const FieldRef& fin)
{
    FieldRef fdy;
    fdy.p = bs.dy;
    fdy.sx = 1;
    fdy.sy = bs.n[0];
    fdy.sz = (long long)bs.n[0] * bs.n[1];
    fdy.sb = fdy.sz * bs.n[2];
    for (int i = kLayers - 1; i >= 0; --i) {
        const LayerDef& d = kNet[i];
        DwArgs w{};
        w.mode_in = d.mode_in;
        if (d.mode_in == IN_FIELD) w.fin = fin;
        else w.in = act(bs, d.in_act);
        if (d.mode_in == IN_UPADD) w.in2 = act(bs, d.skip_act);
        if (d.mode_in == IN_POOL) {
            w.mode_in = IN_PLAIN;  // the forward input: pool(act) as its producer wrote it
            w.in = level(bs, d.level, bs.pooled[d.in_act]);
        }
        if (i == kLayers - 1) {
            w.dz_mode = 2;
            w.fdy = fdy;
        } else {
            w.dz_mode = 0;
            w.dy = dact(bs, i);
            w.y = act(bs, i);
        }
        w.B = bs.B;
        w.D = bs.dims[d.level][0];
        w.H = bs.dims[d.level][1];
        w.W = bs.dims[d.level][2];
        w.part = bs.part;
        launch_dw(w, stream_);
        launch_dw_reduce(bs.part, dw_blocks(w), grads_ + (size_t)i * kLayerParams, stream_);
        if (i == 0) break;
        ConvArgs c{};
        if (i == kLayers - 1) {
            c.mode_in = IN_FIELD;
            c.fin = fdy;
        } else {
            c.mode_in = IN_MASK;
            c.in = dact(bs, i);
            c.in2 = act(bs, i);
        }
        c.mode_out = OUT_LINEAR;
        c.wl = wlt_ + (size_t)i * kPackedHalves;
        const bool plain = d.mode_in == IN_PLAIN;
        c.out = plain ? dact(bs, d.in_act) : level(bs, d.level, bs.dx[d.level]);
        launch_conv(c, stream_);
        if (d.mode_in == IN_POOL) launch_pool_back(c.out, dact(bs, d.in_act), 1, stream_);
        if (d.mode_in == IN_UPADD) launch_upadd_back(c.out, dact(bs, d.skip_act), 0, dact(bs, d.in_act), stream_);
    }
}

void Net::loss(BufSet& bs, int B, long long nv)
{
    const int per_block = (int)((nv + bs.lblocks - 1) / bs.lblocks);
    hipLaunchKernelGGL(k_loss1, dim3(bs.lblocks, B), dim3(256), 0, stream_, bs.y, bs.in, bs.ky, bs.n[0], bs.n[1],
                       bs.n[2], per_block, bs.lpart);
    hipLaunchKernelGGL(k_loss2, dim3(B), dim3(64), 0, stream_, bs.lpart, bs.lblocks, bs.g, B, bs.coef, bs.lossv);
    const long long nt = nv * B;
    hipLaunchKernelGGL(k_loss3, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, stream_, bs.ky, bs.in, bs.coef, nv, B,
                       bs.dy);
}

static FieldRef dense_field(float* p, const int n[3])
{
    FieldRef f;
    f.p = p;
    f.is64 = 0;
    f.sx = 1;
    f.sy = n[0];
    f.sz = (long long)n[0] * n[1];
    f.sb = f.sz * n[2];
    return f;
}

void Net::field_refs(const NNFieldGeom& g, void* out, int out_prec, const void* in, int in_prec, const double* rr,
                     FieldRef& fi, FieldRef& fo)
{
    fi = FieldRef{};
    fi.p = const_cast<void*>(in);
    fi.is64 = in_prec == 64;
    fi.off = g.off;
    fi.sx = 1;
    fi.sy = g.sy;
    fi.sz = g.sz;
    fi.sb = 0;
    fi.scale_sum = rr;
    fi.scale_count = (double)g.n0 * g.n1 * g.n2;
    fo = fi;
    fo.p = out;
    fo.is64 = out_prec == 64;
    fo.scale_sum = nullptr;
}

int Net::apply(hipStream_t s, const NNFieldGeom& g, void* out, int out_prec, const void* in, int in_prec,
               const int* gate, const int* gate2, const double* rr)
{
    int nb;
    return apply_dot(s, g, out, out_prec, in, in_prec, gate, gate2, rr, nullptr, nullptr, 0, &nb, nullptr);
}

int Net::apply_dot(hipStream_t s, const NNFieldGeom& g, void* out, int out_prec, const void* in, int in_prec,
                   const int* gate, const int* gate2, const double* rr, const void* dot, double* dpart, int dcap,
                   int* nblocks, const NNResUpdate* upd)
{
    *nblocks = 0;
    if (upd && upd->r_cur != in) return fail("apply: the updated residual must be the input field");
    const int n[3] = {g.n0, g.n1, g.n2};
    int rc = ensure(inf_, 1, n, false);
    if (rc) return rc;
    const hipStream_t own = stream_;
    stream_ = s;
    FieldRef fi, fo;
    field_refs(g, out, out_prec, in, in_prec, rr, fi, fo);
    // the output layer runs fused at inference (one sample): it can sum out . dot
    const ConvArgs last = layer_args(inf_, kLayers - 1, fi, fo, gate, gate2);
    const long long nb = outf_blocks(last.out.D, last.out.H, last.out.W);
    const bool with_dot = dot && dpart && last.fused && nb <= dcap;
    forward(inf_, fi, fo, gate, gate2, with_dot ? dot : nullptr, with_dot ? dpart : nullptr, upd);
    if (with_dot) *nblocks = (int)nb;
    stream_ = own;
    return hip(hipGetLastError(), "network launch");
}

int Net::time_layer(hipStream_t s, const NNFieldGeom& g, void* out, int out_prec, const void* in, int in_prec,
                    const void* dot, const double* rr, int i, int reps, double* ms, double* bytes, double* flops)
{
    if (i < 0 || i >= kLayers || reps < 1) return fail("time_layer: layer %d / reps %d out of range", i, reps);
    const int n[3] = {g.n0, g.n1, g.n2};
    if (!(inf_.B == 1 && inf_.n[0] == n[0] && inf_.n[1] == n[1] && inf_.n[2] == n[2] && !inf_.train))
        return fail("time_layer: run apply() on this region first");
    FieldRef fi, fo;
    field_refs(g, out, out_prec, in, in_prec, rr, fi, fo);
    const hipStream_t own = stream_;
    stream_ = s;
    ConvArgs a = layer_args(inf_, i, fi, fo, nullptr, nullptr);
    // the fused output layer runs as the Krylov loop runs it: also summing
    // z . dot (a field of z's precision and layout)
    double* dpart = nullptr;
    int rc = 0;
    if (a.fused && dot) {
        rc = hip(hipMalloc(&dpart, sizeof(double) * outf_blocks(a.out.D, a.out.H, a.out.W)), "hipMalloc partials");
        a.dot = dot;
        a.dpart = dpart;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc) rc = hip(hipEventCreate(&e0), "event");
    if (!rc) rc = hip(hipEventCreate(&e1), "event");
    if (!rc) {
        launch_layer(a, i);  // warm-up
        (void)hipEventRecord(e0, s);
        for (int k = 0; k < reps; ++k) launch_layer(a, i);
        (void)hipEventRecord(e1, s);
        rc = hip(hipEventSynchronize(e1), "event sync");
    }
    float t = 0.0f;
    if (!rc) rc = hip(hipEventElapsedTime(&t, e0, e1), "elapsed");
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (dpart) (void)hipFree(dpart);
    stream_ = own;
    if (rc) return rc;
    *ms = t / reps;
    // algorithmic traffic: every input value read once, every output written once
    const LayerDef& d = kNet[i];
    double vox[3];
    for (int l = 0; l < 3; ++l) vox[l] = (double)inf_.dims[l][0] * inf_.dims[l][1] * inf_.dims[l][2];
    const double v = vox[d.level];
    const double act_b = kC * 2.0;
    double rd = 0.0, wr = 0.0;
    switch (d.mode_in) {
    case IN_FIELD: rd = v * (in_prec == 64 ? 8.0 : 4.0); break;
    case IN_POOL: rd = v * act_b; break;  // the producer's pooled copy
    case IN_UPADD: rd = vox[d.level + 1] * act_b + v * act_b; break;
    default: rd = v * act_b; break;
    }
    wr = (i == kLayers - 1) ? v * (out_prec == 64 ? 8.0 : 4.0) : v * act_b;
    if (i == 0) wr = 0.0;  // inference keeps only the lift's pooled copy
    if (i < kActs && pooled_act(i)) wr += vox[d.level + 1] * act_b;
    double fl = 2.0 * kTaps * d.cin * d.cout * v;
    if (i == kLayers - 1) {
        // fused: reads r and the coarse activation, recomputes the lift, and
        // reads the dot operand
        rd = v * (in_prec == 64 ? 8.0 : 4.0) + vox[1] * act_b;
        if (a.dot) rd += v * (out_prec == 64 ? 8.0 : 4.0);
        fl += 2.0 * kTaps * kNet[0].cin * kNet[0].cout * v;
    }
    *bytes = rd + wr;
    *flops = fl;
    return 0;
}

int Net::forward_host(const float* in, int B, const int n[3], float* out)
{
    int rc = ensure(trn_, B, n, true);
    if (rc) return rc;
    const size_t nv = (size_t)B * n[0] * n[1] * n[2];
    rc = hip(hipMemcpyAsync(trn_.in, in, nv * 4, hipMemcpyHostToDevice, stream_), "upload");
    if (rc) return rc;
    forward(trn_, dense_field(trn_.in, n), dense_field(trn_.y, n), nullptr, nullptr);
    rc = hip(hipGetLastError(), "network launch");
    if (!rc) rc = hip(hipMemcpyAsync(out, trn_.y, nv * 4, hipMemcpyDeviceToHost, stream_), "download");
    if (!rc) rc = hip(hipStreamSynchronize(stream_), "sync");
    return rc;
}

int Net::activation_host(int i, float* out)
{
    if (i < 0 || i >= kActs) return fail("activation index %d out of range", i);
    if (!trn_.B) return fail("no forward pass to read from");
    const ActRef a = act(trn_, i);
    const size_t nh = (size_t)a.B * a.D * a.H * a.W * kC;
    std::vector<uint16_t> h(nh);
    int rc = hip(hipMemcpyAsync(h.data(), a.p, nh * 2, hipMemcpyDeviceToHost, stream_), "download");
    if (!rc) rc = hip(hipStreamSynchronize(stream_), "sync");
    if (rc) return rc;
    for (size_t k = 0; k < nh; ++k) {
        const unsigned u = (unsigned)h[k] << 16;
        memcpy(out + k, &u, 4);
    }
    return 0;
}

int Net::loss_grad_host(const float* in, const double* G, int B, const int n[3], double* loss_out, float* grad)
{
    int rc = ensure(trn_, B, n, true);
    if (rc) return rc;
    const long long nv = (long long)n[0] * n[1] * n[2];
    rc = hip(hipMemcpyAsync(trn_.in, in, (size_t)nv * B * 4, hipMemcpyHostToDevice, stream_), "upload");
    if (!rc) rc = hip(hipMemcpyAsync(trn_.g, G, (size_t)B * 8, hipMemcpyHostToDevice, stream_), "upload");
    if (rc) return rc;
    const FieldRef fin = dense_field(trn_.in, n);
    forward(trn_, fin, dense_field(trn_.y, n), nullptr, nullptr);
    loss(trn_, B, nv);
    backward(trn_, fin);
    rc = hip(hipGetLastError(), "network launch");
    std::vector<double> lv(B);
    if (!rc) rc = hip(hipMemcpyAsync(grad, grads_, kParams * sizeof(float), hipMemcpyDeviceToHost, stream_), "download");
    if (!rc) rc = hip(hipMemcpyAsync(lv.data(), trn_.lossv, B * 8, hipMemcpyDeviceToHost, stream_), "download");
    if (!rc) rc = hip(hipStreamSynchronize(stream_), "sync");
    if (rc) return rc;
    *loss_out = std::accumulate(lv.begin(), lv.end(), 0.0) / B;
    return 0;
}

int Net::dataset_reserve(long long npairs, const int n[3])
{
    const long long nv = (long long)n[0] * n[1] * n[2];
    if (data_r_) (void)hipFree(data_r_);
    data_r_ = nullptr;
    int rc = hip(hipMalloc(&data_r_, (size_t)npairs * nv * 4), "hipMalloc training set");
    if (rc) return rc;
    data_n_ = nv;
    data_cap_ = npairs;
    data_pairs_ = npairs;
    data_g_.assign(npairs, 0.0);
    for (int d = 0; d < 3; ++d) data_dims_[d] = n[d];
    return 0;
}

int Net::train(const TrainCfg& c, TrainStats* st)
{
    if (data_pairs_ < 1) return fail("empty training set");
    if (c.batch < 1 || c.steps < 1) return fail("batch and steps must be positive");
    const int B = (int)std::min<long long>(c.batch, data_pairs_);
    int rc = ensure(trn_, B, data_dims_, true);
    if (rc) return rc;
    const long long nv = data_n_;
    std::mt19937_64 rng(c.seed);
    std::vector<long long> perm(data_pairs_);
    std::iota(perm.begin(), perm.end(), 0LL);
    std::shuffle(perm.begin(), perm.end(), rng);
    long long pos = 0;
    std::vector<double> gb(B), lv(B);
    const FieldRef fin = dense_field(trn_.in, data_dims_);
    const FieldRef fout = dense_field(trn_.y, data_dims_);
    const auto t0 = std::chrono::steady_clock::now();
    double first = -1.0, last = -1.0;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int step = 0; step < c.steps; ++step) {
        for (int k = 0; k < B; ++k) {
            if (pos == data_pairs_) {
                std::shuffle(perm.begin(), perm.end(), rng);
                pos = 0;
            }
            const long long i = perm[pos++];
            rc = hip(hipMemcpyAsync(trn_.in + k * nv, data_r_ + i * nv, (size_t)nv * 4, hipMemcpyDeviceToDevice, stream_),
                     "batch copy");
            if (rc) return rc;
            gb[k] = data_g_[i];
        }
        rc = hip(hipMemcpyAsync(trn_.g, gb.data(), (size_t)B * 8, hipMemcpyHostToDevice, stream_), "upload");
        if (rc) return rc;
        forward(trn_, fin, fout, nullptr, nullptr);
        loss(trn_, B, nv);
        backward(trn_, fin);
        ++adam_t_;
        const double frac = c.steps > 1 ? (double)step / (c.steps - 1) : 1.0;
        const float lr = (float)(c.lr_final + 0.5 * (c.lr - c.lr_final) * (1.0 + std::cos(M_PI * frac)));
        const float bc1 = 1.0f - std::pow(b1, (float)adam_t_), bc2 = 1.0f - std::pow(b2, (float)adam_t_);
        hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, stream_, params_, grads_, adam_m_, adam_v_,
                           kParams, lr, b1, b2, eps, bc1, bc2);
        pack_all();
        if (step == 0 || step == c.steps - 1 || (step % 50) == 0) {
            rc = hip(hipMemcpyAsync(lv.data(), trn_.lossv, (size_t)B * 8, hipMemcpyDeviceToHost, stream_), "download");
            if (!rc) rc = hip(hipStreamSynchronize(stream_), "sync");
            if (rc) return rc;
            const double l = std::accumulate(lv.begin(), lv.end(), 0.0) / B;
            if (first < 0) first = l;
            last = l;
            if (!std::isfinite(l)) return fail("training diverged at step %d (loss %g)", step, l);
        }
    }
    rc = hip(hipStreamSynchronize(stream_), "sync");
    if (rc) return rc;
    if (st) {
        st->loss_first = first;
        st->loss_last = last;
        st->steps = c.steps;
        st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return 0;
}

}  // namespace nn
}  // namespace pps
