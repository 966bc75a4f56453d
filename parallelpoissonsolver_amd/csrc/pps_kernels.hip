// pps_kernels.hip -- hand-written gfx950 kernels of the Krylov hot path.
//
// Two kernel skeletons carry every operation:
//  * k_stencil: 2.5-D z-marching tile.  A block of 64x8 threads owns a 64x32
//    column tile (kTX x kTH: each thread four rows) of one x-y slab and marches
//    through kz planes (the longest of 64/32/16/8 that still gives 2048
//    blocks).  Each thread keeps its columns' (k-1, k, k+1) values in
//    registers; the plane-k tile and its one-point halo are staged in LDS
//    (double-buffered, one barrier per plane), so x- and y-neighbours come from
//    LDS and every value of the input field is produced once per plane (this is
//    what lets the CG search direction p = z + beta*p be formed on the fly
//    inside the stencil).  The loads of planes k+1 and k+2 are in flight while
//    plane k is computed, so the barrier never waits on HBM latency.
//    The 7-point operator is the reference's +Laplacian
//    (matrixFreeOperatorA.hpp:33-38) written in difference form,
//    cx((x[i-1]-x)+(x[i+1]-x)) + cy(..) + cz(..), which keeps FP32 accurate
//    for smooth fields.
//  * k_vector: one-shot blocks over consecutive points of an x-row (512 x 1
//    for FP64 operands, 128 threads x 4 points for FP32) for the fused
//    axpy/dot updates: all of a thread's loads are issued before its store,
//    and consecutive blocks stream consecutive addresses.
// A third, k_cheb_pair, runs two Chebyshev sweeps per pass.
//
// Reductions are FP64 in every precision and deterministic: DPP row
// butterflies inside each wavefront, LDS across the wavefronts, one partial
// per block, then a one-block epilogue kernel (after a level of span sums when
// there are many blocks) sums the partials in a fixed order and runs the
// solver's scalar update (alpha, beta, omega, convergence flag) on the device.
// The epilogue runs as its own launch instead of a last-block-arrives scheme:
// on gfx950 the agent-scope release fence that scheme needs in every block
// writes back the XCD's L2, which costs more than the whole reduction.
//
// Blocks are remapped so that consecutive tiles land on the same XCD (the
// hardware deals blocks round-robin over the 8 XCDs, each with its own L2):
// the y-halo rows a tile reads from its neighbour are then L2 hits.
#include "pps_kernels.hpp"

#include <algorithm>
#include <type_traits>

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>

namespace pps {


// tiling of one launch: 64-wide tiles of `th` rows, z-chunks of kz planes
static Geo retile(const Geo& in, int th, int kz)
{
    Geo g = in;
    g.kz = kz > 0 ? kz : 1;
    g.ntx = (g.hi0 - g.lo0 + kTX - 1) / kTX;
    g.nty = (g.hi1 - g.lo1 + th - 1) / th;
    const int ntz = (g.hi2 - g.lo2 + g.kz - 1) / g.kz;
    g.nblocks = g.ntx * g.nty * ntz;
    return g;
}

Geo make_geo(const Layout& L, const Region& R, const double coef[3], int kz_st, int nt)
{
    Geo g{};
    g.plane = L.plane;
    g.px = L.px;
    g.ox = L.ox;
    g.lo0 = R.lo[0]; g.hi0 = R.hi[0];
    g.lo1 = R.lo[1]; g.hi1 = R.hi[1];
    g.lo2 = R.lo[2]; g.hi2 = R.hi[2];
    g.kz_st = kz_st;
    g.nt = nt;
    g.cx = coef[0];
    g.cy = coef[1];
    g.cz = coef[2];
    return retile(g, kTH, kz_st);
}

// stores of the field outputs: non-temporal (NT) when the fields exceed the
// caches.  NT is a template argument so that each kernel body holds one kind
// of store: with a run-time flag the compiler merges the two stores of an
// if/else into one plain store and the non-temporal hint is lost.
template <typename T> struct NoDeduce { using type = T; };
// loads of fields read once per pass (the vector kernels' operands):
// non-temporal in the same mode, so streamed inputs do not evict the caches'
// useful lines (measured at 512^3: FP64 update 5.6-5.7 -> 5.8-6.1 TB/s, FP32
// update 5.4 -> 5.9)
template <bool NT, typename T>
__device__ __forceinline__ T ld(const T* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void put(T* p, typename NoDeduce<T>::type v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// ------------------------------------------------------------ common pieces

__device__ __forceinline__ int xcd_remap(int b, int nb)
{
    const int full = nb & ~7;
    if (b >= full) return b;
    return (b & 7) * (full >> 3) + (b >> 3);
}

// one DPP lane exchange of a double (two 32-bit moves); CTRL is the DPP control
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// sum over the wavefront, the same value in every lane, in a fixed order:
// butterflies inside each row of 16 lanes (quad_perm [1,0,3,2], quad_perm
// [2,3,0,1], row_half_mirror, row_mirror), then the four row sums read out
// and added as (r0 + r1) + (r2 + r3).  DPP moves stay in the VALU, where a
// shuffle is a round trip through the LDS crossbar per step; in a one-shot
// reducing kernel over FP32 fields that round trip was the bottleneck.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_mov<0xB1>(v);
    v += dpp_mov<0x4E>(v);
    v += dpp_mov<0x141>(v);
    v += dpp_mov<0x140>(v);
    const long long b = __double_as_longlong(v);
    double r[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int lo = __builtin_amdgcn_readlane((int)b, 16 * q);
        const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 16 * q);
        r[q] = __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    }
    return (r[0] + r[1]) + (r[2] + r[3]);
}

// sum over the block in a fixed order; result valid in thread 0
template <int NQ, int NTHREADS = kThreads>
__device__ __forceinline__ void block_sum(double (&v)[NQ], double (&tot)[NQ])
{
    constexpr int NW = NTHREADS / 64;
    __shared__ double sm[NQ][NW];
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    const int lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double t = wave_sum(v[q]);
        if (lane == 0) sm[q][w] = t;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double t = 0.0;
#pragma unroll
        for (int ww = 0; ww < NW; ++ww) t += sm[q][ww];
        tot[q] = t;
    }
}

// the block's sums as partial number `slot` of the launch
template <int NQ, int NTHREADS = kThreads>
__device__ __forceinline__ void block_partials(double (&v)[NQ], const RedBuf& rb, int slot)
{
    double tot[NQ];
    block_sum<NQ, NTHREADS>(v, tot);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) rb.partials[(size_t)q * rb.cap + slot] = tot[q];
    }
}

// one block: sums the partials of the preceding launch and runs op.finish.
// Thread t owns partials t, t+1024, ... summed in that order with two
// interleaved accumulators (more loads in flight), then a fixed tree.
constexpr int kFinThreads = 1024;
template <int NQ, class Op>
__global__ __launch_bounds__(kFinThreads) void k_finish(Op op, RedBuf rb, int nblocks)
{
    if (op.gated()) return;
    double a0[NQ], a1[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) a0[q] = a1[q] = 0.0;
    int b = threadIdx.x;
    for (; b + kFinThreads < nblocks; b += 2 * kFinThreads) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            a0[q] += rb.partials[(size_t)q * rb.cap + b];
            a1[q] += rb.partials[(size_t)q * rb.cap + b + kFinThreads];
        }
    }
    if (b < nblocks) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) a0[q] += rb.partials[(size_t)q * rb.cap + b];
    }
    double acc[NQ], tot[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = a0[q] + a1[q];
    block_sum<NQ, kFinThreads>(acc, tot);
    if (threadIdx.x == 0) op.finish(tot);
}

struct Gate {
    const int* gate;
    __device__ __forceinline__ bool gated() const { return gate != nullptr && *gate != 0; }
};

// record the end of one iteration of a main/inner Krylov loop
__device__ __forceinline__ void finish_iteration(Scal* S, double err, bool history)
{
    S->v[S_ERR] = err;
    const int it = S->iter + 1;
    S->iter = it;
    if (history && S->hist != nullptr && it < S->hist_cap) S->hist[it] = err;
    if (err < S->v[S_TOL] || it >= S->maxit) S->done = 1;
}

// --------------------------------------------------------- kernel skeletons

// Stencil op hooks.  A plain op has in(c) (the staged value, a plain load) and
// out(c, v, Av, acc).  An op whose staged value is computed from several
// fields, or whose output reads more fields, declares
//   struct Raw; Raw in_raw(c); T in_value(const Raw&)   -- loads, and their combination
//   struct Ox;  Ox out_raw(c)                            -- the output's extra loads
//   out(c, v, Av, const Raw& centre, const Ox&, acc)
// (Raw defaults to T, Ox to NoRaw).  The kernel issues the raw loads planes
// ahead and combines them only when needed, so no load is waited on as soon
// as it is issued.
struct NoRaw {};
template <class Op, typename T, class = void> struct RawOf { using type = T; };
template <class Op, typename T> struct RawOf<Op, T, std::void_t<typename Op::Raw>> { using type = typename Op::Raw; };
template <class Op, class = void> struct OxOf { using type = NoRaw; };
template <class Op> struct OxOf<Op, std::void_t<typename Op::Ox>> { using type = typename Op::Ox; };
template <class Op, typename T>
constexpr bool kHooks = !std::is_same_v<typename RawOf<Op, T>::type, T> || !std::is_same_v<typename OxOf<Op>::type, NoRaw>;

template <class Op, typename T>
__device__ __forceinline__ typename RawOf<Op, T>::type st_raw(const Op& op, long long c)
{
    if constexpr (std::is_same_v<typename RawOf<Op, T>::type, T>) return op.in(c);
    else return op.in_raw(c);
}
template <class Op, typename T>
__device__ __forceinline__ T st_value(const Op& op, const typename RawOf<Op, T>::type& r)
{
    if constexpr (std::is_same_v<typename RawOf<Op, T>::type, T>) return r;
    else return op.in_value(r);
}
// the output's extra operands are read once per pass: non-temporal loads in
// the streaming mode (the staged field's loads stay cached: neighbouring tiles
// re-read its halo)
template <bool NT, class Op>
__device__ __forceinline__ typename OxOf<Op>::type st_ox(const Op& op, long long c)
{
    if constexpr (std::is_same_v<typename OxOf<Op>::type, NoRaw>) return NoRaw{};
    else return op.template out_raw<NT>(c);
}
template <bool NT, class Op, typename T>
__device__ __forceinline__ void st_out(const Op& op, long long c, T v, T Av, const typename RawOf<Op, T>::type& r,
                                       const typename OxOf<Op>::type& o, double* acc)
{
    if constexpr (kHooks<Op, T>) op.template out<NT>(c, v, Av, r, o, acc);
    else op.template out<NT>(c, v, Av, acc);
}

template <typename T, int NQ, class Op, bool NT>
__global__ __launch_bounds__(kSThreads) void k_stencil(Geo g, Op op, RedBuf rb)
{
    if (op.gated()) return;
    op.load();
    // tile: 64 columns x kTH rows; thread (lx, ly) owns rows ly, ly+kSTY, ...
    __shared__ T lds[2][kTH + 2][kTX + 2];
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int tx0 = b % g.ntx;
    const int rr = b / g.ntx;
    const int ty0 = rr % g.nty;
    const int tz0 = rr / g.nty;
    const int lx = threadIdx.x, ly = threadIdx.y, tid = lx + kTX * ly;
    const int i = g.lo0 + tx0 * kTX + lx;
    const int jt = g.lo1 + ty0 * kTH;  // first row of the tile
    const int k0 = g.lo2 + tz0 * g.kz;
    const int k1 = min(k0 + g.kz, g.hi2);

    // halo producers: 128 threads for the two y-halo rows, 2*kTH for the x-halo columns
    static_assert(2 * kTX + 2 * kTH <= kSThreads, "one thread per halo point of a plane");
    int hr = 0, hc_ = 0, hi = 0, hj = 0;
    bool hv = false;
    if (tid < 2 * kTX) {
        const int col = tid % kTX;
        const bool top = tid >= kTX;
        hr = top ? kTH + 1 : 0;
        hc_ = col + 1;
        hi = g.lo0 + tx0 * kTX + col;
        hj = jt + (top ? kTH : -1);
        hv = true;
    } else if (tid < 2 * kTX + 2 * kTH) {
        const int t = tid - 2 * kTX;
        const int row = t % kTH;
        const bool right = t >= kTH;
        hr = row + 1;
        hc_ = right ? kTX + 1 : 0;
        hi = g.lo0 + tx0 * kTX + (right ? kTX : -1);
        hj = jt + row;
        hv = true;
    }
    const bool hin = hv && hi <= g.hi0 && hj <= g.hi1;
    const long long hbase = (long long)g.ox + hi + (long long)g.px * hj;
    const T cx = (T)g.cx, cy = (T)g.cy, cz = (T)g.cz;

    using Raw = typename RawOf<Op, T>::type;
    using Ox = typename OxOf<Op>::type;
    bool act[kRY], inb[kRY];
    long long base[kRY];
    // per row q: values of planes k-1, k, k+1; raw loads of planes k (centre,
    // for the output), k+1 and k+2; the output's extra loads for planes k, k+1
    T vm[kRY], vc[kRY], vp[kRY];
    Raw rc[kRY], r1[kRY], r2[kRY];
    Ox oc[kRY], on[kRY];
#pragma unroll
    for (int q = 0; q < kRY; ++q) {
        const int j = jt + ly + q * kSTY;
        act[q] = (i < g.hi0) && (j < g.hi1);
        inb[q] = (i <= g.hi0) && (j <= g.hi1);
        base[q] = (long long)g.ox + i + (long long)g.px * j;
        vm[q] = inb[q] ? st_value<Op, T>(op, st_raw<Op, T>(op, base[q] + g.plane * (k0 - 1))) : T(0);
        rc[q] = inb[q] ? st_raw<Op, T>(op, base[q] + g.plane * k0) : Raw{};
        r1[q] = inb[q] ? st_raw<Op, T>(op, base[q] + g.plane * (k0 + 1)) : Raw{};
        if (act[q]) oc[q] = st_ox<NT>(op, base[q] + g.plane * k0);
        vc[q] = inb[q] ? st_value<Op, T>(op, rc[q]) : T(0);
    }
    double acc[NQ > 0 ? NQ : 1];
#pragma unroll
    for (int q = 0; q < (NQ > 0 ? NQ : 1); ++q) acc[q] = 0.0;

    Raw hc = hin ? st_raw<Op, T>(op, hbase + g.plane * k0) : Raw{};
    Raw h1 = (hin && k0 + 1 < k1) ? st_raw<Op, T>(op, hbase + g.plane * (k0 + 1)) : Raw{};
    int buf = 0;
    for (int k = k0; k < k1; ++k) {
        // loads first: plane k+2 (used from the next step on), the output's
        // loads of plane k+1, the halo of plane k+2
        const bool more = k + 1 < k1;
#pragma unroll
        for (int q = 0; q < kRY; ++q) {
            if (inb[q] && more) r2[q] = st_raw<Op, T>(op, base[q] + g.plane * (k + 2));
            if (act[q] && more) on[q] = st_ox<NT>(op, base[q] + g.plane * (k + 1));
        }
        Raw h2{};
        if (hin && k + 2 < k1) h2 = st_raw<Op, T>(op, hbase + g.plane * (k + 2));
        // plane k+1's values: its loads were issued one step ago
#pragma unroll
        for (int q = 0; q < kRY; ++q) vp[q] = inb[q] ? st_value<Op, T>(op, r1[q]) : T(0);
#pragma unroll
        for (int q = 0; q < kRY; ++q) lds[buf][ly + q * kSTY + 1][lx + 1] = vc[q];
        if (hv) lds[buf][hr][hc_] = hin ? st_value<Op, T>(op, hc) : T(0);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kRY; ++q) {
            if (act[q]) {
                const int r = ly + q * kSTY + 1;
                const T v = vc[q];
                const T Av = cx * ((lds[buf][r][lx] - v) + (lds[buf][r][lx + 2] - v)) +
                             cy * ((lds[buf][r - 1][lx + 1] - v) + (lds[buf][r + 1][lx + 1] - v)) +
                             cz * ((vm[q] - v) + (vp[q] - v));
                st_out<NT, Op, T>(op, base[q] + g.plane * k, v, Av, rc[q], oc[q], acc);
            }
            vm[q] = vc[q];
            vc[q] = vp[q];
            rc[q] = r1[q];
            r1[q] = r2[q];
            oc[q] = on[q];
        }
        hc = h1;
        h1 = h2;
        buf ^= 1;
    }
    if constexpr (NQ > 0) block_partials<NQ, kSThreads>(acc, rb, (int)blockIdx.x);
}

// Vector kernels: one-shot blocks on a 3-D grid (segment of an x-row, y, z).
// A block of BT threads covers BT * E consecutive points of one row, thread t
// taking points t, t + BT, ... (coalesced for every e); all of a thread's
// loads are issued before its first store, and consecutive blocks cover
// consecutive addresses, so the blocks in flight stream one compact window of
// memory.  The grid coordinates are the row and segment, so a block spends no
// instructions finding them.  With a few hundred bytes per wave, the per-wave
// prologue, the bytes in flight per thread and the block reduction are what
// separate these kernels from a flat copy: ops that fetch 16 bytes or more per
// point run 512 x 1, ops that fetch 8 bytes or less (FP32 fields) 128 x 4.
// (Measured on a 512^3 update r -= a Ap with r.r and r.(jac r) partials:
// FP64 256 x 1 5.6 TB/s, 512 x 1 5.8, where z-marching 64 x 4 x 32 tiles
// reached 5.1-5.35; FP32 128 x 4 5.2 TB/s where 256 x 1 reached 3.9.)
template <class Op>
struct VecShape {
    static constexpr int E = sizeof(typename Op::R) <= 8 ? 4 : 1;
    static constexpr int BT = E == 1 ? 512 : 128;
};

// segment `seg` of row (j, k): the block's BT * E points
template <bool NT, class Op>
__device__ __forceinline__ void vector_segment(const Geo& g, const Op& op, int seg, int j, int k, double* acc)
{
    constexpr int E = VecShape<Op>::E, BT = VecShape<Op>::BT;
    const int i0 = g.lo0 + seg * (BT * E) + (int)threadIdx.x;
    const long long row = (long long)g.ox + (long long)g.px * j + g.plane * k;
    typename Op::R r[E];
#pragma unroll
    for (int e = 0; e < E; ++e)
        if (i0 + e * BT < g.hi0) r[e] = op.template fetch<NT>(row + i0 + e * BT);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = i0 + e * BT;
        if (i < g.hi0) op.template apply<NT>(row + i, r[e], i, j, k, acc);
    }
}

template <int NQ, class Op, bool NT>
__global__ __launch_bounds__(VecShape<Op>::BT) void k_vector(Geo g, Op op, RedBuf rb)
{
    if (op.gated()) return;
    op.load();
    double acc[NQ > 0 ? NQ : 1];
#pragma unroll
    for (int q = 0; q < (NQ > 0 ? NQ : 1); ++q) acc[q] = 0.0;
    vector_segment<NT>(g, op, (int)blockIdx.x, g.lo1 + (int)blockIdx.y, g.lo2 + (int)blockIdx.z, acc);
    if constexpr (NQ > 0)
        block_partials<NQ, VecShape<Op>::BT>(acc, rb,
                                             (int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)));
}

// The same pass on a fixed grid of at most kStridedBlocks blocks, each taking
// segments blockIdx.x, blockIdx.x + gridDim.x, ...: for passes that are
// usually gated off (a gated one-shot launch still dispatches one block per
// segment, ~0.1 ms at 512^3 for nothing), at some cost when they do run.
constexpr int kStridedBlocks = 2048;
template <int NQ, class Op, bool NT>
__global__ __launch_bounds__(VecShape<Op>::BT) void k_vector_strided(Geo g, Op op, RedBuf rb, int nseg, int ny,
                                                                     long long nb)
{
    if (op.gated()) return;
    op.load();
    double acc[NQ > 0 ? NQ : 1];
#pragma unroll
    for (int q = 0; q < (NQ > 0 ? NQ : 1); ++q) acc[q] = 0.0;
    for (long long b = blockIdx.x; b < nb; b += gridDim.x) {
        const long long rw = b / nseg;
        vector_segment<NT>(g, op, (int)(b - rw * nseg), g.lo1 + (int)(rw % ny), g.lo2 + (int)(rw / ny), acc);
    }
    if constexpr (NQ > 0) block_partials<NQ, VecShape<Op>::BT>(acc, rb, (int)blockIdx.x);
}

// sums of kSpan consecutive block partials, one block per span, fixed order
template <int NQ, class Op>
__global__ __launch_bounds__(kThreads) void k_span(Op op, RedBuf rb, int nblocks)
{
    if (op.gated()) return;
    const int b0 = blockIdx.x * kSpan, b1 = min(nblocks, b0 + kSpan);
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    for (int b = b0 + (int)threadIdx.x; b < b1; b += kThreads) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[q] += rb.partials[(size_t)q * rb.cap + b];
    }
    double tot[NQ];
    block_sum<NQ, kThreads>(acc, tot);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) rb.partials2[(size_t)q * rb.cap2 + blockIdx.x] = tot[q];
    }
}

// the epilogue of a reducing launch of nblocks blocks: one level of span sums
// first when there are many partials and the workspace has room for them
template <int NQ, class Op>
static void launch_finish(hipStream_t s, const Op& op, RedBuf rb, int nblocks)
{
    const int nspan = (nblocks + kSpan - 1) / kSpan;
    if (nblocks > 4 * kSpan && rb.partials2 != nullptr && nspan <= rb.cap2) {
        hipLaunchKernelGGL((k_span<NQ, Op>), dim3(nspan), dim3(kThreads), 0, s, op, rb, nblocks);
        RedBuf r2 = rb;
        r2.partials = rb.partials2;
        r2.cap = rb.cap2;
        hipLaunchKernelGGL((k_finish<NQ, Op>), dim3(1), dim3(kFinThreads), 0, s, op, r2, nspan);
    } else {
        hipLaunchKernelGGL((k_finish<NQ, Op>), dim3(1), dim3(kFinThreads), 0, s, op, rb, nblocks);
    }
}

template <int NQ, class Op>
static bool check_capacity(const Geo& g, const RedBuf& rb)
{
    if (NQ > 0 && g.nblocks > rb.cap) {
        std::fprintf(stderr, "pps: %d blocks exceed the reduction workspace (%d); launch skipped\n", g.nblocks, rb.cap);
        return false;
    }
    return true;
}

// tiling of a stencil launch: the z-chunk set on the context, or (kz_st = 0)
// the longest of 64, 32, 16, 8 planes that still gives 2048 blocks (a longer
// chunk re-reads fewer halo planes; 2048 blocks keep every CU busy to the end)
static Geo stencil_geo(const Geo& g0)
{
    if (g0.kz_st > 0) return retile(g0, kTH, g0.kz_st);
    int kz = 64;
    Geo g = retile(g0, kTH, kz);
    while (kz > 8 && g.nblocks < 2048) {
        kz /= 2;
        g = retile(g0, kTH, kz);
    }
    return g;
}

template <typename T, int NQ, class Op>
static void run_stencil(const Geo& g0, hipStream_t s, const Op& op, RedBuf rb)
{
    const Geo g = stencil_geo(g0);
    if (g.nblocks <= 0 || !check_capacity<NQ, Op>(g, rb)) return;
    if (g.nt) hipLaunchKernelGGL((k_stencil<T, NQ, Op, true>), dim3(g.nblocks), dim3(kTX, kSTY), 0, s, g, op, rb);
    else hipLaunchKernelGGL((k_stencil<T, NQ, Op, false>), dim3(g.nblocks), dim3(kTX, kSTY), 0, s, g, op, rb);
    if constexpr (NQ > 0) launch_finish<NQ>(s, op, rb, g.nblocks);
}

template <int NQ, class Op>
static void run_vector(const Geo& g0, hipStream_t s, const Op& op, RedBuf rb, bool strided = false)
{
    const int nx = g0.hi0 - g0.lo0, ny = g0.hi1 - g0.lo1, nz = g0.hi2 - g0.lo2;
    if (nx <= 0 || ny <= 0 || nz <= 0) return;
    constexpr int W = VecShape<Op>::BT * VecShape<Op>::E;
    const int nseg = (nx + W - 1) / W;
    const long long nb = (long long)nseg * ny * nz;
    if (nb > 0x7FFFFFFFll || ny > 65535 || nz > 65535) {
        std::fprintf(stderr, "pps: a %d x %d x %d region exceeds a vector launch; launch skipped\n", nx, ny, nz);
        return;
    }
    Geo g = g0;
    if (strided) {
        g.nblocks = (int)std::min<long long>(nb, kStridedBlocks);
        if (!check_capacity<NQ, Op>(g, rb)) return;
        const dim3 grid(g.nblocks), block(VecShape<Op>::BT);
        if (g.nt) hipLaunchKernelGGL((k_vector_strided<NQ, Op, true>), grid, block, 0, s, g, op, rb, nseg, ny, nb);
        else hipLaunchKernelGGL((k_vector_strided<NQ, Op, false>), grid, block, 0, s, g, op, rb, nseg, ny, nb);
        if constexpr (NQ > 0) launch_finish<NQ>(s, op, rb, g.nblocks);
        return;
    }
    g.nblocks = (int)nb;
    if (!check_capacity<NQ, Op>(g, rb)) return;
    const dim3 grid(nseg, ny, nz);
    const dim3 block(VecShape<Op>::BT);
    if (g.nt) hipLaunchKernelGGL((k_vector<NQ, Op, true>), grid, block, 0, s, g, op, rb);
    else hipLaunchKernelGGL((k_vector<NQ, Op, false>), grid, block, 0, s, g, op, rb);
    if constexpr (NQ > 0) launch_finish<NQ>(s, op, rb, g.nblocks);
}

static const RedBuf kNoRed{nullptr, nullptr, 0, nullptr, 0};

struct Empty {};

// ------------------------------------------------------------ basic kernels

template <typename T>
struct OpApply : Gate {
    const T* __restrict__ x;
    T* __restrict__ y;
    __device__ void load() {}
    __device__ T in(long long c) const { return x[c]; }
    template <bool NT> __device__ void out(long long c, T, T Av, double*) const { put<NT>(y + c, Av); }
    __device__ void finish(const double*) {}
};

template <typename T>
void launch_apply_A(const Geo& g, hipStream_t s, const T* x, T* y)
{
    OpApply<T> op;
    op.gate = nullptr;
    op.x = x;
    op.y = y;
    run_stencil<T, 0>(g, s, op, kNoRed);
}

template <typename T>
struct OpCopy : Gate {
    const T* __restrict__ x;
    T* __restrict__ y;
    struct R { T a; };
    __device__ void load() {}
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(x + c)}; }
    template <bool NT> __device__ void apply(long long c, const R& r, int, int, int, double*) const { put<NT>(y + c, r.a); }
    __device__ void finish(const double*) {}
};

template <typename T>
void launch_copy(const Geo& g, hipStream_t s, const T* x, T* y)
{
    OpCopy<T> op;
    op.gate = nullptr;
    op.x = x;
    op.y = y;
    run_vector<0>(g, s, op, kNoRed);
}

// plain copy of n contiguous doubles, 16 bytes per lane with non-temporal
// loads and stores: the HBM ceiling the field kernels are measured against
typedef double d2_t __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_copy_flat(const double* __restrict__ x, double* __restrict__ y, long long n)
{
    const long long i = ((long long)blockIdx.x * 256 + threadIdx.x) * 2;
    if (i + 1 < n) {
        const d2_t v = __builtin_nontemporal_load(reinterpret_cast<const d2_t*>(x + i));
        __builtin_nontemporal_store(v, reinterpret_cast<d2_t*>(y + i));
    } else if (i < n) {
        y[i] = x[i];
    }
}

void launch_copy_flat(hipStream_t s, const double* x, double* y, long long n)
{
    const long long blocks = (n + 511) / 512;
    if (blocks > 0) hipLaunchKernelGGL(k_copy_flat, dim3((unsigned)blocks), dim3(256), 0, s, x, y, n);
}

template <typename T>
struct OpDot : Gate {
    const T* __restrict__ a;
    const T* __restrict__ b;
    double* out;
    struct R { T a, b; };
    __device__ void load() {}
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(a + c), ld<NT>(b + c)}; }
    template <bool NT> __device__ void apply(long long, const R& r, int, int, int, double* acc) const { acc[0] += (double)r.a * (double)r.b; }
    __device__ void finish(const double* t) { *out = t[0]; }
};

template <typename T>
void launch_dot(const Geo& g, hipStream_t s, RedBuf rb, const T* a, const T* b, double* out, const int* gate)
{
    OpDot<T> op;
    op.gate = gate;
    op.a = a;
    op.b = b;
    op.out = out;
    run_vector<1>(g, s, op, rb);
}

template <typename T>
struct OpFill : Gate {
    T* __restrict__ a;
    T v;
    using R = Empty;
    __device__ void load() {}
    template <bool NT> __device__ R fetch(long long) const { return {}; }
    template <bool NT> __device__ void apply(long long c, const R&, int, int, int, double*) const { put<NT>(a + c, v); }
    __device__ void finish(const double*) {}
};

template <typename T>
void launch_fill(const Geo& g, hipStream_t s, T* a, T v)
{
    OpFill<T> op;
    op.gate = nullptr;
    op.a = a;
    op.v = v;
    run_vector<0>(g, s, op, kNoRed);
}

template <typename T>
struct OpScaleBox : Gate {
    T* __restrict__ a;
    const Scal* S;
    int mode;
    double nrm;
    struct R { T a; };
    __device__ void load() { nrm = S->v[S_NORM]; }
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(a + c)}; }
    template <bool NT> __device__ void apply(long long c, const R& r, int, int, int, double*) const
    {
        // normalizeProblemToFieldBNorm divides (iterativeSolverBase.hpp:227-231);
        // the solvers' epilogues multiply (baseCG.hpp:248-252)
        if (mode == 0) put<NT>(a + c, (T)((double)r.a / nrm));
        else put<NT>(a + c, (T)((double)r.a * nrm));
    }
    __device__ void finish(const double*) {}
};

template <typename T>
void launch_scale_box(const Geo& gbox, hipStream_t s, T* a, const Scal* S, int mode)
{
    OpScaleBox<T> op;
    op.gate = nullptr;
    op.a = a;
    op.S = S;
    op.mode = mode;
    run_vector<0>(gbox, s, op, kNoRed);
}

template <typename T>
struct OpPointwise : Gate {
    const T* __restrict__ in;
    T* __restrict__ out;
    T c;
    struct R { T a; };
    __device__ void load() {}
    template <bool NT> __device__ R fetch(long long p) const { return {ld<NT>(in + p)}; }
    template <bool NT> __device__ void apply(long long p, const R& r, int, int, int, double*) const { put<NT>(out + p, c * r.a); }
    __device__ void finish(const double*) {}
};

template <typename T>
void launch_pointwise(const Geo& gbox, hipStream_t s, const T* in, T* out, double c, const int* gate)
{
    OpPointwise<T> op;
    op.gate = gate;
    op.in = in;
    op.out = out;
    op.c = (T)c;
    run_vector<0>(gbox, s, op, kNoRed);
}

template <typename TS, typename TD>
struct OpCast : Gate {
    const TS* __restrict__ src;
    TD* __restrict__ dst;
    struct R { TS a; };
    __device__ void load() {}
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(src + c)}; }
    template <bool NT> __device__ void apply(long long c, const R& r, int, int, int, double*) const { put<NT>(dst + c, (TD)r.a); }
    __device__ void finish(const double*) {}
};

template <typename TS, typename TD>
void launch_cast(const Geo& gbox, hipStream_t s, const TS* src, TD* dst, const int* gate)
{
    OpCast<TS, TD> op;
    op.gate = gate;
    op.src = src;
    op.dst = dst;
    run_vector<0>(gbox, s, op, kNoRed);
}

// ---------------------------------------------------- residual and b-norm

// r = b - A x (iterativeSolverBase.hpp:257-268)
struct OpResidual : Gate {
    const double* __restrict__ x;
    const double* __restrict__ b;
    double* __restrict__ r;
    Scal* S;
    int mode;
    double jac;
    __device__ void load() { jac = S->v[S_JAC]; }
    struct Ox { double b; };
    __device__ double in(long long c) const { return x[c]; }
    template <bool NT> __device__ Ox out_raw(long long c) const { return {ld<NT>(b + c)}; }
    template <bool NT> __device__ void out(long long c, double, double Av, const double&, const Ox& o, double* acc) const
    {
        const double rv = o.b - Av;
        put<NT>(r + c, rv);
        acc[0] += rv * rv;
        acc[1] += rv * (jac * rv);
    }
    __device__ void finish(const double* t)
    {
        const double e = sqrt(t[0]);
        if (mode == 0) {
            S->v[S_ERR0] = e;
            S->v[S_ERR] = e;
            S->v[S_ERR_OP] = e;
            S->v[S_RZ] = t[1];
            S->v[S_RR] = t[0];
            S->v[S_BETA] = 0.0;
            S->v[S_RHO0] = 1.0;
            S->v[S_ALPHA] = 0.0;  // no pending x update before the first iteration
            S->iter = 0;
            if (S->hist != nullptr && S->hist_cap > 0) S->hist[0] = e;
            S->done = (e < S->v[S_TOL]) ? 2 : 0;
        } else if (mode == 1) {
            S->v[S_ERR_OP] = e;
        } else {
            // residual replacement after a recursive stop: r now holds the true
            // residual of the complete x; the loop resumes from it unless it is
            // below the tolerance or the iteration cap is reached
            S->v[S_ERR_OP] = e;
            S->v[S_RR] = t[0];
            S->v[S_ALPHA] = 0.0;  // x is complete: no pending update
            if (mode == 3) S->v[S_RZ] = t[1];   // CG, pointwise preconditioner: r.(jac r) of the new r
            if (mode == 4) S->v[S_RHO0] = t[0]; // BiCGSTAB restarted with r0 = r: rho = r0.r
            S->done = (e < S->v[S_TOL] || S->iter >= S->maxit) ? 1 : 0;
        }
    }
};

void launch_residual(const Geo& g, hipStream_t s, RedBuf rb, const double* x, const double* b, double* r, Scal* S, int mode)
{
    OpResidual op;
    op.gate = nullptr;
    op.x = x;
    op.b = b;
    op.r = r;
    op.S = S;
    op.mode = mode;
    run_stencil<double, 2>(g, s, op, rb);
}

// r = b - A x and ||r|| into one device double: the residual above without a
// solver's scalar block, for callers that bring their own fields and partials
struct OpResidualNorm : Gate {
    const double* __restrict__ x;
    const double* __restrict__ b;
    double* __restrict__ r;
    double* norm;
    __device__ void load() {}
    struct Ox { double b; };
    __device__ double in(long long c) const { return x[c]; }
    template <bool NT> __device__ Ox out_raw(long long c) const { return {ld<NT>(b + c)}; }
    template <bool NT> __device__ void out(long long c, double, double Av, const double&, const Ox& o, double* acc) const
    {
        const double rv = o.b - Av;
        put<NT>(r + c, rv);
        acc[0] += rv * rv;
    }
    __device__ void finish(const double* t) { *norm = sqrt(t[0]); }
};

int launch_residual_norm(const Geo& g, hipStream_t s, RedBuf rb, const double* x, const double* b, double* r, double* out)
{
    const Geo gt = stencil_geo(g);
    if (gt.nblocks > rb.cap) return -1;
    OpResidualNorm op;
    op.gate = nullptr;
    op.x = x;
    op.b = b;
    op.r = r;
    op.norm = out;
    run_stencil<double, 1>(g, s, op, rb);
    return 0;
}

// boundary adjustment of b at one point (adjustFieldBForDirichletNeumanBCs,
// iterativeSolverBase.hpp:429-534): Dirichlet faces subtract the boundary value
// of x over ds^2 on the first interior layer, Neumann faces add -+ 2 du/dn / ds
// on the boundary layer; applied in the reference's face order x- x+ y- y+ z- z+
struct BAdjust {
    const double* __restrict__ x;
    const double* flux[6];
    int bc[6], ld[6], fng[6][2];
    double ds[3];
    long long stride[3];
    __device__ long long fpos(int d, int i, int j, int k) const
    {
        const int f = 2 * d;
        if (d == 0) return j + (long long)fng[f][0] * k;
        if (d == 1) return i + (long long)fng[f][0] * k;
        return i + (long long)fng[f][0] * j;
    }
    __device__ double adjust(double v, long long c, int i, int j, int k) const
    {
        const int idx[3] = {i, j, k};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double dd = ds[d] * ds[d];
            if (bc[2 * d] == 0) {
                if (idx[d] == ld[2 * d] + 1) v -= x[c - stride[d]] / dd;
            } else if (bc[2 * d] == 1 && flux[2 * d] != nullptr) {
                if (idx[d] == ld[2 * d]) v += 2 * flux[2 * d][fpos(d, i, j, k)] / ds[d];
            }
            if (bc[2 * d + 1] == 0) {
                if (idx[d] == ld[2 * d + 1] - 2) v -= x[c + stride[d]] / dd;
            } else if (bc[2 * d + 1] == 1 && flux[2 * d + 1] != nullptr) {
                if (idx[d] == ld[2 * d + 1] - 1) v -= 2 * flux[2 * d + 1][fpos(d, i, j, k)] / ds[d];
            }
        }
        return v;
    }
};

static BAdjust make_badjust(const Geo& g, const double* x, const int bc[6], const int lim_data[6],
                            const double* const flux[6], const int fng[6][2], const double ds[3])
{
    BAdjust a;
    a.x = x;
    for (int f = 0; f < 6; ++f) {
        a.flux[f] = flux[f];
        a.bc[f] = bc[f];
        a.ld[f] = lim_data[f];
        a.fng[f][0] = fng[f][0];
        a.fng[f][1] = fng[f][1];
    }
    for (int d = 0; d < 3; ++d) a.ds[d] = ds[d];
    a.stride[0] = 1;
    a.stride[1] = g.px;
    a.stride[2] = g.plane;
    return a;
}

// ||b_adj||^2 over the solver region (iterativeSolverBase.hpp:171-201)
struct OpBNorm : Gate {
    const double* __restrict__ b;
    BAdjust adj;
    Scal* S;
    struct R { double b; };
    __device__ void load() {}
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(b + c)}; }
    template <bool NT> __device__ void apply(long long c, const R& r, int i, int j, int k, double* acc) const
    {
        const double v = adj.adjust(r.b, c, i, j, k);
        acc[0] += v * v;
    }
    __device__ void finish(const double* t) { S->v[S_NORM] = sqrt(t[0]); }
};

void launch_bnorm(const Geo& g, hipStream_t s, RedBuf rb, const double* x, const double* b, const int bc[6],
                  const int lim_data[6], const int lim_solver[6], const double* const flux[6], const int fng[6][2],
                  const double ds[3], Scal* S)
{
    (void)lim_solver;
    OpBNorm op;
    op.gate = nullptr;
    op.b = b;
    op.adj = make_badjust(g, x, bc, lim_data, flux, fng, ds);
    op.S = S;
    run_vector<1>(g, s, op, rb);
}

// the adjusted right-hand side written in place over the data region (the
// main-loop Chebyshev works on this copy; chebyshevIteration.hpp:61-67)
struct OpAdjustB : Gate {
    double* __restrict__ b;
    BAdjust adj;
    struct R { double b; };
    __device__ void load() {}
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(b + c)}; }
    template <bool NT> __device__ void apply(long long c, const R& r, int i, int j, int k, double*) const { put<NT>(b + c, adj.adjust(r.b, c, i, j, k)); }
    __device__ void finish(const double*) {}
};

void adjust_b_field(const Geo& gdata, hipStream_t s, const double* x, double* b, const int bc[6], const int lim_data[6],
                    const double* const flux[6], const int fng[6][2], const double ds[3])
{
    OpAdjustB op;
    op.gate = nullptr;
    op.b = b;
    op.adj = make_badjust(gdata, x, bc, lim_data, flux, fng, ds);
    run_vector<0>(gdata, s, op, kNoRed);
}

// ------------------------------------------------------- Neumann ghost fill

// iterativeSolverBase.hpp:62-169 (orderNeumanBcs == 2): the guard point of a
// Neumann face takes the value of the data point mirrored through the face
// point; data fields add -+ 2 ds du/dn / normFieldB_.
template <typename T>
__global__ __launch_bounds__(256) void k_ghost(Layout L, Faces F, T* field, int with_flux, const Scal* S, const int* gate)
{
    if (gate != nullptr && *gate != 0) return;
    const int slot = blockIdx.y;
    const int f = F.face[slot];
    const int d = f >> 1, side = f & 1;
    const int d1 = d == 0 ? 1 : 0, d2 = d == 2 ? 1 : 2;
    const int n1 = F.lim_data[2 * d1 + 1] - F.lim_data[2 * d1];
    const int n2 = F.lim_data[2 * d2 + 1] - F.lim_data[2 * d2];
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n1 * n2) return;
    int idx[3];
    idx[d1] = F.lim_data[2 * d1] + (int)(t % n1);
    idx[d2] = F.lim_data[2 * d2] + (int)(t / n1);
    idx[d] = side == 0 ? F.lim_data[2 * d] : F.lim_data[2 * d + 1] - 1;
    const long long e = d == 0 ? 1 : (d == 1 ? (long long)L.px : L.plane);
    const long long p = L.at(idx[0], idx[1], idx[2]);
    const long long guard = side == 0 ? p - e : p + e;
    const long long data = side == 0 ? p + e : p - e;
    if (with_flux && F.flux[slot] != nullptr) {
        const double nrm = S != nullptr ? S->v[S_NORM] : 1.0;
        const double q = F.flux[slot][idx[d1] + (long long)F.fng[f][0] * idx[d2]];
        const double v = (double)field[data];
        field[guard] = (T)(side == 0 ? v - 2 * F.ds[d] * q / nrm : v + 2 * F.ds[d] * q / nrm);
    } else {
        field[guard] = field[data];
    }
}

template <typename T>
void launch_ghost(hipStream_t s, const Layout& L, const Faces& F, T* field, int with_flux, const Scal* S, const int* gate)
{
    if (F.n == 0) return;
    int maxpts = 0;
    for (int q = 0; q < F.n; ++q) {
        const int f = F.face[q], d = f >> 1;
        const int d1 = d == 0 ? 1 : 0, d2 = d == 2 ? 1 : 2;
        const int pts = (F.lim_data[2 * d1 + 1] - F.lim_data[2 * d1]) * (F.lim_data[2 * d2 + 1] - F.lim_data[2 * d2]);
        if (pts > maxpts) maxpts = pts;
    }
    dim3 grid((maxpts + 255) / 256, F.n);
    hipLaunchKernelGGL(k_ghost<T>, grid, dim3(256), 0, s, L, F, field, with_flux, S, gate);
}

// ------------------------------------------------------------------- CG

// p_new = a*src + beta*p_old, Ap = A p_new, p.Ap ; alpha = rz / pAp  (baseCG.hpp:126-150,197-207)
// The reference updates x and r together after alpha is known (baseCG.hpp:154-165);
// here x's update for iteration k-1 rides along with iteration k's stencil
// pass, which reads p_{k-1} anyway: the same x after every iteration, one
// fewer field pass per iteration.  The last update runs after the loop.
template <typename T>
struct OpCgA : Gate {
    const T* __restrict__ src;
    const T* __restrict__ pin;
    T* __restrict__ pout;
    T* __restrict__ Ap;
    T* __restrict__ x;
    Scal* S;
    bool pointwise;
    int slot;  // >= 0: deferred x updates (x untouched; alpha to S->ahist[slot])
    T a, beta, alpha_prev;
    __device__ void load()
    {
        a = pointwise ? (T)S->v[S_JAC] : T(1);
        beta = (T)S->v[S_BETA];
        alpha_prev = (T)S->v[S_ALPHA];
    }
    struct Raw { T src, pold; };  // p_old at the point also feeds the folded x update
    struct Ox { T x; };
    __device__ Raw in_raw(long long c) const { return {src[c], pin[c]}; }
    __device__ T in_value(const Raw& r) const { return a * r.src + beta * r.pold; }
    template <bool NT> __device__ Ox out_raw(long long c) const { return {slot < 0 ? ld<NT>(x + c) : T(0)}; }
    template <bool NT> __device__ void out(long long c, T vc, T Av, const Raw& r, const Ox& o, double* acc) const
    {
        put<NT>(pout + c, vc);
        put<NT>(Ap + c, Av);
        if (slot < 0) put<NT>(x + c, o.x + alpha_prev * r.pold);
        acc[0] += (double)vc * (double)Av;
    }
    __device__ void finish(const double* t)
    {
        S->v[S_PAP] = t[0];
        S->v[S_ALPHA] = S->v[S_RZ] / t[0];
        if (slot >= 0) S->ahist[slot] = S->v[S_ALPHA];
    }
};

// pin/pout double-buffer the search direction: neighbouring tiles rebuild
// p_new on their halo from p_old, so p_old must stay intact during the launch.
template <typename T>
void launch_cg_A2(const Geo& g, hipStream_t s, RedBuf rb, const T* src, const T* pin, T* pout, T* Ap, T* x, Scal* S,
                  bool pointwise, int ring_slot)
{
    OpCgA<T> op;
    op.gate = &S->done;
    op.src = src;
    op.pin = pin;
    op.pout = pout;
    op.Ap = Ap;
    op.x = x;
    op.S = S;
    op.pointwise = pointwise;
    op.slot = ring_slot;
    run_stencil<T, 1>(g, s, op, rb);
}

template <typename T>
struct OpCgFinalX : Gate {
    T* __restrict__ x;
    const T* __restrict__ p;
    const Scal* S;
    T alpha;
    struct R { T x, p; };
    __device__ void load() { alpha = (T)S->v[S_ALPHA]; }
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(x + c), ld<NT>(p + c)}; }
    template <bool NT> __device__ void apply(long long c, const R& v, int, int, int, double*) const { put<NT>(x + c, v.x + alpha * v.p); }
    __device__ void finish(const double*) {}
};

template <typename T>
void launch_cg_final_x(const Geo& g, hipStream_t s, T* x, const T* p, const Scal* S)
{
    OpCgFinalX<T> op;
    op.gate = nullptr;
    op.x = x;
    op.p = p;
    op.S = S;
    run_vector<0>(g, s, op, kNoRed);
}

// r -= alpha Ap, r.r, r.(jac r)  (baseCG.hpp:154-194)
template <typename T>
struct OpCgB : Gate {
    T* __restrict__ r;
    const T* __restrict__ Ap;
    Scal* S;
    bool pointwise, history;
    T alpha, jac;
    struct R { T r, ap; };
    __device__ void load()
    {
        alpha = (T)S->v[S_ALPHA];
        jac = (T)S->v[S_JAC];
    }
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(r + c), ld<NT>(Ap + c)}; }
    template <bool NT> __device__ void apply(long long c, const R& v, int, int, int, double* acc) const
    {
        const T rv = v.r - alpha * v.ap;
        put<NT>(r + c, rv);
        acc[0] += (double)rv * (double)rv;
        acc[1] += (double)rv * (double)(jac * rv);
    }
    __device__ void finish(const double* t)
    {
        if (pointwise) {
            S->v[S_BETA] = t[1] / S->v[S_RZ];
            S->v[S_RZ] = t[1];
        }
        S->v[S_RR] = t[0];
        finish_iteration(S, sqrt(t[0]), history);
    }
};

template <typename T>
void launch_cg_B(const Geo& g, hipStream_t s, RedBuf rb, T* r, const T* Ap, Scal* S, bool pointwise, bool history)
{
    OpCgB<T> op;
    op.gate = &S->done;
    op.r = r;
    op.Ap = Ap;
    op.S = S;
    op.pointwise = pointwise;
    op.history = history;
    run_vector<2>(g, s, op, rb);
}

template <typename T>
struct OpCgRz : Gate {
    const T* __restrict__ r;
    const T* __restrict__ z;
    Scal* S;
    struct R { T r, z; };
    __device__ void load() {}
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(r + c), ld<NT>(z + c)}; }
    template <bool NT> __device__ void apply(long long, const R& v, int, int, int, double* acc) const { acc[0] += (double)v.r * (double)v.z; }
    __device__ void finish(const double* t)
    {
        S->v[S_BETA] = S->iter == 0 ? 0.0 : t[0] / S->v[S_RZ];
        S->v[S_RZ] = t[0];
    }
};

template <typename T>
void launch_cg_rz(const Geo& g, hipStream_t s, RedBuf rb, const T* r, const T* z, Scal* S)
{
    OpCgRz<T> op;
    op.gate = &S->done;
    op.r = r;
    op.z = z;
    op.S = S;
    run_vector<1>(g, s, op, rb);
}

// ------------------------------------------------------------- BiCGSTAB

// AMp = A (amp*Mp), r0.AMp ; alpha = rho0 / r0.AMp   (BiCGSTAB.hpp:142-164)
template <typename T>
struct OpBiK1 : Gate {
    const T* __restrict__ Mp;
    T* __restrict__ AMp;
    const T* __restrict__ r0;
    Scal* S;
    T amp;
    __device__ void load() {}
    struct Raw { T m; };
    struct Ox { T r0; };
    __device__ Raw in_raw(long long c) const { return {Mp[c]}; }
    __device__ T in_value(const Raw& r) const { return amp * r.m; }
    template <bool NT> __device__ Ox out_raw(long long c) const { return {ld<NT>(r0 + c)}; }
    template <bool NT> __device__ void out(long long c, T, T Av, const Raw&, const Ox& o, double* acc) const
    {
        put<NT>(AMp + c, Av);
        acc[0] += (double)o.r0 * (double)Av;
    }
    __device__ void finish(const double* t) { S->v[S_ALPHA] = S->v[S_RHO0] / t[0]; }
};

template <typename T>
void launch_bicg_k1(const Geo& g, hipStream_t s, RedBuf rb, const T* Mp, double amp, T* AMp, const T* r0, Scal* S)
{
    OpBiK1<T> op;
    op.gate = &S->done;
    op.Mp = Mp;
    op.AMp = AMp;
    op.r0 = r0;
    op.S = S;
    op.amp = (T)amp;
    run_stencil<T, 1>(g, s, op, rb);
}

template <typename T>
struct OpBiK2 : Gate {
    T* __restrict__ r;
    const T* __restrict__ AMp;
    const Scal* S;
    T alpha;
    struct R { T r, a; };
    __device__ void load() { alpha = (T)S->v[S_ALPHA]; }
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(r + c), ld<NT>(AMp + c)}; }
    template <bool NT> __device__ void apply(long long c, const R& v, int, int, int, double*) const { put<NT>(r + c, v.r - alpha * v.a); }
    __device__ void finish(const double*) {}
};

template <typename T>
void launch_bicg_k2(const Geo& g, hipStream_t s, T* r, const T* AMp, const Scal* S)
{
    OpBiK2<T> op;
    op.gate = &S->done;
    op.r = r;
    op.AMp = AMp;
    op.S = S;
    run_vector<0>(g, s, op, kNoRed);
}

// Az = A (az*z), r.Az, Az.Az ; omega   (BiCGSTAB.hpp:189-225)
template <typename T>
struct OpBiK3 : Gate {
    const T* __restrict__ z;
    T* __restrict__ Az;
    const T* __restrict__ r;
    Scal* S;
    T az;
    __device__ void load() {}
    struct Raw { T z; };
    struct Ox { T r; };
    __device__ Raw in_raw(long long c) const { return {z[c]}; }
    __device__ T in_value(const Raw& v) const { return az * v.z; }
    template <bool NT> __device__ Ox out_raw(long long c) const { return {ld<NT>(r + c)}; }
    template <bool NT> __device__ void out(long long c, T, T Av, const Raw&, const Ox& o, double* acc) const
    {
        put<NT>(Az + c, Av);
        acc[0] += (double)o.r * (double)Av;
        acc[1] += (double)Av * (double)Av;
    }
    __device__ void finish(const double* t) { S->v[S_OMEGA] = t[0] / t[1]; }
};

template <typename T>
void launch_bicg_k3(const Geo& g, hipStream_t s, RedBuf rb, const T* z, double az, T* Az, const T* r, Scal* S)
{
    OpBiK3<T> op;
    op.gate = &S->done;
    op.z = z;
    op.Az = Az;
    op.r = r;
    op.S = S;
    op.az = (T)az;
    run_stencil<T, 2>(g, s, op, rb);
}

// x += alpha Mp + omega z ; r -= omega Az ; r0.r, r.r ; beta   (BiCGSTAB.hpp:227-259)
template <typename T>
struct OpBiK45 : Gate {
    T* __restrict__ x;
    const T* __restrict__ Mp;
    const T* __restrict__ z;
    T* __restrict__ r;
    const T* __restrict__ Az;
    const T* __restrict__ r0;
    Scal* S;
    T amp, az, alpha, omega;
    bool history;
    struct R { T x, mp, z, r, az, r0; };
    __device__ void load()
    {
        alpha = (T)S->v[S_ALPHA];
        omega = (T)S->v[S_OMEGA];
    }
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(x + c), ld<NT>(Mp + c), ld<NT>(z + c), ld<NT>(r + c), ld<NT>(Az + c), ld<NT>(r0 + c)}; }
    template <bool NT> __device__ void apply(long long c, const R& v, int, int, int, double* acc) const
    {
        put<NT>(x + c, v.x + alpha * (amp * v.mp) + omega * (az * v.z));
        const T rv = v.r - omega * v.az;
        put<NT>(r + c, rv);
        acc[0] += (double)v.r0 * (double)rv;
        acc[1] += (double)rv * (double)rv;
    }
    __device__ void finish(const double* t)
    {
        const double rho1 = t[0];
        S->v[S_BETA] = rho1 / S->v[S_RHO0] * S->v[S_ALPHA] / S->v[S_OMEGA];
        S->v[S_RHO0] = rho1;
        finish_iteration(S, sqrt(t[1]), history);
    }
};

template <typename T>
void launch_bicg_k45(const Geo& g, hipStream_t s, RedBuf rb, T* x, const T* Mp, double amp, const T* z, double az, T* r,
                     const T* Az, const T* r0, Scal* S, bool history)
{
    OpBiK45<T> op;
    op.gate = &S->done;
    op.x = x;
    op.Mp = Mp;
    op.z = z;
    op.r = r;
    op.Az = Az;
    op.r0 = r0;
    op.S = S;
    op.amp = (T)amp;
    op.az = (T)az;
    op.history = history;
    run_vector<2>(g, s, op, rb);
}

// p = r + beta (p - omega AMp)   (BiCGSTAB.hpp:262-272)
template <typename T>
struct OpBiK6 : Gate {
    T* __restrict__ p;
    const T* __restrict__ r;
    const T* __restrict__ AMp;
    const Scal* S;
    T beta, omega;
    struct R { T p, r, a; };
    __device__ void load()
    {
        beta = (T)S->v[S_BETA];
        omega = (T)S->v[S_OMEGA];
    }
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(p + c), ld<NT>(r + c), ld<NT>(AMp + c)}; }
    template <bool NT> __device__ void apply(long long c, const R& v, int, int, int, double*) const { put<NT>(p + c, v.r + beta * (v.p - omega * v.a)); }
    __device__ void finish(const double*) {}
};

template <typename T>
void launch_bicg_k6(const Geo& g, hipStream_t s, T* p, const T* r, const T* AMp, const Scal* S)
{
    OpBiK6<T> op;
    op.gate = &S->done;
    op.p = p;
    op.r = r;
    op.AMp = AMp;
    op.S = S;
    run_vector<0>(g, s, op, kNoRed);
}

// ------------------------------------------------------------ Chebyshev

// Z = c0 B ; Y = c1 (2B + c0 A B)   (chebyshevIteration.hpp:79-90)
template <typename T>
struct OpChebFirst : Gate {
    const T* __restrict__ B;
    T* __restrict__ Z;
    T* __restrict__ Y;
    T c0, c1;
    __device__ void load() {}
    __device__ T in(long long c) const { return B[c]; }
    template <bool NT> __device__ void out(long long c, T vc, T Av, double*) const
    {
        put<NT>(Z + c, c0 * vc);
        put<NT>(Y + c, c1 * (T(2) * vc + c0 * Av));
    }
    __device__ void finish(const double*) {}
};

template <typename T>
void launch_cheb_first(const Geo& g, hipStream_t s, const T* B, T* Z, T* Y, double c0, double c1, const int* gate)
{
    OpChebFirst<T> op;
    op.gate = gate;
    op.B = B;
    op.Z = Z;
    op.Y = Y;
    op.c0 = (T)c0;
    op.c1 = (T)c1;
    run_stencil<T, 0>(g, s, op, kNoRed);
}

template <typename T, typename TO>
struct OpChebFirstOut : Gate {
    const T* __restrict__ B;
    TO* __restrict__ X;
    T c0, c1;
    int which;
    double sign;
    __device__ void load() {}
    __device__ T in(long long c) const { return B[c]; }
    template <bool NT> __device__ void out(long long c, T vc, T Av, double*) const
    {
        const T y = which == 0 ? c0 * vc : c1 * (T(2) * vc + c0 * Av);
        put<NT>(X + c, (TO)(sign * (double)y));
    }
    __device__ void finish(const double*) {}
};

template <typename T, typename TO>
void launch_cheb_first_out(const Geo& g, hipStream_t s, const T* B, TO* X, double c0, double c1, int which, double sign,
                           const int* gate)
{
    OpChebFirstOut<T, TO> op;
    op.gate = gate;
    op.B = B;
    op.X = X;
    op.c0 = (T)c0;
    op.c1 = (T)c1;
    op.which = which;
    op.sign = sign;
    run_stencil<T, 0>(g, s, op, kNoRed);
}

// W = rho (2 sigma Y + 2/delta (B + A Y) - rho_old Z)   (chebyshevIteration.hpp:103-113)
template <typename T, typename TO>
struct OpChebSweep : Gate {
    const T* __restrict__ Y;
    const T* __restrict__ B;
    const T* __restrict__ Z;
    TO* __restrict__ W;
    T a, b, cz;
    double sign;
    __device__ void load() {}
    struct Ox { T b, z; };
    __device__ T in(long long c) const { return Y[c]; }
    template <bool NT> __device__ Ox out_raw(long long c) const { return {ld<NT>(B + c), ld<NT>(Z + c)}; }
    template <bool NT> __device__ void out(long long c, T vc, T Av, const T&, const Ox& o, double*) const
    {
        const T w = a * vc + b * (o.b + Av) - cz * o.z;
        if constexpr (sizeof(TO) == sizeof(T)) put<NT>(W + c, sign < 0 ? -w : w);
        else put<NT>(W + c, (TO)(sign * (double)w));
    }
    __device__ void finish(const double*) {}
};

template <typename T, typename TO>
void launch_cheb_sweep(const Geo& g, hipStream_t s, const T* Y, const T* B, const T* Z, TO* W, double a, double b,
                       double c, double sign, const int* gate)
{
    OpChebSweep<T, TO> op;
    op.gate = gate;
    op.Y = Y;
    op.B = B;
    op.Z = Z;
    op.W = W;
    op.a = (T)a;
    op.b = (T)b;
    op.cz = (T)c;
    op.sign = sign;
    run_stencil<T, 0>(g, s, op, kNoRed);
}

// ---------------------------------------------- fused pair of Chebyshev sweeps
// Two consecutive sweeps of the three-term recurrence in one pass over the grid
// (temporal blocking of chebyshevIteration.hpp:103-113 / the alpaka
// Chebyshev2Kernel, kernelsAlpakaChebyshev.hpp:191-273):
//   W1 = a1 Y  + b1 (B + A Y)  - c1 Z      sweep c    (Y = y_c, Z = y_{c-1})
//   W2 = a2 W1 + b2 (B + A W1) - c2 Y      sweep c+1
// with the Neumann ghosts of W1 reset between the two as the unfused sweeps
// reset those of their input (ghost = the data point two cells inward).  A
// block owns a 64 x 32 tile of W2 and marches z; it computes W1 on the tile
// plus a one-point ring (the neighbours W2 needs), from Y on the tile plus a
// two-point ring.  Per plane: the Y plane is staged in LDS (x-y neighbours),
// each point's z history lives in the registers of the thread that owns it
// (tile points: 4 rows per thread; ring points: one per thread), W1 of plane
// s goes to LDS and W2 of plane s-1 comes out of it.  HBM traffic per pair:
// Y, Z, B read once (plus the rings and the two halo planes), W1 and W2
// written once: 5 fields in place of the 8 of two separate sweeps.
#ifndef PPS_CHEB_WAVES
#define PPS_CHEB_WAVES 8
#endif
constexpr int kCpW = PPS_CHEB_WAVES;       // wavefronts per block (tile rows ly + kCpW q)
constexpr int kCpT = kTX * kCpW;           // threads per block
constexpr int kCpRY = 4;                   // tile rows per thread
constexpr int kCpTY = kCpW * kCpRY;        // tile height
constexpr int kCpE1X = kTX + 2, kCpE1Y = kCpTY + 2;  // W1 region: tile + 1
constexpr int kCpE2X = kTX + 4, kCpE2Y = kCpTY + 4;  // Y region: tile + 2
constexpr int kCpRing1 = 2 * kCpE1X + 2 * kCpTY;      // W1 ring points (164)
constexpr int kCpRing2 = 2 * kCpE2X + 2 * kCpE1Y;     // outer Y ring points (172)
static_assert(kCpRing1 <= kCpT && kCpRing2 <= kCpT, "ring points need one thread each");

// ring1 point t -> (row, col) of the W1 region; ring2 point t -> (row, col) of the Y region
__device__ __forceinline__ void cp_ring1(int t, int& r, int& c)
{
    if (t < kCpE1X) { r = 0; c = t; }
    else if (t < 2 * kCpE1X) { r = kCpE1Y - 1; c = t - kCpE1X; }
    else if (t < 2 * kCpE1X + kCpTY) { r = 1 + t - 2 * kCpE1X; c = 0; }
    else { r = 1 + t - 2 * kCpE1X - kCpTY; c = kCpE1X - 1; }
}
__device__ __forceinline__ void cp_ring2(int t, int& r, int& c)
{
    if (t < kCpE2X) { r = 0; c = t; }
    else if (t < 2 * kCpE2X) { r = kCpE2Y - 1; c = t - kCpE2X; }
    else if (t < 2 * kCpE2X + kCpE1Y) { r = 1 + t - 2 * kCpE2X; c = 0; }
    else { r = 1 + t - 2 * kCpE2X - kCpE1Y; c = kCpE2X - 1; }
}

template <typename T, typename TO, bool NT, bool LAST>
__global__ __launch_bounds__(kCpT) void k_cheb_pair(Geo g, ChebPairGeo cg, const T* __restrict__ Y,
                                                        const T* __restrict__ Z, const T* __restrict__ B,
                                                        T* __restrict__ W1, TO* __restrict__ W2, T a1, T b1, T c1,
                                                        T a2, T b2, T c2, double sign, const int* gate)
{
    if (gate != nullptr && *gate != 0) return;
    __shared__ T ysl[2][kCpE2Y][kCpE2X];
    __shared__ T wsl[2][kCpE1Y][kCpE1X];
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int tx0 = b % g.ntx, rr = b / g.ntx;
    const int ty0 = rr % g.nty, tz0 = rr / g.nty;
    const int i0 = g.lo0 + tx0 * kTX, j0 = g.lo1 + ty0 * kCpTY;
    const int k0 = g.lo2 + tz0 * g.kz, k1 = min(k0 + g.kz, g.hi2);
    const int tid = threadIdx.x + kTX * threadIdx.y;
    const int lx = threadIdx.x, ly = threadIdx.y;
    const T cx = (T)g.cx, cy = (T)g.cy, czz = (T)g.cz;
    auto solver = [&](int i, int j, int k) {
        return i >= g.lo0 && i < g.hi0 && j >= g.lo1 && j < g.hi1 && k >= g.lo2 && k < g.hi2;
    };
    auto inbox = [&](int i, int j, int k) {
        return i >= 0 && i < cg.ng[0] && j >= 0 && j < cg.ng[1] && k >= 0 && k < cg.ng[2];
    };
    auto at = [&](int i, int j, int k) { return (long long)g.ox + i + (long long)g.px * j + g.plane * k; };
    auto yld = [&](int i, int j, int k) { return inbox(i, j, k) ? Y[at(i, j, k)] : T(0); };

    // tile points: (i0 + lx, j0 + ly + kCpW q)
    const int ic = i0 + lx;
    T ym[kCpRY], yc[kCpRY], yp[kCpRY], yn[kCpRY], zc[kCpRY], zn[kCpRY], bc[kCpRY], bn[kCpRY], bp[kCpRY];
    T wm[kCpRY], wc[kCpRY], wn[kCpRY];
    // W1 ring point (one per thread t < kCpRing1) and outer Y ring point (t < kCpRing2)
    int r1r = 0, r1c = 0, r2r = 0, r2c = 0;
    const bool has1 = tid < kCpRing1, has2 = tid < kCpRing2;
    if (has1) cp_ring1(tid, r1r, r1c);
    if (has2) cp_ring2(tid, r2r, r2c);
    const int i1 = i0 - 1 + r1c, j1 = j0 - 1 + r1r;  // ring1 point
    const int i2 = i0 - 2 + r2c, j2 = j0 - 2 + r2r;  // ring2 point
    T qm = 0, qc = 0, qp = 0, qn = 0, qz = 0, qzn = 0, qb = 0, qbn = 0;
    T oc = 0, on = 0;

    const int s0 = k0 - 1;
#pragma unroll
    for (int q = 0; q < kCpRY; ++q) {
        const int j = j0 + ly + kCpW * q;
        ym[q] = yld(ic, j, s0 - 1);
        yc[q] = yld(ic, j, s0);
        yp[q] = yld(ic, j, s0 + 1);
        const bool sv = solver(ic, j, s0);
        zc[q] = sv ? Z[at(ic, j, s0)] : T(0);
        bc[q] = sv ? B[at(ic, j, s0)] : T(0);
        bp[q] = T(0);
        wm[q] = wc[q] = wn[q] = T(0);
        yn[q] = zn[q] = bn[q] = T(0);
    }
    if (has1) {
        qm = yld(i1, j1, s0 - 1);
        qc = yld(i1, j1, s0);
        qp = yld(i1, j1, s0 + 1);
        if (solver(i1, j1, s0)) {
            qz = Z[at(i1, j1, s0)];
            qb = B[at(i1, j1, s0)];
        }
    }
    if (has2) oc = yld(i2, j2, s0);

    for (int s = s0; s <= k1; ++s) {
        const int buf = s & 1;
        // Y of plane s into LDS
#pragma unroll
        for (int q = 0; q < kCpRY; ++q) ysl[buf][ly + kCpW * q + 2][lx + 2] = yc[q];
        if (has1) ysl[buf][r1r + 1][r1c + 1] = qc;
        if (has2) ysl[buf][r2r][r2c] = oc;
        // loads for the next step: Y of plane s+2 (s+1 for the outer ring), Z and B of plane s+1
        const bool ny = s + 2 <= k1 + 1, nzb = s + 1 <= k1;
#pragma unroll
        for (int q = 0; q < kCpRY; ++q) {
            const int j = j0 + ly + kCpW * q;
            if (ny) yn[q] = yld(ic, j, s + 2);
            const bool sv = nzb && solver(ic, j, s + 1);
            zn[q] = sv ? Z[at(ic, j, s + 1)] : T(0);
            bn[q] = sv ? B[at(ic, j, s + 1)] : T(0);
        }
        if (has1) {
            if (ny) qn = yld(i1, j1, s + 2);
            const bool sv = nzb && solver(i1, j1, s + 1);
            qzn = sv ? Z[at(i1, j1, s + 1)] : T(0);
            qbn = sv ? B[at(i1, j1, s + 1)] : T(0);
        }
        if (has2 && nzb) on = yld(i2, j2, s + 1);
        __syncthreads();
        // W1 of plane s on the tile and its ring (0 off the solver region)
#pragma unroll
        for (int q = 0; q < kCpRY; ++q) {
            const int j = j0 + ly + kCpW * q;
            const int r = ly + kCpW * q + 2, c = lx + 2;
            T w = T(0);
            if (solver(ic, j, s)) {
                const T v = yc[q];
                const T Av = cx * ((ysl[buf][r][c - 1] - v) + (ysl[buf][r][c + 1] - v)) +
                             cy * ((ysl[buf][r - 1][c] - v) + (ysl[buf][r + 1][c] - v)) +
                             czz * ((ym[q] - v) + (yp[q] - v));
                w = a1 * v + b1 * (bc[q] + Av) - c1 * zc[q];
                if constexpr (!LAST)
                    if (s >= k0 && s < k1) put<NT>(W1 + at(ic, j, s), w);
            }
            wn[q] = w;
            wsl[buf][r - 1][c - 1] = w;
        }
        if (has1) {
            T w = T(0);
            if (solver(i1, j1, s)) {
                const int r = r1r + 1, c = r1c + 1;
                const T v = qc;
                const T Av = cx * ((ysl[buf][r][c - 1] - v) + (ysl[buf][r][c + 1] - v)) +
                             cy * ((ysl[buf][r - 1][c] - v) + (ysl[buf][r + 1][c] - v)) +
                             czz * ((qm - v) + (qp - v));
                w = a1 * v + b1 * (qb + Av) - c1 * qz;
            }
            wsl[buf][r1r][r1c] = w;
        }
        __syncthreads();
        // Neumann x/y ghosts of W1: the data point two cells inward
        if (cg.neu & 15) {
            const bool zin = s >= cg.lo_d[2] && s < cg.hi_d[2];
            auto mirror = [&](int i, int j, int r, int c) {
                if (!zin) return;
                const bool yin = j >= cg.lo_d[1] && j < cg.hi_d[1], xin = i >= cg.lo_d[0] && i < cg.hi_d[0];
                if ((cg.neu & 1) && i == cg.lo_d[0] - 1 && yin) wsl[buf][r][c] = wsl[buf][r][c + 2];
                else if ((cg.neu & 2) && i == cg.hi_d[0] && yin) wsl[buf][r][c] = wsl[buf][r][c - 2];
                else if ((cg.neu & 4) && j == cg.lo_d[1] - 1 && xin) wsl[buf][r][c] = wsl[buf][r + 2][c];
                else if ((cg.neu & 8) && j == cg.hi_d[1] && xin) wsl[buf][r][c] = wsl[buf][r - 2][c];
            };
#pragma unroll
            for (int q = 0; q < kCpRY; ++q) mirror(ic, j0 + ly + kCpW * q, ly + kCpW * q + 1, lx + 1);
            if (has1) mirror(i1, j1, r1r, r1c);
        }
        // W2 of plane s-1 on the tile (W1 of plane s-1 is in the other LDS buffer)
        if (s - 1 >= k0) {
            const int p = s - 1, pb = buf ^ 1;
            const bool zlo = (cg.neu & 16) && p == cg.lo_d[2], zhi = (cg.neu & 32) && p == cg.hi_d[2] - 1;
#pragma unroll
            for (int q = 0; q < kCpRY; ++q) {
                const int j = j0 + ly + kCpW * q;
                if (!solver(ic, j, p)) continue;
                const int r = ly + kCpW * q + 1, c = lx + 1;
                const T v = wc[q];
                const T m = zlo ? wn[q] : wm[q];
                const T u = zhi ? wm[q] : wn[q];
                const T Av = cx * ((wsl[pb][r][c - 1] - v) + (wsl[pb][r][c + 1] - v)) +
                             cy * ((wsl[pb][r - 1][c] - v) + (wsl[pb][r + 1][c] - v)) + czz * ((m - v) + (u - v));
                const T w = a2 * v + b2 * (bp[q] + Av) - c2 * ym[q];
                if constexpr (sizeof(TO) == sizeof(T)) put<NT>(W2 + at(ic, j, p), (TO)(sign < 0 ? -w : w));
                else put<NT>(W2 + at(ic, j, p), (TO)(sign * (double)w));
            }
        }
        // advance one plane
#pragma unroll
        for (int q = 0; q < kCpRY; ++q) {
            ym[q] = yc[q];
            yc[q] = yp[q];
            yp[q] = yn[q];
            zc[q] = zn[q];
            bp[q] = bc[q];
            bc[q] = bn[q];
            wm[q] = wc[q];
            wc[q] = wn[q];
        }
        qm = qc;
        qc = qp;
        qp = qn;
        qz = qzn;
        qb = qbn;
        oc = on;
    }
}

template <typename T, typename TO>
void launch_cheb_pair(const Geo& g0, const ChebPairGeo& cg, hipStream_t s, const T* Y, const T* Z, const T* B, T* W1,
                      TO* W2, const double a[2], const double b[2], const double c[2], double sign, bool last,
                      const int* gate)
{
    Geo g = g0;
    g.ntx = (g.hi0 - g.lo0 + kTX - 1) / kTX;
    g.nty = (g.hi1 - g.lo1 + kCpTY - 1) / kCpTY;
    // z-chunk: as set, or the longest of 64, 32, 16 planes giving 1024 blocks
    // (two rounds of the two blocks per CU that fit; halo planes cost 2 / kz)
    const int nz = g.hi2 - g.lo2;
    g.kz = cg.kz;
    if (g.kz <= 0) {
        g.kz = 64;
        while (g.kz > 16 && (long long)g.ntx * g.nty * ((nz + g.kz - 1) / g.kz) < 1024) g.kz /= 2;
    }
    const int ntz = (nz + g.kz - 1) / g.kz;
    const long long nb = (long long)g.ntx * g.nty * ntz;
    if (g.ntx <= 0 || g.nty <= 0 || ntz <= 0) return;
    if (nb > 0x7FFFFFFFll) {
        std::fprintf(stderr, "pps: fused Chebyshev grid too large; launch skipped\n");
        return;
    }
    const dim3 grid((unsigned)nb), block(kTX, kCpW);
    const T A1 = (T)a[0], B1 = (T)b[0], C1 = (T)c[0], A2 = (T)a[1], B2 = (T)b[1], C2 = (T)c[1];
    if (last) {
        if (g.nt) hipLaunchKernelGGL((k_cheb_pair<T, TO, true, true>), grid, block, 0, s, g, cg, Y, Z, B, W1, W2, A1, B1, C1, A2, B2, C2, sign, gate);
        else hipLaunchKernelGGL((k_cheb_pair<T, TO, false, true>), grid, block, 0, s, g, cg, Y, Z, B, W1, W2, A1, B1, C1, A2, B2, C2, sign, gate);
    } else if constexpr (std::is_same_v<T, TO>) {
        if (g.nt) hipLaunchKernelGGL((k_cheb_pair<T, T, true, false>), grid, block, 0, s, g, cg, Y, Z, B, W1, W2, A1, B1, C1, A2, B2, C2, sign, gate);
        else hipLaunchKernelGGL((k_cheb_pair<T, T, false, false>), grid, block, 0, s, g, cg, Y, Z, B, W1, W2, A1, B1, C1, A2, B2, C2, sign, gate);
    } else {
        std::fprintf(stderr, "pps: a fused Chebyshev pair changes precision only on its last output; launch skipped\n");
    }
}

// ------------------------------------------------------------------ FCG

// end of an FCG iteration once ||r||^2 of the updated residual is known
__device__ inline void update_finish(Scal* S, double rr, bool history)
{
    S->v[S_PAP_OLD] = S->v[S_PAP];
    S->v[S_RR] = rr;
    const double err = sqrt(rr);
    finish_iteration(S, err, history);
    // stall monitor: every stall_window iterations the residual must have
    // dropped by the factor v[S_AUX1], else the fallback preconditioner takes over
    if (S->stall_window > 0 && !S->fallback && S->iter % S->stall_window == 0) {
        if (err > S->v[S_AUX1] * S->v[S_AUX2]) S->fallback = 1;
        S->v[S_AUX2] = err;
    }
}

// gamma = z.Apold and beta.  When the preconditioner sums z.Apold itself (the
// learned one, from its output kernel), only the finish runs, gated off once
// the fallback preconditioner has taken over; the pass over z and Apold then
// runs only while the fallback is on.
template <typename T>
struct OpFcgGamma : Gate {
    const T* __restrict__ z;
    const T* __restrict__ Apold;
    Scal* S;
    const int* skip_if = nullptr;  // gated while *skip_if != 0
    const int* run_if = nullptr;   // gated while *run_if == 0
    __device__ __forceinline__ bool gated() const
    {
        return Gate::gated() || (skip_if && *skip_if) || (run_if && !*run_if);
    }
    struct R { T z, a; };
    __device__ void load() {}
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(z + c), ld<NT>(Apold + c)}; }
    template <bool NT> __device__ void apply(long long, const R& v, int, int, int, double* acc) const { acc[0] += (double)v.z * (double)v.a; }
    __device__ void finish(const double* t)
    {
        S->v[S_GAM] = t[0];
        // restart after a discarded step: the old direction may hold non-finite values
        S->v[S_BETA] = (S->iter == 0 || S->restart) ? 0.0 : t[0] / S->v[S_PAP_OLD];
    }
};

template <typename T>
void launch_fcg_gamma(const Geo& g, hipStream_t s, RedBuf rb, const T* z, const T* Apold, Scal* S, bool fallback_only)
{
    OpFcgGamma<T> op;
    op.gate = &S->done;
    op.z = z;
    op.Apold = Apold;
    op.S = S;
    if (fallback_only) op.run_if = &S->fallback;
    run_vector<1>(g, s, op, rb, fallback_only);
}

template <typename T>
void launch_fcg_gamma_finish(hipStream_t s, RedBuf rb, int nblocks, Scal* S)
{
    OpFcgGamma<T> op;
    op.gate = &S->done;
    op.z = nullptr;
    op.Apold = nullptr;
    op.S = S;
    op.skip_if = &S->fallback;
    if (nblocks <= 0 || nblocks > rb.cap) return;
    hipLaunchKernelGGL((k_finish<1, OpFcgGamma<T>>), dim3(1), dim3(kFinThreads), 0, s, op, rb, nblocks);
}

// p = z - beta pold ; Ap = A p ; p.Ap, p.r ; alpha = p.r / p.Ap ; and the
// previous iteration's x += alpha_prev pold (folded in: this pass reads pold
// anyway), unless that step was discarded (restart).
// FUSED: the residual update r -= alpha Ap has no pass of its own; the next
// iteration's preconditioner forms it while reading r.  This pass then also
// sums r.Ap, Ap.Ap and r.r, and its finish takes over the update's:
// ||r - alpha Ap||^2 = r.r - 2 alpha r.Ap + alpha^2 Ap.Ap, the convergence
// test and the stall monitor.
template <typename T, bool FUSED = false>
struct OpFcgP : Gate {
    const T* __restrict__ z;
    const T* __restrict__ pold;
    T* __restrict__ p;
    T* __restrict__ Ap;
    const T* __restrict__ r;
    T* __restrict__ x;
    Scal* S;
    bool history;  // FUSED: record the residual history
    T beta, alpha_prev;
    bool upd;  // apply the previous step to x
    int slot;  // >= 0: deferred x updates (x untouched; alpha to S->ahist[slot])
    __device__ void load()
    {
        beta = (T)S->v[S_BETA];
        alpha_prev = (T)S->v[S_ALPHA];
        upd = slot < 0 && !S->restart && alpha_prev != (T)0;
    }
    // beta = 0 takes z as it is, and pold is read only when it is used, so a
    // restart never multiplies a non-finite pold
    struct Raw { T z, pold; };
    struct Ox { T r, x; };
    __device__ Raw in_raw(long long c) const { return {z[c], (beta == (T)0 && !upd) ? (T)0 : pold[c]}; }
    __device__ T in_value(const Raw& v) const { return beta == (T)0 ? v.z : v.z - beta * v.pold; }
    template <bool NT> __device__ Ox out_raw(long long c) const { return {ld<NT>(r + c), upd ? ld<NT>(x + c) : (T)0}; }
    template <bool NT> __device__ void out(long long c, T vc, T Av, const Raw& v, const Ox& o, double* acc) const
    {
        put<NT>(p + c, vc);
        put<NT>(Ap + c, Av);
        if (upd) put<NT>(x + c, o.x + alpha_prev * v.pold);
        acc[0] += (double)vc * (double)Av;
        acc[1] += (double)vc * (double)o.r;
        if constexpr (FUSED) {
            acc[2] += (double)o.r * (double)Av;
            acc[3] += (double)Av * (double)Av;
            acc[4] += (double)o.r * (double)o.r;
        }
    }
    __device__ void finish(const double* t)
    {
        S->v[S_PAP] = t[0];
        const double a = t[1] / t[0];
        double rr_next = 0.0;
        // a non-finite step (the preconditioner returned inf/NaN or a null
        // direction): discard it, switch to the fallback preconditioner, restart
        if (!isfinite(a)) {
            S->restart = 1;
            S->fallback = 1;
            S->v[S_ALPHA] = 0.0;
            if constexpr (FUSED) rr_next = t[4];  // r stays
        } else {
            S->restart = 0;
            S->v[S_ALPHA] = a;
            // the expansion cancels when one step removes almost all of r: its
            // error is a few eps * r.r, so below 1e-8 r.r it is noise.  Clamping
            // there over-estimates the next residual (the loop runs on, and the
            // next pass sums r.r exactly) and never gives the next network input
            // a zero scale
            if constexpr (FUSED) rr_next = fmax(t[4] - 2.0 * a * t[2] + a * a * t[3], 1e-8 * t[4]);
        }
        if (slot >= 0) S->ahist[slot] = S->v[S_ALPHA];
        if constexpr (FUSED) update_finish(S, rr_next, history);
    }
};

template <typename T>
void launch_fcg_p(const Geo& g, hipStream_t s, RedBuf rb, const T* z, const T* pold, T* p, T* Ap, const T* r, T* x,
                  Scal* S)
{
    OpFcgP<T> op;
    op.gate = &S->done;
    op.z = z;
    op.pold = pold;
    op.p = p;
    op.Ap = Ap;
    op.r = r;
    op.x = x;
    op.S = S;
    op.history = false;
    op.slot = -1;
    run_stencil<T, 2>(g, s, op, rb);
}

template <typename T>
void launch_fcg_p_fused(const Geo& g, hipStream_t s, RedBuf rb, const T* z, const T* pold, T* p, T* Ap, const T* r,
                        T* x, Scal* S, bool history, int ring_slot)
{
    OpFcgP<T, true> op;
    op.gate = &S->done;
    op.z = z;
    op.pold = pold;
    op.p = p;
    op.Ap = Ap;
    op.r = r;
    op.x = x;
    op.S = S;
    op.history = history;
    op.slot = ring_slot;
    run_stencil<T, 5>(g, s, op, rb);
}

// deferred x updates: x += ahist[j] * ring[j] over the slots in order, the
// slots with alpha 0 (unused, or a discarded step whose direction may hold
// non-finite values) neither read nor added
template <typename T>
struct OpFcgFlush : Gate {
    T* __restrict__ x;
    const T* ring[kFcgRing];
    const Scal* S;
    T a[kFcgRing];
    unsigned used;
    struct R { T x, p[kFcgRing]; };
    __device__ void load()
    {
        used = 0u;
#pragma unroll
        for (int j = 0; j < kFcgRing; ++j) {
            a[j] = (T)S->ahist[j];
            if (S->ahist[j] != 0.0) used |= 1u << j;
        }
    }
    template <bool NT> __device__ R fetch(long long c) const
    {
        R v;
        v.x = ld<NT>(x + c);
#pragma unroll
        for (int j = 0; j < kFcgRing; ++j) v.p[j] = ((used >> j) & 1u) ? ld<NT>(ring[j] + c) : (T)0;
        return v;
    }
    template <bool NT> __device__ void apply(long long c, const R& v, int, int, int, double*) const
    {
        T acc = v.x;
#pragma unroll
        for (int j = 0; j < kFcgRing; ++j)
            if ((used >> j) & 1u) acc = acc + a[j] * v.p[j];
        put<NT>(x + c, acc);
    }
    __device__ void finish(const double*) {}
};

__global__ void k_fcg_ring_reset(Scal* S)
{
    if (threadIdx.x < kFcgRing) S->ahist[threadIdx.x] = 0.0;
}

void launch_fcg_ring_reset(hipStream_t s, Scal* S)
{
    hipLaunchKernelGGL(k_fcg_ring_reset, dim3(1), dim3(64), 0, s, S);
}

template <typename T>
void launch_fcg_flush(const Geo& g, hipStream_t s, T* x, T* const ring[kFcgRing], Scal* S)
{
    OpFcgFlush<T> op;
    op.gate = nullptr;
    op.x = x;
    for (int j = 0; j < kFcgRing; ++j) op.ring[j] = ring[j];
    op.S = S;
    run_vector<0>(g, s, op, kNoRed);
    launch_fcg_ring_reset(s, S);
}

// the last step of an FCG solve: x += alpha p unless it was discarded
template <typename T>
struct OpFcgFinalX : Gate {
    T* __restrict__ x;
    const T* __restrict__ p;
    const Scal* S;
    T alpha;
    int skip;
    struct R { T x, p; };
    __device__ void load()
    {
        alpha = (T)S->v[S_ALPHA];
        skip = S->restart;
    }
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(x + c), skip ? (T)0 : ld<NT>(p + c)}; }
    template <bool NT> __device__ void apply(long long c, const R& v, int, int, int, double*) const
    {
        if (!skip) put<NT>(x + c, v.x + alpha * v.p);
    }
    __device__ void finish(const double*) {}
};

template <typename T>
void launch_fcg_final_x(const Geo& g, hipStream_t s, T* x, const T* p, const Scal* S)
{
    OpFcgFinalX<T> op;
    op.gate = nullptr;
    op.x = x;
    op.p = p;
    op.S = S;
    run_vector<0>(g, s, op, kNoRed);
}

template <typename T>
struct OpFcgUpdate : Gate {
    T* __restrict__ r;
    const T* __restrict__ Ap;
    Scal* S;
    bool history;
    T alpha;
    int skip;
    struct R { T r, ap; };
    __device__ void load()
    {
        alpha = (T)S->v[S_ALPHA];
        skip = S->restart;
    }
    // r -= alpha Ap; x takes alpha p in the next stencil pass (or in the final
    // x kernel after the last iteration)
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(r + c), skip ? (T)0 : ld<NT>(Ap + c)}; }
    template <bool NT> __device__ void apply(long long c, const R& v, int, int, int, double* acc) const
    {
        T rv = v.r;
        if (!skip) {
            rv = v.r - alpha * v.ap;
            put<NT>(r + c, rv);
        }
        acc[0] += (double)rv * (double)rv;
    }
    __device__ void finish(const double* t) { update_finish(S, t[0], history); }
};

template <typename T>
void launch_fcg_update(const Geo& g, hipStream_t s, RedBuf rb, T* r, const T* Ap, Scal* S, bool history)
{
    OpFcgUpdate<T> op;
    op.gate = &S->done;
    op.r = r;
    op.Ap = Ap;
    op.S = S;
    op.history = history;
    run_vector<1>(g, s, op, rb);
}

// --------------------------------------------------- iterative refinement

struct OpIrRes : Gate {
    const double* __restrict__ x;
    const double* __restrict__ b;
    float* __restrict__ r32;
    Scal* So;
    Scal* SI;
    double itol, itol0;  // inner tolerance: first outer step, later ones
    double adapt;        // > 0: no less than adapt * tol / ||r||
    __device__ void load() {}
    struct Ox { double b; };
    __device__ double in(long long c) const { return x[c]; }
    template <bool NT> __device__ Ox out_raw(long long c) const { return {ld<NT>(b + c)}; }
    template <bool NT> __device__ void out(long long c, double, double Av, const double&, const Ox& o, double* acc) const
    {
        const double rv = o.b - Av;
        const float rf = (float)rv;
        put<NT>(r32 + c, rf);
        acc[0] += rv * rv;
        acc[1] += (double)rf * (double)rf;
    }
    __device__ void finish(const double* t)
    {
        const double e = sqrt(t[0]);
        So->v[S_ERR] = e;
        So->v[S_ERR_OP] = e;
        const int o = So->outer;
        if (So->hist != nullptr && o < So->hist_cap) So->hist[o] = e;
        if (o == 0) So->v[S_ERR0] = e;
        So->done = e < So->v[S_TOL] ? 1 : 0;
        SI->v[S_INNER_NORM] = sqrt(t[1]);
        const double base = o == 0 ? itol0 : itol;
        SI->v[S_TOL] = (adapt > 0.0 && e > 0.0) ? fmax(base, adapt * So->v[S_TOL] / e) : base;
    }
};

void launch_ir_residual(const Geo& g, hipStream_t s, RedBuf rb, const double* x, const double* b, float* r32, Scal* So,
                        Scal* SI, double itol, double itol0, double adapt)
{
    OpIrRes op;
    op.gate = nullptr;
    op.x = x;
    op.b = b;
    op.r32 = r32;
    op.So = So;
    op.SI = SI;
    op.itol = itol;
    op.itol0 = itol0;
    op.adapt = adapt;
    run_stencil<double, 2>(g, s, op, rb);
}

struct OpIrUpdate : Gate {
    double* __restrict__ x;
    const float* __restrict__ d;
    const Scal* SI;
    float sc;
    struct R { double x; float d; };
    __device__ void load() { sc = (float)SI->v[S_INNER_NORM]; }
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(x + c), ld<NT>(d + c)}; }
    template <bool NT> __device__ void apply(long long c, const R& v, int, int, int, double*) const { put<NT>(x + c, v.x + (double)(v.d * sc)); }
    __device__ void finish(const double*) {}
};

void launch_ir_update(const Geo& g, hipStream_t s, double* x, const float* d, const Scal* SI)
{
    OpIrUpdate op;
    op.gate = nullptr;
    op.x = x;
    op.d = d;
    op.SI = SI;
    run_vector<0>(g, s, op, kNoRed);
}

// inner solve start: the non-main normalisation of BaseCG (x = 0, b /= ||b||,
// r0 = b) applied to r32 in place
struct OpIrInit : Gate {
    float* __restrict__ r;
    float* __restrict__ x;
    float* __restrict__ p;
    Scal* SI;
    bool pointwise;
    float inv, jac;
    struct R { float r; };
    __device__ void load()
    {
        inv = (float)(1.0 / SI->v[S_INNER_NORM]);
        jac = (float)SI->v[S_JAC];
    }
    template <bool NT> __device__ R fetch(long long c) const { return {ld<NT>(r + c)}; }
    template <bool NT> __device__ void apply(long long c, const R& v, int, int, int, double* acc) const
    {
        const float rv = v.r * inv;
        put<NT>(r + c, rv);
        put<NT>(x + c, 0.0f);
        put<NT>(p + c, 0.0f);
        acc[0] += (double)rv * (double)rv;
        acc[1] += (double)rv * (double)(jac * rv);
    }
    __device__ void finish(const double* t)
    {
        const double e = sqrt(t[0]);
        SI->v[S_ERR0] = e;
        SI->v[S_ERR] = e;
        SI->v[S_RZ] = t[1];
        SI->v[S_RR] = t[0];
        SI->v[S_BETA] = 0.0;
        SI->v[S_PAP_OLD] = 1.0;
        SI->v[S_ALPHA] = 0.0;
        SI->v[S_AUX2] = 1e300;
        SI->fallback = 0;
        SI->restart = 0;
        SI->iter = 0;
        SI->done = e < SI->v[S_TOL] ? 2 : 0;
    }
};

void launch_ir_inner_init(const Geo& g, hipStream_t s, RedBuf rb, float* r, float* x, float* p, Scal* SI, bool pointwise)
{
    OpIrInit op;
    op.gate = nullptr;
    op.r = r;
    op.x = x;
    op.p = p;
    op.SI = SI;
    op.pointwise = pointwise;
    run_vector<2>(g, s, op, rb);
}

// ------------------------------------------------------ instantiations

#define PPS_INST_T(T)                                                                                          \
    template void launch_apply_A<T>(const Geo&, hipStream_t, const T*, T*);                                    \
    template void launch_copy<T>(const Geo&, hipStream_t, const T*, T*);                                       \
    template void launch_dot<T>(const Geo&, hipStream_t, RedBuf, const T*, const T*, double*, const int*);      \
    template void launch_scale_box<T>(const Geo&, hipStream_t, T*, const Scal*, int);                          \
    template void launch_fill<T>(const Geo&, hipStream_t, T*, T);                                              \
    template void launch_ghost<T>(hipStream_t, const Layout&, const Faces&, T*, int, const Scal*, const int*);  \
    template void launch_cg_A2<T>(const Geo&, hipStream_t, RedBuf, const T*, const T*, T*, T*, T*, Scal*, bool,  \
                                  int);                                                                            \
    template void launch_cg_final_x<T>(const Geo&, hipStream_t, T*, const T*, const Scal*);                      \
    template void launch_cg_B<T>(const Geo&, hipStream_t, RedBuf, T*, const T*, Scal*, bool, bool);            \
    template void launch_cg_rz<T>(const Geo&, hipStream_t, RedBuf, const T*, const T*, Scal*);                 \
    template void launch_pointwise<T>(const Geo&, hipStream_t, const T*, T*, double, const int*);              \
    template void launch_bicg_k1<T>(const Geo&, hipStream_t, RedBuf, const T*, double, T*, const T*, Scal*);   \
    template void launch_bicg_k2<T>(const Geo&, hipStream_t, T*, const T*, const Scal*);                       \
    template void launch_bicg_k3<T>(const Geo&, hipStream_t, RedBuf, const T*, double, T*, const T*, Scal*);   \
    template void launch_bicg_k45<T>(const Geo&, hipStream_t, RedBuf, T*, const T*, double, const T*, double, T*, \
                                     const T*, const T*, Scal*, bool);                                         \
    template void launch_bicg_k6<T>(const Geo&, hipStream_t, T*, const T*, const T*, const Scal*);             \
    template void launch_cheb_first<T>(const Geo&, hipStream_t, const T*, T*, T*, double, double, const int*); \
    template void launch_fcg_gamma<T>(const Geo&, hipStream_t, RedBuf, const T*, const T*, Scal*, bool);       \
    template void launch_fcg_gamma_finish<T>(hipStream_t, RedBuf, int, Scal*);                                  \
    template void launch_fcg_p<T>(const Geo&, hipStream_t, RedBuf, const T*, const T*, T*, T*, const T*, T*, Scal*); \
    template void launch_fcg_p_fused<T>(const Geo&, hipStream_t, RedBuf, const T*, const T*, T*, T*, const T*, T*,    \
                                        Scal*, bool, int);                                                         \
    template void launch_fcg_final_x<T>(const Geo&, hipStream_t, T*, const T*, const Scal*);                      \
    template void launch_fcg_flush<T>(const Geo&, hipStream_t, T*, T* const[kFcgRing], Scal*);                    \
    template void launch_fcg_update<T>(const Geo&, hipStream_t, RedBuf, T*, const T*, Scal*, bool);

PPS_INST_T(double)
PPS_INST_T(float)

template void launch_cheb_pair<double, double>(const Geo&, const ChebPairGeo&, hipStream_t, const double*,
                                               const double*, const double*, double*, double*, const double[2],
                                               const double[2], const double[2], double, bool, const int*);
template void launch_cheb_pair<float, float>(const Geo&, const ChebPairGeo&, hipStream_t, const float*, const float*,
                                             const float*, float*, float*, const double[2], const double[2],
                                             const double[2], double, bool, const int*);
template void launch_cheb_pair<float, double>(const Geo&, const ChebPairGeo&, hipStream_t, const float*, const float*,
                                              const float*, float*, double*, const double[2], const double[2],
                                              const double[2], double, bool, const int*);
template void launch_cheb_sweep<double, double>(const Geo&, hipStream_t, const double*, const double*, const double*,
                                                double*, double, double, double, double, const int*);
template void launch_cheb_sweep<float, float>(const Geo&, hipStream_t, const float*, const float*, const float*, float*,
                                              double, double, double, double, const int*);
template void launch_cheb_sweep<float, double>(const Geo&, hipStream_t, const float*, const float*, const float*,
                                               double*, double, double, double, double, const int*);
template void launch_cheb_first_out<double, double>(const Geo&, hipStream_t, const double*, double*, double, double, int,
                                                    double, const int*);
template void launch_cheb_first_out<float, float>(const Geo&, hipStream_t, const float*, float*, double, double, int,
                                                  double, const int*);
template void launch_cheb_first_out<float, double>(const Geo&, hipStream_t, const float*, double*, double, double, int,
                                                   double, const int*);
template void launch_cast<double, float>(const Geo&, hipStream_t, const double*, float*, const int*);
template void launch_cast<float, double>(const Geo&, hipStream_t, const float*, double*, const int*);

}  // namespace pps
