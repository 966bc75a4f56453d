// pps_kernels.hip -- hand-written gfx950 kernels of the Krylov hot path.
//
// Two kernel skeletons carry every operation:
//  * k_stencil: 2.5-D z-marching tile.  A block of 64x4 threads owns a 64x4
//    column tile of one x-y slab and marches through kz planes.  Each thread
//    keeps its column's (k-1, k, k+1) values in registers; the plane-k tile and
//    its one-point halo are staged in LDS (double-buffered, one barrier per
//    plane), so x- and y-neighbours come from LDS and every value of the input
//    field is produced once per plane (this is what lets the CG search
//    direction p = z + beta*p be formed on the fly inside the stencil).  The
//    column value and the halo of plane k+1 are loaded while plane k is
//    computed, so the barrier never waits on HBM latency.
//    The 7-point operator is the reference's +Laplacian
//    (matrixFreeOperatorA.hpp:33-38) written in difference form,
//    cx((x[i-1]-x)+(x[i+1]-x)) + cy(..) + cz(..), which keeps FP32 accurate
//    for smooth fields.
//  * k_vector: the same tiling without halo, for fused axpy/dot updates; each
//    thread issues the loads of kU planes before it computes and stores any of
//    them, so a wavefront keeps kU x (loads per point) requests in flight.
//
// Reductions are FP64 in every precision and deterministic: wavefront
// shuffles, LDS across the four wavefronts, one partial per block, then a
// one-block epilogue kernel sums the partials in a fixed order and runs the
// solver's scalar update (alpha, beta, omega, convergence flag) on the device.
// The epilogue runs as its own launch instead of a last-block-arrives scheme:
// on gfx950 the agent-scope release fence that scheme needs in every block
// writes back the XCD's L2, which costs more than the whole reduction.
//
// Blocks are remapped so that consecutive tiles land on the same XCD (the
// hardware deals blocks round-robin over the 8 XCDs, each with its own L2):
// the y-halo rows a tile reads from its neighbour are then L2 hits.
#include "pps_kernels.hpp"

#include <type_traits>

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>

namespace pps {

constexpr int kU = 4;  // planes whose loads a vector-kernel thread issues together

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

Geo make_geo(const Layout& L, const Region& R, const double coef[3], int kz_st, int kz_vec, int nt)
{
    Geo g{};
    g.plane = L.plane;
    g.px = L.px;
    g.ox = L.ox;
    g.lo0 = R.lo[0]; g.hi0 = R.hi[0];
    g.lo1 = R.lo[1]; g.hi1 = R.hi[1];
    g.lo2 = R.lo[2]; g.hi2 = R.hi[2];
    g.kz_st = kz_st;
    g.kz_vec = kz_vec;
    g.nt = nt;
    g.cx = coef[0];
    g.cy = coef[1];
    g.cz = coef[2];
    return retile(g, kTY, kz_vec);
}

// stores of streamed outputs: non-temporal when the fields exceed the caches
template <typename T>
__device__ __forceinline__ void put(T* p, T v, int nt)
{
    if (nt) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// ------------------------------------------------------------ common pieces

__device__ __forceinline__ int xcd_remap(int b, int nb)
{
    const int full = nb & ~7;
    if (b >= full) return b;
    return (b & 7) * (full >> 3) + (b >> 3);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
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

template <int NQ>
__device__ __forceinline__ void block_partials(double (&v)[NQ], const RedBuf& rb)
{
    double tot[NQ];
    block_sum<NQ>(v, tot);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) rb.partials[(size_t)q * rb.cap + blockIdx.x] = tot[q];
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
template <class Op>
__device__ __forceinline__ typename OxOf<Op>::type st_ox(const Op& op, long long c)
{
    if constexpr (std::is_same_v<typename OxOf<Op>::type, NoRaw>) return NoRaw{};
    else return op.out_raw(c);
}
template <class Op, typename T>
__device__ __forceinline__ void st_out(const Op& op, long long c, T v, T Av, const typename RawOf<Op, T>::type& r,
                                       const typename OxOf<Op>::type& o, double* acc)
{
    if constexpr (kHooks<Op, T>) op.out(c, v, Av, r, o, acc);
    else op.out(c, v, Av, acc);
}

template <typename T, int NQ, class Op>
__global__ __launch_bounds__(kThreads) void k_stencil(Geo g, Op op, RedBuf rb)
{
    if (op.gated()) return;
    op.load();
    // tile: 64 columns x kTH rows; thread (lx, ly) owns rows ly, ly+kTY, ...
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
        const int j = jt + ly + q * kTY;
        act[q] = (i < g.hi0) && (j < g.hi1);
        inb[q] = (i <= g.hi0) && (j <= g.hi1);
        base[q] = (long long)g.ox + i + (long long)g.px * j;
        vm[q] = inb[q] ? st_value<Op, T>(op, st_raw<Op, T>(op, base[q] + g.plane * (k0 - 1))) : T(0);
        rc[q] = inb[q] ? st_raw<Op, T>(op, base[q] + g.plane * k0) : Raw{};
        r1[q] = inb[q] ? st_raw<Op, T>(op, base[q] + g.plane * (k0 + 1)) : Raw{};
        if (act[q]) oc[q] = st_ox(op, base[q] + g.plane * k0);
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
            if (act[q] && more) on[q] = st_ox(op, base[q] + g.plane * (k + 1));
        }
        Raw h2{};
        if (hin && k + 2 < k1) h2 = st_raw<Op, T>(op, hbase + g.plane * (k + 2));
        // plane k+1's values: its loads were issued one step ago
#pragma unroll
        for (int q = 0; q < kRY; ++q) vp[q] = inb[q] ? st_value<Op, T>(op, r1[q]) : T(0);
#pragma unroll
        for (int q = 0; q < kRY; ++q) lds[buf][ly + q * kTY + 1][lx + 1] = vc[q];
        if (hv) lds[buf][hr][hc_] = hin ? st_value<Op, T>(op, hc) : T(0);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kRY; ++q) {
            if (act[q]) {
                const int r = ly + q * kTY + 1;
                const T v = vc[q];
                const T Av = cx * ((lds[buf][r][lx] - v) + (lds[buf][r][lx + 2] - v)) +
                             cy * ((lds[buf][r - 1][lx + 1] - v) + (lds[buf][r + 1][lx + 1] - v)) +
                             cz * ((vm[q] - v) + (vp[q] - v));
                st_out<Op, T>(op, base[q] + g.plane * k, v, Av, rc[q], oc[q], acc);
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
    if constexpr (NQ > 0) block_partials<NQ>(acc, rb);
}

template <int NQ, class Op>
__global__ __launch_bounds__(kThreads) void k_vector(Geo g, Op op, RedBuf rb)
{
    if (op.gated()) return;
    op.load();
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int tx0 = b % g.ntx;
    const int rr = b / g.ntx;
    const int ty0 = rr % g.nty;
    const int tz0 = rr / g.nty;
    const int i = g.lo0 + tx0 * kTX + threadIdx.x;
    const int j = g.lo1 + ty0 * kTY + threadIdx.y;
    const int k0 = g.lo2 + tz0 * g.kz;
    const int k1 = min(k0 + g.kz, g.hi2);
    double acc[NQ > 0 ? NQ : 1];
#pragma unroll
    for (int q = 0; q < (NQ > 0 ? NQ : 1); ++q) acc[q] = 0.0;
    if (i < g.hi0 && j < g.hi1) {
        const long long c0 = (long long)g.ox + i + (long long)g.px * j;
        int k = k0;
        for (; k + kU <= k1; k += kU) {
            typename Op::R r[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) r[u] = op.fetch(c0 + g.plane * (k + u));
#pragma unroll
            for (int u = 0; u < kU; ++u) op.apply(c0 + g.plane * (k + u), r[u], i, j, k + u, acc);
        }
        for (; k < k1; ++k) {
            const typename Op::R r = op.fetch(c0 + g.plane * k);
            op.apply(c0 + g.plane * k, r, i, j, k, acc);
        }
    }
    if constexpr (NQ > 0) block_partials<NQ>(acc, rb);
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

template <typename T, int NQ, class Op>
static void run_stencil(const Geo& g0, hipStream_t s, const Op& op, RedBuf rb)
{
    const Geo g = retile(g0, kTH, g0.kz_st);
    if (g.nblocks <= 0 || !check_capacity<NQ, Op>(g, rb)) return;
    hipLaunchKernelGGL((k_stencil<T, NQ, Op>), dim3(g.nblocks), dim3(kTX, kTY), 0, s, g, op, rb);
    if constexpr (NQ > 0) hipLaunchKernelGGL((k_finish<NQ, Op>), dim3(1), dim3(kFinThreads), 0, s, op, rb, g.nblocks);
}

template <int NQ, class Op>
static void run_vector(const Geo& g0, hipStream_t s, const Op& op, RedBuf rb)
{
    const Geo g = retile(g0, kTY, g0.kz_vec);
    if (g.nblocks <= 0 || !check_capacity<NQ, Op>(g, rb)) return;
    hipLaunchKernelGGL((k_vector<NQ, Op>), dim3(g.nblocks), dim3(kTX, kTY), 0, s, g, op, rb);
    if constexpr (NQ > 0) hipLaunchKernelGGL((k_finish<NQ, Op>), dim3(1), dim3(kFinThreads), 0, s, op, rb, g.nblocks);
}

static const RedBuf kNoRed{nullptr, nullptr, 0};

struct Empty {};

// ------------------------------------------------------------ basic kernels

template <typename T>
struct OpApply : Gate {
    const T* __restrict__ x;
    T* __restrict__ y;
    __device__ void load() {}
    __device__ T in(long long c) const { return x[c]; }
    __device__ void out(long long c, T, T Av, double*) const { y[c] = Av; }
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
    __device__ R fetch(long long c) const { return {x[c]}; }
    __device__ void apply(long long c, const R& r, int, int, int, double*) const { y[c] = r.a; }
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

template <typename T>
struct OpDot : Gate {
    const T* __restrict__ a;
    const T* __restrict__ b;
    double* out;
    struct R { T a, b; };
    __device__ void load() {}
    __device__ R fetch(long long c) const { return {a[c], b[c]}; }
    __device__ void apply(long long, const R& r, int, int, int, double* acc) const { acc[0] += (double)r.a * (double)r.b; }
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
    __device__ R fetch(long long) const { return {}; }
    __device__ void apply(long long c, const R&, int, int, int, double*) const { a[c] = v; }
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
    __device__ R fetch(long long c) const { return {a[c]}; }
    __device__ void apply(long long c, const R& r, int, int, int, double*) const
    {
        // normalizeProblemToFieldBNorm divides (iterativeSolverBase.hpp:227-231);
        // the solvers' epilogues multiply (baseCG.hpp:248-252)
        if (mode == 0) a[c] = (T)((double)r.a / nrm);
        else a[c] = (T)((double)r.a * nrm);
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
    __device__ R fetch(long long p) const { return {in[p]}; }
    __device__ void apply(long long p, const R& r, int, int, int, double*) const { out[p] = c * r.a; }
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
    __device__ R fetch(long long c) const { return {src[c]}; }
    __device__ void apply(long long c, const R& r, int, int, int, double*) const { dst[c] = (TD)r.a; }
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
    __device__ Ox out_raw(long long c) const { return {b[c]}; }
    __device__ void out(long long c, double, double Av, const double&, const Ox& o, double* acc) const
    {
        const double rv = o.b - Av;
        r[c] = rv;
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
    __device__ Ox out_raw(long long c) const { return {b[c]}; }
    __device__ void out(long long c, double, double Av, const double&, const Ox& o, double* acc) const
    {
        const double rv = o.b - Av;
        r[c] = rv;
        acc[0] += rv * rv;
    }
    __device__ void finish(const double* t) { *norm = sqrt(t[0]); }
};

int launch_residual_norm(const Geo& g, hipStream_t s, RedBuf rb, const double* x, const double* b, double* r, double* out)
{
    const Geo gt = retile(g, kTH, g.kz_st);
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
    __device__ R fetch(long long c) const { return {b[c]}; }
    __device__ void apply(long long c, const R& r, int i, int j, int k, double* acc) const
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
    __device__ R fetch(long long c) const { return {b[c]}; }
    __device__ void apply(long long c, const R& r, int i, int j, int k, double*) const { b[c] = adj.adjust(r.b, c, i, j, k); }
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
    int nt;
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
    __device__ Ox out_raw(long long c) const { return {x[c]}; }
    __device__ void out(long long c, T vc, T Av, const Raw& r, const Ox& o, double* acc) const
    {
        put(pout + c, vc, nt);
        put(Ap + c, Av, nt);
        put(x + c, o.x + alpha_prev * r.pold, nt);
        acc[0] += (double)vc * (double)Av;
    }
    __device__ void finish(const double* t)
    {
        S->v[S_PAP] = t[0];
        S->v[S_ALPHA] = S->v[S_RZ] / t[0];
    }
};

// pin/pout double-buffer the search direction: neighbouring tiles rebuild
// p_new on their halo from p_old, so p_old must stay intact during the launch.
template <typename T>
void launch_cg_A2(const Geo& g, hipStream_t s, RedBuf rb, const T* src, const T* pin, T* pout, T* Ap, T* x, Scal* S,
                  bool pointwise)
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
    op.nt = g.nt;
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
    __device__ R fetch(long long c) const { return {x[c], p[c]}; }
    __device__ void apply(long long c, const R& v, int, int, int, double*) const { x[c] = v.x + alpha * v.p; }
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
    int nt;
    T alpha, jac;
    struct R { T r, ap; };
    __device__ void load()
    {
        alpha = (T)S->v[S_ALPHA];
        jac = (T)S->v[S_JAC];
    }
    __device__ R fetch(long long c) const { return {r[c], Ap[c]}; }
    __device__ void apply(long long c, const R& v, int, int, int, double* acc) const
    {
        const T rv = v.r - alpha * v.ap;
        put(r + c, rv, nt);
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
    op.nt = g.nt;
    run_vector<2>(g, s, op, rb);
}

template <typename T>
struct OpCgRz : Gate {
    const T* __restrict__ r;
    const T* __restrict__ z;
    Scal* S;
    struct R { T r, z; };
    __device__ void load() {}
    __device__ R fetch(long long c) const { return {r[c], z[c]}; }
    __device__ void apply(long long, const R& v, int, int, int, double* acc) const { acc[0] += (double)v.r * (double)v.z; }
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
    __device__ Ox out_raw(long long c) const { return {r0[c]}; }
    __device__ void out(long long c, T, T Av, const Raw&, const Ox& o, double* acc) const
    {
        AMp[c] = Av;
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
    __device__ R fetch(long long c) const { return {r[c], AMp[c]}; }
    __device__ void apply(long long c, const R& v, int, int, int, double*) const { r[c] = v.r - alpha * v.a; }
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
    __device__ Ox out_raw(long long c) const { return {r[c]}; }
    __device__ void out(long long c, T, T Av, const Raw&, const Ox& o, double* acc) const
    {
        Az[c] = Av;
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
    __device__ R fetch(long long c) const { return {x[c], Mp[c], z[c], r[c], Az[c], r0[c]}; }
    __device__ void apply(long long c, const R& v, int, int, int, double* acc) const
    {
        x[c] = v.x + alpha * (amp * v.mp) + omega * (az * v.z);
        const T rv = v.r - omega * v.az;
        r[c] = rv;
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
    __device__ R fetch(long long c) const { return {p[c], r[c], AMp[c]}; }
    __device__ void apply(long long c, const R& v, int, int, int, double*) const { p[c] = v.r + beta * (v.p - omega * v.a); }
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
    __device__ void out(long long c, T vc, T Av, double*) const
    {
        Z[c] = c0 * vc;
        Y[c] = c1 * (T(2) * vc + c0 * Av);
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
    __device__ void out(long long c, T vc, T Av, double*) const
    {
        const T y = which == 0 ? c0 * vc : c1 * (T(2) * vc + c0 * Av);
        X[c] = (TO)(sign * (double)y);
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
    __device__ Ox out_raw(long long c) const { return {B[c], Z[c]}; }
    __device__ void out(long long c, T vc, T Av, const T&, const Ox& o, double*) const
    {
        const T w = a * vc + b * (o.b + Av) - cz * o.z;
        if constexpr (sizeof(TO) == sizeof(T)) W[c] = sign < 0 ? -w : w;
        else W[c] = (TO)(sign * (double)w);
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
    __device__ R fetch(long long c) const { return {z[c], Apold[c]}; }
    __device__ void apply(long long, const R& v, int, int, int, double* acc) const { acc[0] += (double)v.z * (double)v.a; }
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
    run_vector<1>(g, s, op, rb);
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
    __device__ void load()
    {
        beta = (T)S->v[S_BETA];
        alpha_prev = (T)S->v[S_ALPHA];
        upd = !S->restart && alpha_prev != (T)0;
    }
    // beta = 0 takes z as it is, and pold is read only when it is used, so a
    // restart never multiplies a non-finite pold
    struct Raw { T z, pold; };
    struct Ox { T r, x; };
    __device__ Raw in_raw(long long c) const { return {z[c], (beta == (T)0 && !upd) ? (T)0 : pold[c]}; }
    __device__ T in_value(const Raw& v) const { return beta == (T)0 ? v.z : v.z - beta * v.pold; }
    __device__ Ox out_raw(long long c) const { return {r[c], upd ? x[c] : (T)0}; }
    __device__ void out(long long c, T vc, T Av, const Raw& v, const Ox& o, double* acc) const
    {
        p[c] = vc;
        Ap[c] = Av;
        if (upd) x[c] = o.x + alpha_prev * v.pold;
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
    run_stencil<T, 2>(g, s, op, rb);
}

template <typename T>
void launch_fcg_p_fused(const Geo& g, hipStream_t s, RedBuf rb, const T* z, const T* pold, T* p, T* Ap, const T* r,
                        T* x, Scal* S, bool history)
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
    run_stencil<T, 5>(g, s, op, rb);
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
    __device__ R fetch(long long c) const { return {x[c], skip ? (T)0 : p[c]}; }
    __device__ void apply(long long c, const R& v, int, int, int, double*) const
    {
        if (!skip) x[c] = v.x + alpha * v.p;
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
    __device__ R fetch(long long c) const { return {r[c], skip ? (T)0 : Ap[c]}; }
    __device__ void apply(long long c, const R& v, int, int, int, double* acc) const
    {
        T rv = v.r;
        if (!skip) {
            rv = v.r - alpha * v.ap;
            r[c] = rv;
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
    __device__ void load() {}
    struct Ox { double b; };
    __device__ double in(long long c) const { return x[c]; }
    __device__ Ox out_raw(long long c) const { return {b[c]}; }
    __device__ void out(long long c, double, double Av, const double&, const Ox& o, double* acc) const
    {
        const double rv = o.b - Av;
        const float rf = (float)rv;
        r32[c] = rf;
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
        SI->v[S_TOL] = o == 0 ? itol0 : itol;
    }
};

void launch_ir_residual(const Geo& g, hipStream_t s, RedBuf rb, const double* x, const double* b, float* r32, Scal* So,
                        Scal* SI, double itol, double itol0)
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
    run_stencil<double, 2>(g, s, op, rb);
}

struct OpIrUpdate : Gate {
    double* __restrict__ x;
    const float* __restrict__ d;
    const Scal* SI;
    float sc;
    struct R { double x; float d; };
    __device__ void load() { sc = (float)SI->v[S_INNER_NORM]; }
    __device__ R fetch(long long c) const { return {x[c], d[c]}; }
    __device__ void apply(long long c, const R& v, int, int, int, double*) const { x[c] = v.x + (double)(v.d * sc); }
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
    __device__ R fetch(long long c) const { return {r[c]}; }
    __device__ void apply(long long c, const R& v, int, int, int, double* acc) const
    {
        const float rv = v.r * inv;
        r[c] = rv;
        x[c] = 0.0f;
        p[c] = 0.0f;
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
    template void launch_cg_A2<T>(const Geo&, hipStream_t, RedBuf, const T*, const T*, T*, T*, T*, Scal*, bool); \
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
                                        Scal*, bool);                                                              \
    template void launch_fcg_final_x<T>(const Geo&, hipStream_t, T*, const T*, const Scal*);                      \
    template void launch_fcg_update<T>(const Geo&, hipStream_t, RedBuf, T*, const T*, Scal*, bool);

PPS_INST_T(double)
PPS_INST_T(float)

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
