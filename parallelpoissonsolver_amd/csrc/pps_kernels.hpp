// pps_kernels.hpp -- device data layout, per-solver device scalars and the
// launchers of the Krylov-path kernels (implemented in pps_kernels.hip).
//
// Data layout in HBM.  A field of the reference's dense guarded box
// (n0+2)(n1+2)(n2+2) (iterativeSolverBase.hpp:34-35) is stored pitched:
// element (i,j,k) of the guarded index space lives at
//      ox + i + px*j + plane*k,   plane = px*(n1+2)
// where ox puts the first solver-region column on a 128-byte boundary and px
// (a multiple of 128 bytes) keeps every row aligned, so each 64-lane
// wavefront reads whole cache lines of one x-row.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pps {

constexpr int kTX = 64;   // tile width in x: one wavefront per row
constexpr int kTY = 4;    // wavefronts of the 256-thread blocks (span sums, the reductions' default)
constexpr int kThreads = kTX * kTY;
// stencil tile: kSTY wavefronts per block, kRY rows per thread (build-time
// knobs for tile-shape sweeps)
#ifndef PPS_STENCIL_ROWS
#define PPS_STENCIL_ROWS 4
#endif
#ifndef PPS_STENCIL_WAVES
#define PPS_STENCIL_WAVES 8
#endif
constexpr int kRY = PPS_STENCIL_ROWS;
constexpr int kSTY = PPS_STENCIL_WAVES;
constexpr int kSThreads = kTX * kSTY;
constexpr int kTH = kSTY * kRY;  // stencil tile height (rows)
constexpr int kMaxRedQ = 5;
constexpr int kMinKz = 4;   // smallest z-chunk; sizes the reduction workspace

struct Layout {
    int ng[3];        // guarded extent per direction
    int px;           // row pitch (elements)
    int ox;           // column offset of guarded index 0
    long long plane;  // px * ng[1]
    long long total;  // elements allocated
    __host__ __device__ long long at(int i, int j, int k) const { return ox + i + (long long)px * j + plane * k; }
};

// [lo, hi) guarded index ranges
struct Region {
    int lo[3], hi[3];
    long long count() const { return (long long)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]); }
};

// Scalars of one Krylov solver instance, kept in device memory so that no
// iteration needs a host round trip; kernels read them at entry and the last
// block of each reduction writes them.
enum ScalIdx {
    S_ALPHA = 0, S_BETA, S_OMEGA, S_RHO0, S_RZ, S_PAP, S_PAP_OLD, S_ERR, S_ERR0, S_NORM,
    S_TOL, S_JAC, S_GAM, S_ERR_OP, S_INNER_NORM, S_DSCALE, S_RR, S_AUX0, S_AUX1, S_AUX2,
    S_COUNT
};

// FCG with deferred x updates: the steps of a ring of kFcgRing directions are
// added to x in one pass (launch_fcg_flush)
constexpr int kFcgRing = 16;

struct alignas(256) Scal {
    double v[S_COUNT];
    double ahist[kFcgRing];  // FCG ring: alpha of the step whose direction is in slot j (0: none, or discarded)
    int iter;        // completed iterations
    int done;        // 0 running, 1 converged or capped, 2 early exit before the loop
    int maxit;
    int hist_cap;    // capacity of hist (entries), 0 = no history
    double* hist;    // residual history (errorFromIterationHistory_)
    int outer;       // IR outer steps
    int fallback;    // FCG: 1 once the plain preconditioner replaced the learned one
    int restart;     // FCG: 1 = discard this iteration's direction (non-finite step) and restart
    int stall_window;  // FCG: iterations per stall check (0 = off)
};

// Deterministic grid reduction workspace: every block of a reducing launch
// stores its partial sums; a one-block epilogue kernel sums them in a fixed
// order and runs the solver's scalar update.  Launches with many blocks first
// sum fixed spans of kSpan partials into partials2 (one block per span), so
// the epilogue reads a few hundred values.
constexpr int kSpan = 2048;
struct RedBuf {
    double* partials;   // kMaxRedQ * cap
    unsigned* counter;
    int cap;            // max blocks per launch
    double* partials2;  // kMaxRedQ * cap2 span sums (null: single-level sums only)
    int cap2;
};

// Grid geometry handed to a launch: region to process and tiling.
struct Geo {
    long long plane;
    int px, ox;
    int lo0, hi0, lo1, hi1, lo2, hi2;
    int ntx, nty, kz;   // tiling of one launch (set by the launcher)
    int nblocks;
    int kz_st;          // z-chunk of stencil launches
    int nt;             // store field outputs non-temporally (working set beyond the caches)
    double cx, cy, cz;  // 1/ds^2 per direction
};

// kz_st: z-chunk length of the stencil kernels; nt: non-temporal stores
Geo make_geo(const Layout& L, const Region& R, const double coef[3], int kz_st, int nt);

// Neumann faces of the current grid, for the ghost-fill kernel.
struct Faces {
    int n;             // number of Neumann faces
    int face[6];       // face ids (x- x+ y- y+ z- z+)
    const double* flux[6];  // du/dn planes (device) or null
    int lim_data[6];
    double ds[3];
    int fng[6][2];     // plane extents of each face (guarded)
};

// ---------------------------------------------------------------- launchers
// Every loop kernel takes `gate`: when *gate != 0 the whole grid returns at
// once (the solver has converged), so iterations can be queued ahead of the
// host's convergence check without changing the iterate sequence.

// y = A x on the solver region
template <typename T>
void launch_apply_A(const Geo& g, hipStream_t s, const T* x, T* y);
// copy of the solver region (bandwidth reference)
template <typename T>
void launch_copy(const Geo& g, hipStream_t s, const T* x, T* y);
// plain copy of n contiguous doubles (16-byte non-temporal lanes; x and y
// 16-byte aligned): the measured HBM ceiling
void launch_copy_flat(hipStream_t s, const double* x, double* y, long long n);
// sum a.b over the region into *out (device)
template <typename T>
void launch_dot(const Geo& g, hipStream_t s, RedBuf rb, const T* a, const T* b, double* out, const int* gate);

// r = b - A x; S[S_ERR0 or S_ERR_OP] = ||r||, S_RZ = r.(jac r); mode 0: start of solve
// (history[0], iteration reset, early-exit flag), mode 1: final true residual;
// modes 2-4: residual replacement after a recursive stop (S_RR, S_ALPHA = 0, done
// = 0 unless below tolerance or at the iteration cap), 3 also sets S_RZ (CG with a
// pointwise preconditioner), 4 S_RHO0 = ||r||^2 (BiCGSTAB restarted with r0 = r)
void launch_residual(const Geo& g, hipStream_t s, RedBuf rb, const double* x, const double* b, double* r, Scal* S, int mode);

// r = b - A x on g's region and ||r|| into *out (device); the partials go to
// rb.partials.  Returns -1 (nothing launched) when the launch would need more
// than rb.cap blocks
int launch_residual_norm(const Geo& g, hipStream_t s, RedBuf rb, const double* x, const double* b, double* r, double* out);

// ||b_adj||^2 of the boundary-adjusted rhs over the solver region -> S_NORM = sqrt
void launch_bnorm(const Geo& g, hipStream_t s, RedBuf rb, const double* x, const double* b, const int bc[6],
                  const int lim_data[6], const int lim_solver[6], const double* const flux[6], const int fng[6][2],
                  const double ds[3], Scal* S);
// b <- adjusted b over the data region, in place (reads the Dirichlet values of x
// one layer outside the adjusted point; in-place is safe because every adjusted
// point only reads x and its own b)
void adjust_b_field(const Geo& gdata, hipStream_t s, const double* x, double* b, const int bc[6], const int lim_data[6],
                    const double* const flux[6], const int fng[6][2], const double ds[3]);
// field over the whole guarded box: mode 0 divide by S_NORM, mode 1 multiply by S_NORM
template <typename T>
void launch_scale_box(const Geo& gbox, hipStream_t s, T* a, const Scal* S, int mode);
// set a field to a constant over a region
template <typename T>
void launch_fill(const Geo& g, hipStream_t s, T* a, T v);
// Neumann ghost fill; with flux: data-field form using S_NORM of `S` (or 1 if S null)
template <typename T>
void launch_ghost(hipStream_t s, const Layout& L, const Faces& F, T* field, int with_flux, const Scal* S,
                  const int* gate);

// ---- CG (plain / pointwise-preconditioned / buffer-preconditioned) ----
// p_new = a*src + beta*p_old (a = S_JAC or 1), Ap = A p_new, pAp -> alpha = S_RZ / pAp,
// and the previous iteration's solution update x += alpha_prev p_old on the
// tile's own points.  pin/pout must be distinct buffers (halo tiles rebuild
// p_new from p_old).
template <typename T>
void launch_cg_A2(const Geo& g, hipStream_t s, RedBuf rb, const T* src, const T* pin, T* pout, T* Ap, T* x, Scal* S,
                  bool pointwise, int ring_slot = -1);  // ring_slot >= 0: as launch_fcg_p_fused's
// r -= alpha Ap, rr (and r.(jac r) when pointwise) -> err, beta, iter, done
template <typename T>
void launch_cg_B(const Geo& g, hipStream_t s, RedBuf rb, T* r, const T* Ap, Scal* S, bool pointwise, bool history);
// x += S[S_ALPHA] p: the last iteration's solution update, after the loop
template <typename T>
void launch_cg_final_x(const Geo& g, hipStream_t s, T* x, const T* p, const Scal* S);
// rz_new = r.z -> beta = rz_new / rz (0 at iteration 0), rz = rz_new (buffer-preconditioned CG)
template <typename T>
void launch_cg_rz(const Geo& g, hipStream_t s, RedBuf rb, const T* r, const T* z, Scal* S);

// ---- pointwise preconditioners: out = c * in over the whole box ----
template <typename T>
void launch_pointwise(const Geo& gbox, hipStream_t s, const T* in, T* out, double c, const int* gate);

// ---- BiCGSTAB (BiCGSTAB.hpp:131-292) ----
template <typename T>
void launch_bicg_k1(const Geo& g, hipStream_t s, RedBuf rb, const T* Mp, double amp, T* AMp, const T* r0, Scal* S);
template <typename T>
void launch_bicg_k2(const Geo& g, hipStream_t s, T* r, const T* AMp, const Scal* S);
template <typename T>
void launch_bicg_k3(const Geo& g, hipStream_t s, RedBuf rb, const T* z, double az, T* Az, const T* r, Scal* S);
template <typename T>
void launch_bicg_k45(const Geo& g, hipStream_t s, RedBuf rb, T* x, const T* Mp, double amp, const T* z, double az,
                     T* r, const T* Az, const T* r0, Scal* S, bool history);
template <typename T>
void launch_bicg_k6(const Geo& g, hipStream_t s, T* p, const T* r, const T* AMp, const Scal* S);

// ---- Chebyshev (chebyshevIteration.hpp:78-128) ----
// coefficients for each launch are computed on the host (they depend only on
// the grid's eigenvalue bounds): first: Z = c0*B, Y = c1*(2B + c0*AB)
template <typename T>
void launch_cheb_first(const Geo& g, hipStream_t s, const T* B, T* Z, T* Y, double c0, double c1, const int* gate);
// W = a*Y + b*(B + AY) - c*Z  (output type TO may differ; sign -1 writes -W)
template <typename T, typename TO>
void launch_cheb_sweep(const Geo& g, hipStream_t s, const T* Y, const T* B, const T* Z, TO* W, double a, double b,
                       double c, double sign, const int* gate);
// X = sign*c0*B (degree-0/1 outputs) and X = sign*c1*(2B + c0*AB)
// two fused sweeps (c, c+1) of the Chebyshev recurrence: W1 = y_{c+1} (not
// written when `last`), W2 = y_{c+2} (sign-applied and converted to TO when
// `last`, the stage output).  Y = y_c needs its Neumann ghosts set; Z = y_{c-1}.
struct ChebPairGeo {
    int lo_d[3], hi_d[3];  // data region (the Neumann ghosts' mirror rule)
    int ng[3];             // guarded box (loads outside it read 0)
    int neu;               // bit f set: face f (x- x+ y- y+ z- z+) is Neumann
    int kz;                // z-chunk of one block
};
template <typename T, typename TO>
void launch_cheb_pair(const Geo& g, const ChebPairGeo& cg, hipStream_t s, const T* Y, const T* Z, const T* B, T* W1,
                      TO* W2, const double a[2], const double b[2], const double c[2], double sign, bool last,
                      const int* gate);
template <typename T, typename TO>
void launch_cheb_first_out(const Geo& g, hipStream_t s, const T* B, TO* X, double c0, double c1, int which,
                           double sign, const int* gate);
// dst = (TD) src over the whole box
template <typename TS, typename TD>
void launch_cast(const Geo& gbox, hipStream_t s, const TS* src, TD* dst, const int* gate);

// ---- FCG (truncation 1) ----
template <typename T>
void launch_fcg_gamma(const Geo& g, hipStream_t s, RedBuf rb, const T* z, const T* Apold, Scal* S, bool fallback_only);
// finish of gamma from per-block partials of z.Apold already in rb (written by
// the preconditioner); gated off while S->fallback is set
template <typename T>
void launch_fcg_gamma_finish(hipStream_t s, RedBuf rb, int nblocks, Scal* S);
template <typename T>
void launch_fcg_p(const Geo& g, hipStream_t s, RedBuf rb, const T* z, const T* pold, T* p, T* Ap, const T* r, T* x,
                  Scal* S);
// launch_fcg_p when the preconditioner applies the residual update itself
// (r of the next iteration = r - alpha Ap, formed while it reads r): also ends
// the iteration (||r||^2 of the updated residual, convergence, stall monitor)
// ring_slot >= 0: x is not touched; the step's alpha goes to
// S->ahist[ring_slot] (0 when the step is discarded) for launch_fcg_flush
template <typename T>
void launch_fcg_p_fused(const Geo& g, hipStream_t s, RedBuf rb, const T* z, const T* pold, T* p, T* Ap, const T* r,
                        T* x, Scal* S, bool history, int ring_slot = -1);
template <typename T>
void launch_fcg_final_x(const Geo& g, hipStream_t s, T* x, const T* p, const Scal* S);
// x += sum over the ring's slots j = 0 .. kFcgRing-1 (in that order) of
// S->ahist[j] * ring[j], skipping the slots whose alpha is 0; then clears the
// alphas.  Adding the steps in the order they were taken gives x bit for bit
// as the per-iteration updates do.
template <typename T>
void launch_fcg_flush(const Geo& g, hipStream_t s, T* x, T* const ring[kFcgRing], Scal* S);
// clears S->ahist (a solve's start)
void launch_fcg_ring_reset(hipStream_t s, Scal* S);
template <typename T>
void launch_fcg_update(const Geo& g, hipStream_t s, RedBuf rb, T* r, const T* Ap, Scal* S, bool history);

// ---- mixed-precision iterative refinement ----
// r = b - A x (FP64), r32 = (float) r; So: S_ERR = ||r||, history[outer];
// SI: S_INNER_NORM = ||r32||
// r32 = b - A x (FP64, narrowed), the outer residual norm and convergence
// test; sets the inner stage's tolerance to itol0 before the first inner
// solve and to itol before the later ones, or (adapt > 0) to no less than
// adapt * tol / ||r||, the reduction the outer tolerance still needs
void launch_ir_residual(const Geo& g, hipStream_t s, RedBuf rb, const double* x, const double* b, float* r32, Scal* So,
                        Scal* SI, double itol, double itol0, double adapt);
// x += ||r32|| * d on the solver region (the inner solve's un-normalisation)
void launch_ir_update(const Geo& g, hipStream_t s, double* x, const float* d, const Scal* SI);
// inner start: r = r32 / ||r32||, x = 0, p = 0, err0, rz (pointwise jac) -> SI
void launch_ir_inner_init(const Geo& g, hipStream_t s, RedBuf rb, float* r, float* x, float* p, Scal* SI, bool pointwise);

}  // namespace pps

namespace pps {
// ---- seeded right-hand sides (pps_rhs.hip) ----
void launch_rhs_normal(hipStream_t s, const Layout& L, const Region& R, const int np[3], uint64_t seed, double* b,
                       double* scratch);
void launch_rhs_gauss(hipStream_t s, const Layout& L, const Region& R, const double* src_dev, int nsrc,
                      const double origin[3], const double ds[3], double* b);
void launch_rhs_spectrum(hipStream_t s, const Layout& L, const Region& R, const int np[3], int K, const double* C,
                         const double* Sx, const double* Sy, const double* Sz, double* T1, double* T2, int accumulate,
                         double* b);
}  // namespace pps
