// pps_host.hip -- host side of libpps: the context behind the C ABI, the
// solver chain (main solver + nested preconditioners, the template chain of
// the reference's inputParam.hpp), and the launch control of the Krylov loops.
//
// Launch control.  All Krylov scalars live on the device, so an iteration is a
// fixed sequence of kernel launches with no host round trip.  The host queues
// iterations in batches and keeps two batches in flight: after queuing batch
// k+1 it waits for the event of batch k and reads the solver's convergence
// flag that batch k copied to pinned memory.  Every loop kernel is gated on
// that flag, so iterations queued past convergence retire as empty launches
// and the iterate sequence, iteration count and stopping rule are exactly
// those of the reference loop (baseCG.hpp:115-228, BiCGSTAB.hpp:131-292).
#include "../../include/pps_abi.h"
#include "pps_kernels.hpp"
#include "pps_nn.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace pps {

static thread_local std::string g_create_error;

struct Field {
    void* p = nullptr;
    int prec = 64;
    double* d() const { return static_cast<double*>(p); }
    float* f() const { return static_cast<float*>(p); }
};

struct Ctx;

// -------------------------------------------------------------- solvers

struct Stage {
    Ctx& C;
    pps_stage cfg;
    int level;
    Stage* pre = nullptr;     // preconditioner stage (null = identity)
    Scal* S = nullptr;        // device scalars
    long long prec_iters = 0; // iterations spent in this stage as a preconditioner
    long long calls = 0;
    Stage(Ctx& c, const pps_stage& s, int l) : C(c), cfg(s), level(l) {}
    virtual ~Stage() = default;
    int prec() const { return cfg.precision == 32 ? 32 : 64; }
    // pointwise preconditioner: out = c * in
    virtual bool pointwise(double* c) const { return false; }
    // needs host control per application (Krylov preconditioners)
    virtual bool host_driven() const { return false; }
    // Preconditioner iterations one application adds to
    // getNumIterationPreconditionerFinal.  Applications of a stage that is not
    // host driven are queued ahead of the caller's convergence test and gated,
    // so the caller adds calls * iters_per_call() once it knows how many of
    // its applications ran.
    virtual long long iters_per_call() const { return 0; }
    // operator()(out, in) as a preconditioner; kernels gated on *gate
    virtual int apply(Field out, Field in, const int* gate) = 0;
    // apply() that may also sum out . dot into the context's reduction
    // partials: *nblocks = their count, or 0 when the stage does not
    virtual int apply_dot(Field out, Field in, const int* gate, Field dot, int* nblocks)
    {
        (void)dot;
        *nblocks = 0;
        return apply(out, in, gate);
    }
    // whether apply_update() is available: the stage forms its input as the
    // updated residual in = r_prev - alpha ap_prev (r_prev while *skip) and
    // writes it to `in`, so the Krylov loop needs no update pass of its own
    virtual bool forms_update() const { return false; }
    virtual int apply_update(Field out, Field in, Field r_prev, Field ap_prev, const double* alpha, const int* skip,
                             const int* gate, Field dot, int* nblocks)
    {
        (void)out; (void)in; (void)r_prev; (void)ap_prev; (void)alpha; (void)skip; (void)gate; (void)dot;
        *nblocks = 0;
        return PPS_ERR_UNSUPPORTED;
    }
    // operator()(x, b) as the main solver on the context's X and B
    virtual int solve_main() { return PPS_ERR_UNSUPPORTED; }
};

struct Ctx {
    pps_grid grid{};
    int np[3]{}, ng[3]{};
    Layout L64{}, L32{};
    Region solver{}, data{}, box{};
    int lim_data[6]{}, lim_solver[6]{};
    double coef[3]{};
    double eig_local[2]{}, eig_global[2]{};
    hipStream_t stream = nullptr;
    hipEvent_t ev_batch[2]{};
    hipEvent_t ev_t0 = nullptr, ev_t1 = nullptr;
    RedBuf rb{};
    Faces faces{};
    double* flux_d[6]{};
    int fng[6][2]{};
    Field fields[PPS_NUM_FIELDS];
    std::map<std::string, Field> work;
    Scal* scal_d = nullptr;  // pool of device scalar blocks
    int nscal = 16;
    int* flag_h = nullptr;   // pinned: [2][4] (iter, done, outer, pad) per in-flight batch
    Scal* scal_h = nullptr;  // pinned copy of one Scal
    double* hist_d = nullptr;
    int hist_cap = 0;
    int hist_len = 0;
    int kz_st = 0;    // z-chunk of stencil launches (0: chosen per launch)
    int nt = 0;       // non-temporal stores of field outputs
    int nt_mode = 0;  // 0: automatic (nt_auto), 1: always, -1: never
    int nt_auto = 0;  // the fields' working set exceeds the last-level cache
    int cheb_pairs = 1;  // Chebyshev sweeps run two per pass (k_cheb_pair)
    int deferred_x = 1;  // Krylov stages keep a ring of directions where the device holds it
    int cheb_kz = 0;     // z-chunk of the fused Chebyshev pair (0: chosen per launch)
    std::string err;
    std::shared_ptr<NNPrecond> nn;  // learned preconditioner of PPS_NN stages (pps_nn_attach)
    int fallback_switches = 0;      // FCG stages that switched to the fallback, this solve

    int fail(int code, const char* fmt, ...)
    {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
    int hip(hipError_t e, const char* what)
    {
        if (e == hipSuccess) return PPS_OK;
        return fail(e == hipErrorOutOfMemory ? PPS_ERR_NOMEM : PPS_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    }
    const Layout& lay(int prec) const { return prec == 32 ? L32 : L64; }
    Geo geo(const Region& R, int prec) const { return make_geo(lay(prec), R, coef, kz_st, nt); }
    Geo gsolver(int prec) const { return geo(solver, prec); }
    Geo gbox(int prec) const { return geo(box, prec); }
    Geo gdata(int prec) const { return geo(data, prec); }
    ChebPairGeo cheb_pair_geo() const
    {
        ChebPairGeo cg{};
        for (int d = 0; d < 3; ++d) {
            cg.lo_d[d] = lim_data[2 * d];
            cg.hi_d[d] = lim_data[2 * d + 1];
            cg.ng[d] = ng[d];
        }
        for (int f = 0; f < 6; ++f)
            if (grid.bcs_type[f] == 1) cg.neu |= 1 << f;
        cg.kz = cheb_kz;
        return cg;
    }
    // the fused pair needs the Neumann mirror point (two cells inward) inside the data region
    bool cheb_pairs_ok() const { return cheb_pairs && np[0] >= 2 && np[1] >= 2 && np[2] >= 2; }

    int alloc_field(Field& f, int prec)
    {
        const size_t bytes = (size_t)lay(prec).total * (prec == 32 ? 4 : 8);
        f.prec = prec;
        int rc = hip(hipMalloc(&f.p, bytes), "hipMalloc field");
        if (rc) return rc;
        return hip(hipMemsetAsync(f.p, 0, bytes, stream), "hipMemsetAsync field");
    }
    int get_work(const std::string& key, int prec, Field& out)
    {
        const std::string k = key + (prec == 32 ? "/32" : "/64");
        auto it = work.find(k);
        if (it != work.end()) {
            out = it->second;
            return PPS_OK;
        }
        Field f;
        int rc = alloc_field(f, prec);
        if (rc) return rc;
        work[k] = f;
        out = f;
        return PPS_OK;
    }
    // the ring of search directions of a Krylov stage with deferred x updates
    // (`key`, slot 0 = first, slot kFcgRing-1 = last, the others work fields):
    // kFcgRing slots, or 0 when the device cannot hold the other kFcgRing-2
    // fields with 8 more to spare (the stage then updates x every iteration)
    int ring(const std::string& key, int prec, Field first, Field last, Field* out, int* slots)
    {
        *slots = 0;
        if (!deferred_x) return PPS_OK;
        bool have = true;
        for (int j = 1; j + 1 < kFcgRing; ++j)
            have = have && work.count(key + ".ring" + std::to_string(j) + (prec == 32 ? "/32" : "/64")) > 0;
        if (!have) {
            size_t fr = 0, tot = 0;
            const size_t bytes = (size_t)lay(prec).total * (prec == 32 ? 4 : 8);
            if (hipMemGetInfo(&fr, &tot) != hipSuccess || fr < (size_t)(kFcgRing - 2 + 8) * bytes) return PPS_OK;
        }
        out[0] = first;
        out[kFcgRing - 1] = last;
        for (int j = 1; j + 1 < kFcgRing; ++j) {
            int rc = get_work(key + ".ring" + std::to_string(j), prec, out[j]);
            if (rc) return rc;
        }
        *slots = kFcgRing;
        return PPS_OK;
    }
    void zero(Field f)
    {
        if (f.prec == 32) launch_fill<float>(gbox(32), stream, f.f(), 0.0f);
        else launch_fill<double>(gbox(64), stream, f.d(), 0.0);
    }
    void ghost(Field f, bool with_flux, const Scal* S, const int* gate)
    {
        if (faces.n == 0) return;
        if (f.prec == 32) launch_ghost<float>(stream, L32, faces, f.f(), 0, nullptr, gate);
        else launch_ghost<double>(stream, L64, faces, f.d(), with_flux ? 1 : 0, S, gate);
    }
    Scal* scal(int i) const { return scal_d + i; }
};

// tiny scalar-update kernel: keeps scalar bookkeeping on the stream, no
// host/device round trip and no memset nodes
enum { SC_SET = 0, SC_RESET_ITER = 1, SC_OUTER_INC = 2, SC_SET_DONE = 3 };
__global__ void k_scal_op(Scal* S, int op, int idx, double v)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (op == SC_SET) S->v[idx] = v;
    else if (op == SC_RESET_ITER) {
        S->iter = 0;
        S->done = 0;
        S->outer = 0;
    } else if (op == SC_OUTER_INC) S->outer += 1;
    else if (op == SC_SET_DONE) S->done = (int)v;
}

static void scal_op(Ctx& C, Scal* S, int op, int idx = 0, double v = 0.0)
{
    hipLaunchKernelGGL(k_scal_op, dim3(1), dim3(64), 0, C.stream, S, op, idx, v);
}

// initialise a scalar block from the host (stream ordered through pinned staging)
static int scal_init(Ctx& C, Scal* S, double tol, int maxit, double jac, double* hist, int hist_cap,
                     int stall_window = 0, double stall_ratio = 0.0)
{
    int rc = C.hip(hipStreamSynchronize(C.stream), "sync");
    if (rc) return rc;
    Scal h{};
    h.v[S_TOL] = tol;
    h.v[S_JAC] = jac;
    h.v[S_NORM] = 1.0;
    h.v[S_BETA] = 0.0;
    h.v[S_RHO0] = 1.0;
    h.v[S_PAP_OLD] = 1.0;
    h.v[S_AUX1] = stall_ratio;
    h.v[S_AUX2] = 1e300;
    h.stall_window = stall_window > 0 ? stall_window : 0;
    h.maxit = maxit;
    h.hist = hist;
    h.hist_cap = hist_cap;
    *C.scal_h = h;
    rc = C.hip(hipMemcpyAsync(S, C.scal_h, sizeof(Scal), hipMemcpyHostToDevice, C.stream), "scal init");
    if (rc) return rc;
    return C.hip(hipStreamSynchronize(C.stream), "sync");
}

static int scal_read(Ctx& C, const Scal* S, Scal& out)
{
    int rc = C.hip(hipMemcpyAsync(C.scal_h, S, sizeof(Scal), hipMemcpyDeviceToHost, C.stream), "scal read");
    if (rc) return rc;
    rc = C.hip(hipStreamSynchronize(C.stream), "sync");
    if (rc) return rc;
    out = *C.scal_h;
    return PPS_OK;
}

// Runs `body` (one iteration's launches) until S->done, two batches in flight.
template <class Body>
static int run_loop(Ctx& C, Scal* S, int maxit, bool host_each, Body body)
{
    if (maxit <= 0) return PPS_OK;  // while (iter < maxIteration) never runs
    // Batch size: 4 doubling, then sized from the measured time per iteration
    // so that a batch holds about kBatchMs of device work (at most 64
    // iterations).  Iterations queued past convergence retire as gated empty
    // launches of a few microseconds each; with expensive iterations (the
    // learned preconditioner: milliseconds) a 64-iteration batch would leave
    // up to ~100 of them, with cheap ones (small grids) the host needs the
    // large batches to stay ahead.
    constexpr double kBatchMs = 5.0;
    using clock = std::chrono::steady_clock;
    int batch = host_each ? 1 : 4;
    int queued = 0;
    int slot = 0;
    bool have_prev = false;
    int n_slot[2] = {0, 0};  // iterations of the batch behind each event
    clock::time_point t_done{};
    bool have_done = false;
    for (;;) {
        const int nb = batch;
        for (int b = 0; b < nb; ++b) {
            int rc = body();
            if (rc) return rc;
        }
        queued += nb;
        n_slot[slot] = nb;
        int rc = C.hip(hipMemcpyAsync(C.flag_h + 4 * slot, &S->iter, 2 * sizeof(int), hipMemcpyDeviceToHost, C.stream),
                       "flag copy");
        if (rc) return rc;
        rc = C.hip(hipEventRecord(C.ev_batch[slot], C.stream), "event record");
        if (rc) return rc;
        rc = C.hip(hipGetLastError(), "kernel launch");
        if (rc) return rc;
        const int waitslot = host_each ? slot : slot ^ 1;
        double ms_per_it = 0.0;
        if (host_each || have_prev) {
            rc = C.hip(hipEventSynchronize(C.ev_batch[waitslot]), "event sync");
            if (rc) return rc;
            if (C.flag_h[4 * waitslot + 1] != 0) break;
            // with two batches in flight the device ran the waited batch
            // between the two completions the host saw
            const clock::time_point now = clock::now();
            if (have_done && n_slot[waitslot] > 0)
                ms_per_it = std::chrono::duration<double, std::milli>(now - t_done).count() / n_slot[waitslot];
            t_done = now;
            have_done = true;
        }
        have_prev = true;
        slot ^= 1;
        if (!host_each) {
            if (ms_per_it > 0.0)
                batch = std::max(1, std::min(64, (int)std::ceil(kBatchMs / ms_per_it)));
            else
                batch = std::min(batch * 2, 64);
        }
        if (queued > maxit + 2 * 64) break;  // every gate is closed by now
    }
    return C.hip(hipStreamSynchronize(C.stream), "sync");
}

// normalizeProblemToFieldBNorm (iterativeSolverBase.hpp:171-234) on device
static void normalize(Ctx& C, Scal* S, bool main, Field x, Field b)
{
    if (main) {
        launch_bnorm(C.gsolver(64), C.stream, C.rb, x.d(), b.d(), C.grid.bcs_type, C.lim_data, C.lim_solver,
                     C.flux_d, C.fng, C.grid.ds, S);
    } else {
        // plain ||b|| of the preconditioner input; FP64 accumulation
        if (b.prec == 32) launch_dot<float>(C.gsolver(32), C.stream, C.rb, b.f(), b.f(), &S->v[S_AUX0], nullptr);
        else launch_dot<double>(C.gsolver(64), C.stream, C.rb, b.d(), b.d(), &S->v[S_AUX0], nullptr);
    }
}

__global__ void k_sqrt_norm(Scal* S)
{
    if (threadIdx.x == 0) S->v[S_NORM] = sqrt(S->v[S_AUX0]);
}

static void scale_box(Ctx& C, Field f, const Scal* S, int mode)
{
    if (f.prec == 32) launch_scale_box<float>(C.gbox(32), C.stream, f.f(), S, mode);
    else launch_scale_box<double>(C.gbox(64), C.stream, f.d(), S, mode);
}

// ------------------------------------------------------------------ None / Jacobi

struct NoneStage : Stage {
    using Stage::Stage;
    bool pointwise(double* c) const override
    {
        *c = 1.0;
        return true;
    }
    int apply(Field out, Field in, const int* gate) override
    {
        // NoneSolver copies the whole guarded array (noneSolver.hpp:24-27)
        if (in.prec == 32) launch_pointwise<float>(C.gbox(32), C.stream, in.f(), out.f(), 1.0, gate);
        else launch_pointwise<double>(C.gbox(64), C.stream, in.d(), out.d(), 1.0, gate);
        return PPS_OK;
    }
};

struct JacobiStage : Stage {
    using Stage::Stage;
    double scale() const
    {
        const double* ds = C.grid.ds;
        return 1.0 / (-2.0 * (1.0 / (ds[0] * ds[0]) + 1.0 / (ds[1] * ds[1]) + 1.0 / (ds[2] * ds[2])));
    }
    bool pointwise(double* c) const override
    {
        *c = scale();
        return true;
    }
    int apply(Field out, Field in, const int* gate) override
    {
        if (in.prec == 32) launch_pointwise<float>(C.gbox(32), C.stream, in.f(), out.f(), scale(), gate);
        else launch_pointwise<double>(C.gbox(64), C.stream, in.d(), out.d(), scale(), gate);
        return PPS_OK;
    }
};

// ---------------------------------------------------------------- Chebyshev

struct ChebStage : Stage {
    double theta = 0, delta = 0, sigma = 0;
    Field Y, Z, W, V, Bc;  // V: fourth iterate buffer of fused pairs; Bc: FP32 copy of the input
    ChebStage(Ctx& c, const pps_stage& s, int l) : Stage(c, s, l)
    {
        // chebyshevIteration.hpp:22-26 (global, rescaled); alpaka local form
        if (cfg.cheb_local) {
            theta = (C.eig_local[0] + C.eig_local[1]) * 0.5;
            delta = (C.eig_local[0] - C.eig_local[1]) * 0.5;
        } else {
            theta = (C.eig_global[0] * cfg.cheb_rescale_min + C.eig_global[1] * cfg.cheb_rescale_max) * 0.5 *
                    (1.0 + cfg.cheb_eps);
            delta = (C.eig_global[0] * cfg.cheb_rescale_min - C.eig_global[1] * cfg.cheb_rescale_max) * 0.5;
        }
        sigma = theta / delta;
    }
    int buffers()
    {
        const int P = prec();
        const std::string k = "cheb" + std::to_string(level);
        int rc = C.get_work(k + ".Y", P, Y);
        if (!rc) rc = C.get_work(k + ".Z", P, Z);
        if (!rc) rc = C.get_work(k + ".W", P, W);
        if (!rc && C.cheb_pairs_ok()) rc = C.get_work(k + ".V", P, V);
        if (!rc && P == 32) rc = C.get_work(k + ".B", 32, Bc);
        return rc;
    }
    template <typename T, typename TO>
    void sweeps(const T* B, TO* X, const int* gate, Field Yf, Field Zf, Field Wf, Field Vf)
    {
        // y1 = B/theta, y2 = 2 rho1/delta (2B + A B/theta), y_{c+1} from the
        // three-term recurrence; the reference returns -W after its final
        // rotation, which is -y_{m-1} (chebyshevIteration.hpp:93-128): the two
        // sweeps computing y_m and y_{m+1} do not reach the output and are skipped.
        // Sweeps c = 2 .. m-2 run two per pass (k_cheb_pair) where the grid
        // allows it, one first when their number is odd.
        const int m = cfg.max_iter;
        const Geo g = C.gsolver(sizeof(T) == 4 ? 32 : 64);
        double rhoOld = 1 / sigma;
        double rhoCurr = 1 / (2 * sigma - rhoOld);
        const double c0 = 1 / theta, c1 = 2 * rhoCurr / delta;
        if (m <= 1) {
            launch_fill<TO>(C.gsolver(sizeof(TO) == 4 ? 32 : 64), C.stream, X, TO(0));
            return;
        }
        if (m == 2 || m == 3) {
            launch_cheb_first_out<T, TO>(g, C.stream, B, X, c0, c1, m == 2 ? 0 : 1, -1.0, gate);
            return;
        }
        T* Yp = static_cast<T*>(Yf.p);
        T* Zp = static_cast<T*>(Zf.p);
        T* Wp = static_cast<T*>(Wf.p);
        T* Vp = static_cast<T*>(Vf.p);
        const bool pairs = C.cheb_pairs_ok() && Vp != nullptr;
        const ChebPairGeo cg = C.cheb_pair_geo();
        launch_cheb_first<T>(g, C.stream, B, Zp, Yp, c0, c1, gate);
        Field yf;
        yf.prec = sizeof(T) == 4 ? 32 : 64;
        double a[2], b[2], cz[2];
        auto next = [&](int q) {
            rhoOld = rhoCurr;
            rhoCurr = 1 / (2 * sigma - rhoOld);
            a[q] = rhoCurr * 2 * sigma;
            b[q] = rhoCurr * 2 / delta;
            cz[q] = rhoCurr * rhoOld;
        };
        for (int c = 2; c <= m - 2;) {
            const int left = m - 2 - c + 1;
            yf.p = Yp;
            C.ghost(yf, false, nullptr, gate);
            if (pairs && left % 2 == 0) {
                next(0);
                next(1);
                if (left > 2) {
                    launch_cheb_pair<T, T>(g, cg, C.stream, Yp, Zp, B, Wp, Vp, a, b, cz, 1.0, false, gate);
                    T* y = Yp;  // y_{c+2} -> Y, y_{c+1} -> Z
                    T* z = Zp;
                    Yp = Vp;
                    Zp = Wp;
                    Wp = y;
                    Vp = z;
                } else {
                    launch_cheb_pair<T, TO>(g, cg, C.stream, Yp, Zp, B, Wp, X, a, b, cz, -1.0, true, gate);
                }
                c += 2;
            } else {
                next(0);
                if (c < m - 2) {
                    launch_cheb_sweep<T, T>(g, C.stream, Yp, B, Zp, Wp, a[0], b[0], cz[0], 1.0, gate);
                    T* t = Zp;  // swap(Z,Y); swap(W,Y)
                    Zp = Yp;
                    Yp = Wp;
                    Wp = t;
                } else {
                    launch_cheb_sweep<T, TO>(g, C.stream, Yp, B, Zp, X, a[0], b[0], cz[0], -1.0, gate);
                }
                c += 1;
            }
        }
    }
    int apply(Field out, Field in, const int* gate) override
    {
        int rc = buffers();
        if (rc) return rc;
        calls++;
        // resetNeumanBCs on the input (chebyshevIteration.hpp:74), in place
        if (prec() == 64 && in.prec == 64) {
            C.ghost(in, false, nullptr, gate);
            if (out.prec == 64) sweeps<double, double>(in.d(), out.d(), gate, Y, Z, W, V);
            else return C.fail(PPS_ERR_UNSUPPORTED, "FP64 Chebyshev into an FP32 field");
        } else {
            // CastPrecisionFieldKernel (kernelsAlpakaChebyshev.hpp:7-18) then FP32 sweeps
            const float* B = nullptr;
            if (in.prec == 64) {
                C.ghost(in, false, nullptr, gate);
                launch_cast<double, float>(C.gbox(32), C.stream, in.d(), Bc.f(), gate);
                B = Bc.f();
            } else {
                C.ghost(in, false, nullptr, gate);
                B = in.f();
            }
            if (out.prec == 64) sweeps<float, double>(B, out.d(), gate, Y, Z, W, V);
            else sweeps<float, float>(B, out.f(), gate, Y, Z, W, V);
        }
        return PPS_OK;
    }
    int solve_main() override;
    // maxIteration per call (BiCGstabAlpaka.hpp:214,290 add the Chebyshev
    // stage's numIterationFinal_ = maxIteration, chebyshevIterationAlpaka.hpp:282)
    long long iters_per_call() const override { return cfg.max_iter; }
};

// ------------------------------------------------------------------ Krylov base

// stall test of a flexible Krylov stage: configured on its PPS_NN
// preconditioner stage if it has one, else on the stage itself
static const pps_stage& stall_cfg(const Stage* k)
{
    if (k->pre && k->pre->cfg.kind == PPS_NN) return k->pre->cfg;
    return k->cfg;
}

struct KrylovStage : Stage {
    using Stage::Stage;
    bool host_driven() const override { return true; }
    bool pre_pointwise(double* c) const
    {
        if (!pre) {
            *c = 1.0;
            return true;
        }
        return pre->pointwise(c);
    }
    bool pre_host() const { return pre && pre->host_driven(); }
    // `calls` applications of a queued (not host-driven) preconditioner ran
    void count_pre(long long calls)
    {
        if (pre && !pre->host_driven()) pre->prec_iters += calls * pre->iters_per_call();
    }
    int pre_apply(Field out, Field in, const int* gate)
    {
        if (!pre) {
            NoneStage none(C, pps_stage{}, level + 1);
            return none.apply(out, in, gate);
        }
        return pre->apply(out, in, gate);
    }
    int pre_apply_dot(Field out, Field in, const int* gate, Field dot, int* nblocks)
    {
        *nblocks = 0;
        if (!pre) return pre_apply(out, in, gate);
        return pre->apply_dot(out, in, gate, dot, nblocks);
    }
    // common prologue of the main/inner loop: returns 1 if the loop must not run
    int prologue(bool main, Field x, Field b, Field r, int* early)
    {
        if (main) C.ghost(x, true, S, nullptr);
        else C.zero(x);
        normalize(C, S, main, x, b);
        if (!main) hipLaunchKernelGGL(k_sqrt_norm, dim3(1), dim3(64), 0, C.stream, S);
        scale_box(C, x, S, 0);
        scale_box(C, b, S, 0);
        C.ghost(x, main, S, nullptr);
        launch_residual(C.gsolver(64), C.stream, C.rb, x.d(), b.d(), r.d(), S, 0);
        Scal h;
        int rc = scal_read(C, S, h);
        if (rc) return rc;
        *early = (h.done == 2);
        return PPS_OK;
    }
    void epilogue(bool main, Field x, Field b, Field r)
    {
        if (main) {
            C.ghost(x, true, S, nullptr);
            launch_residual(C.gsolver(64), C.stream, C.rb, x.d(), b.d(), r.d(), S, 1);
        }
        scale_box(C, x, S, 1);
        scale_box(C, b, S, 1);
    }
    // true-residual finish of a main solve (pps_stage::true_finish): r = b - A x
    // of the complete x in FP64 (launch_residual modes 2-4); h->done tells
    // whether the solve ends here
    int replace_residual(Field x, Field b, Field r, int mode, Scal* h)
    {
        C.ghost(x, true, S, nullptr);
        launch_residual(C.gsolver(64), C.stream, C.rb, x.d(), b.d(), r.d(), S, mode);
        return scal_read(C, S, *h);
    }
};

// ------------------------------------------------------------------------ CG

struct CGStage : KrylovStage {
    using KrylovStage::KrylovStage;
    // BaseCG::operator() (baseCG.hpp:44-260), FP64 vectors
    int run(bool main, Field x, Field b)
    {
        if (prec() != 64 || x.prec != 64) return C.fail(PPS_ERR_UNSUPPORTED, "CG stage %d: FP32 vectors only inside PPS_IR", level);
        const std::string k = "cg" + std::to_string(level);
        Field r, p0, p1, Ap, z;
        int rc = C.get_work(k + ".r", 64, r);
        if (!rc) rc = C.get_work(k + ".p0", 64, p0);
        if (!rc) rc = C.get_work(k + ".p1", 64, p1);
        if (!rc) rc = C.get_work(k + ".Ap", 64, Ap);
        double c = 1.0;
        const bool pw = pre_pointwise(&c);
        if (!rc && !pw) rc = C.get_work(k + ".z", 64, z);
        if (rc) return rc;
        C.zero(r);
        C.zero(p0);
        C.zero(p1);
        C.zero(Ap);
        if (!pw) C.zero(z);
        rc = scal_init(C, S, cfg.tol, cfg.max_iter, pw ? c : 1.0, main ? C.hist_d : nullptr, main ? C.hist_cap : 0);
        if (rc) return rc;
        int early = 0;
        rc = prologue(main, x, b, r, &early);
        if (rc) return rc;
        if (early) return PPS_OK;  // baseCG.hpp:103-107: x, b stay normalised
        const Geo g = C.gsolver(64);
        if (!pw) {
            rc = pre_apply(z, r, &S->done);
            if (rc) return rc;
            launch_cg_rz<double>(g, C.stream, C.rb, r.d(), z.d(), S);
        }
        Field p[2] = {p0, p1};
        // where the device holds a ring of directions, x's updates are deferred
        // to one pass per nring iterations (iteration q's direction in slot q
        // mod nring; slot 0 = p1 and the last slot = p0 are zero, so the first
        // iteration reads zeros as its old direction)
        Field ring[kFcgRing];
        double* ringp[kFcgRing] = {};
        int nring = 0;
        rc = C.ring(k, 64, p1, p0, ring, &nring);
        if (rc) return rc;
        for (int j = 0; j < nring; ++j) ringp[j] = ring[j].d();
        if (nring) launch_fcg_ring_reset(C.stream, S);
        int q = 0;
        int cur = 0;
        auto body = [&]() -> int {
            // iteration k reads p[k&1], writes p[(k+1)&1], and applies
            // iteration k-1's x update on the way (with the ring: reads slot
            // k-1, writes slot k, and leaves x to the flush)
            if (nring) {
                const int qs = q % nring, qo = (q + nring - 1) % nring;
                launch_cg_A2<double>(g, C.stream, C.rb, pw ? r.d() : z.d(), ring[qo].d(), ring[qs].d(), Ap.d(), x.d(),
                                     S, pw, qs);
                if (qs == nring - 1) launch_fcg_flush<double>(g, C.stream, x.d(), ringp, S);
                ++q;
            } else {
                launch_cg_A2<double>(g, C.stream, C.rb, pw ? r.d() : z.d(), p[cur].d(), p[cur ^ 1].d(), Ap.d(), x.d(), S,
                                     pw);
            }
            cur ^= 1;
            launch_cg_B<double>(g, C.stream, C.rb, r.d(), Ap.d(), S, pw, main);
            if (!pw) {
                int e = pre_apply(z, r, &S->done);
                if (e) return e;
                launch_cg_rz<double>(g, C.stream, C.rb, r.d(), z.d(), S);
            }
            return PPS_OK;
        };
        Scal h;
        // preconditioner applications: the one before the loop, then one per
        // iteration except the stopping one (gated), one per replacement
        long long pre_calls = pw ? 0 : 1;
        int it0 = 0;
        for (int repl = 0;; ++repl) {
            rc = run_loop(C, S, cfg.max_iter, pre_host(), body);
            if (rc) return rc;
            // the last executed iteration K-1 left x += alpha_{K-1} p_{K-1} pending; p_{K-1} is p[K&1]
            rc = scal_read(C, S, h);
            if (rc) return rc;
            if (!pw) pre_calls += std::max(0, h.iter - it0 - 1);
            it0 = h.iter;
            if (nring) launch_fcg_flush<double>(g, C.stream, x.d(), ringp, S);
            else if (h.iter > 0) launch_cg_final_x<double>(g, C.stream, x.d(), p[h.iter & 1].d(), S);
            if (!main || repl >= cfg.true_finish) break;
            // true-residual finish: r = b - A x of the complete x replaces the
            // recursive residual; the search direction p[K&1] and beta stay
            rc = replace_residual(x, b, r, pw ? 3 : 2, &h);
            if (rc) return rc;
            if (h.done) break;
            if (!pw) {
                rc = pre_apply(z, r, &S->done);
                if (rc) return rc;
                launch_cg_rz<double>(g, C.stream, C.rb, r.d(), z.d(), S);
                ++pre_calls;
            }
            cur = h.iter & 1;
            q = h.iter;  // queued iterations past the stop advanced q without running
        }
        count_pre(pre_calls);
        epilogue(main, x, b, r);
        return PPS_OK;
    }
    int solve_main() override { return run(true, C.fields[PPS_FIELD_X], C.fields[PPS_FIELD_B]); }
    int apply(Field out, Field in, const int*) override
    {
        calls++;
        int rc = run(false, out, in);
        Scal h;
        if (!rc) rc = scal_read(C, S, h);
        if (!rc) prec_iters += h.iter;
        return rc;
    }
};

// ------------------------------------------------------------------ BiCGSTAB

struct BiCGStage : KrylovStage {
    using KrylovStage::KrylovStage;
    // BiCGSTAB::operator() (BiCGSTAB.hpp:55-322) as the K1..K6 kernels of
    // BiCGstabAlpaka.hpp:210-450 with K4/K5 fused
    int run(bool main, Field x, Field b)
    {
        if (prec() != 64 || x.prec != 64) return C.fail(PPS_ERR_UNSUPPORTED, "BiCGSTAB stage %d: FP64 only", level);
        const std::string k = "bicg" + std::to_string(level);
        Field r, r0, p, AMp, Az, Mp, z;
        int rc = C.get_work(k + ".r", 64, r);
        if (!rc) rc = C.get_work(k + ".r0", 64, r0);
        if (!rc) rc = C.get_work(k + ".p", 64, p);
        if (!rc) rc = C.get_work(k + ".AMp", 64, AMp);
        if (!rc) rc = C.get_work(k + ".Az", 64, Az);
        double c = 1.0;
        const bool pw = pre_pointwise(&c);
        if (!rc && !pw) rc = C.get_work(k + ".Mp", 64, Mp);
        if (!rc && !pw) rc = C.get_work(k + ".z", 64, z);
        if (rc) return rc;
        for (Field f : {r, r0, p, AMp, Az}) C.zero(f);
        if (!pw) {
            C.zero(Mp);
            C.zero(z);
        }
        rc = scal_init(C, S, cfg.tol, cfg.max_iter, pw ? c : 1.0, main ? C.hist_d : nullptr, main ? C.hist_cap : 0);
        if (rc) return rc;
        int early = 0;
        rc = prologue(main, x, b, r, &early);
        if (rc) return rc;
        if (early) return PPS_OK;
        const Geo g = C.gsolver(64);
        launch_copy<double>(g, C.stream, r.d(), p.d());
        launch_copy<double>(g, C.stream, r.d(), r0.d());
        const int* gate = &S->done;
        auto body = [&]() -> int {
            if (pw) {
                C.ghost(p, false, nullptr, gate);
                launch_bicg_k1<double>(g, C.stream, C.rb, p.d(), c, AMp.d(), r0.d(), S);
            } else {
                int e = pre_apply(Mp, p, gate);
                if (e) return e;
                C.ghost(Mp, false, nullptr, gate);
                launch_bicg_k1<double>(g, C.stream, C.rb, Mp.d(), 1.0, AMp.d(), r0.d(), S);
            }
            launch_bicg_k2<double>(g, C.stream, r.d(), AMp.d(), S);
            if (pw) {
                C.ghost(r, false, nullptr, gate);
                launch_bicg_k3<double>(g, C.stream, C.rb, r.d(), c, Az.d(), r.d(), S);
                launch_bicg_k45<double>(g, C.stream, C.rb, x.d(), p.d(), c, r.d(), c, r.d(), Az.d(), r0.d(), S, main);
            } else {
                int e = pre_apply(z, r, gate);
                if (e) return e;
                C.ghost(z, false, nullptr, gate);
                launch_bicg_k3<double>(g, C.stream, C.rb, z.d(), 1.0, Az.d(), r.d(), S);
                launch_bicg_k45<double>(g, C.stream, C.rb, x.d(), Mp.d(), 1.0, z.d(), 1.0, r.d(), Az.d(), r0.d(), S, main);
            }
            launch_bicg_k6<double>(g, C.stream, p.d(), r.d(), AMp.d(), S);
            return PPS_OK;
        };
        for (int repl = 0;; ++repl) {
            rc = run_loop(C, S, cfg.max_iter, pre_host(), body);
            if (rc) return rc;
            if (!main || repl >= cfg.true_finish) break;
            // true-residual finish: restart from r = b - A x with r0 = p = r
            Scal h;
            rc = replace_residual(x, b, r, 4, &h);
            if (rc) return rc;
            if (h.done) break;
            launch_copy<double>(g, C.stream, r.d(), p.d());
            launch_copy<double>(g, C.stream, r.d(), r0.d());
        }
        {
            // two applications in every iteration that ran (BiCGSTAB.hpp:170,190)
            Scal h;
            rc = scal_read(C, S, h);
            if (rc) return rc;
            if (!pw) count_pre(2ll * h.iter);
        }
        // resetNeumanBCs<isMainLoop,true>(fieldX) (BiCGSTAB.hpp:301) before the true residual
        C.ghost(x, main, S, nullptr);
        epilogue(main, x, b, r);
        return PPS_OK;
    }
    int solve_main() override { return run(true, C.fields[PPS_FIELD_X], C.fields[PPS_FIELD_B]); }
    int apply(Field out, Field in, const int*) override
    {
        calls++;
        int rc = run(false, out, in);
        Scal h;
        if (!rc) rc = scal_read(C, S, h);
        if (!rc) prec_iters += h.iter;
        return rc;
    }
};

// ------------------------------------------------------------------------ FCG

struct FCGStage : KrylovStage {
    using KrylovStage::KrylovStage;
    // Flexible CG, truncation 1 (Notay 2000); prologue/epilogue as BaseCG.
    // T = double (main) or float (inner solve of PPS_IR, see IRStage).
    int run(bool main, Field x, Field b)
    {
        if (prec() != 64 || x.prec != 64) return C.fail(PPS_ERR_UNSUPPORTED, "FCG stage %d: FP32 vectors only inside PPS_IR", level);
        const std::string k = "fcg" + std::to_string(level);
        Field r, z, p0, p1, A0, A1;
        int rc = C.get_work(k + ".r", 64, r);
        if (!rc) rc = C.get_work(k + ".z", 64, z);
        if (!rc) rc = C.get_work(k + ".p0", 64, p0);
        if (!rc) rc = C.get_work(k + ".p1", 64, p1);
        if (!rc) rc = C.get_work(k + ".A0", 64, A0);
        if (!rc) rc = C.get_work(k + ".A1", 64, A1);
        if (rc) return rc;
        for (Field f : {r, z, p0, p1, A0, A1}) C.zero(f);
        const pps_stage& sc = stall_cfg(this);
        rc = scal_init(C, S, cfg.tol, cfg.max_iter, 1.0, main ? C.hist_d : nullptr, main ? C.hist_cap : 0,
                       sc.stall_window, sc.stall_ratio);
        if (rc) return rc;
        int early = 0;
        rc = prologue(main, x, b, r, &early);
        if (rc) return rc;
        if (early) return PPS_OK;
        const Geo g = C.gsolver(64);
        Field p[2] = {p0, p1}, Ap[2] = {A0, A1};
        // a preconditioner that forms the residual update itself (the learned
        // one): r alternates between two fields and the update pass drops out
        const bool fused_r = pre && pre->forms_update();
        Field rr[2] = {r, r};
        if (fused_r && (rc = C.get_work(k + ".r2", 64, rr[1]))) return rc;
        // and its directions go round a ring where the device holds one,
        // reaching x in one pass per nring iterations (iteration q's in slot
        // q mod nring)
        Field ring[kFcgRing];
        double* ringp[kFcgRing] = {};
        int nring = 0;
        if (fused_r) {
            rc = C.ring(k, 64, p0, p1, ring, &nring);
            if (rc) return rc;
            for (int j = 0; j < nring; ++j) ringp[j] = ring[j].d();
            if (nring) launch_fcg_ring_reset(C.stream, S);
        }
        int q = 0;
        int cur = 0;  // p[cur] = p_new, p[cur^1] = p_old
        auto body = [&]() -> int {
            if (fused_r) {
                // iteration k: r_k = r_{k-1} - alpha_{k-1} Ap_{k-1} into rr[cur ^ 1]
                int nb = 0;
                int e = pre->apply_update(z, rr[cur ^ 1], rr[cur], Ap[cur ^ 1], &S->v[S_ALPHA], &S->restart, &S->done,
                                          Ap[cur ^ 1], &nb);
                if (e) return e;
                if (nb > 0) launch_fcg_gamma_finish<double>(C.stream, C.rb, nb, S);
                launch_fcg_gamma<double>(g, C.stream, C.rb, z.d(), Ap[cur ^ 1].d(), S, nb > 0);
                if (nring) {
                    const int qs = q % nring, qo = (q + nring - 1) % nring;
                    launch_fcg_p_fused<double>(g, C.stream, C.rb, z.d(), ring[qo].d(), ring[qs].d(), Ap[cur].d(),
                                               rr[cur ^ 1].d(), x.d(), S, main, qs);
                    if (qs == nring - 1) launch_fcg_flush<double>(g, C.stream, x.d(), ringp, S);
                    ++q;
                } else {
                    launch_fcg_p_fused<double>(g, C.stream, C.rb, z.d(), p[cur ^ 1].d(), p[cur].d(), Ap[cur].d(),
                                               rr[cur ^ 1].d(), x.d(), S, main);
                }
                cur ^= 1;
                return PPS_OK;
            }
            // the learned preconditioner sums z . Ap_old in its output kernel
            int nb = 0;
            int e = pre_apply_dot(z, r, &S->done, Ap[cur ^ 1], &nb);
            if (e) return e;
            if (nb > 0) launch_fcg_gamma_finish<double>(C.stream, C.rb, nb, S);
            launch_fcg_gamma<double>(g, C.stream, C.rb, z.d(), Ap[cur ^ 1].d(), S, nb > 0);
            launch_fcg_p<double>(g, C.stream, C.rb, z.d(), p[cur ^ 1].d(), p[cur].d(), Ap[cur].d(), r.d(), x.d(), S);
            launch_fcg_update<double>(g, C.stream, C.rb, r.d(), Ap[cur].d(), S, main);
            cur ^= 1;
            return PPS_OK;
        };
        Scal h;
        for (int repl = 0;; ++repl) {
            rc = run_loop(C, S, cfg.max_iter, pre_host(), body);
            if (rc) return rc;
            rc = scal_read(C, S, h);
            if (rc) return rc;
            // iteration k wrote p[k & 1]; its x update waits for the next stencil pass
            // (or, with the ring, for its flush)
            if (nring) launch_fcg_flush<double>(g, C.stream, x.d(), ringp, S);
            else if (h.iter > 0) launch_fcg_final_x<double>(g, C.stream, x.d(), p[(h.iter - 1) & 1].d(), S);
            if (!main || repl >= cfg.true_finish) break;
            // true-residual finish: r = b - A x of the complete x, written where
            // iteration K reads its residual (rr[K & 1] when the preconditioner
            // forms the update: with alpha = 0 it forms r_K = r); the direction stays
            rc = replace_residual(x, b, fused_r ? rr[h.iter & 1] : r, 2, &h);
            if (rc) return rc;
            if (h.done) break;
            cur = h.iter & 1;
            q = h.iter;  // queued iterations past the stop advanced q without running
        }
        C.fallback_switches += h.fallback;
        count_pre(h.iter);  // one application per iteration that ran
        epilogue(main, x, b, r);
        return PPS_OK;
    }
    int solve_main() override { return run(true, C.fields[PPS_FIELD_X], C.fields[PPS_FIELD_B]); }
    int apply(Field out, Field in, const int*) override
    {
        calls++;
        int rc = run(false, out, in);
        Scal h;
        if (!rc) rc = scal_read(C, S, h);
        if (!rc) prec_iters += h.iter;
        return rc;
    }
};

// --------------------------------------------------------- learned preconditioner

// the fallback preconditioner (Jacobi) of an NN stage: runs only once the
// calling FCG stage has set `fallback`
template <typename T>
__global__ void k_fallback(T* out, const T* in, Layout L, Region R, T c, const int* done, const int* fallback)
{
    if (*done || !*fallback) return;
    const int n0 = R.hi[0] - R.lo[0], n1 = R.hi[1] - R.lo[1], n2 = R.hi[2] - R.lo[2];
    const long long n = (long long)n0 * n1 * n2;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % n0);
        const long long q = t / n0;
        const long long o = L.at(R.lo[0] + i, R.lo[1] + (int)(q % n1), R.lo[2] + (int)(q / n1));
        out[o] = c * in[o];
    }
}

// the same with the residual update folded in: in = r_prev - alpha ap_prev
// (r_prev while *skip) is formed, written and preconditioned
template <typename T>
__global__ void k_fallback_upd(T* out, T* in, const T* r_prev, const T* ap_prev, const double* alpha, const int* skip,
                               Layout L, Region R, T c, const int* done, const int* fallback)
{
    if (*done || !*fallback) return;
    const T al = (T)*alpha;
    const bool sk = *skip != 0;
    const int n0 = R.hi[0] - R.lo[0], n1 = R.hi[1] - R.lo[1], n2 = R.hi[2] - R.lo[2];
    const long long n = (long long)n0 * n1 * n2;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % n0);
        const long long q = t / n0;
        const long long o = L.at(R.lo[0] + i, R.lo[1] + (int)(q % n1), R.lo[2] + (int)(q / n1));
        const T r = sk ? r_prev[o] : r_prev[o] - al * ap_prev[o];
        in[o] = r;
        out[o] = c * r;
    }
}

// a preconditioner is called with gate = &caller->done of the Krylov stage it
// preconditions; that stage's scalar block holds ||r||^2 and the fallback flag
static Scal* caller_of(const int* gate)
{
    return reinterpret_cast<Scal*>(reinterpret_cast<char*>(const_cast<int*>(gate)) - offsetof(Scal, done));
}

// FP64 -> FP32 copy of a whole field allocation (FP32 and FP64 fields share one
// element layout)
__global__ void k_narrow(const double* __restrict__ a, float* __restrict__ b, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}
__global__ void k_widen(const float* __restrict__ a, double* __restrict__ b, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = (double)a[i];
}

struct NNStage : Stage {
    using Stage::Stage;
    long long iters_per_call() const override { return 1; }
    int apply(Field out, Field in, const int* gate) override
    {
        int nb;
        return apply_dot(out, in, gate, Field{}, &nb);
    }
    int apply_dot(Field out, Field in, const int* gate, Field dot, int* nblocks) override
    {
        return run(out, in, gate, dot, nullptr, nblocks);
    }
    bool forms_update() const override { return true; }
    int apply_update(Field out, Field in, Field r_prev, Field ap_prev, const double* alpha, const int* skip,
                     const int* gate, Field dot, int* nblocks) override
    {
        if (r_prev.prec != in.prec || ap_prev.prec != in.prec) return C.fail(PPS_ERR_ARG, "PPS_NN: update precision differs");
        const NNResUpdate u{r_prev.p, ap_prev.p, in.p, alpha, skip};
        return run(out, in, gate, dot, &u, nblocks);
    }
    // dot.p set: the output kernel also writes per-block partials of out . dot
    // to the context's reduction workspace (gated like the network itself).
    // u set: the network's lift (or the fallback) forms `in` from the residual
    // update first
    int run(Field out, Field in, const int* gate, Field dot, const NNResUpdate* u, int* nblocks)
    {
        *nblocks = 0;
        if (!C.nn) return C.fail(PPS_ERR_STATE, "PPS_NN stage without a network (pps_nn_attach)");
        if (!gate) return C.fail(PPS_ERR_STATE, "PPS_NN stage called outside a Krylov stage");
        if (out.prec != in.prec) return C.fail(PPS_ERR_ARG, "PPS_NN: input and output precision differ");
        calls++;
        Scal* caller = caller_of(gate);
        const Layout& L = C.lay(in.prec);
        NNFieldGeom g;
        g.off = L.at(C.solver.lo[0], C.solver.lo[1], C.solver.lo[2]);
        g.sy = L.px;
        g.sz = L.plane;
        g.n0 = C.solver.hi[0] - C.solver.lo[0];
        g.n1 = C.solver.hi[1] - C.solver.lo[1];
        g.n2 = C.solver.hi[2] - C.solver.lo[2];
        if (dot.p && dot.prec != out.prec) return C.fail(PPS_ERR_ARG, "PPS_NN: dot field precision differs");
        if (C.nn->apply_dot(C.stream, g, out.p, out.prec, in.p, in.prec, &caller->done, &caller->fallback,
                            &caller->v[S_RR], dot.p, C.rb.partials, C.rb.cap, nblocks, u))
            return C.fail(PPS_ERR_HIP, "learned preconditioner: %s", C.nn->error());
        const double* ds = C.grid.ds;
        const double jac = 1.0 / (-2.0 * (1.0 / (ds[0] * ds[0]) + 1.0 / (ds[1] * ds[1]) + 1.0 / (ds[2] * ds[2])));
        if (u) {
            if (in.prec == 32)
                hipLaunchKernelGGL(k_fallback_upd<float>, dim3(2048), dim3(256), 0, C.stream, out.f(), in.f(),
                                   (const float*)u->r_prev, (const float*)u->ap_prev, u->alpha, u->skip, L, C.solver,
                                   (float)jac, &caller->done, &caller->fallback);
            else
                hipLaunchKernelGGL(k_fallback_upd<double>, dim3(2048), dim3(256), 0, C.stream, out.d(), in.d(),
                                   (const double*)u->r_prev, (const double*)u->ap_prev, u->alpha, u->skip, L, C.solver,
                                   jac, &caller->done, &caller->fallback);
        } else if (in.prec == 32) {
            hipLaunchKernelGGL(k_fallback<float>, dim3(2048), dim3(256), 0, C.stream, out.f(), in.f(), L, C.solver,
                               (float)jac, &caller->done, &caller->fallback);
        } else {
            hipLaunchKernelGGL(k_fallback<double>, dim3(2048), dim3(256), 0, C.stream, out.d(), in.d(), L, C.solver, jac,
                               &caller->done, &caller->fallback);
        }
        return PPS_OK;
    }
};

// ------------------------------------------------- mixed-precision refinement

struct IRStage : KrylovStage {
    using KrylovStage::KrylovStage;
    long long inner_total = 0;
    int outer_steps = 0;
    // FP32 inner solve of A d = r32 (CG or FCG, non-main semantics of BaseCG:
    // x = 0, b normalised by its norm, relative tolerance), gated on SI->done
    int inner(Stage* in, Field d, Field r32)
    {
        Scal* SI = in->S;
        const std::string k = "ir" + std::to_string(level);
        Field p0, p1, A0, A1, z;
        int rc = C.get_work(k + ".p0", 32, p0);
        if (!rc) rc = C.get_work(k + ".p1", 32, p1);
        if (!rc) rc = C.get_work(k + ".A0", 32, A0);
        if (!rc) rc = C.get_work(k + ".A1", 32, A1);
        if (rc) return rc;
        Stage* M = in->pre;
        double c = 1.0;
        const bool pw = (M == nullptr) || M->pointwise(&c);
        const bool flexible = in->cfg.kind == PPS_FCG;
        if ((!pw || flexible) && (rc = C.get_work(k + ".z", 32, z))) return rc;
        // FCG with a preconditioner that forms the residual update itself (the
        // learned one): r alternates between r32 and a second field, written by
        // the preconditioner, and the update pass drops out of the loop
        const bool fused_r = flexible && M && M->forms_update();
        Field rr[2] = {r32, r32};
        if (fused_r && (rc = C.get_work(k + ".r32b", 32, rr[1]))) return rc;
        // SI was initialised by solve_main; the outer residual kernel has just
        // written SI[S_INNER_NORM], and ir_inner_init resets iter/done
        const Geo g = C.gsolver(32);
        Field p[2] = {p0, p1}, Ap[2] = {A0, A1};
        // with the learned preconditioner the directions go round a ring and
        // reach d in one pass per kFcgRing iterations (iteration q's in slot
        // q mod kFcgRing): the p-stencil pass neither reads nor writes d
        // (as the plain inner CG's, where the device holds the ring; slot 0 =
        // p1 and the last slot = p0, both zero, so CG's first iteration reads
        // zeros as its old direction)
        Field ring[kFcgRing];
        float* ringp[kFcgRing] = {};
        int nring = 0;
        if (fused_r || !flexible) {
            rc = C.ring(k, 32, p1, p0, ring, &nring);
            if (rc) return rc;
            for (int j = 0; j < nring; ++j) ringp[j] = ring[j].f();
        }
        launch_ir_inner_init(g, C.stream, C.rb, r32.f(), d.f(), p0.f(), SI, pw);
        if (nring) launch_fcg_ring_reset(C.stream, SI);
        int q = 0;
        // pin p1 and the old Ap to zero as well: iteration 0 multiplies them by beta = 0
        launch_fill<float>(g, C.stream, p1.f(), 0.0f);
        launch_fill<float>(g, C.stream, A0.f(), 0.0f);
        launch_fill<float>(g, C.stream, A1.f(), 0.0f);
        int cur = 0;
        const bool host_each = M && M->host_driven();
        auto applyM = [&](const int* gate) -> int {
            if (M) return M->apply(z, r32, gate);
            launch_pointwise<float>(C.gbox(32), C.stream, r32.f(), z.f(), 1.0, gate);
            return PPS_OK;
        };
        // the learned preconditioner sums z . Ap_old in its output kernel
        auto applyM_dot = [&](const int* gate, Field dot, int* nb) -> int {
            *nb = 0;
            if (M) return M->apply_dot(z, r32, gate, dot, nb);
            return applyM(gate);
        };
        if (!flexible && !pw) {
            rc = applyM(&SI->done);
            if (rc) return rc;
            launch_cg_rz<float>(g, C.stream, C.rb, r32.f(), z.f(), SI);
        }
        rc = run_loop(C, SI, in->cfg.max_iter, host_each, [&]() -> int {
            if (fused_r) {
                // iteration k: r_k = r_{k-1} - alpha_{k-1} Ap_{k-1} into rr[cur ^ 1]
                int nb = 0;
                int e = M->apply_update(z, rr[cur ^ 1], rr[cur], Ap[cur ^ 1], &SI->v[S_ALPHA], &SI->restart, &SI->done,
                                        Ap[cur ^ 1], &nb);
                if (e) return e;
                if (nb > 0) launch_fcg_gamma_finish<float>(C.stream, C.rb, nb, SI);
                launch_fcg_gamma<float>(g, C.stream, C.rb, z.f(), Ap[cur ^ 1].f(), SI, nb > 0);
                if (nring) {
                    const int qs = q % nring, qo = (q + nring - 1) % nring;
                    launch_fcg_p_fused<float>(g, C.stream, C.rb, z.f(), ring[qo].f(), ring[qs].f(), Ap[cur].f(),
                                              rr[cur ^ 1].f(), d.f(), SI, false, qs);
                    if (qs == nring - 1) launch_fcg_flush<float>(g, C.stream, d.f(), ringp, SI);
                    ++q;
                } else {
                    launch_fcg_p_fused<float>(g, C.stream, C.rb, z.f(), p[cur ^ 1].f(), p[cur].f(), Ap[cur].f(),
                                              rr[cur ^ 1].f(), d.f(), SI, false);
                }
                cur ^= 1;
            } else if (flexible) {
                int nb = 0;
                int e = applyM_dot(&SI->done, Ap[cur ^ 1], &nb);
                if (e) return e;
                if (nb > 0) launch_fcg_gamma_finish<float>(C.stream, C.rb, nb, SI);
                launch_fcg_gamma<float>(g, C.stream, C.rb, z.f(), Ap[cur ^ 1].f(), SI, nb > 0);
                launch_fcg_p<float>(g, C.stream, C.rb, z.f(), p[cur ^ 1].f(), p[cur].f(), Ap[cur].f(), r32.f(), d.f(), SI);
                launch_fcg_update<float>(g, C.stream, C.rb, r32.f(), Ap[cur].f(), SI, false);
                cur ^= 1;
            } else {
                if (nring) {
                    const int qs = q % nring, qo = (q + nring - 1) % nring;
                    launch_cg_A2<float>(g, C.stream, C.rb, pw ? r32.f() : z.f(), ring[qo].f(), ring[qs].f(), Ap[0].f(),
                                        d.f(), SI, pw, qs);
                    if (qs == nring - 1) launch_fcg_flush<float>(g, C.stream, d.f(), ringp, SI);
                    ++q;
                } else {
                    launch_cg_A2<float>(g, C.stream, C.rb, pw ? r32.f() : z.f(), p[cur].f(), p[cur ^ 1].f(), Ap[0].f(),
                                        d.f(), SI, pw);
                }
                cur ^= 1;
                launch_cg_B<float>(g, C.stream, C.rb, r32.f(), Ap[0].f(), SI, pw, false);
                if (!pw) {
                    int e = applyM(&SI->done);
                    if (e) return e;
                    launch_cg_rz<float>(g, C.stream, C.rb, r32.f(), z.f(), SI);
                }
            }
            return PPS_OK;
        });
        if (rc) return rc;
        Scal h;
        rc = scal_read(C, SI, h);
        if (rc) return rc;
        if (nring) launch_fcg_flush<float>(g, C.stream, d.f(), ringp, SI);  // the ring's last steps
        else if (!flexible && h.iter > 0) launch_cg_final_x<float>(g, C.stream, d.f(), p[h.iter & 1].f(), SI);
        else if (flexible && h.iter > 0) launch_fcg_final_x<float>(g, C.stream, d.f(), p[(h.iter - 1) & 1].f(), SI);
        inner_total += h.iter;
        if (flexible) C.fallback_switches += h.fallback;
        in->prec_iters += h.iter;
        in->calls++;
        // applications of M that ran: CG applies it once before its loop and in
        // every iteration but the stopping one, FCG once per iteration
        if (M && !M->host_driven()) {
            const long long calls = flexible ? h.iter : (pw ? 0 : 1 + std::max(0, h.iter - 1));
            M->prec_iters += calls * M->iters_per_call();
        }
        return PPS_OK;
    }
    int solve_main() override
    {
        if (!pre || (pre->cfg.kind != PPS_CG && pre->cfg.kind != PPS_FCG) || pre->prec() != 32)
            return C.fail(PPS_ERR_UNSUPPORTED, "PPS_IR needs an FP32 CG or FCG inner stage");
        Field x = C.fields[PPS_FIELD_X], b = C.fields[PPS_FIELD_B];
        Field r32, d;
        int rc = C.get_work("ir.r32", 32, r32);
        if (!rc) rc = C.get_work("ir.d", 32, d);
        if (rc) return rc;
        C.zero(r32);
        C.zero(d);
        inner_total = 0;
        outer_steps = 0;
        rc = scal_init(C, S, cfg.tol, cfg.max_iter, 1.0, C.hist_d, C.hist_cap);
        if (rc) return rc;
        Scal* SI = pre->S;
        double c = 1.0;
        const bool pw = (pre->pre == nullptr) || pre->pre->pointwise(&c);
        const pps_stage& sc = stall_cfg(pre);
        rc = scal_init(C, SI, pre->cfg.tol, pre->cfg.max_iter, pw ? c : 1.0, nullptr, 0, sc.stall_window,
                       sc.stall_ratio);
        if (rc) return rc;
        C.ghost(x, true, S, nullptr);
        normalize(C, S, true, x, b);
        scale_box(C, x, S, 0);
        scale_box(C, b, S, 0);
        C.ghost(x, true, S, nullptr);
        const Geo g64 = C.gsolver(64);
        const double itol = pre->cfg.tol, itol0 = cfg.ir_first_tol > 0.0 ? cfg.ir_first_tol : itol;
        launch_ir_residual(g64, C.stream, C.rb, x.d(), b.d(), r32.f(), S, SI, itol, itol0, cfg.ir_adapt);
        Scal h;
        rc = scal_read(C, S, h);
        if (rc) return rc;
        if (h.done) return PPS_OK;  // early exit as in BaseCG
        while (outer_steps < cfg.max_iter) {
            rc = inner(pre, d, r32);
            if (rc) return rc;
            launch_ir_update(g64, C.stream, x.d(), d.f(), SI);
            C.ghost(x, true, S, nullptr);
            scal_op(C, S, SC_OUTER_INC);
            launch_ir_residual(g64, C.stream, C.rb, x.d(), b.d(), r32.f(), S, SI, itol, itol0, cfg.ir_adapt);
            outer_steps++;
            rc = scal_read(C, S, h);
            if (rc) return rc;
            if (h.done) break;
        }
        scale_box(C, x, S, 1);
        scale_box(C, b, S, 1);
        return PPS_OK;
    }
    int apply(Field, Field, const int*) override { return C.fail(PPS_ERR_UNSUPPORTED, "PPS_IR is a main solver only"); }
};

// Chebyshev as the main solver (isMainLoop = true, chebyshevIteration.hpp:61-67,132-139)
__global__ void k_adjust_copy(double* dst, const double* src, long long n)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

int ChebStage::solve_main()
{
    int rc = buffers();
    if (rc) return rc;
    if (prec() != 64) return C.fail(PPS_ERR_UNSUPPORTED, "main Chebyshev runs in FP64");
    Field x = C.fields[PPS_FIELD_X], b = C.fields[PPS_FIELD_B];
    Field Bt, r;
    rc = C.get_work("chebmain.B", 64, Bt);
    if (!rc) rc = C.get_work("chebmain.r", 64, r);
    if (rc) return rc;
    rc = scal_init(C, S, cfg.tol, cfg.max_iter, 1.0, nullptr, 0);
    if (rc) return rc;
    const long long n = C.L64.total;
    hipLaunchKernelGGL(k_adjust_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, C.stream, Bt.d(), b.d(), n);
    adjust_b_field(C.gdata(64), C.stream, x.d(), Bt.d(), C.grid.bcs_type, C.lim_data, C.flux_d, C.fng, C.grid.ds);
    C.ghost(Bt, false, nullptr, nullptr);
    sweeps<double, double>(Bt.d(), x.d(), nullptr, Y, Z, W, V);
    C.ghost(x, true, S, nullptr);
    launch_residual(C.gsolver(64), C.stream, C.rb, x.d(), b.d(), r.d(), S, 1);
    scal_op(C, S, SC_SET, S_ERR, 0.0);
    return PPS_OK;
}

}  // namespace pps

// ====================================================================== C ABI

using namespace pps;

// main: stage 0 solves (pps_solve); otherwise stage 0 is applied once as a
// preconditioner (pps_precondition)
static int build_chain(Ctx& C, const pps_stage* chain, int n, std::vector<std::unique_ptr<Stage>>& st, bool main = true)
{
    if (!chain || n < 1 || n > 8) return C.fail(PPS_ERR_ARG, "solver chain needs 1..8 stages");
    for (int i = 0; i < n; ++i) {
        pps_stage s = chain[i];
        if (s.precision != 32) s.precision = 64;
        std::unique_ptr<Stage> p;
        switch (s.kind) {
        case PPS_NONE: p.reset(new NoneStage(C, s, i)); break;
        case PPS_JACOBI: p.reset(new JacobiStage(C, s, i)); break;
        case PPS_CG: p.reset(new CGStage(C, s, i)); break;
        case PPS_BICGSTAB: p.reset(new BiCGStage(C, s, i)); break;
        case PPS_CHEBYSHEV: p.reset(new ChebStage(C, s, i)); break;
        case PPS_FCG: p.reset(new FCGStage(C, s, i)); break;
        case PPS_IR: p.reset(new IRStage(C, s, i)); break;
        case PPS_NN: p.reset(new NNStage(C, s, i)); break;
        default: return C.fail(PPS_ERR_ARG, "unknown stage kind %d", s.kind);
        }
        if (i >= C.nscal) return C.fail(PPS_ERR_ARG, "too many stages");
        p->S = C.scal(i);
        if (s.max_iter < 0) return C.fail(PPS_ERR_ARG, "stage %d: max_iter < 0", i);
        st.push_back(std::move(p));
    }
    for (int i = 0; i + 1 < n; ++i) st[i]->pre = st[i + 1].get();
    const int k0 = chain[0].kind;
    if (main && (k0 == PPS_NONE || k0 == PPS_JACOBI || k0 == PPS_NN))
        return C.fail(PPS_ERR_ARG, "stage 0 must be a solver (CG, BiCGSTAB, Chebyshev, FCG, IR)");
    if (!main && (k0 == PPS_IR || k0 == PPS_NN))
        return C.fail(PPS_ERR_ARG, "stage 0 of a preconditioner chain must be None, Jacobi, CG, BiCGSTAB, Chebyshev "
                                   "or FCG (the learned preconditioner: pps_nn_apply_field)");
    return PPS_OK;
}

static Ctx* as_ctx(void* p) { return static_cast<Ctx*>(p); }

extern "C" {

int pps_abi_version(void) { return PPS_ABI_VERSION; }

const char* pps_last_error(void* ctx)
{
    if (!ctx) return g_create_error.c_str();
    return as_ctx(ctx)->err.c_str();
}

int pps_create(const pps_grid* grid, void** out)
{
    if (!out || !grid) {
        g_create_error = "null argument";
        return PPS_ERR_ARG;
    }
    *out = nullptr;
    if (grid->dim != 3) {
        g_create_error = "DIM must be 3 in this build";
        return PPS_ERR_UNSUPPORTED;
    }
    if (grid->nranks[0] * grid->nranks[1] * grid->nranks[2] != 1 || grid->my_rank != 0) {
        g_create_error = "one GPU: the rank grid must be 1x1x1";
        return PPS_ERR_UNSUPPORTED;
    }
    for (int d = 0; d < 3; ++d) {
        if (grid->guards[d] != 1) {
            g_create_error = "guards must be {1,1,1} for the 7-point stencil";
            return PPS_ERR_ARG;
        }
        if (grid->npglobal[d] < 3 || !(grid->ds[d] > 0)) {
            g_create_error = "npglobal must be >= 3 and ds > 0";
            return PPS_ERR_ARG;
        }
    }
    for (int f = 0; f < 6; ++f)
        if (grid->bcs_type[f] != 0 && grid->bcs_type[f] != 1) {
            g_create_error = "bcsType must be 0 (Dirichlet) or 1 (Neumann)";
            return PPS_ERR_ARG;
        }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_create_error = "no HIP device";
        return PPS_ERR_NOGPU;
    }
    std::unique_ptr<Ctx> C(new Ctx());
    C->grid = *grid;
    for (int d = 0; d < 3; ++d) {
        C->np[d] = grid->npglobal[d];
        C->ng[d] = grid->npglobal[d] + 2;
        C->coef[d] = 1.0 / (grid->ds[d] * grid->ds[d]);
        C->lim_data[2 * d] = 1;
        C->lim_data[2 * d + 1] = grid->npglobal[d] + 1;
    }
    // blockGrid.hpp:208-222 and :301-340 for one rank
    std::memcpy(C->lim_solver, C->lim_data, sizeof C->lim_solver);
    for (int d = 0; d < 3; ++d) {
        if (grid->bcs_type[2 * d] == 0) C->lim_solver[2 * d] += 1;
        if (grid->bcs_type[2 * d + 1] == 0) C->lim_solver[2 * d + 1] -= 1;
    }
    const double PI = 3.141592653589793;
    double emin = 0, emax = 0, gmin = 0, gmax = 0;
    for (int d = 0; d < 3; ++d) {
        const int cnt = C->lim_solver[2 * d + 1] - C->lim_solver[2 * d];
        const double s1 = std::sin(1 * PI / 2 / (cnt + 1)), s2 = std::sin(cnt * PI / 2 / (cnt + 1));
        const double dd = grid->ds[d] * grid->ds[d];
        emin += 4 * s1 * s1 / dd;
        emax += 4 * s2 * s2 / dd;
        int ni = grid->npglobal[d];
        if (grid->bcs_type[2 * d] == 0) ni -= 1;
        if (grid->bcs_type[2 * d + 1] == 0) ni -= 1;
        const double g1 = std::sin(1 * PI / 2 / (ni + 1)), g2 = std::sin(ni * PI / 2 / (ni + 1));
        gmin += 4 * g1 * g1 / dd;
        gmax += 4 * g2 * g2 / dd;
    }
    C->eig_local[0] = emin;
    C->eig_local[1] = emax;
    C->eig_global[0] = gmin;
    C->eig_global[1] = gmax;
    for (int d = 0; d < 3; ++d) {
        C->solver.lo[d] = C->lim_solver[2 * d];
        C->solver.hi[d] = C->lim_solver[2 * d + 1];
        C->data.lo[d] = C->lim_data[2 * d];
        C->data.hi[d] = C->lim_data[2 * d + 1];
        C->box.lo[d] = 0;
        C->box.hi[d] = C->ng[d];
    }
    for (int prec : {64, 32}) {
        Layout& L = prec == 32 ? C->L32 : C->L64;
        // One element layout for every precision: FP32 and FP64 fields share
        // offsets (mixed-precision kernels index both with one offset).  The row
        // pitch is the guarded width rounded up to 32 elements and the first
        // solver column sits on a 32-element boundary, so every row's interior
        // starts 128-B (FP32) / 256-B (FP64) aligned; a row's trailing guard
        // point may spill into the next row's leading pad columns, which no row
        // uses (rows never overlap because px >= ng[0]).
        const int A = 32;
        for (int d = 0; d < 3; ++d) L.ng[d] = C->ng[d];
        L.ox = A - C->lim_solver[0];
        L.px = ((C->ng[0] + A - 1) / A) * A;
        L.plane = (long long)L.px * C->ng[1];
        L.total = L.plane * C->ng[2] + L.ox + C->ng[0] + A;
    }
    Ctx& X = *C;
    if (X.hip(hipStreamCreateWithFlags(&X.stream, hipStreamNonBlocking), "stream")) {
        g_create_error = X.err;
        return PPS_ERR_HIP;
    }
    int rc = PPS_OK;
    rc = rc ? rc : X.hip(hipEventCreateWithFlags(&X.ev_batch[0], hipEventDisableTiming), "event");
    rc = rc ? rc : X.hip(hipEventCreateWithFlags(&X.ev_batch[1], hipEventDisableTiming), "event");
    rc = rc ? rc : X.hip(hipEventCreate(&X.ev_t0), "event");
    rc = rc ? rc : X.hip(hipEventCreate(&X.ev_t1), "event");
    // reduction workspace: partials for the largest launch (a vector kernel
    // over the whole box, or a stencil at the smallest z-chunk allowed), and
    // the span sums of the two-level epilogue
    const Geo gbig = make_geo(X.L64, X.box, X.coef, kMinKz, 0);
    const long long nvec = (long long)((X.ng[0] + kThreads - 1) / kThreads) * X.ng[1] * X.ng[2];
    X.rb.cap = (int)std::max<long long>(gbig.nblocks, nvec) + 8;
    X.rb.cap2 = X.rb.cap / kSpan + 8;
    rc = rc ? rc : X.hip(hipMalloc(&X.rb.partials2, sizeof(double) * kMaxRedQ * X.rb.cap2), "span partials");
    // a solve streams 4-8 fields per iteration: once a field passes 32 MiB
    // that working set exceeds the 256 MiB last-level cache, and the outputs
    // are stored non-temporally
    X.nt_auto = (X.L64.total * 8 > (32ll << 20)) ? 1 : 0;
    X.nt = X.nt_auto;
    rc = rc ? rc : X.hip(hipMalloc(&X.rb.partials, sizeof(double) * kMaxRedQ * X.rb.cap), "partials");
    rc = rc ? rc : X.hip(hipMalloc(&X.rb.counter, 256), "counter");
    rc = rc ? rc : X.hip(hipMemsetAsync(X.rb.counter, 0, 256, X.stream), "counter");
    rc = rc ? rc : X.hip(hipMalloc(&X.scal_d, sizeof(Scal) * X.nscal), "scalars");
    rc = rc ? rc : X.hip(hipMemsetAsync(X.scal_d, 0, sizeof(Scal) * X.nscal, X.stream), "scalars");
    rc = rc ? rc : X.hip(hipHostMalloc(&X.flag_h, 64, hipHostMallocDefault), "pinned flags");
    rc = rc ? rc : X.hip(hipHostMalloc(&X.scal_h, sizeof(Scal), hipHostMallocDefault), "pinned scalars");
    X.hist_cap = 1 << 20;
    rc = rc ? rc : X.hip(hipMalloc(&X.hist_d, sizeof(double) * X.hist_cap), "history");
    rc = rc ? rc : X.alloc_field(X.fields[PPS_FIELD_X], 64);
    rc = rc ? rc : X.alloc_field(X.fields[PPS_FIELD_B], 64);
    // Neumann faces
    X.faces.n = 0;
    for (int f = 0; f < 6; ++f) {
        const int d = f / 2;
        const int d1 = d == 0 ? 1 : 0, d2 = d == 2 ? 1 : 2;
        X.fng[f][0] = X.ng[d1];
        X.fng[f][1] = X.ng[d2];
        if (grid->bcs_type[f] == 1 && !rc) {
            const size_t bytes = sizeof(double) * X.ng[d1] * X.ng[d2];
            rc = X.hip(hipMalloc(&X.flux_d[f], bytes), "flux");
            if (!rc) rc = X.hip(hipMemsetAsync(X.flux_d[f], 0, bytes, X.stream), "flux");
            const int q = X.faces.n++;
            X.faces.face[q] = f;
            X.faces.flux[q] = X.flux_d[f];
        }
    }
    for (int f = 0; f < 6; ++f) {
        X.faces.lim_data[f] = X.lim_data[f];
        X.faces.fng[f][0] = X.fng[f][0];
        X.faces.fng[f][1] = X.fng[f][1];
    }
    for (int d = 0; d < 3; ++d) X.faces.ds[d] = grid->ds[d];
    if (!rc) rc = X.hip(hipStreamSynchronize(X.stream), "sync");
    if (rc) {
        g_create_error = X.err;
        pps_destroy(C.release());
        return rc;
    }
    *out = C.release();
    return PPS_OK;
}

void pps_destroy(void* ctx)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return;
    if (C->stream) (void)hipStreamSynchronize(C->stream);
    C->nn.reset();
    for (auto& kv : C->work) (void)hipFree(kv.second.p);
    for (auto& f : C->fields)
        if (f.p) (void)hipFree(f.p);
    for (auto* p : C->flux_d)
        if (p) (void)hipFree(p);
    if (C->rb.partials) (void)hipFree(C->rb.partials);
    if (C->rb.partials2) (void)hipFree(C->rb.partials2);
    if (C->rb.counter) (void)hipFree(C->rb.counter);
    if (C->scal_d) (void)hipFree(C->scal_d);
    if (C->hist_d) (void)hipFree(C->hist_d);
    if (C->flag_h) (void)hipHostFree(C->flag_h);
    if (C->scal_h) (void)hipHostFree(C->scal_h);
    for (auto e : C->ev_batch)
        if (e) (void)hipEventDestroy(e);
    if (C->ev_t0) (void)hipEventDestroy(C->ev_t0);
    if (C->ev_t1) (void)hipEventDestroy(C->ev_t1);
    if (C->stream) (void)hipStreamDestroy(C->stream);
    delete C;
}

int pps_get_extent(void* ctx, int ng[3])
{
    Ctx* C = as_ctx(ctx);
    if (!C || !ng) return PPS_ERR_ARG;
    for (int d = 0; d < 3; ++d) ng[d] = C->ng[d];
    return PPS_OK;
}

static int field_of(Ctx& C, int id, Field& f)
{
    if (id < 0 || id >= PPS_NUM_FIELDS) return C.fail(PPS_ERR_ARG, "bad field id %d", id);
    if (!C.fields[id].p) {
        int rc = C.alloc_field(C.fields[id], 64);
        if (rc) return rc;
    }
    f = C.fields[id];
    return PPS_OK;
}

int pps_upload(void* ctx, int field, const double* host)
{
    Ctx* C = as_ctx(ctx);
    if (!C || !host) return PPS_ERR_ARG;
    Field f;
    int rc = field_of(*C, field, f);
    if (rc) return rc;
    hipMemcpy3DParms p{};
    p.srcPtr = make_hipPitchedPtr(const_cast<double*>(host), sizeof(double) * C->ng[0], C->ng[0], C->ng[1]);
    p.dstPtr = make_hipPitchedPtr(f.d() + C->L64.ox, sizeof(double) * C->L64.px, C->ng[0], C->ng[1]);
    p.extent = make_hipExtent(sizeof(double) * C->ng[0], C->ng[1], C->ng[2]);
    p.kind = hipMemcpyHostToDevice;
    rc = C->hip(hipMemcpy3DAsync(&p, C->stream), "upload");
    if (rc) return rc;
    return C->hip(hipStreamSynchronize(C->stream), "sync");
}

int pps_download(void* ctx, int field, double* host)
{
    Ctx* C = as_ctx(ctx);
    if (!C || !host) return PPS_ERR_ARG;
    Field f;
    int rc = field_of(*C, field, f);
    if (rc) return rc;
    hipMemcpy3DParms p{};
    p.srcPtr = make_hipPitchedPtr(f.d() + C->L64.ox, sizeof(double) * C->L64.px, C->ng[0], C->ng[1]);
    p.dstPtr = make_hipPitchedPtr(host, sizeof(double) * C->ng[0], C->ng[0], C->ng[1]);
    p.extent = make_hipExtent(sizeof(double) * C->ng[0], C->ng[1], C->ng[2]);
    p.kind = hipMemcpyDeviceToHost;
    rc = C->hip(hipMemcpy3DAsync(&p, C->stream), "download");
    if (rc) return rc;
    return C->hip(hipStreamSynchronize(C->stream), "sync");
}

int pps_fill(void* ctx, int field, double value)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    Field f;
    int rc = field_of(*C, field, f);
    if (rc) return rc;
    launch_fill<double>(C->gbox(64), C->stream, f.d(), value);
    return C->hip(hipStreamSynchronize(C->stream), "fill");
}

int pps_copy_field(void* ctx, int dst, int src)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    Field fd, fs;
    int rc = field_of(*C, dst, fd);
    if (!rc) rc = field_of(*C, src, fs);
    if (rc) return rc;
    rc = C->hip(hipMemcpyAsync(fd.p, fs.p, sizeof(double) * C->L64.total, hipMemcpyDeviceToDevice, C->stream), "copy");
    if (rc) return rc;
    return C->hip(hipStreamSynchronize(C->stream), "sync");
}

int pps_set_neumann_flux(void* ctx, int face, const double* host)
{
    Ctx* C = as_ctx(ctx);
    if (!C || !host || face < 0 || face > 5) return PPS_ERR_ARG;
    if (!C->flux_d[face]) return C->fail(PPS_ERR_ARG, "face %d is not a Neumann face", face);
    const size_t bytes = sizeof(double) * C->fng[face][0] * C->fng[face][1];
    int rc = C->hip(hipMemcpyAsync(C->flux_d[face], host, bytes, hipMemcpyHostToDevice, C->stream), "flux upload");
    if (rc) return rc;
    return C->hip(hipStreamSynchronize(C->stream), "sync");
}

int pps_solve(void* ctx, const pps_stage* chain, int nstages, pps_result* res)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    std::vector<std::unique_ptr<Stage>> st;
    int rc = build_chain(*C, chain, nstages, st);
    if (rc) return rc;
    C->err.clear();
    C->fallback_switches = 0;
    rc = C->hip(hipEventRecord(C->ev_t0, C->stream), "event");
    if (rc) return rc;
    rc = st[0]->solve_main();
    if (rc) return rc;
    rc = C->hip(hipEventRecord(C->ev_t1, C->stream), "event");
    if (!rc) rc = C->hip(hipEventSynchronize(C->ev_t1), "event sync");
    if (!rc) rc = C->hip(hipGetLastError(), "kernel");
    if (rc) return rc;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, C->ev_t0, C->ev_t1);
    Scal h;
    rc = scal_read(*C, st[0]->S, h);
    if (rc) return rc;
    if (res) {
        std::memset(res, 0, sizeof *res);
        const int kind = chain[0].kind;
        res->seconds = ms * 1e-3;
        res->norm_b = h.v[S_NORM];
        if (kind == PPS_CHEBYSHEV) {
            res->iterations = chain[0].max_iter;
            res->error_from_iteration = h.v[S_ERR_OP];
            res->error_compute_operator = h.v[S_ERR_OP];
            res->history_len = 0;
        } else if (kind == PPS_IR) {
            auto* ir = static_cast<IRStage*>(st[0].get());
            res->iterations = (int)ir->inner_total;
            res->outer_iterations = ir->outer_steps;
            res->error_from_iteration = h.v[S_ERR];
            res->error_compute_operator = h.v[S_ERR_OP];
            res->history_len = std::min(h.outer + 1, C->hist_cap);
        } else {
            res->iterations = h.iter;
            res->error_from_iteration = h.done == 2 ? h.v[S_ERR0] : h.v[S_ERR];
            res->error_compute_operator = h.v[S_ERR_OP];
            res->history_len = std::min(h.iter + 1, C->hist_cap);
        }
        long long pi = 0;
        for (size_t i = 1; i < st.size(); ++i) pi += st[i]->prec_iters;
        res->precond_iterations = pi;
        res->fallback_switches = C->fallback_switches;
        C->hist_len = res->history_len;
    }
    return PPS_OK;
}

int pps_precondition(void* ctx, const pps_stage* chain, int nstages, int out, int in, long long* iterations)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    std::vector<std::unique_ptr<Stage>> st;
    int rc = build_chain(*C, chain, nstages, st, false);
    if (rc) return rc;
    C->err.clear();
    Field fo, fi;
    rc = field_of(*C, out, fo);
    if (!rc) rc = field_of(*C, in, fi);
    if (rc) return rc;
    if (out == in) return C->fail(PPS_ERR_ARG, "pps_precondition: out and in must differ");
    rc = st[0]->apply(fo, fi, nullptr);
    if (rc) return rc;
    if (!st[0]->host_driven()) st[0]->prec_iters += st[0]->iters_per_call();
    rc = C->hip(hipStreamSynchronize(C->stream), "sync");
    if (!rc) rc = C->hip(hipGetLastError(), "kernel");
    if (rc) return rc;
    if (iterations) {
        long long pi = 0;
        for (auto& s : st) pi += s->prec_iters;
        *iterations = pi;
    }
    return PPS_OK;
}

int pps_get_history(void* ctx, double* out, int cap)
{
    Ctx* C = as_ctx(ctx);
    if (!C || !out || cap < 0) return PPS_ERR_ARG;
    const int n = std::min(cap, C->hist_len);
    if (n == 0) return 0;
    int rc = C->hip(hipMemcpy(out, C->hist_d, sizeof(double) * n, hipMemcpyDeviceToHost), "history");
    return rc ? rc : n;
}

int pps_synchronize(void* ctx)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    return C->hip(hipStreamSynchronize(C->stream), "sync");
}

int pps_apply_A(void* ctx, int y, int x)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    Field fy, fx;
    int rc = field_of(*C, y, fy);
    if (!rc) rc = field_of(*C, x, fx);
    if (rc) return rc;
    launch_apply_A<double>(C->gsolver(64), C->stream, fx.d(), fy.d());
    return C->hip(hipStreamSynchronize(C->stream), "apply_A");
}

int pps_ghost_fill(void* ctx, int field, int with_flux, double norm)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    Field f;
    int rc = field_of(*C, field, f);
    if (rc) return rc;
    Scal* S = C->scal(C->nscal - 1);
    rc = scal_init(*C, S, 0.0, 0, 1.0, nullptr, 0);
    if (rc) return rc;
    scal_op(*C, S, SC_SET, S_NORM, norm);
    C->ghost(f, with_flux != 0, S, nullptr);
    return C->hip(hipStreamSynchronize(C->stream), "ghost");
}

int pps_residual(void* ctx, int r, int x, int b, double* norm)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    Field fr, fx, fb;
    int rc = field_of(*C, r, fr);
    if (!rc) rc = field_of(*C, x, fx);
    if (!rc) rc = field_of(*C, b, fb);
    if (rc) return rc;
    Scal* S = C->scal(C->nscal - 1);
    rc = scal_init(*C, S, 0.0, 0, 1.0, nullptr, 0);
    if (rc) return rc;
    launch_residual(C->gsolver(64), C->stream, C->rb, fx.d(), fb.d(), fr.d(), S, 1);
    Scal h;
    rc = scal_read(*C, S, h);
    if (rc) return rc;
    if (norm) *norm = h.v[S_ERR_OP];
    return PPS_OK;
}

}  // extern "C"

// ------------------------------------------- allocation-free residual variant
//
// pps_residual_noalloc keeps its reduction partials and its result slot in
// this file-scope device array instead of context memory, so it calls no
// allocator.  Each GPU holds its own copy until the process ends; every call
// on one GPU shares it (the last call's contents stay), so calls that overlap
// on one GPU race.  Host code never passes the array's name to a kernel or a
// copy: that name is a host-side stand-in, not a device address.  It uses the
// address hipGetSymbolAddress returns for the current device, cached per
// device at that device's first call.
// 1 MiB of partials (one per block of one launch) and a 256-B result slot
constexpr int kScratchPartials = 1 << 17;
static __device__ __align__(256) double scratch[kScratchPartials + 32];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// the current device's address of `scratch` (0 on success)
int scratch_for_current_device(double** out, std::string& err)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        err = std::string("hipGetDevice: ") + hipGetErrorString(e);
        return PPS_ERR_NOGPU;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        err = "device index beyond the scratch table";
        return PPS_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return PPS_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return PPS_OK;
}
}  // namespace

extern "C" {

int pps_residual_noalloc(const pps_grid* grid, const double* x, const double* b, double* r, double* norm_dev,
                         double* norm_host, void* stream)
{
    g_create_error.clear();
    if (!grid || !x || !b || !r) {
        g_create_error = "pps_residual_noalloc: null argument";
        return PPS_ERR_ARG;
    }
    if (grid->dim != 3 || grid->nranks[0] * grid->nranks[1] * grid->nranks[2] != 1 || grid->my_rank != 0) {
        g_create_error = "pps_residual_noalloc: DIM 3 on a 1x1x1 rank grid only";
        return PPS_ERR_UNSUPPORTED;
    }
    Layout L{};
    Region R{};
    double coef[3];
    for (int d = 0; d < 3; ++d) {
        if (grid->guards[d] != 1 || grid->npglobal[d] < 3 || !(grid->ds[d] > 0)) {
            g_create_error = "pps_residual_noalloc: guards must be 1, npglobal >= 3 and ds > 0";
            return PPS_ERR_ARG;
        }
        for (int s = 0; s < 2; ++s)
            if (grid->bcs_type[2 * d + s] != 0 && grid->bcs_type[2 * d + s] != 1) {
                g_create_error = "pps_residual_noalloc: bcsType must be 0 or 1";
                return PPS_ERR_ARG;
            }
        L.ng[d] = grid->npglobal[d] + 2;
        coef[d] = 1.0 / (grid->ds[d] * grid->ds[d]);
        // BlockGrid's solver limits for one rank (blockGrid.hpp:208-222)
        R.lo[d] = 1 + (grid->bcs_type[2 * d] == 0);
        R.hi[d] = grid->npglobal[d] + 1 - (grid->bcs_type[2 * d + 1] == 0);
    }
    // the reference's dense guarded host layout, x fastest (iterativeSolverBase.hpp:34-35)
    L.px = L.ng[0];
    L.ox = 0;
    L.plane = (long long)L.ng[0] * L.ng[1];
    L.total = L.plane * L.ng[2];
    // z-chunk: 16 planes, longer when a launch would need more blocks than
    // the scratch holds partials
    const long long tiles = (long long)((R.hi[0] - R.lo[0] + kTX - 1) / kTX) * ((R.hi[1] - R.lo[1] + kTH - 1) / kTH);
    if (tiles > kScratchPartials) {
        g_create_error = "pps_residual_noalloc: an x-y plane needs more than 131072 tiles of " + std::to_string(kTX) +
                         "x" + std::to_string(kTH) + " points (the scratch holds one partial per block)";
        return PPS_ERR_UNSUPPORTED;
    }
    const int nz = R.hi[2] - R.lo[2];
    int kz = 16;
    while (tiles * ((nz + kz - 1) / kz) > kScratchPartials) kz *= 2;
    double* sp = nullptr;
    int rc = scratch_for_current_device(&sp, g_create_error);
    if (rc) return rc;
    RedBuf rb{sp, nullptr, kScratchPartials};
    double* out = norm_dev ? norm_dev : sp + kScratchPartials;  // else the result slot past the partials
    const Geo g = make_geo(L, R, coef, kz, (L.total * 8 > (32ll << 20)) ? 1 : 0);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (launch_residual_norm(g, s, rb, x, b, r, out) != 0) {
        g_create_error = "pps_residual_noalloc: launch exceeds the scratch partials";
        return PPS_ERR_UNSUPPORTED;
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && norm_host) {
        e = hipMemcpyAsync(norm_host, out, sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (e != hipSuccess) {
        g_create_error = std::string("pps_residual_noalloc: ") + hipGetErrorString(e);
        return PPS_ERR_HIP;
    }
    return PPS_OK;
}

int pps_dot(void* ctx, int a, int b, double* out)
{
    Ctx* C = as_ctx(ctx);
    if (!C || !out) return PPS_ERR_ARG;
    Field fa, fb;
    int rc = field_of(*C, a, fa);
    if (!rc) rc = field_of(*C, b, fb);
    if (rc) return rc;
    Scal* S = C->scal(C->nscal - 1);
    launch_dot<double>(C->gsolver(64), C->stream, C->rb, fa.d(), fb.d(), &S->v[S_AUX0], nullptr);
    Scal h;
    rc = scal_read(*C, S, h);
    if (rc) return rc;
    *out = h.v[S_AUX0];
    return PPS_OK;
}

int pps_norm_b_adjusted(void* ctx, double* out)
{
    Ctx* C = as_ctx(ctx);
    if (!C || !out) return PPS_ERR_ARG;
    Scal* S = C->scal(C->nscal - 1);
    launch_bnorm(C->gsolver(64), C->stream, C->rb, C->fields[PPS_FIELD_X].d(), C->fields[PPS_FIELD_B].d(),
                 C->grid.bcs_type, C->lim_data, C->lim_solver, C->flux_d, C->fng, C->grid.ds, S);
    Scal h;
    int rc = scal_read(*C, S, h);
    if (rc) return rc;
    *out = h.v[S_NORM];
    return PPS_OK;
}

int pps_set_tiling(void* ctx, int kz_stencil, int nontemporal)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    if (kz_stencil > 0 && (kz_stencil < kMinKz || kz_stencil > 4096))
        return C->fail(PPS_ERR_ARG, "kz must be in [%d, 4096]", kMinKz);
    if (nontemporal < -1 || nontemporal > 1) return C->fail(PPS_ERR_ARG, "nontemporal must be -1, 0 or 1");
    if (kz_stencil > 0) C->kz_st = kz_stencil;
    if (kz_stencil < 0) C->kz_st = 0;
    C->nt_mode = nontemporal;
    C->nt = nontemporal == 0 ? C->nt_auto : (nontemporal > 0 ? 1 : 0);
    return PPS_OK;
}

int pps_set_deferred_x(void* ctx, int on)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    if (on < 0 || on > 1) return C->fail(PPS_ERR_ARG, "on must be 0 or 1");
    C->deferred_x = on;
    return PPS_OK;
}

int pps_set_cheb_pairs(void* ctx, int pairs, int kz)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    if (pairs < 0 || pairs > 1) return C->fail(PPS_ERR_ARG, "pairs must be 0 or 1");
    if (kz > 4096) return C->fail(PPS_ERR_ARG, "kz must be at most 4096");
    C->cheb_pairs = pairs;
    if (kz > 0) C->cheb_kz = kz;
    if (kz < 0) C->cheb_kz = 0;
    return PPS_OK;
}

int pps_time_kernel(void* ctx, int kernel, int reps, double* ms, double* bytes)
{
    Ctx* C = as_ctx(ctx);
    if (!C || reps < 1 || !ms || !bytes) return PPS_ERR_ARG;
    const int P = (kernel == PPS_K_CG_STENCIL_F32 || kernel == PPS_K_CG_UPDATE_F32 || kernel == PPS_K_FCG_STENCIL_F32)
                      ? 32 : 64;
    const double w = P == 32 ? 4.0 : 8.0;
    const double n = (double)C->solver.count();
    Field a, b, c, d, e, f;
    int rc = C->get_work("time.a", P, a);
    if (!rc) rc = C->get_work("time.b", P, b);
    if (!rc) rc = C->get_work("time.c", P, c);
    if (!rc) rc = C->get_work("time.d", P, d);
    if (!rc) rc = C->get_work("time.e", P, e);
    if (!rc) rc = C->get_work("time.f", P, f);
    if (rc) return rc;
    Scal* S = C->scal(C->nscal - 2);
    rc = scal_init(*C, S, -1.0, 1 << 30, 1.0, nullptr, 0);
    if (rc) return rc;
    scal_op(*C, S, SC_SET, S_RZ, 1.0);
    scal_op(*C, S, SC_SET, S_BETA, 0.5);
    scal_op(*C, S, SC_SET, S_ALPHA, 1e-3);
    if (P == 64) {
        launch_fill<double>(C->gsolver(64), C->stream, a.d(), 1.0);
        launch_fill<double>(C->gsolver(64), C->stream, b.d(), 0.5);
        launch_fill<double>(C->gsolver(64), C->stream, c.d(), 0.25);
        launch_fill<double>(C->gsolver(64), C->stream, f.d(), 0.125);
    } else {
        launch_fill<float>(C->gsolver(32), C->stream, a.f(), 1.0f);
        launch_fill<float>(C->gsolver(32), C->stream, b.f(), 0.5f);
        launch_fill<float>(C->gsolver(32), C->stream, c.f(), 0.25f);
        launch_fill<float>(C->gsolver(32), C->stream, f.f(), 0.125f);
    }
    const Geo g = C->gsolver(P);
    auto launch = [&](int it) {
        switch (kernel) {
        case PPS_K_CG_STENCIL:
            launch_cg_A2<double>(g, C->stream, C->rb, a.d(), (it & 1) ? c.d() : b.d(), (it & 1) ? b.d() : c.d(), d.d(),
                                 e.d(), S, true, 0);
            break;
        case PPS_K_CG_STENCIL_F32:
            launch_cg_A2<float>(g, C->stream, C->rb, a.f(), (it & 1) ? c.f() : b.f(), (it & 1) ? b.f() : c.f(), d.f(),
                                e.f(), S, true, 0);
            break;
        case PPS_K_CG_UPDATE: launch_cg_B<double>(g, C->stream, C->rb, a.d(), c.d(), S, true, false); break;
        case PPS_K_CG_UPDATE_F32: launch_cg_B<float>(g, C->stream, C->rb, a.f(), c.f(), S, true, false); break;
        case PPS_K_RESIDUAL: launch_residual(g, C->stream, C->rb, a.d(), b.d(), c.d(), S, 1); break;
        case PPS_K_COPY: launch_copy<double>(g, C->stream, a.d(), e.d()); break;
        case PPS_K_CHEB_SWEEP: launch_cheb_sweep<double, double>(g, C->stream, a.d(), b.d(), c.d(), e.d(), 0.1, 0.2, 0.3, 1.0, nullptr); break;
        case PPS_K_FCG_STENCIL:
            launch_fcg_p_fused<double>(g, C->stream, C->rb, a.d(), (it & 1) ? c.d() : b.d(), (it & 1) ? b.d() : c.d(),
                                       d.d(), f.d(), e.d(), S, false, 0);
            break;
        case PPS_K_FCG_STENCIL_F32:
            launch_fcg_p_fused<float>(g, C->stream, C->rb, a.f(), (it & 1) ? c.f() : b.f(), (it & 1) ? b.f() : c.f(),
                                      d.f(), f.f(), e.f(), S, false, 0);
            break;
        case PPS_K_COPY_FLAT: launch_copy_flat(C->stream, a.d(), e.d(), C->L64.total); break;
        case PPS_K_CHEB_PAIR: {
            const double ca[2] = {0.1, 0.15}, cb[2] = {0.2, 0.25}, cc[2] = {0.3, 0.35};
            launch_cheb_pair<double, double>(g, C->cheb_pair_geo(), C->stream, (it & 1) ? d.d() : a.d(), b.d(), c.d(),
                                             (it & 1) ? e.d() : d.d(), (it & 1) ? a.d() : e.d(), ca, cb, cc, 1.0, false,
                                             nullptr);
            break;
        }
        }
    };
    double per = 0;
    switch (kernel) {
    case PPS_K_CG_STENCIL: case PPS_K_CG_STENCIL_F32: per = 4; break;   // read r, p ; write p, Ap
    case PPS_K_CG_UPDATE: case PPS_K_CG_UPDATE_F32: per = 3; break;     // read r, Ap ; write r
    case PPS_K_RESIDUAL: per = 3; break;                                // read x, b ; write r
    case PPS_K_COPY: per = 2; break;
    case PPS_K_CHEB_SWEEP: per = 4; break;                              // read Y, B, Z ; write W
    case PPS_K_FCG_STENCIL: case PPS_K_FCG_STENCIL_F32: per = 5; break; // read z, p_old, r ; write p, Ap
    case PPS_K_COPY_FLAT: per = 2.0 * (double)C->L64.total / n; break;  // the whole allocation, read and written
    case PPS_K_CHEB_PAIR: per = 5; break;                               // read Y, Z, B ; write W1, W2 (two sweeps)
    default: return C->fail(PPS_ERR_ARG, "unknown kernel id %d", kernel);
    }
    for (int i = 0; i < 3; ++i) launch(i);  // warm-up
    rc = C->hip(hipEventRecord(C->ev_t0, C->stream), "event");
    for (int i = 0; i < reps; ++i) launch(i);
    if (!rc) rc = C->hip(hipEventRecord(C->ev_t1, C->stream), "event");
    if (!rc) rc = C->hip(hipEventSynchronize(C->ev_t1), "event sync");
    if (!rc) rc = C->hip(hipGetLastError(), "kernel");
    if (rc) return rc;
    float t = 0;
    (void)hipEventElapsedTime(&t, C->ev_t0, C->ev_t1);
    *ms = t / reps;
    *bytes = per * n * w;
    return PPS_OK;
}

}  // extern "C"

// ------------------------------------------------------------ right-hand sides

namespace {

uint64_t mix64h(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double u01h(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    const uint64_t r = mix64h(mix64h(mix64h(seed) ^ stream) ^ ctr);
    return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
double normalh(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    const double a = u01h(seed, stream, 2u * ctr), b = u01h(seed, stream, 2u * ctr + 1u);
    return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
}

// temporary device buffer freed at scope exit (after a stream sync)
struct TmpDev {
    void* p = nullptr;
    ~TmpDev()
    {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" {

int pps_rhs_normal(void* ctx, uint64_t seed)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    TmpDev t;
    int rc = C->hip(hipMalloc(&t.p, sizeof(double) * 4096), "rhs scratch");
    if (rc) return rc;
    launch_rhs_normal(C->stream, C->L64, C->solver, C->np, seed, C->fields[PPS_FIELD_B].d(), static_cast<double*>(t.p));
    return C->hip(hipStreamSynchronize(C->stream), "rhs_normal");
}

int pps_rhs_gauss(void* ctx, uint64_t seed, int nsrc, double wmin, double wmax)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    if (nsrc < 1 || nsrc > 256) return C->fail(PPS_ERR_ARG, "nsrc must be in [1,256]");
    // centre uniform in [0,1]^3, width uniform in [wmin,wmax], amplitude N(0,1)
    std::vector<double> tab(5 * (size_t)nsrc);
    for (int s = 0; s < nsrc; ++s) {
        tab[s] = u01h(seed, 2u, 4u * (uint64_t)s + 0u);
        tab[nsrc + s] = u01h(seed, 2u, 4u * (uint64_t)s + 1u);
        tab[2 * nsrc + s] = u01h(seed, 2u, 4u * (uint64_t)s + 2u);
        tab[3 * nsrc + s] = wmin + (wmax - wmin) * u01h(seed, 2u, 4u * (uint64_t)s + 3u);
        tab[4 * nsrc + s] = normalh(seed, 4u, (uint64_t)s);
    }
    TmpDev t;
    int rc = C->hip(hipMalloc(&t.p, sizeof(double) * tab.size()), "gauss table");
    if (!rc) rc = C->hip(hipMemcpy(t.p, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice), "gauss table");
    if (rc) return rc;
    launch_rhs_gauss(C->stream, C->L64, C->solver, static_cast<double*>(t.p), nsrc, C->grid.origin, C->grid.ds,
                     C->fields[PPS_FIELD_B].d());
    return C->hip(hipStreamSynchronize(C->stream), "rhs_gauss");
}

int pps_rhs_spectrum(void* ctx, uint64_t seed, int K, int accumulate)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    if (K < 1 || K > 256) return C->fail(PPS_ERR_ARG, "kmax must be in [1,256]");
    const double PI = 3.141592653589793;
    std::vector<double> coef((size_t)K * K * K);
    for (int kz = 1; kz <= K; ++kz)
        for (int ky = 1; ky <= K; ++ky)
            for (int kx = 1; kx <= K; ++kx) {
                const uint64_t idx = (uint64_t)(kx - 1) + (uint64_t)K * ((uint64_t)(ky - 1) + (uint64_t)K * (uint64_t)(kz - 1));
                const double k2 = (double)kx * kx + (double)ky * ky + (double)kz * kz;
                coef[idx] = normalh(seed, 3u, idx) / k2;
            }
    std::vector<double> S[3];
    for (int d = 0; d < 3; ++d) {
        const int n = C->np[d];
        S[d].resize((size_t)K * n);
        for (int k = 1; k <= K; ++k)
            for (int i = 0; i < n; ++i) S[d][(k - 1) + (size_t)K * i] = std::sin(k * PI * (double)i / (n - 1));
    }
    TmpDev dc, dsx, dsy, dsz, t1, t2;
    int rc = C->hip(hipMalloc(&dc.p, sizeof(double) * coef.size()), "spec");
    if (!rc) rc = C->hip(hipMalloc(&dsx.p, sizeof(double) * S[0].size()), "spec");
    if (!rc) rc = C->hip(hipMalloc(&dsy.p, sizeof(double) * S[1].size()), "spec");
    if (!rc) rc = C->hip(hipMalloc(&dsz.p, sizeof(double) * S[2].size()), "spec");
    if (!rc) rc = C->hip(hipMalloc(&t1.p, sizeof(double) * (size_t)K * K * C->np[2]), "spec");
    if (!rc) rc = C->hip(hipMalloc(&t2.p, sizeof(double) * (size_t)K * C->np[1] * C->np[2]), "spec");
    if (!rc) rc = C->hip(hipMemcpy(dc.p, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice), "spec");
    if (!rc) rc = C->hip(hipMemcpy(dsx.p, S[0].data(), sizeof(double) * S[0].size(), hipMemcpyHostToDevice), "spec");
    if (!rc) rc = C->hip(hipMemcpy(dsy.p, S[1].data(), sizeof(double) * S[1].size(), hipMemcpyHostToDevice), "spec");
    if (!rc) rc = C->hip(hipMemcpy(dsz.p, S[2].data(), sizeof(double) * S[2].size(), hipMemcpyHostToDevice), "spec");
    if (rc) return rc;
    launch_rhs_spectrum(C->stream, C->L64, C->solver, C->np, K, static_cast<double*>(dc.p), static_cast<double*>(dsx.p),
                        static_cast<double*>(dsy.p), static_cast<double*>(dsz.p), static_cast<double*>(t1.p),
                        static_cast<double*>(t2.p), accumulate, C->fields[PPS_FIELD_B].d());
    return C->hip(hipStreamSynchronize(C->stream), "rhs_spectrum");
}


// ------------------------------------------------------- learned preconditioner

struct NNHandle {
    std::shared_ptr<nn::Net> net;
};

static NNHandle* as_nn(void* p) { return static_cast<NNHandle*>(p); }
static thread_local std::string g_nn_error;

static int nn_fail(NNHandle* H, int code)
{
    g_nn_error = H ? H->net->error() : "null network handle";
    return code;
}

// e = u* - x over the whole box
__global__ void k_nn_sub(double* e, const double* us, const double* x, long long n)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) e[t] = us[t] - x[t];
}

// dense copy of the solver region, scaled: dst[i + n0 (j + n1 k)] = s * f(lo + (i,j,k))
__global__ void k_nn_gather(const double* f, Layout L, Region R, double s, float* dst)
{
    const int n0 = R.hi[0] - R.lo[0], n1 = R.hi[1] - R.lo[1], n2 = R.hi[2] - R.lo[2];
    const long long n = (long long)n0 * n1 * n2;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % n0);
        const long long q = t / n0;
        dst[t] = (float)(s * f[L.at(R.lo[0] + i, R.lo[1] + (int)(q % n1), R.lo[2] + (int)(q / n1))]);
    }
}

const char* pps_nn_last_error(void* nn)
{
    (void)nn;
    return g_nn_error.c_str();
}

int pps_nn_create(uint64_t seed, void** out)
{
    if (!out) return PPS_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_nn_error = "no GPU visible";
        return PPS_ERR_NOGPU;
    }
    auto* H = new NNHandle;
    H->net = std::make_shared<nn::Net>();
    if (H->net->init(seed)) {
        g_nn_error = H->net->error();
        delete H;
        return PPS_ERR_HIP;
    }
    *out = H;
    return PPS_OK;
}

void pps_nn_destroy(void* nn) { delete as_nn(nn); }

int pps_nn_num_params(void* nn, long long* stored, long long* effective)
{
    if (!nn) return PPS_ERR_ARG;
    if (stored) *stored = nn::kParams;
    if (effective) *effective = nn::Net::effective_params();
    return PPS_OK;
}

int pps_nn_get_params(void* nn, float* host, long long n)
{
    NNHandle* H = as_nn(nn);
    if (!H || !host) return PPS_ERR_ARG;
    return H->net->get_params(host, n) ? nn_fail(H, PPS_ERR_ARG) : PPS_OK;
}

int pps_nn_set_params(void* nn, const float* host, long long n)
{
    NNHandle* H = as_nn(nn);
    if (!H || !host) return PPS_ERR_ARG;
    return H->net->set_params(host, n) ? nn_fail(H, PPS_ERR_ARG) : PPS_OK;
}

int pps_nn_forward(void* nn, const float* in, int batch, const int n[3], float* out)
{
    NNHandle* H = as_nn(nn);
    if (!H || !in || !out || !n || batch < 1 || n[0] < 1 || n[1] < 1 || n[2] < 1) return PPS_ERR_ARG;
    return H->net->forward_host(in, batch, n, out) ? nn_fail(H, PPS_ERR_HIP) : PPS_OK;
}

int pps_nn_activation(void* nn, int index, float* out)
{
    NNHandle* H = as_nn(nn);
    if (!H || !out) return PPS_ERR_ARG;
    return H->net->activation_host(index, out) ? nn_fail(H, PPS_ERR_ARG) : PPS_OK;
}

int pps_nn_loss_grad(void* nn, const float* in, const double* g, int batch, const int n[3], double* loss, float* grad)
{
    NNHandle* H = as_nn(nn);
    if (!H || !in || !g || !loss || !grad || !n || batch < 1 || n[0] < 1 || n[1] < 1 || n[2] < 1) return PPS_ERR_ARG;
    return H->net->loss_grad_host(in, g, batch, n, loss, grad) ? nn_fail(H, PPS_ERR_HIP) : PPS_OK;
}

int pps_nn_generate(void* nn, const pps_nn_train_cfg* cfg, long long* pairs)
{
    NNHandle* H = as_nn(nn);
    if (!H || !cfg) return PPS_ERR_ARG;
    const int n = cfg->n;
    if (n < 4 || n > 512 || cfg->nrhs < 1 || cfg->snapshots < 1 || cfg->kmax < 1) {
        g_nn_error = "pps_nn_generate: need 4 <= n <= 512, nrhs >= 1, snapshots >= 1, kmax >= 1";
        return PPS_ERR_ARG;
    }
    pps_grid g{};
    g.dim = 3;
    for (int d = 0; d < 3; ++d) {
        g.nranks[d] = 1;
        g.npglobal[d] = n + 2;
        g.ds[d] = 1.0 / (n + 1);
        g.guards[d] = 1;
    }
    void* cx = nullptr;
    int rc = pps_create(&g, &cx);
    if (rc) {
        g_nn_error = std::string("pps_nn_generate: ") + pps_last_error(nullptr);
        return rc;
    }
    std::unique_ptr<void, void (*)(void*)> guard(cx, pps_destroy);
    Ctx* C = as_ctx(cx);
    // scratch fields are allocated on first use: the raw kernels below write T1/T2
    for (int id : {PPS_FIELD_T0, PPS_FIELD_T1, PPS_FIELD_T2, PPS_FIELD_T3}) {
        Field f;
        if ((rc = field_of(*C, id, f))) {
            g_nn_error = std::string("pps_nn_generate: ") + C->err;
            return rc;
        }
    }
    const int dims[3] = {n, n, n};
    const long long npairs = (long long)cfg->nrhs * cfg->snapshots;
    if (H->net->dataset_reserve(npairs, dims)) return nn_fail(H, PPS_ERR_NOMEM);
    std::vector<double>& G = H->net->dataset_g();
    const double h = 1.0 / (n + 1);
    const double nvox = (double)n * n * n;
    pps_stage cg[2]{};
    cg[0].kind = PPS_CG;
    cg[0].precision = 64;
    cg[1].kind = PPS_NONE;
    cg[1].precision = 64;
    pps_result res{};
    long long k = 0;
    auto fail_ctx = [&](int code) {
        g_nn_error = std::string("pps_nn_generate: ") + C->err;
        return code;
    };
    for (int s = 0; s < cfg->nrhs; ++s) {
        // converged solution u* of a k^-2 spectrum (or white-noise) right-hand side;
        // the white ones are spread evenly over the sequence
        const bool white = (int)((s + 1) * cfg->white_fraction) > (int)(s * cfg->white_fraction);
        if ((rc = pps_fill(cx, PPS_FIELD_X, 0.0)) || (rc = pps_fill(cx, PPS_FIELD_B, 0.0)) ||
            (rc = white ? pps_rhs_normal(cx, cfg->seed + (uint64_t)s) : pps_rhs_spectrum(cx, cfg->seed + (uint64_t)s, cfg->kmax, 0)) ||
            (rc = pps_copy_field(cx, PPS_FIELD_T3, PPS_FIELD_B)))
            return fail_ctx(rc);
        cg[0].tol = 1e-11;
        cg[0].max_iter = 100000;
        if ((rc = pps_solve(cx, cg, 2, &res))) return fail_ctx(rc);
        const int kstar = res.iterations;
        if ((rc = pps_copy_field(cx, PPS_FIELD_T0, PPS_FIELD_X))) return fail_ctx(rc);
        for (int j = 0; j < cfg->snapshots; ++j, ++k) {
            // CG iterate x_j of the same problem from x = 0 (j = 0: x = 0, r = b)
            const int it = j == 0 ? 0 : 1 + (int)((long long)(j - 1) * (9LL * kstar) / (10LL * std::max(1, cfg->snapshots - 1)));
            if ((rc = pps_fill(cx, PPS_FIELD_X, 0.0)) || (rc = pps_copy_field(cx, PPS_FIELD_B, PPS_FIELD_T3)))
                return fail_ctx(rc);
            if (it > 0) {
                cg[0].tol = 0.0;
                cg[0].max_iter = it;
                if ((rc = pps_solve(cx, cg, 2, &res))) return fail_ctx(rc);
            }
            double rn = 0.0, er = 0.0;
            if ((rc = pps_residual(cx, PPS_FIELD_T1, PPS_FIELD_X, PPS_FIELD_B, &rn))) return fail_ctx(rc);
            const long long tot = C->L64.total;
            hipLaunchKernelGGL(k_nn_sub, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, C->stream,
                               C->fields[PPS_FIELD_T2].d(), C->fields[PPS_FIELD_T0].d(), C->fields[PPS_FIELD_X].d(), tot);
            if ((rc = pps_dot(cx, PPS_FIELD_T2, PPS_FIELD_T1, &er))) return fail_ctx(rc);
            const double rho = rn / std::sqrt(nvox);
            if (!(rho > 0.0)) {
                g_nn_error = "pps_nn_generate: zero residual snapshot";
                return PPS_ERR_STATE;
            }
            hipLaunchKernelGGL(k_nn_gather, dim3(1024), dim3(256), 0, C->stream, C->fields[PPS_FIELD_T1].d(), C->L64,
                               C->solver, 1.0 / rho, H->net->dataset_sample(k));
            // energy norm of the normalised error in index-space units: -(e.r) / (h^2 rho^2)
            G[k] = -er / (h * h * rho * rho);
        }
    }
    if ((rc = C->hip(hipStreamSynchronize(C->stream), "sync"))) return fail_ctx(rc);
    if (pairs) *pairs = npairs;
    return PPS_OK;
}

int pps_nn_dataset(void* nn, long long index, float* r, double* g)
{
    NNHandle* H = as_nn(nn);
    if (!H || index < 0 || index >= H->net->dataset_size()) return PPS_ERR_ARG;
    if (r && H->net->hip(hipMemcpy(r, H->net->dataset_sample(index), (size_t)H->net->sample_size() * sizeof(float),
                                   hipMemcpyDeviceToHost),
                         "download sample"))
        return nn_fail(H, PPS_ERR_HIP);
    if (g) *g = H->net->dataset_g()[index];
    return PPS_OK;
}

int pps_nn_train(void* nn, const pps_nn_train_cfg* cfg, pps_nn_train_stats* st)
{
    NNHandle* H = as_nn(nn);
    if (!H || !cfg) return PPS_ERR_ARG;
    nn::TrainCfg c{cfg->batch, cfg->steps, (float)cfg->lr, (float)cfg->lr_final, cfg->seed};
    nn::TrainStats s{};
    if (H->net->train(c, &s)) return nn_fail(H, PPS_ERR_STATE);
    if (st) {
        st->pairs = H->net->dataset_size();
        st->steps = s.steps;
        st->seconds = s.seconds;
        st->loss_first = s.loss_first;
        st->loss_last = s.loss_last;
    }
    return PPS_OK;
}

int pps_nn_apply_field(void* ctx, int out, int in)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    if (!C->nn) return C->fail(PPS_ERR_STATE, "no network attached (pps_nn_attach)");
    Field fo, fi;
    int rc = field_of(*C, out, fo);
    if (!rc) rc = field_of(*C, in, fi);
    if (rc) return rc;
    if (out == in) return C->fail(PPS_ERR_ARG, "pps_nn_apply_field: out and in must differ");
    // a scratch scalar block stands in for the calling Krylov stage: ||in||^2, open gates
    Scal* S = C->scal(C->nscal - 1);
    rc = scal_init(*C, S, 0.0, 0, 1.0, nullptr, 0);
    if (rc) return rc;
    launch_dot<double>(C->gsolver(64), C->stream, C->rb, fi.d(), fi.d(), &S->v[S_RR], nullptr);
    pps_stage st{};
    st.kind = PPS_NN;
    NNStage stage(*C, st, 1);
    rc = stage.apply(fo, fi, &S->done);
    if (rc) return rc;
    return C->hip(hipStreamSynchronize(C->stream), "sync");
}

int pps_nn_apply_field_prec(void* ctx, int out, int in, int prec, int dot, double* gamma)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    if (prec != 32 && prec != 64) return C->fail(PPS_ERR_ARG, "pps_nn_apply_field_prec: precision must be 32 or 64");
    if (!C->nn) return C->fail(PPS_ERR_STATE, "no network attached (pps_nn_attach)");
    if (dot >= 0 && !gamma) return C->fail(PPS_ERR_ARG, "pps_nn_apply_field_prec: dot field without gamma");
    Field fo, fi, fd;
    int rc = field_of(*C, out, fo);
    if (!rc) rc = field_of(*C, in, fi);
    if (!rc && dot >= 0) rc = field_of(*C, dot, fd);
    if (rc) return rc;
    if (out == in || out == dot) return C->fail(PPS_ERR_ARG, "pps_nn_apply_field_prec: out must differ from in and dot");
    // ||in||^2 of the FP64 field stands in for the calling Krylov stage's, as in pps_nn_apply_field
    Scal* S = C->scal(C->nscal - 1);
    rc = scal_init(*C, S, 0.0, 0, 1.0, nullptr, 0);
    if (rc) return rc;
    launch_dot<double>(C->gsolver(64), C->stream, C->rb, fi.d(), fi.d(), &S->v[S_RR], nullptr);
    const long long n = C->L64.total;
    const unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 65536);
    Field ai = fi, ao = fo, ad = fd;
    if (prec == 32) {
        // FP32 copies, as the mixed-precision inner solve hands its fields to the network
        if ((rc = C->get_work("nnapply.in", 32, ai)) || (rc = C->get_work("nnapply.out", 32, ao))) return rc;
        if (dot >= 0 && (rc = C->get_work("nnapply.dot", 32, ad))) return rc;
        hipLaunchKernelGGL(k_narrow, dim3(nb), dim3(256), 0, C->stream, fi.d(), ai.f(), n);
        if (dot >= 0) hipLaunchKernelGGL(k_narrow, dim3(nb), dim3(256), 0, C->stream, fd.d(), ad.f(), n);
    }
    pps_stage st{};
    st.kind = PPS_NN;
    NNStage stage(*C, st, 1);
    int nblocks = 0;
    rc = stage.apply_dot(ao, ai, &S->done, dot >= 0 ? ad : Field{}, &nblocks);
    if (rc) return rc;
    if (prec == 32) hipLaunchKernelGGL(k_widen, dim3(nb), dim3(256), 0, C->stream, ao.f(), fo.d(), n);
    if (dot < 0) return C->hip(hipStreamSynchronize(C->stream), "sync");
    if (nblocks <= 0) return C->fail(PPS_ERR_STATE, "pps_nn_apply_field_prec: the output layer formed no dot partials");
    std::vector<double> part((size_t)nblocks);
    if ((rc = C->hip(hipMemcpyAsync(part.data(), C->rb.partials, sizeof(double) * nblocks, hipMemcpyDeviceToHost,
                                    C->stream), "copy partials")))
        return rc;
    if ((rc = C->hip(hipStreamSynchronize(C->stream), "sync"))) return rc;
    double g = 0.0;
    for (double v : part) g += v;
    *gamma = g;
    return PPS_OK;
}

int pps_nn_apply_update(void* ctx, int out, int r_prev, int ap_prev, int r_cur, double alpha, double rr, int prec)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    if (prec != 32 && prec != 64) return C->fail(PPS_ERR_ARG, "pps_nn_apply_update: precision must be 32 or 64");
    if (!C->nn) return C->fail(PPS_ERR_STATE, "no network attached (pps_nn_attach)");
    if (!(rr > 0.0)) return C->fail(PPS_ERR_ARG, "pps_nn_apply_update: rr must be positive");
    Field fo, fp, fa, fc;
    int rc = field_of(*C, out, fo);
    if (!rc) rc = field_of(*C, r_prev, fp);
    if (!rc) rc = field_of(*C, ap_prev, fa);
    if (!rc) rc = field_of(*C, r_cur, fc);
    if (rc) return rc;
    if (out == r_prev || out == ap_prev || out == r_cur || r_cur == r_prev || r_cur == ap_prev)
        return C->fail(PPS_ERR_ARG, "pps_nn_apply_update: out and r_cur must differ from the other fields");
    // the FCG stage's scalars: step length alpha and ||r_cur||^2 (which the
    // stage's own reduction supplies in the solve), no restart
    Scal* S = C->scal(C->nscal - 1);
    rc = scal_init(*C, S, 0.0, 0, 1.0, nullptr, 0);
    if (rc) return rc;
    C->scal_h->v[S_ALPHA] = alpha;
    C->scal_h->v[S_RR] = rr;
    if ((rc = C->hip(hipMemcpyAsync(S, C->scal_h, sizeof(Scal), hipMemcpyHostToDevice, C->stream), "scal")))
        return rc;
    const long long n = C->L64.total;
    const unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 65536);
    Field ao = fo, ap = fp, aa = fa, ac = fc;
    if (prec == 32) {
        if ((rc = C->get_work("nnupd.out", 32, ao)) || (rc = C->get_work("nnupd.rp", 32, ap)) ||
            (rc = C->get_work("nnupd.ap", 32, aa)) || (rc = C->get_work("nnupd.rc", 32, ac)))
            return rc;
        hipLaunchKernelGGL(k_narrow, dim3(nb), dim3(256), 0, C->stream, fp.d(), ap.f(), n);
        hipLaunchKernelGGL(k_narrow, dim3(nb), dim3(256), 0, C->stream, fa.d(), aa.f(), n);
        hipLaunchKernelGGL(k_narrow, dim3(nb), dim3(256), 0, C->stream, fc.d(), ac.f(), n);
    }
    pps_stage st{};
    st.kind = PPS_NN;
    NNStage stage(*C, st, 1);
    int nblocks = 0;
    rc = stage.apply_update(ao, ac, ap, aa, &S->v[S_ALPHA], &S->restart, &S->done, Field{}, &nblocks);
    if (rc) return rc;
    if (prec == 32) {
        hipLaunchKernelGGL(k_widen, dim3(nb), dim3(256), 0, C->stream, ao.f(), fo.d(), n);
        hipLaunchKernelGGL(k_widen, dim3(nb), dim3(256), 0, C->stream, ac.f(), fc.d(), n);
    }
    return C->hip(hipStreamSynchronize(C->stream), "sync");
}

int pps_nn_time_layer(void* ctx, int layer, int prec, int reps, double* ms, double* bytes, double* flops)
{
    Ctx* C = as_ctx(ctx);
    if (!C || !ms || !bytes || !flops) return PPS_ERR_ARG;
    if (prec != 32 && prec != 64) return C->fail(PPS_ERR_ARG, "pps_nn_time_layer: precision must be 32 or 64");
    // an apply on T0 <- B sets up the activations the layer reads
    int rc = pps_nn_apply_field(ctx, PPS_FIELD_T0, PPS_FIELD_B);
    if (rc) return rc;
    Field fo, fi, fd;
    if ((rc = field_of(*C, PPS_FIELD_T0, fo)) || (rc = field_of(*C, PPS_FIELD_B, fi))) return rc;
    // the output layer's dot operand (Ap_old in FCG): a field of its own, a copy of B
    if ((rc = C->get_work("nntime.dot", prec, fd))) return rc;
    const long long n = C->L64.total;
    const unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 65536);
    if (prec == 32) {
        // FP32 copies of B and an FP32 output, as the mixed-precision inner solve hands them over
        Field fi32, fo32;
        if ((rc = C->get_work("nntime.in", 32, fi32)) || (rc = C->get_work("nntime.out", 32, fo32))) return rc;
        hipLaunchKernelGGL(k_narrow, dim3(nb), dim3(256), 0, C->stream, fi.d(), fi32.f(), n);
        hipLaunchKernelGGL(k_narrow, dim3(nb), dim3(256), 0, C->stream, fi.d(), fd.f(), n);
        fi = fi32;
        fo = fo32;
    } else if ((rc = C->hip(hipMemcpyAsync(fd.p, fi.p, n * 8, hipMemcpyDeviceToDevice, C->stream), "copy"))) {
        return rc;
    }
    Scal* S = C->scal(C->nscal - 1);
    const Layout& L = C->lay(prec);
    NNFieldGeom g;
    g.off = L.at(C->solver.lo[0], C->solver.lo[1], C->solver.lo[2]);
    g.sy = L.px;
    g.sz = L.plane;
    g.n0 = C->solver.hi[0] - C->solver.lo[0];
    g.n1 = C->solver.hi[1] - C->solver.lo[1];
    g.n2 = C->solver.hi[2] - C->solver.lo[2];
    auto* net = dynamic_cast<nn::Net*>(C->nn.get());
    if (!net) return C->fail(PPS_ERR_STATE, "no network attached");
    if (net->time_layer(C->stream, g, fo.p, prec, fi.p, prec, fd.p, &S->v[S_RR], layer, reps, ms, bytes, flops))
        return C->fail(PPS_ERR_ARG, "%s", net->error());
    return PPS_OK;
}

int pps_nn_attach(void* ctx, void* nn)
{
    Ctx* C = as_ctx(ctx);
    if (!C) return PPS_ERR_ARG;
    C->nn = nn ? std::static_pointer_cast<NNPrecond>(as_nn(nn)->net) : nullptr;
    return PPS_OK;
}

}  // extern "C"
