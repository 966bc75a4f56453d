/*
 * pps_abi.h -- C ABI of the MI355X Poisson Krylov library (libpps.so).
 *
 * The library replaces the device side of the reference's solver entry call
 *     T_Solver solver(blockGrid, exactSolutionAndBCs, communicator, alpakaHelper);
 *     solver(fieldX, fieldB);
 * (solverPoissonMPI_alpaka/src/main.cpp:90,104; BiCGstabAlpaka.hpp:68-475) and
 * of the CPU form solver.setProblem(x,b); solver(x,b,operatorA)
 * (solverPoissonMPI_CPU/src/main.cpp:83,94,99).  The reference has no C ABI:
 * these entry points are what an FFI (ctypes, or the C++ facade in
 * include/pps/) binds for that path.  Plain pointers and sizes only.
 *
 * Conventions
 *  - every call returns 0 on success or a negative PPS_ERR_* code; the text of
 *    the last error is pps_last_error(ctx) (ctx may be NULL for pps_create);
 *  - host fields are dense guarded arrays of (n0+2)(n1+2)(n2+2) doubles, x
 *    fastest, the reference host layout (iterativeSolverBase.hpp:34-35);
 *  - the context owns its device memory and one HIP stream; calls on one
 *    context are not thread safe; contexts are independent;
 *  - one GPU: the rank grid must be 1x1x1 (the MPI halo exchange of
 *    communicationMPI.hpp becomes the on-device ghost fill).
 */
#ifndef PPS_ABI_H
#define PPS_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PPS_ABI_VERSION 1

enum {
    PPS_OK = 0,
    PPS_ERR_ARG = -1,          /* invalid argument                                 */
    PPS_ERR_HIP = -2,          /* HIP runtime error                                */
    PPS_ERR_NOMEM = -3,        /* device allocation failed                         */
    PPS_ERR_UNSUPPORTED = -4,  /* configuration outside what this build implements */
    PPS_ERR_STATE = -5,        /* call out of order (e.g. NN not trained)          */
    PPS_ERR_NOGPU = -6         /* no HIP device                                    */
};

/* Grid description: the arguments of BlockGrid's constructor
 * (solverPoissonMPI_CPU/include/blockGrid.hpp:15). */
typedef struct pps_grid {
    int dim;            /* DIM; this build implements 3                            */
    int nranks[3];      /* rank grid; must be {1,1,1} on one GPU                   */
    int my_rank;        /* 0                                                       */
    int npglobal[3];    /* data points per direction, incl. Dirichlet layers       */
    double ds[3];       /* mesh spacing                                            */
    double origin[3];   /* coordinates of the first data point                     */
    int guards[3];      /* ghost width; the 7-point stencil needs {1,1,1}          */
    int bcs_type[6];    /* 0 Dirichlet, 1 Neumann; order x- x+ y- y+ z- z+          */
    double bcs_value[6];/* carried for signature parity; values come from fields   */
} pps_grid;

/* fields addressable through the ABI (FP64, device resident) */
enum {
    PPS_FIELD_X = 0,    /* solution / initial guess with Dirichlet values          */
    PPS_FIELD_B = 1,    /* right-hand side                                          */
    PPS_FIELD_T0 = 2,   /* scratch fields for the kernel-level entry points         */
    PPS_FIELD_T1 = 3,
    PPS_FIELD_T2 = 4,
    PPS_FIELD_T3 = 5,
    PPS_NUM_FIELDS = 6
};

/* solver-chain stage kinds: the template chain of inputParam.hpp:22-31 */
enum {
    PPS_NONE = 0,        /* NoneSolver (noneSolver.hpp:24-27)                      */
    PPS_CG = 1,          /* BaseCG (baseCG.hpp:44-260)                             */
    PPS_BICGSTAB = 2,    /* BiCGSTAB (BiCGSTAB.hpp:55-322; BiCGstabAlpaka.hpp)     */
    PPS_CHEBYSHEV = 3,   /* ChebyshevIteration (chebyshevIteration.hpp:48-140)     */
    PPS_JACOBI = 4,      /* diagonal preconditioner (no reference counterpart)     */
    PPS_FCG = 5,         /* flexible CG, truncation 1 (no reference counterpart)   */
    PPS_IR = 6,          /* FP64 outer refinement of an FP32 inner stage (new)     */
    PPS_NN = 7           /* learned multilevel preconditioner (new)                */
};

/* One stage of the solver chain.  Stage 0 is the main solver
 * (isMainLoop = true), stage s+1 preconditions stage s. */
typedef struct pps_stage {
    int kind;            /* PPS_*                                                   */
    int precision;       /* 64 or 32: vector precision (T_data / T_data_chebyshev)  */
    double tol;          /* tolerance * tollScalingFactor (baseCG.hpp:22)           */
    int max_iter;        /* maxIteration; Chebyshev degree chebyshevMax             */
    double cheb_eps;     /* epsilon (solverSetup.hpp)                               */
    double cheb_rescale_min; /* rescaleEigMin                                       */
    double cheb_rescale_max; /* rescaleEigMax                                       */
    int cheb_local;      /* alpaka `local` eigenvalue variant                       */
    int stall_window;    /* FCG over PPS_NN: iterations per stall test (0 = never fall back); set on either stage */
    double stall_ratio;  /* fall back to Jacobi when the residual drop over the window is above this factor */
    double ir_first_tol; /* PPS_IR: the inner stage's tolerance for the first outer step (0: its own tol).  The
                            first FP32 correction cannot bring the true residual much below kappa * eps32, so a
                            looser first inner solve loses little; no reference counterpart (the reference has
                            no refinement) */
    int true_finish;     /* CG, BiCGSTAB, FCG as the main solver: 0 stops on the recursive residual, the
                            reference's rule (baseCG.hpp:224, BiCGSTAB.hpp:290).  n > 0: at that stop r = b - A x
                            is recomputed in FP64; while its norm is at or above tol (and iterations remain), r
                            is replaced by it and the loop resumes (CG, FCG keep their search direction,
                            BiCGSTAB restarts with r0 = r), at most n times.  The recursive residual drifts from
                            the true one as iterations accumulate; this makes the true residual reach tol. */
    double ir_adapt;     /* PPS_IR: 0 keeps the inner tolerance of every outer step.  a in (0, 1): an outer step
                            asks the inner stage for no more reduction than the outer tolerance still needs,
                            times a: max(inner tolerance, a * tol / e), e the outer residual before the step (the
                            last step of a solve typically needs less than the inner tolerance).  No reference
                            counterpart (the reference has no refinement) */
} pps_stage;

typedef struct pps_result {
    int iterations;                 /* getNumIterationFinal                         */
    int outer_iterations;           /* PPS_IR outer refinement steps                */
    long long precond_iterations;   /* getNumIterationPreconditionerFinal           */
    double error_from_iteration;    /* getErrorFromIteration (recursive residual)   */
    double error_compute_operator;  /* getErrorComputeOperator (true residual)      */
    double norm_b;                  /* normalisation factor ||b_adj||               */
    double seconds;                 /* device time of the solve (HIP events)        */
    int fallback_switches;          /* learned preconditioner -> fallback switches   */
    int history_len;                /* entries available from pps_get_history       */
} pps_result;

/* ---- context ----------------------------------------------------------- */
int pps_create(const pps_grid* grid, void** ctx);
void pps_destroy(void* ctx);
const char* pps_last_error(void* ctx);
int pps_abi_version(void);
/* guarded extent (n+2 per direction) and the device row pitch in elements */
int pps_get_extent(void* ctx, int ng[3]);

/* ---- fields (replaces the host<->device copies of BiCGstabAlpaka.hpp:95-112,473-474) */
int pps_upload(void* ctx, int field, const double* host_guarded);
int pps_download(void* ctx, int field, double* host_guarded);
int pps_fill(void* ctx, int field, double value);
/* device-to-device copy of a whole field */
int pps_copy_field(void* ctx, int dst, int src);
/* du/dn data of a Neumann face (trueSolutionDdir on the face's data points);
 * plane of the two other directions' guarded extents, lower direction fastest */
int pps_set_neumann_flux(void* ctx, int face, const double* host_plane);

/* ---- the solve (BaseCG / BiCGSTAB / ChebyshevIteration operator()) ------ */
int pps_solve(void* ctx, const pps_stage* chain, int nstages, pps_result* res);
/* one preconditioner application out = M(in): the preconditioner protocol
 * void operator()(Buf& out, Buf& in) (BiCGstabAlpaka.hpp:213-214,
 * chebyshevIterationAlpaka.hpp:119; CPU form BiCGSTAB.hpp:170,190).  chain[0]
 * is applied once (None, Jacobi, CG, BiCGSTAB, Chebyshev or FCG; Chebyshev
 * with precision 32 is the T_data_chebyshev = float form), chain[1..] nest as
 * in pps_solve; `in` has its Neumann ghosts reset in place, as the reference
 * preconditioners do.  *iterations (may be NULL) = iterations spent. */
int pps_precondition(void* ctx, const pps_stage* chain, int nstages, int out, int in, long long* iterations);
/* residual history of the last solve (errorFromIterationHistory_); returns the
 * number of entries copied (<= cap) or a negative error */
int pps_get_history(void* ctx, double* out, int cap);
/* wait for the context's stream */
int pps_synchronize(void* ctx);

/* ---- kernel-level entry points (parity tests, halo/ghost semantics) ----- */
/* y = A x on the solver region (matrixFreeOperatorA.hpp:22-39) */
int pps_apply_A(void* ctx, int y, int x);
/* ghost fill: Neumann mirror (iterativeSolverBase.hpp:62-169); with_flux adds
 * the -+2 ds flux / norm term of data fields.  Replaces CommunicatorMPI on one GPU. */
int pps_ghost_fill(void* ctx, int field, int with_flux, double norm);
/* r = b - A x on the solver region, returns ||r|| (iterativeSolverBase.hpp:236-280) */
int pps_residual(void* ctx, int r, int x, int b, double* norm);
/* pps_residual without a context and without allocating: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * r = b - A x on the solver region of `grid` (BlockGrid limits, one rank) and
 * ||r|| over it, on caller-owned device arrays in the reference's dense
 * guarded layout ((n0+2)(n1+2)(n2+2) doubles, x fastest); r is written on the
 * solver region only.  ||r|| goes to the device double *norm_dev and/or the
 * host double *norm_host (either may be NULL; a host result synchronises
 * `stream`).  stream: a hipStream_t, NULL for the null stream.  The reduction
 * partials live in a static device array of the library (1 MiB per GPU, kept
 * until the process ends): an x-y plane may need at most 131072 tiles of
 * 64 x 32 points, larger grids are refused with PPS_ERR_UNSUPPORTED.  Calls
 * that overlap on one GPU share that array and race.  Errors: pps_last_error(NULL). */
int pps_residual_noalloc(const pps_grid* grid, const double* x, const double* b, double* r, double* norm_dev,
                         double* norm_host, void* stream);
/* sum a.b over the solver region, deterministic single-pass grid reduction */
int pps_dot(void* ctx, int a, int b, double* out);
/* ||b_adj|| of the boundary-adjusted right-hand side (iterativeSolverBase.hpp:171-234) */
int pps_norm_b_adjusted(void* ctx, double* out);

/* ---- seeded right-hand sides, written into PPS_FIELD_B as b = -f -------- */
int pps_rhs_normal(void* ctx, uint64_t seed);
int pps_rhs_gauss(void* ctx, uint64_t seed, int nsrc, double wmin, double wmax);
/* accumulate: 0 overwrite, 1 add to the current b */
int pps_rhs_spectrum(void* ctx, uint64_t seed, int kmax, int accumulate);

/* ---- measurement ------------------------------------------------------- */
/* kernel ids for pps_time_kernel */
enum {
    PPS_K_CG_STENCIL = 0,   /* fused p = z + beta p, Ap = A p, p.Ap (x's updates deferred to a
                               pass per 16 iterations, as CG runs it where the device holds
                               the directions' ring) */
    PPS_K_CG_UPDATE = 1,    /* r -= alpha Ap, r.r, r.z                    */
    PPS_K_RESIDUAL = 2,     /* r = b - A x, r.r                           */
    PPS_K_COPY = 3,         /* plain copy of the solver region (reference) */
    PPS_K_CHEB_SWEEP = 4,   /* Chebyshev sweep                            */
    PPS_K_CG_STENCIL_F32 = 5,
    PPS_K_CG_UPDATE_F32 = 6,
    PPS_K_FCG_STENCIL = 7,     /* FCG p-stencil as the learned FCG runs it: p = z - beta p_old,
                                  Ap, p.Ap, p.r, r.Ap, Ap.Ap, r.r (x's updates are deferred to a
                                  pass per 16 iterations) */
    PPS_K_FCG_STENCIL_F32 = 8,
    PPS_K_COPY_FLAT = 9,       /* plain 1-D copy of a whole field allocation (16-byte
                                  non-temporal lanes): the measured HBM ceiling */
    PPS_K_CHEB_PAIR = 10       /* two fused Chebyshev sweeps: read Y, Z, B ; write y_{c+1}, y_{c+2} */
};
/* Average device time (ms) of `reps` back-to-back launches of one kernel on
 * the context stream, measured with HIP events; *bytes = algorithmic bytes
 * moved per launch. */
int pps_time_kernel(void* ctx, int kernel, int reps, double* ms, double* bytes);
/* Launch shape: the z-chunk of the stencil kernels (0 keeps the current one,
 * -1 returns to the automatic choice: the longest of 64/32/16/8 planes that
 * gives 2048 blocks) and the stores of the field outputs: nontemporal = 0
 * automatic (non-temporal once a field exceeds 32 MiB), 1 always
 * non-temporal, -1 never. */
int pps_set_tiling(void* ctx, int kz_stencil, int nontemporal);
/* Chebyshev sweeps: pairs = 1 runs them two per pass (one fused kernel reads
 * y_c, y_{c-1} and B once and writes y_{c+1} and y_{c+2}; the default), 0 one
 * sweep per pass; kz = z-chunk of the fused kernel (0 keeps the current one,
 * -1 returns to the automatic choice: the longest of 64/32/16 planes giving
 * 1024 blocks).
 * No counterpart in the reference: its sweeps are one kernel each
 * (kernelsAlpakaChebyshev.hpp:191-273). */
int pps_set_cheb_pairs(void* ctx, int pairs, int kz);

/* CG and FCG with the learned preconditioner keep their search directions in
 * a ring of 16 fields and add the steps to x in one pass per 16 iterations
 * (x bit for bit as with a per-iteration update, fewer bytes per iteration),
 * where the device holds the ring.  on = 0 updates x every iteration instead;
 * 1 (the default) uses the ring.  No counterpart in the reference. */
int pps_set_deferred_x(void* ctx, int on);

/* ---- learned preconditioner (new: the reference has no counterpart) ------
 * A 3-level convolutional network (16 channels, 3^3 kernels, ReLU, BF16
 * weights/activations, FP32 accumulation) used as the PPS_NN stage below a
 * PPS_FCG stage.  Trained on the GPU by its own forward/backward kernels on
 * (residual, error) pairs of plain CG iterates. */
typedef struct pps_nn_train_cfg {
    int n;                /* training grid: n^3 interior points, Dirichlet unit cube  */
    int nrhs;             /* right-hand sides (k^-2 power-law spectra, seeds seed+s)  */
    int snapshots;        /* CG iterates per right-hand side (pairs = nrhs*snapshots) */
    int kmax;             /* spectrum modes per direction                             */
    unsigned long long seed;
    int batch;            /* pairs per Adam step                                      */
    int steps;            /* Adam steps                                               */
    double lr, lr_final;  /* cosine learning-rate schedule                            */
    double white_fraction;/* share of right-hand sides drawn as white noise N(0,1)
                             instead of the k^-2 spectrum (rough residuals)           */
} pps_nn_train_cfg;

typedef struct pps_nn_train_stats {
    long long pairs;
    int steps;
    double seconds;       /* training time (generation excluded)                      */
    double loss_first, loss_last;  /* mean 1 - cos^2 in the energy norm (0 = exact)   */
} pps_nn_train_stats;

/* network with He-initialised weights (std::mt19937_64 seeded by `seed`) */
int pps_nn_create(uint64_t seed, void** nn);
void pps_nn_destroy(void* nn);
const char* pps_nn_last_error(void* nn);
/* stored FP32 parameters (14 layers x (27*16*16 + 16)) and the effective count */
int pps_nn_num_params(void* nn, long long* stored, long long* effective);
int pps_nn_get_params(void* nn, float* host, long long n);
int pps_nn_set_params(void* nn, const float* host, long long n);
/* network output for a dense batch [batch][n2][n1][n0] of inputs */
int pps_nn_forward(void* nn, const float* in, int batch, const int n[3], float* out);
/* activation `index` (0..12) of the last forward pass, [batch][z][y][x][16] */
int pps_nn_activation(void* nn, int index, float* out);
/* training loss of a batch (in = normalised residuals, g = their energy norms)
 * and its gradient with respect to all stored parameters */
int pps_nn_loss_grad(void* nn, const float* in, const double* g, int batch, const int n[3], double* loss,
                     float* grad);
/* training set: plain FP64 CG on the cfg grid; per right-hand side the
 * converged u* and `snapshots` iterates x_k; pairs (r_k / rms(r_k), energy
 * norm of e_k = u* - x_k) kept on the device */
int pps_nn_generate(void* nn, const pps_nn_train_cfg* cfg, long long* pairs);
/* one pair of the training set (r: n^3 floats, may be null) */
int pps_nn_dataset(void* nn, long long index, float* r, double* g);
/* Adam on the training set */
int pps_nn_train(void* nn, const pps_nn_train_cfg* cfg, pps_nn_train_stats* stats);
/* out = M(in) with the attached network on FP64 fields, as a PPS_NN stage
 * applies it (input scaled to unit RMS over the solver region) */
int pps_nn_apply_field(void* ctx, int out, int in);
/* out = M(in) as pps_nn_apply_field, with the network run on FP64 fields
 * (prec 64) or on FP32 copies as the mixed-precision inner solve runs it
 * (prec 32; the output is widened into `out`).  With dot >= 0 the output
 * layer also sums gamma = out . dot over the solver region in its epilogue,
 * as FCG's gamma = z . Ap_old (FP32 copy of dot for prec 32); the block
 * partials are added in block order on the host.  No reference counterpart:
 * the reference has no learned preconditioner. */
int pps_nn_apply_field_prec(void* ctx, int out, int in, int prec, int dot, double* gamma);
/* out = M(r_cur) with r_cur = r_prev - alpha * ap_prev formed by the
 * network's first layer, as FCG with the learned preconditioner runs it (the
 * residual update folded into the lift); r_cur is written on the solver
 * region.  rr = ||r_cur||^2 over the solver region (in the solve it comes
 * from the stage's own reduction).  prec 64: FP64 fields; prec 32: FP32
 * copies of the fields, with out and r_cur widened back over the whole
 * field.  No reference counterpart: the reference has no learned
 * preconditioner. */
int pps_nn_apply_update(void* ctx, int out, int r_prev, int ap_prev, int r_cur, double alpha, double rr, int prec);
/* average device time (HIP events, context stream) of `reps` launches of
 * network layer `layer` (0..13) at the context's grid, input B in FP64 fields
 * (prec 64) or in FP32 copies as the mixed-precision inner solve uses them
 * (prec 32); the layer's algorithmic HBM bytes and FLOPs per launch.  The
 * output layer runs as FCG runs it, also summing z . (a field of z's layout) */
int pps_nn_time_layer(void* ctx, int layer, int prec, int reps, double* ms, double* bytes, double* flops);
/* PPS_NN stages of ctx use this network (nn = NULL detaches) */
int pps_nn_attach(void* ctx, void* nn);

#ifdef __cplusplus
}
#endif
#endif
