/*
 * pps_oracle.h -- CPU oracle for the Krylov hot path of the Poisson solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * CPU solver (headers under solverPoissonMPI_CPU/include) on one sub-domain, plus the
 * parts the MI355X library adds (Jacobi, FCG, FP32-inner/FP64-outer iterative
 * refinement).  Only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark may load it, and only as the checker.  The GPU library never calls
 * into it.
 *
 * Layout: every field is a dense guarded array of ng[0]*ng[1]*ng[2] values,
 * x fastest, index i + ng[0]*j + ng[0]*ng[1]*k, ng = npglobal + 2 (one guard
 * layer), exactly as the reference host arrays
 * (iterativeSolverBase.hpp:34-35, blockGrid.hpp:172-182).
 *
 * Parity pins (each function's check lives in tests/):
 *   og_grid_init          vs the reference's BlockGrid limits/eigenvalues and closed forms
 *   og_apply_A            polynomial exactness, DST eigenvectors, dense brute force
 *   og_set_problem        vs the reference driver's fields (fixtures)
 *   og_solve (CG, BiCGSTAB, Chebyshev, None chains)
 *                         vs the compiled reference (golden fixtures: iterations,
 *                         histories, solutions), 1-rank and multi-rank
 *   og_solve (Jacobi)     Jacobi == scaled identity: same iterates as None
 *   og_solve (FCG)        reduces to CG for a fixed SPD preconditioner; residual drop
 *   og_solve (IR)         true FP64 residual below tolerance; FP32 CG vs FP64 CG
 *   og_precondition (Chebyshev, precision 32)
 *                         the FP64 Chebyshev (itself pinned by the fixtures) to FP32
 *                         rounding, the DST closed form of the polynomial, and a
 *                         solve preconditioned by it converging like the FP64 one
 *   og_rhs_*              statistics of the recipes (mean, variance, closed forms)
 */
#ifndef PPS_ORACLE_H
#define PPS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct og_grid {
    int np[3];          /* npglobal: data points per direction               */
    int ng[3];          /* guarded extent, np + 2                             */
    long sj, sk, ntot;  /* strides and guarded size                           */
    double ds[3], origin[3];
    int bc[6];          /* 0 Dirichlet, 1 Neumann; order x- x+ y- y+ z- z+    */
    int lim_data[6];    /* [lo,hi) per direction, blockGrid.hpp:184-206       */
    int lim_solver[6];  /* blockGrid.hpp:208-222                              */
    double eig_local[2], eig_global[2]; /* blockGrid.hpp:301-340             */
} og_grid;

/* solver-chain stage kinds */
enum {
    OG_NONE = 0,      /* NoneSolver: identity copy (noneSolver.hpp:24-27)          */
    OG_CG = 1,        /* BaseCG (baseCG.hpp:44-260)                               */
    OG_BICGSTAB = 2,  /* BiCGSTAB (BiCGSTAB.hpp:55-322)                           */
    OG_CHEB = 3,      /* ChebyshevIteration (chebyshevIteration.hpp:48-140)       */
    OG_JACOBI = 4,    /* diagonal scaling by 1/diag(A); no reference counterpart  */
    OG_FCG = 5,       /* flexible CG, truncation 1; no reference counterpart      */
    OG_IR = 6,        /* FP64 outer refinement around an FP32 inner stage         */
    OG_USER = 7       /* FP32 callback preconditioner (learned preconditioner)    */
};

typedef void (*og_user_fn)(void* user, const og_grid* g, const float* in, float* out);

typedef struct og_stage {
    int kind;
    double tol;         /* effective tolerance = tolerance * tollScalingFactor          */
    int maxit;          /* maxIteration (main/inner iterations; Chebyshev degree m)     */
    double cheb_eps, cheb_rmin, cheb_rmax; /* epsilon, rescaleEigMin, rescaleEigMax     */
    int cheb_local;                        /* 1: local eigenvalues, no rescale (alpaka)   */
    og_user_fn user_fn;                    /* OG_USER */
    void* user;
    double ir_first_tol;                   /* OG_IR: inner tolerance of the first outer step (0: the inner's) */
    int precision;                         /* OG_CHEB under an FP64 caller: 32 runs the alpaka
                                              T_data_chebyshev = float form, else FP64          */
    double ir_adapt;                       /* OG_IR: > 0: an outer step's inner tolerance is no less
                                              than ir_adapt * tol / (outer residual)            */
} og_stage;

typedef struct og_result {
    int iterations;                  /* getNumIterationFinal                  */
    long precond_iterations;         /* total inner iterations                */
    int outer_iterations;            /* IR outer steps                        */
    double error_from_iteration;     /* recursive residual (normalised)       */
    double error_compute_operator;   /* true residual b - A x (normalised)    */
    double norm_b;                   /* normalisation factor ||b_adj||        */
    double* history;                 /* caller buffer of history_cap values, or NULL */
    int history_cap;
    int history_len;
} og_result;

int og_grid_init(og_grid* g, const int np[3], const double ds[3], const double origin[3], const int bc[6]);

/* one point of the reference stencil (matrixFreeOperatorA.hpp:33-38) */
double og_stencil_point(const og_grid* g, const double* x, long idx);
/* y = A x on the solver region; other points of y untouched */
void og_apply_A(const og_grid* g, const double* x, double* y);
void og_apply_A_f32(const og_grid* g, const float* x, float* y);

/* built-in manufactured problems: 0 the reference's test problem, 1 sin*sin*sin,
 * 2 zero boundary data.  Sets x's Dirichlet layers and b on the data region
 * (iterativeSolverBase.hpp:51-55,537-603) and fills flux[f] (size ng_a*ng_b,
 * guarded plane, lower direction fastest) with du/dn data of Neumann faces. */
void og_set_problem(const og_grid* g, int kind, double* x, double* b, double* const flux[6]);
/* face plane size of face f (guarded extent of the two other directions) */
long og_face_size(const og_grid* g, int f);

/* Neumann ghost reset (iterativeSolverBase.hpp:62-169); flux may be NULL
 * (helper-field form); with flux, ghost = mirror -+ 2 ds flux / norm */
void og_reset_neumann(const og_grid* g, double* field, const double* const flux[6], double norm);

/* b-adjustment for boundary values (iterativeSolverBase.hpp:429-534) */
void og_adjust_b(const og_grid* g, const double* x, double* b, const double* const flux[6]);

/* r = b - A x on the solver region, returns ||r|| (iterativeSolverBase.hpp:236-280) */
double og_residual(const og_grid* g, const double* x, const double* b, double* r);

/* Full solve of chain[0] as the main solver with chain[1..] as nested
 * preconditioners.  x, b: guarded arrays (x holds Dirichlet values and the
 * initial guess; both are normalised and restored like the reference does).
 * flux: Neumann data per face or NULL.  Returns 0 on success. */
int og_solve(const og_grid* g, const og_stage* chain, int nstages, double* x, double* b,
             const double* const flux[6], og_result* res);

/* one preconditioner application out = M(in) of chain[0] (chain[1..] nested);
 * `in` gets its Neumann ghosts reset in place.  Returns 0 on success. */
int og_precondition(const og_grid* g, const og_stage* chain, int nstages, double* out, double* in, og_result* res);

/* seeded right-hand sides written as b = -f on the solver region (pps_rhs.h) */
void og_rhs_normal(const og_grid* g, uint64_t seed, double* b);
void og_rhs_gauss(const og_grid* g, uint64_t seed, int nsrc, double wmin, double wmax, double* b);
void og_rhs_spectrum(const og_grid* g, uint64_t seed, int K, double* b);
/* the recipes' random parameters (pps_rhs.h): the splitmix64 finaliser, the
 * normal draw, one Gaussian source, one spectrum coefficient */
uint64_t og_mix64(uint64_t z);
double og_rng_normal(uint64_t seed, uint64_t stream, uint64_t ctr);
void og_rhs_gauss_source(uint64_t seed, int s, double wmin, double wmax, double c[3], double* w, double* a);
double og_rhs_spec_coef(uint64_t seed, int K, int kx, int ky, int kz);

#ifdef __cplusplus
}
#endif
#endif
