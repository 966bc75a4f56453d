/*
 * pps_oracle.c -- CPU oracle for the Krylov hot path (TEST INFRASTRUCTURE ONLY).
 *
 * Plain-C restatement of the reference CPU solver on one sub-domain
 * (headers under solverPoissonMPI_CPU/include).  Loop orders, operation orders and
 * summation orders follow the reference so that, on one rank, iterates agree
 * with the compiled reference to the last bit.  Functions without a reference
 * counterpart (Jacobi, FCG, iterative refinement, FP32 inner solvers) are
 * marked as such and pinned by the invariants listed in pps_oracle.h.
 */
#include "pps_oracle.h"
#include "pps_rhs.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OG_PI 3.141592653589793

/* ------------------------------------------------------------------ grid */

/* BlockGrid for one rank: blockGrid.hpp:160-222 (limits), :301-340 (eigenvalues) */
int og_grid_init(og_grid* g, const int np[3], const double ds[3], const double origin[3], const int bc[6])
{
    memset(g, 0, sizeof *g);
    for (int d = 0; d < 3; ++d) {
        if (np[d] < 3) return -1;
        g->np[d] = np[d];
        g->ng[d] = np[d] + 2;
        g->ds[d] = ds[d];
        g->origin[d] = origin[d];
        g->lim_data[2 * d] = 1;
        g->lim_data[2 * d + 1] = np[d] + 1;
    }
    for (int f = 0; f < 6; ++f) g->bc[f] = bc[f];
    g->sj = g->ng[0];
    g->sk = (long)g->ng[0] * g->ng[1];
    g->ntot = g->sk * g->ng[2];
    memcpy(g->lim_solver, g->lim_data, sizeof g->lim_solver);
    for (int d = 0; d < 3; ++d) { /* one rank: every face is a boundary face */
        if (bc[2 * d] == 0) g->lim_solver[2 * d] += 1;
        if (bc[2 * d + 1] == 0) g->lim_solver[2 * d + 1] -= 1;
    }
    double emin = 0, emax = 0;
    for (int d = 0; d < 3; ++d) {
        const int lo = g->lim_solver[2 * d], hi = g->lim_solver[2 * d + 1];
        const double s1 = sin(1 * OG_PI / 2 / (hi - lo + 1));
        const double s2 = sin((hi - lo) * OG_PI / 2 / (hi - lo + 1));
        emin += 4 * s1 * s1 / (ds[d] * ds[d]);
        emax += 4 * s2 * s2 / (ds[d] * ds[d]);
    }
    g->eig_local[0] = emin;
    g->eig_local[1] = emax;
    emin = 0;
    emax = 0;
    for (int d = 0; d < 3; ++d) {
        int ni = np[d];
        if (bc[2 * d] == 0) ni -= 1;
        if (bc[2 * d + 1] == 0) ni -= 1;
        const double s1 = sin(1 * OG_PI / 2 / (ni + 1));
        const double s2 = sin(ni * OG_PI / 2 / (ni + 1));
        emin += 4 * s1 * s1 / (ds[d] * ds[d]);
        emax += 4 * s2 * s2 / (ds[d] * ds[d]);
    }
    g->eig_global[0] = emin;
    g->eig_global[1] = emax;
    return 0;
}

static inline long og_idx(const og_grid* g, int i, int j, int k) { return i + g->sj * j + g->sk * k; }

/* ------------------------------------------------------------- operator */

/* matrixFreeOperatorA.hpp:33-38 */
double og_stencil_point(const og_grid* g, const double* x, long c)
{
    const long sj = g->sj, sk = g->sk;
    return (x[c - 1] - 2 * x[c] + x[c + 1]) / (g->ds[0] * g->ds[0]) +
           (x[c - sj] - 2 * x[c] + x[c + sj]) / (g->ds[1] * g->ds[1]) +
           (x[c - sk] - 2 * x[c] + x[c + sk]) / (g->ds[2] * g->ds[2]);
}

static inline float og_stencil_point_f(const og_grid* g, const float* x, long c)
{
    const long sj = g->sj, sk = g->sk;
    const float dx2 = (float)g->ds[0] * (float)g->ds[0];
    const float dy2 = (float)g->ds[1] * (float)g->ds[1];
    const float dz2 = (float)g->ds[2] * (float)g->ds[2];
    return (x[c - 1] - 2 * x[c] + x[c + 1]) / dx2 + (x[c - sj] - 2 * x[c] + x[c + sj]) / dy2 +
           (x[c - sk] - 2 * x[c] + x[c + sk]) / dz2;
}

#define OG_FOR_SOLVER(g, c)                                                  \
    for (int k_ = (g)->lim_solver[4]; k_ < (g)->lim_solver[5]; ++k_)         \
        for (int j_ = (g)->lim_solver[2]; j_ < (g)->lim_solver[3]; ++j_)     \
            for (long c = og_idx((g), (g)->lim_solver[0], j_, k_),           \
                      c##_end = og_idx((g), (g)->lim_solver[1], j_, k_);     \
                 c < c##_end; ++c)

void og_apply_A(const og_grid* g, const double* x, double* y)
{
    OG_FOR_SOLVER(g, c) y[c] = og_stencil_point(g, x, c);
}

void og_apply_A_f32(const og_grid* g, const float* x, float* y)
{
    OG_FOR_SOLVER(g, c) y[c] = og_stencil_point_f(g, x, c);
}

/* ------------------------------------------------------------- problems */

static double og_coord(const og_grid* g, int d, int idx) { return g->origin[d] + (idx - g->lim_data[2 * d]) * g->ds[d]; }

static double prob_u(int kind, double x, double y, double z)
{
    if (kind == 0) return sin(x) + cos(y) + 3 * sin(z) + x * x * y * z + x * x + 10;
    if (kind == 1) return sin(OG_PI * x) * sin(OG_PI * y) * sin(OG_PI * z);
    return 0;
}
static double prob_f(int kind, double x, double y, double z)
{
    if (kind == 0) return -sin(x) - cos(y) - 3 * sin(z) + 2 * y * z + 2;
    if (kind == 1) return -3 * OG_PI * OG_PI * sin(OG_PI * x) * sin(OG_PI * y) * sin(OG_PI * z);
    return 0;
}
static double prob_du(int kind, double x, double y, double z, int dir)
{
    if (kind == 0) {
        if (dir == 0) return cos(x) + 2 * x * y * z + 2 * x;
        if (dir == 1) return -sin(y) + x * x * z;
        return 3 * cos(z) + x * x * y;
    }
    if (kind == 1) {
        if (dir == 0) return OG_PI * cos(OG_PI * x) * sin(OG_PI * y) * sin(OG_PI * z);
        if (dir == 1) return OG_PI * sin(OG_PI * x) * cos(OG_PI * y) * sin(OG_PI * z);
        return OG_PI * sin(OG_PI * x) * sin(OG_PI * y) * cos(OG_PI * z);
    }
    return 0;
}

long og_face_size(const og_grid* g, int f)
{
    const int d = f / 2;
    const int a = d == 0 ? 1 : 0, b = d == 2 ? 1 : 2;
    return (long)g->ng[a] * g->ng[b];
}

/* position of point (i,j,k) inside the face plane of direction d */
static inline long og_face_pos(const og_grid* g, int d, int i, int j, int k)
{
    if (d == 0) return j + (long)g->ng[1] * k;
    if (d == 1) return i + (long)g->ng[0] * k;
    return i + (long)g->ng[0] * j;
}

/* iterativeSolverBase.hpp:51-55 (setProblem), :537-555 (b), :557-603 (Dirichlet x) */
void og_set_problem(const og_grid* g, int kind, double* x, double* b, double* const flux[6])
{
    const int* L = g->lim_data;
    for (int d = 0; d < 3; ++d)
        for (int side = 0; side < 2; ++side) {
            const int f = 2 * d + side;
            int lim[6];
            memcpy(lim, L, sizeof lim);
            if (side == 0) lim[2 * d + 1] = lim[2 * d] + 1;
            else lim[2 * d] = lim[2 * d + 1] - 1;
            for (int k = lim[4]; k < lim[5]; ++k)
                for (int j = lim[2]; j < lim[3]; ++j)
                    for (int i = lim[0]; i < lim[1]; ++i) {
                        const double X = og_coord(g, 0, i), Y = og_coord(g, 1, j), Z = og_coord(g, 2, k);
                        if (g->bc[f] == 0 && x) x[og_idx(g, i, j, k)] = prob_u(kind, X, Y, Z);
                        if (g->bc[f] == 1 && flux && flux[f]) flux[f][og_face_pos(g, d, i, j, k)] = prob_du(kind, X, Y, Z, d);
                    }
        }
    if (b)
        for (int k = L[4]; k < L[5]; ++k)
            for (int j = L[2]; j < L[3]; ++j)
                for (int i = L[0]; i < L[1]; ++i)
                    b[og_idx(g, i, j, k)] = prob_f(kind, og_coord(g, 0, i), og_coord(g, 1, j), og_coord(g, 2, k));
}

/* iterativeSolverBase.hpp:62-169, orderNeumanBcs == 2 */
void og_reset_neumann(const og_grid* g, double* field, const double* const flux[6], double norm)
{
    for (int d = 0; d < 3; ++d) {
        const long e = d == 0 ? 1 : (d == 1 ? g->sj : g->sk);
        for (int side = 0; side < 2; ++side) {
            const int f = 2 * d + side;
            if (g->bc[f] != 1) continue;
            int lim[6];
            memcpy(lim, g->lim_data, sizeof lim);
            if (side == 0) lim[2 * d + 1] = lim[2 * d] + 1;
            else lim[2 * d] = lim[2 * d + 1] - 1;
            for (int k = lim[4]; k < lim[5]; ++k)
                for (int j = lim[2]; j < lim[3]; ++j)
                    for (int i = lim[0]; i < lim[1]; ++i) {
                        const long p = og_idx(g, i, j, k);
                        const long guard = side == 0 ? p - e : p + e;
                        const long data = side == 0 ? p + e : p - e;
                        if (flux) {
                            const double q = flux[f][og_face_pos(g, d, i, j, k)];
                            if (side == 0) field[guard] = field[data] - 2 * g->ds[d] * q / norm;
                            else field[guard] = field[data] + 2 * g->ds[d] * q / norm;
                        } else {
                            field[guard] = field[data];
                        }
                    }
        }
    }
}

static void og_reset_neumann_f(const og_grid* g, float* field)
{
    for (int d = 0; d < 3; ++d) {
        const long e = d == 0 ? 1 : (d == 1 ? g->sj : g->sk);
        for (int side = 0; side < 2; ++side) {
            if (g->bc[2 * d + side] != 1) continue;
            int lim[6];
            memcpy(lim, g->lim_data, sizeof lim);
            if (side == 0) lim[2 * d + 1] = lim[2 * d] + 1;
            else lim[2 * d] = lim[2 * d + 1] - 1;
            for (int k = lim[4]; k < lim[5]; ++k)
                for (int j = lim[2]; j < lim[3]; ++j)
                    for (int i = lim[0]; i < lim[1]; ++i) {
                        const long p = og_idx(g, i, j, k);
                        field[side == 0 ? p - e : p + e] = field[side == 0 ? p + e : p - e];
                    }
        }
    }
}

/* iterativeSolverBase.hpp:429-534, orderNeumanBcs == 2 */
void og_adjust_b(const og_grid* g, const double* x, double* b, const double* const flux[6])
{
    for (int d = 0; d < 3; ++d) {
        const long e = d == 0 ? 1 : (d == 1 ? g->sj : g->sk);
        for (int side = 0; side < 2; ++side) {
            const int f = 2 * d + side;
            int lim[6];
            memcpy(lim, g->lim_data, sizeof lim);
            if (side == 0) lim[2 * d + 1] = lim[2 * d] + 1;
            else lim[2 * d] = lim[2 * d + 1] - 1;
            for (int k = lim[4]; k < lim[5]; ++k)
                for (int j = lim[2]; j < lim[3]; ++j)
                    for (int i = lim[0]; i < lim[1]; ++i) {
                        const long p = og_idx(g, i, j, k);
                        if (g->bc[f] == 0) {
                            b[side == 0 ? p + e : p - e] -= x[p] / (g->ds[d] * g->ds[d]);
                        } else if (g->bc[f] == 1) {
                            const double q = flux ? flux[f][og_face_pos(g, d, i, j, k)] : 0.0;
                            if (side == 0) b[p] += 2 * q / g->ds[d];
                            else b[p] -= 2 * q / g->ds[d];
                        }
                    }
        }
    }
}

/* r = b - A x on the solver region (iterativeSolverBase.hpp:257-268) */
double og_residual(const og_grid* g, const double* x, const double* b, double* r)
{
    double s = 0.0;
    OG_FOR_SOLVER(g, c)
    {
        r[c] = b[c] - og_stencil_point(g, x, c);
        s += r[c] * r[c];
    }
    return sqrt(s);
}

/* ------------------------------------------------------------- solvers */

typedef struct og_state {
    /* FP64 work vectors */
    double *p, *r, *Ap, *z, *r0, *Mp, *AMp, *Az, *Y, *Z, *W, *Btmp, *pold, *Apold, *xin, *bin;
    /* FP32 work vectors */
    float *fp, *fr, *fAp, *fz, *fY, *fZ, *fW, *fpold, *fApold, *fin, *fout;
    double norm; /* normFieldB_, starts at 1 (iterativeSolverBase.hpp:38) */
    int iters;
    long prec_iters;
    double err_it, err_op;
    double tol_override; /* > 0: tolerance of the next FP32 inner solve (set by og_ir) */
} og_state;

typedef struct og_run {
    const og_grid* g;
    const og_stage* chain;
    int nst;
    const double* const* flux;
    og_state* st;
    og_result* res;
} og_run;

static double* zalloc(const og_grid* g) { return (double*)calloc((size_t)g->ntot, sizeof(double)); }
static float* zallocf(const og_grid* g) { return (float*)calloc((size_t)g->ntot, sizeof(float)); }
#define OG_ENSURE(ptr, g) \
    do {                  \
        if (!(ptr)) (ptr) = zalloc(g); \
    } while (0)
#define OG_ENSUREF(ptr, g) \
    do {                   \
        if (!(ptr)) (ptr) = zallocf(g); \
    } while (0)

static void hist_set(og_run* R, int i, double v)
{
    og_result* res = R->res;
    if (!res || !res->history || i < 0 || i >= res->history_cap) return;
    res->history[i] = v;
    if (i + 1 > res->history_len) res->history_len = i + 1;
}

static void apply_stage(og_run* R, int s, double* out, double* in);
static void main_stage(og_run* R, int s, double* x, double* b);
static void apply_stage_f(og_run* R, int s, float* out, float* in);

/* normalizeProblemToFieldBNorm (iterativeSolverBase.hpp:171-234) */
static void og_normalize(og_run* R, og_state* S, int is_main, double* x, double* b)
{
    const og_grid* g = R->g;
    S->norm = 1;
    double sum = 0.0;
    if (is_main) {
        double* tmp = (double*)malloc(sizeof(double) * (size_t)g->ntot);
        memcpy(tmp, b, sizeof(double) * (size_t)g->ntot);
        og_adjust_b(g, x, tmp, R->flux);
        OG_FOR_SOLVER(g, c) sum += tmp[c] * tmp[c];
        free(tmp);
    } else {
        OG_FOR_SOLVER(g, c) sum += b[c] * b[c];
    }
    S->norm = sqrt(sum);
    for (long i = 0; i < g->ntot; ++i) {
        x[i] /= S->norm;
        b[i] /= S->norm;
    }
}

/* computeErrorOperatorA with its ghost refresh (iterativeSolverBase.hpp:236-280) */
static double og_error_operator(og_run* R, og_state* S, int is_main, double* x, const double* b, double* r)
{
    og_reset_neumann(R->g, x, is_main ? R->flux : NULL, S->norm);
    return og_residual(R->g, x, b, r);
}

static void og_unnormalize(og_run* R, og_state* S, double* x, double* b)
{
    for (long i = 0; i < R->g->ntot; ++i) {
        x[i] *= S->norm;
        b[i] *= S->norm;
    }
    S->norm = 1;
}

/* BaseCG::operator() (baseCG.hpp:44-260) */
static void og_cg(og_run* R, int s, int is_main, double* x, double* b)
{
    const og_grid* g = R->g;
    og_state* S = &R->st[s];
    const double tol = R->chain[s].tol;
    const int maxit = R->chain[s].maxit;
    OG_ENSURE(S->p, g); OG_ENSURE(S->r, g); OG_ENSURE(S->Ap, g); OG_ENSURE(S->z, g);
    memset(S->p, 0, sizeof(double) * g->ntot);
    memset(S->r, 0, sizeof(double) * g->ntot);
    memset(S->Ap, 0, sizeof(double) * g->ntot);
    memset(S->z, 0, sizeof(double) * g->ntot);
    double *p = S->p, *r = S->r, *Ap = S->Ap, *z = S->z;

    if (is_main) og_reset_neumann(g, x, R->flux, S->norm);
    else memset(x, 0, sizeof(double) * g->ntot);
    og_normalize(R, S, is_main, x, b);

    S->err_op = og_error_operator(R, S, is_main, x, b, r);
    if (is_main) hist_set(R, 0, S->err_op);
    if (S->err_op < tol) { /* early exit keeps x, b normalised (baseCG.hpp:103-107) */
        S->iters = 0;
        return;
    }
    apply_stage(R, s + 1, z, r);
    memcpy(p, z, sizeof(double) * g->ntot);

    int iter = 0;
    double alpha = 1, beta = 1;
    while (iter < maxit) {
        double s1 = 0.0, s2 = 0.0;
        OG_FOR_SOLVER(g, c)
        {
            Ap[c] = og_stencil_point(g, p, c);
            s1 += r[c] * z[c];
            s2 += p[c] * Ap[c];
        }
        alpha = s1 / s2;
        OG_FOR_SOLVER(g, c)
        {
            x[c] = x[c] + alpha * p[c];
            r[c] = r[c] - alpha * Ap[c];
        }
        apply_stage(R, s + 1, z, r);
        double s3 = 0.0, srr = 0.0;
        OG_FOR_SOLVER(g, c)
        {
            s3 += r[c] * z[c];
            srr += r[c] * r[c];
        }
        beta = s3 / s1;
        S->err_it = sqrt(srr);
        OG_FOR_SOLVER(g, c) p[c] = z[c] + beta * p[c];
        iter++;
        if (is_main) hist_set(R, iter, S->err_it);
        if (S->err_it < tol) break;
    }
    S->iters = iter;
    if (is_main) S->err_op = og_error_operator(R, S, is_main, x, b, r);
    og_unnormalize(R, S, x, b);
}

/* BiCGSTAB::operator() (BiCGSTAB.hpp:55-322) */
static void og_bicgstab(og_run* R, int s, int is_main, double* x, double* b)
{
    const og_grid* g = R->g;
    og_state* S = &R->st[s];
    const double tol = R->chain[s].tol;
    const int maxit = R->chain[s].maxit;
    const size_t nb = sizeof(double) * (size_t)g->ntot;
    OG_ENSURE(S->p, g); OG_ENSURE(S->r, g); OG_ENSURE(S->r0, g); OG_ENSURE(S->Mp, g);
    OG_ENSURE(S->AMp, g); OG_ENSURE(S->z, g); OG_ENSURE(S->Az, g);
    double *p = S->p, *r = S->r, *r0 = S->r0, *Mp = S->Mp, *AMp = S->AMp, *z = S->z, *Az = S->Az;
    memset(p, 0, nb); memset(r, 0, nb); memset(r0, 0, nb); memset(Mp, 0, nb);
    memset(AMp, 0, nb); memset(z, 0, nb); memset(Az, 0, nb);

    if (is_main) og_reset_neumann(g, x, R->flux, S->norm);
    else memset(x, 0, nb);
    og_normalize(R, S, is_main, x, b);
    S->err_op = og_error_operator(R, S, is_main, x, b, r);
    if (is_main) hist_set(R, 0, S->err_op);
    if (S->err_op < tol) {
        S->iters = 0;
        return;
    }
    memcpy(p, r, nb);
    memcpy(r0, r, nb);
    double rho0 = 1, rho1 = 1, alpha = 1, omega = 1, beta = 1;
    int iter = 0;
    while (iter < maxit) {
        apply_stage(R, s + 1, Mp, p);
        og_reset_neumann(g, Mp, NULL, 1.0);
        double s2 = 0.0;
        OG_FOR_SOLVER(g, c)
        {
            AMp[c] = og_stencil_point(g, Mp, c);
            s2 += r0[c] * AMp[c];
        }
        alpha = rho0 / s2;
        OG_FOR_SOLVER(g, c) r[c] = r[c] - alpha * AMp[c];
        apply_stage(R, s + 1, z, r);
        og_reset_neumann(g, z, NULL, 1.0);
        OG_FOR_SOLVER(g, c) Az[c] = og_stencil_point(g, z, c);
        double s1 = 0.0;
        s2 = 0.0;
        OG_FOR_SOLVER(g, c)
        {
            s1 += r[c] * Az[c];
            s2 += Az[c] * Az[c];
        }
        omega = s1 / s2;
        for (long i = 0; i < g->ntot; ++i) x[i] = x[i] + alpha * Mp[i] + omega * z[i];
        s1 = 0.0;
        s2 = 0.0;
        OG_FOR_SOLVER(g, c)
        {
            r[c] = r[c] - omega * Az[c];
            s1 += r0[c] * r[c];
            s2 += r[c] * r[c];
        }
        S->err_it = sqrt(s2);
        rho1 = s1;
        beta = rho1 / rho0 * alpha / omega;
        rho0 = rho1;
        OG_FOR_SOLVER(g, c) p[c] = r[c] + beta * (p[c] - omega * AMp[c]);
        iter++;
        if (is_main) hist_set(R, iter, S->err_it);
        if (S->err_it < tol) break;
    }
    S->iters = iter;
    og_reset_neumann(g, x, is_main ? R->flux : NULL, S->norm);
    if (is_main) S->err_op = og_error_operator(R, S, is_main, x, b, r);
    og_unnormalize(R, S, x, b);
}

/* Chebyshev coefficients: chebyshevIteration.hpp:22-26 (global, rescaled) or
 * the alpaka local form chebyshevIterationAlpaka.hpp:33-36 */
static void og_cheb_coefs(const og_run* R, int s, double* theta, double* delta, double* sigma)
{
    const og_stage* st = &R->chain[s];
    const og_grid* g = R->g;
    if (st->cheb_local) {
        *theta = (g->eig_local[0] + g->eig_local[1]) * 0.5;
        *delta = (g->eig_local[0] - g->eig_local[1]) * 0.5;
    } else {
        *theta = (g->eig_global[0] * st->cheb_rmin + g->eig_global[1] * st->cheb_rmax) * 0.5 * (1.0 + st->cheb_eps);
        *delta = (g->eig_global[0] * st->cheb_rmin - g->eig_global[1] * st->cheb_rmax) * 0.5;
    }
    *sigma = *theta / *delta;
}

/* ChebyshevIteration::operator() (chebyshevIteration.hpp:48-140) */
static void og_cheb(og_run* R, int s, int is_main, double* X, double* Bin)
{
    const og_grid* g = R->g;
    og_state* S = &R->st[s];
    const int m = R->chain[s].maxit;
    double theta, delta, sigma;
    og_cheb_coefs(R, s, &theta, &delta, &sigma);
    OG_ENSURE(S->Y, g); OG_ENSURE(S->Z, g); OG_ENSURE(S->W, g);
    double rhoOld = 1 / sigma;
    double rhoCurr = 1 / (2 * sigma - rhoOld);
    double* B = Bin;
    if (is_main) {
        OG_ENSURE(S->Btmp, g);
        memcpy(S->Btmp, Bin, sizeof(double) * g->ntot);
        B = S->Btmp;
        og_adjust_b(g, X, B, R->flux);
    }
    og_reset_neumann(g, B, NULL, 1.0);
    OG_FOR_SOLVER(g, c)
    {
        S->Z[c] = B[c] / theta;
        S->Y[c] = 2 * rhoCurr / delta * (2 * B[c] + og_stencil_point(g, B, c) / theta);
    }
    for (int it = 2; it <= m; ++it) {
        rhoOld = rhoCurr;
        rhoCurr = 1 / (2 * sigma - rhoOld);
        og_reset_neumann(g, S->Y, NULL, 1.0);
        double *Y = S->Y, *Z = S->Z, *W = S->W;
        OG_FOR_SOLVER(g, c) W[c] = rhoCurr * (2 * sigma * Y[c] + 2 / delta * (B[c] + og_stencil_point(g, Y, c)) - rhoOld * Z[c]);
        double* t = S->Z; S->Z = S->Y; S->Y = t; /* swap(fieldZ, fieldY) */
        t = S->W; S->W = S->Y; S->Y = t;         /* swap(fieldW, fieldY) */
    }
    OG_FOR_SOLVER(g, c) X[c] = (-1) * S->W[c];
    if (is_main) {
        S->err_op = og_error_operator(R, S, 1, X, Bin, S->Btmp);
        S->err_it = S->err_op;
        S->iters = m;
    } else {
        S->prec_iters += m;
    }
}

/* Diagonal (Jacobi) preconditioner: out = in / diag(A) over the whole array.
 * No reference counterpart; for this constant-coefficient operator it is the
 * NoneSolver copy (noneSolver.hpp:24-27) scaled by -1/(2 sum 1/ds^2). */
static double og_jacobi_scale(const og_grid* g)
{
    return 1.0 / (-2.0 * (1.0 / (g->ds[0] * g->ds[0]) + 1.0 / (g->ds[1] * g->ds[1]) + 1.0 / (g->ds[2] * g->ds[2])));
}

/* Flexible CG with truncation 1 (Notay, SIAM J. Sci. Comput. 22 (2000)):
 * p_k = z_k - (z_k.Ap_{k-1} / p_{k-1}.Ap_{k-1}) p_{k-1}, alpha = p.r / p.Ap.
 * No reference counterpart.  Prologue/epilogue (normalisation, true residual,
 * history) follow BaseCG (baseCG.hpp:71-107, 230-259). */
static void og_fcg(og_run* R, int s, int is_main, double* x, double* b)
{
    const og_grid* g = R->g;
    og_state* S = &R->st[s];
    const double tol = R->chain[s].tol;
    const int maxit = R->chain[s].maxit;
    const size_t nb = sizeof(double) * (size_t)g->ntot;
    OG_ENSURE(S->p, g); OG_ENSURE(S->r, g); OG_ENSURE(S->Ap, g); OG_ENSURE(S->z, g);
    OG_ENSURE(S->pold, g); OG_ENSURE(S->Apold, g);
    memset(S->p, 0, nb); memset(S->r, 0, nb); memset(S->Ap, 0, nb); memset(S->z, 0, nb);
    memset(S->pold, 0, nb); memset(S->Apold, 0, nb);
    double* r = S->r;
    if (is_main) og_reset_neumann(g, x, R->flux, S->norm);
    else memset(x, 0, nb);
    og_normalize(R, S, is_main, x, b);
    S->err_op = og_error_operator(R, S, is_main, x, b, r);
    if (is_main) hist_set(R, 0, S->err_op);
    if (S->err_op < tol) {
        S->iters = 0;
        return;
    }
    double pApOld = 1.0;
    int iter = 0;
    while (iter < maxit) {
        double *p = S->p, *Ap = S->Ap, *z = S->z, *pold = S->pold, *Apold = S->Apold;
        apply_stage(R, s + 1, z, r);
        if (iter > 0) {
            double gam = 0.0;
            OG_FOR_SOLVER(g, c) gam += z[c] * Apold[c];
            const double beta = gam / pApOld;
            OG_FOR_SOLVER(g, c) p[c] = z[c] - beta * pold[c];
        } else {
            OG_FOR_SOLVER(g, c) p[c] = z[c];
        }
        double pAp = 0.0, pr = 0.0;
        OG_FOR_SOLVER(g, c)
        {
            Ap[c] = og_stencil_point(g, p, c);
            pAp += p[c] * Ap[c];
            pr += p[c] * r[c];
        }
        const double alpha = pr / pAp;
        double rr = 0.0;
        OG_FOR_SOLVER(g, c)
        {
            x[c] = x[c] + alpha * p[c];
            r[c] = r[c] - alpha * Ap[c];
            rr += r[c] * r[c];
        }
        S->err_it = sqrt(rr);
        pApOld = pAp;
        double* t = S->p; S->p = S->pold; S->pold = t;
        t = S->Ap; S->Ap = S->Apold; S->Apold = t;
        iter++;
        if (is_main) hist_set(R, iter, S->err_it);
        if (S->err_it < tol) break;
    }
    S->iters = iter;
    if (is_main) S->err_op = og_error_operator(R, S, is_main, x, b, r);
    og_unnormalize(R, S, x, b);
}

/* ----------------------------------------------- FP32 inner solvers (new) */

/* FP32 vectors, FP64 dot products and scalars.  Same structure as the
 * non-main BaseCG (baseCG.hpp:44-260 with isMainLoop false).  No reference
 * counterpart (the reference's FP32 hook is the Chebyshev cast only). */
static void og_cg_f(og_run* R, int s, float* x, float* b, int flexible)
{
    const og_grid* g = R->g;
    og_state* S = &R->st[s];
    const double tol = S->tol_override > 0.0 ? S->tol_override : R->chain[s].tol;
    const int maxit = R->chain[s].maxit;
    const size_t nb = sizeof(float) * (size_t)g->ntot;
    OG_ENSUREF(S->fp, g); OG_ENSUREF(S->fr, g); OG_ENSUREF(S->fAp, g); OG_ENSUREF(S->fz, g);
    OG_ENSUREF(S->fpold, g); OG_ENSUREF(S->fApold, g);
    memset(S->fp, 0, nb); memset(S->fr, 0, nb); memset(S->fAp, 0, nb); memset(S->fz, 0, nb);
    memset(S->fpold, 0, nb); memset(S->fApold, 0, nb);
    float* r = S->fr;
    memset(x, 0, nb);
    double sum = 0.0;
    OG_FOR_SOLVER(g, c) sum += (double)b[c] * (double)b[c];
    S->norm = sqrt(sum);
    const float inv = (float)(1.0 / S->norm);
    for (long i = 0; i < g->ntot; ++i) b[i] *= inv;
    /* x = 0: r = b on the solver region */
    double rr0 = 0.0;
    OG_FOR_SOLVER(g, c)
    {
        r[c] = b[c];
        rr0 += (double)r[c] * (double)r[c];
    }
    S->err_op = sqrt(rr0);
    int iter = 0;
    if (S->err_op >= tol) {
        double rz_old = 1.0, pApOld = 1.0;
        while (iter < maxit) {
            float *p = S->fp, *Ap = S->fAp, *z = S->fz, *pold = S->fpold, *Apold = S->fApold;
            apply_stage_f(R, s + 1, z, r);
            double rz = 0.0;
            if (!flexible) OG_FOR_SOLVER(g, c) rz += (double)r[c] * (double)z[c];
            if (iter == 0) {
                OG_FOR_SOLVER(g, c) p[c] = z[c];
            } else if (flexible) {
                double gam = 0.0;
                OG_FOR_SOLVER(g, c) gam += (double)z[c] * (double)Apold[c];
                const float beta = (float)(gam / pApOld);
                OG_FOR_SOLVER(g, c) p[c] = z[c] - beta * pold[c];
            } else {
                const float beta = (float)(rz / rz_old);
                OG_FOR_SOLVER(g, c) p[c] = z[c] + beta * pold[c];
            }
            double pAp = 0.0, pr = 0.0;
            OG_FOR_SOLVER(g, c)
            {
                Ap[c] = og_stencil_point_f(g, p, c);
                pAp += (double)p[c] * (double)Ap[c];
                if (flexible) pr += (double)p[c] * (double)r[c];
            }
            /* CG: alpha = r.z / p.Ap ; FCG: alpha = p.r / p.Ap */
            const float alpha = (float)((flexible ? pr : rz) / pAp);
            rz_old = rz;
            double rr = 0.0;
            OG_FOR_SOLVER(g, c)
            {
                x[c] = x[c] + alpha * p[c];
                r[c] = r[c] - alpha * Ap[c];
                rr += (double)r[c] * (double)r[c];
            }
            S->err_it = sqrt(rr);
            pApOld = pAp;
            float* t = S->fp; S->fp = S->fpold; S->fpold = t;
            t = S->fAp; S->fAp = S->fApold; S->fApold = t;
            iter++;
            if (S->err_it < tol) break;
        }
    }
    S->iters = iter;
    const float nf = (float)S->norm;
    for (long i = 0; i < g->ntot; ++i) {
        x[i] *= nf;
        b[i] *= nf;
    }
    S->norm = 1;
}

static void og_cheb_f(og_run* R, int s, float* X, float* B)
{
    const og_grid* g = R->g;
    og_state* S = &R->st[s];
    const int m = R->chain[s].maxit;
    double theta, delta, sigma;
    og_cheb_coefs(R, s, &theta, &delta, &sigma);
    OG_ENSUREF(S->fY, g); OG_ENSUREF(S->fZ, g); OG_ENSUREF(S->fW, g);
    double rhoOld = 1 / sigma, rhoCurr = 1 / (2 * sigma - rhoOld);
    og_reset_neumann_f(g, B);
    const float it = (float)(1 / theta);
    float c1 = (float)(2 * rhoCurr / delta);
    OG_FOR_SOLVER(g, c)
    {
        S->fZ[c] = B[c] * it;
        S->fY[c] = c1 * (2 * B[c] + og_stencil_point_f(g, B, c) * it);
    }
    for (int k = 2; k <= m; ++k) {
        rhoOld = rhoCurr;
        rhoCurr = 1 / (2 * sigma - rhoOld);
        og_reset_neumann_f(g, S->fY);
        const float a = (float)(rhoCurr * 2 * sigma), bb = (float)(rhoCurr * 2 / delta), cz = (float)(rhoCurr * rhoOld);
        float *Y = S->fY, *Z = S->fZ, *W = S->fW;
        OG_FOR_SOLVER(g, c) W[c] = a * Y[c] + bb * (B[c] + og_stencil_point_f(g, Y, c)) - cz * Z[c];
        float* t = S->fZ; S->fZ = S->fY; S->fY = t;
        t = S->fW; S->fW = S->fY; S->fY = t;
    }
    OG_FOR_SOLVER(g, c) X[c] = -S->fW[c];
    S->prec_iters += m;
}

/* The alpaka Chebyshev preconditioner with T_data_chebyshev = float under an
 * FP64 caller (chebyshevIterationAlpaka.hpp:119-196, 291-296): B's Neumann
 * ghosts reset in FP64 (:127), CastPrecisionFieldKernel over the whole box
 * (kernelsAlpakaChebyshev.hpp:7-18), Chebyshev1Kernel (:115-184) and m-1
 * sweeps of Chebyshev2Kernel (:191-273, jumpCheb = 0) in FP32 with the
 * coefficients and the rho recurrence evaluated in FP32 as the members and
 * kernel locals are (chebyshevIterationAlpaka.hpp:23-30,123-124,157-158,
 * 597-606), Y's ghosts reset in FP32 before each sweep, X = -1 * W on the
 * solver region (AssignFieldWith1FieldKernel, :23-44).  delta keeps the alpaka
 * sign (positive); ds[0] is x here where the alpaka Vec is z-y-x. */
static void og_cheb_alp32(og_run* R, int s, double* X, double* B64)
{
    const og_grid* g = R->g;
    og_state* S = &R->st[s];
    const og_stage* st = &R->chain[s];
    const int m = st->maxit;
    float theta, delta;
    if (st->cheb_local) {
        theta = (float)((g->eig_local[0] + g->eig_local[1]) * 0.5);
        delta = (float)((g->eig_local[1] - g->eig_local[0]) * 0.5);
    } else {
        theta = (float)((g->eig_global[0] * st->cheb_rmin + g->eig_global[1] * st->cheb_rmax) * 0.5 * (1.0 + st->cheb_eps));
        delta = (float)((g->eig_global[1] * st->cheb_rmax - g->eig_global[0] * st->cheb_rmin) * 0.5);
    }
    const float sigma = theta / delta;
    OG_ENSUREF(S->fin, g); OG_ENSUREF(S->fY, g); OG_ENSUREF(S->fZ, g); OG_ENSUREF(S->fW, g);
    og_reset_neumann(g, B64, NULL, 1.0);
    float* B = S->fin;
    for (long i = 0; i < g->ntot; ++i) B[i] = (float)B64[i];
    const long sj = g->sj, sk = g->sk;
    const float r0 = (float)(g->ds[0] * g->ds[0]), r1 = (float)(g->ds[1] * g->ds[1]), r2 = (float)(g->ds[2] * g->ds[2]);
    float rhoOld = 1 / sigma;
    float rhoCurr = 1 / (2 * sigma - rhoOld);
    {
        const float f0 = 2 * rhoCurr / delta * (1 / r0 / theta);
        const float f1 = 2 * rhoCurr / delta * (1 / r1 / theta);
        const float f2 = 2 * rhoCurr / delta * (1 / r2 / theta);
        const float fc0 = 2 * rhoCurr / delta * (2 - 2 * (1 / r0 + 1 / r1 + 1 / r2) / theta);
        OG_FOR_SOLVER(g, c)
        {
            S->fZ[c] = B[c] / theta;
            float tmp = B[c] * fc0 + (B[c - 1] + B[c + 1]) * f0;
            tmp += (B[c - sj] + B[c + sj]) * f1 + (B[c - sk] + B[c + sk]) * f2;
            S->fY[c] = tmp;
        }
    }
    for (int it = 2; it <= m; ++it) {
        rhoOld = rhoCurr;
        rhoCurr = 1 / (2 * sigma - rhoOld);
        og_reset_neumann_f(g, S->fY);
        const float f0 = rhoCurr * 2 / delta * (1 / r0);
        const float f1 = rhoCurr * 2 / delta * (1 / r1);
        const float f2 = rhoCurr * 2 / delta * (1 / r2);
        const float fB = rhoCurr * 2 / delta;
        const float fZ = -rhoCurr * rhoOld;
        const float fc0 = rhoCurr * (2 * sigma - 4 / delta * (1 / r0 + 1 / r1 + 1 / r2));
        const float *Y = S->fY, *Z = S->fZ;
        float* W = S->fW;
        OG_FOR_SOLVER(g, c)
        {
            float tmp = Y[c] * fc0 + (Y[c - 1] + Y[c + 1]) * f0;
            tmp += (Y[c - sj] + Y[c + sj]) * f1 + (Y[c - sk] + Y[c + sk]) * f2;
            tmp += B[c] * fB;
            tmp += Z[c] * fZ;
            W[c] = tmp;
        }
        float* t = S->fZ; S->fZ = S->fY; S->fY = t; /* swap(bufZ, bufY) */
        t = S->fW; S->fW = S->fY; S->fY = t;        /* swap(bufW, bufY) */
    }
    const float minus1 = -1.0f;
    OG_FOR_SOLVER(g, c) X[c] = (double)(minus1 * S->fW[c]);
    S->prec_iters += m;
}

static void apply_stage_f(og_run* R, int s, float* out, float* in)
{
    const og_grid* g = R->g;
    const int kind = s < R->nst ? R->chain[s].kind : OG_NONE;
    switch (kind) {
    case OG_JACOBI: {
        const float sc = (float)og_jacobi_scale(g);
        for (long i = 0; i < g->ntot; ++i) out[i] = in[i] * sc;
        break;
    }
    case OG_CHEB: og_cheb_f(R, s, out, in); break;
    case OG_USER:
        memset(out, 0, sizeof(float) * g->ntot);
        R->chain[s].user_fn(R->chain[s].user, g, in, out);
        R->st[s].prec_iters += 1;
        break;
    default: memcpy(out, in, sizeof(float) * g->ntot); break;
    }
}

/* Iterative refinement: FP64 outer residual r = b - A x and update x += d,
 * FP32 inner solve A d = r (chain[s+1], CG or FCG).  No reference counterpart;
 * prologue/epilogue follow BaseCG's main loop. */
static void og_ir(og_run* R, int s, double* x, double* b)
{
    const og_grid* g = R->g;
    og_state* S = &R->st[s];
    const double tol = R->chain[s].tol;
    const int maxouter = R->chain[s].maxit;
    OG_ENSURE(S->r, g);
    memset(S->r, 0, sizeof(double) * g->ntot);
    og_state* SI = &R->st[s + 1];
    OG_ENSUREF(SI->fin, g); OG_ENSUREF(SI->fout, g);
    og_reset_neumann(g, x, R->flux, S->norm);
    og_normalize(R, S, 1, x, b);
    S->err_op = og_error_operator(R, S, 1, x, b, S->r);
    hist_set(R, 0, S->err_op);
    if (S->err_op < tol) {
        S->iters = 0;
        return;
    }
    int outer = 0;
    long inner_total = 0;
    double err = S->err_op;
    while (outer < maxouter && err >= tol) {
        memset(SI->fin, 0, sizeof(float) * g->ntot);
        OG_FOR_SOLVER(g, c) SI->fin[c] = (float)S->r[c];
        {
            /* the inner tolerance: the first step's own (ir_first_tol), else the
             * inner stage's; with ir_adapt, no less than the reduction the outer
             * tolerance still needs, times ir_adapt */
            const double base = (outer == 0 && R->chain[s].ir_first_tol > 0.0) ? R->chain[s].ir_first_tol
                                                                                : R->chain[s + 1].tol;
            const double adapt = R->chain[s].ir_adapt;
            SI->tol_override = (adapt > 0.0 && err > 0.0) ? fmax(base, adapt * tol / err) : base;
        }
        og_cg_f(R, s + 1, SI->fout, SI->fin, R->chain[s + 1].kind == OG_FCG);
        inner_total += SI->iters;
        OG_FOR_SOLVER(g, c) x[c] += (double)SI->fout[c];
        err = og_error_operator(R, S, 1, x, b, S->r);
        outer++;
        hist_set(R, outer, err);
    }
    S->iters = (int)inner_total;
    S->err_it = err;
    S->err_op = err;
    if (R->res) R->res->outer_iterations = outer;
    og_unnormalize(R, S, x, b);
}

/* preconditioner call: operator()(out, in) of stage s */
static void apply_stage(og_run* R, int s, double* out, double* in)
{
    const og_grid* g = R->g;
    const int kind = s < R->nst ? R->chain[s].kind : OG_NONE;
    og_state* S = &R->st[s];
    switch (kind) {
    case OG_NONE: memcpy(out, in, sizeof(double) * g->ntot); break;
    case OG_JACOBI: {
        const double sc = og_jacobi_scale(g);
        for (long i = 0; i < g->ntot; ++i) out[i] = in[i] * sc;
        break;
    }
    case OG_CG: og_cg(R, s, 0, out, in); S->prec_iters += S->iters; break;
    case OG_BICGSTAB: og_bicgstab(R, s, 0, out, in); S->prec_iters += S->iters; break;
    case OG_CHEB:
        if (R->chain[s].precision == 32) og_cheb_alp32(R, s, out, in);
        else og_cheb(R, s, 0, out, in);
        break;
    case OG_FCG: og_fcg(R, s, 0, out, in); S->prec_iters += S->iters; break;
    case OG_USER: {
        OG_ENSUREF(S->fin, g); OG_ENSUREF(S->fout, g);
        for (long i = 0; i < g->ntot; ++i) S->fin[i] = (float)in[i];
        memset(S->fout, 0, sizeof(float) * g->ntot);
        R->chain[s].user_fn(R->chain[s].user, g, S->fin, S->fout);
        memset(out, 0, sizeof(double) * g->ntot);
        OG_FOR_SOLVER(g, c) out[c] = (double)S->fout[c];
        S->prec_iters += 1;
        break;
    }
    default: memcpy(out, in, sizeof(double) * g->ntot); break;
    }
}

static void main_stage(og_run* R, int s, double* x, double* b)
{
    switch (R->chain[s].kind) {
    case OG_CG: og_cg(R, s, 1, x, b); break;
    case OG_BICGSTAB: og_bicgstab(R, s, 1, x, b); break;
    case OG_CHEB: og_cheb(R, s, 1, x, b); break;
    case OG_FCG: og_fcg(R, s, 1, x, b); break;
    case OG_IR: og_ir(R, s, x, b); break;
    default: break;
    }
}

static void free_state(og_state* S)
{
    double* d[] = {S->p, S->r, S->Ap, S->z, S->r0, S->Mp, S->AMp, S->Az, S->Y, S->Z, S->W, S->Btmp, S->pold, S->Apold, S->xin, S->bin};
    for (size_t i = 0; i < sizeof d / sizeof d[0]; ++i) free(d[i]);
    float* f[] = {S->fp, S->fr, S->fAp, S->fz, S->fY, S->fZ, S->fW, S->fpold, S->fApold, S->fin, S->fout};
    for (size_t i = 0; i < sizeof f / sizeof f[0]; ++i) free(f[i]);
}

int og_solve(const og_grid* g, const og_stage* chain, int nstages, double* x, double* b,
             const double* const flux[6], og_result* res)
{
    if (nstages < 1 || nstages > 8) return -1;
    if (chain[0].kind == OG_IR && (nstages < 2 || (chain[1].kind != OG_CG && chain[1].kind != OG_FCG))) return -2;
    og_state st[9];
    memset(st, 0, sizeof st);
    for (int s = 0; s < 9; ++s) st[s].norm = 1.0;
    og_run R = {g, chain, nstages, flux, st, res};
    if (res) res->history_len = 0;
    main_stage(&R, 0, x, b);
    if (res) {
        res->iterations = st[0].iters;
        res->error_from_iteration = st[0].err_it;
        res->error_compute_operator = st[0].err_op;
        long pi = 0;
        for (int s = 1; s < nstages; ++s) pi += st[s].prec_iters;
        res->precond_iterations = pi;
    }
    for (int s = 0; s < 9; ++s) free_state(&st[s]);
    return 0;
}

/* One preconditioner application out = M(in) of chain[0] with chain[1..]
 * nested: the solvers' preconditioner protocol operator()(out, in)
 * (BiCGstabAlpaka.hpp:213-214; BiCGSTAB.hpp:170,190).  `in` has its Neumann
 * ghosts reset in place, as the reference's preconditioners do. */
int og_precondition(const og_grid* g, const og_stage* chain, int nstages, double* out, double* in, og_result* res)
{
    if (nstages < 1 || nstages > 8) return -1;
    if (chain[0].kind == OG_IR) return -2;
    og_state st[9];
    memset(st, 0, sizeof st);
    for (int s = 0; s < 9; ++s) st[s].norm = 1.0;
    og_run R = {g, chain, nstages, NULL, st, res};
    if (res) res->history_len = 0;
    apply_stage(&R, 0, out, in);
    if (res) {
        res->iterations = st[0].iters;
        long pi = 0;
        for (int s = 0; s < nstages; ++s) pi += st[s].prec_iters;
        res->precond_iterations = pi;
    }
    for (int s = 0; s < 9; ++s) free_state(&st[s]);
    return 0;
}

/* ------------------------------------------------------------- RHS recipes */

void og_rhs_normal(const og_grid* g, uint64_t seed, double* b)
{
    double sum = 0.0;
    long cnt = 0;
    const int* L = g->lim_data;
    OG_FOR_SOLVER(g, c)
    {
        const long k = c / g->sk, j = (c % g->sk) / g->sj, i = c % g->sj;
        const uint64_t idx = (uint64_t)(i - L[0]) + (uint64_t)g->np[0] * ((uint64_t)(j - L[2]) + (uint64_t)g->np[1] * (uint64_t)(k - L[4]));
        const double f = pps_rng_normal(seed, PPS_RHS_STREAM_POINT, idx);
        b[c] = -f;
        sum += f;
        cnt++;
    }
    const double mean = sum / (double)cnt;
    OG_FOR_SOLVER(g, c) b[c] += mean;
}

void og_rhs_gauss(const og_grid* g, uint64_t seed, int nsrc, double wmin, double wmax, double* b)
{
    double *cx = malloc(sizeof(double) * nsrc), *cy = malloc(sizeof(double) * nsrc), *cz = malloc(sizeof(double) * nsrc);
    double *w = malloc(sizeof(double) * nsrc), *a = malloc(sizeof(double) * nsrc);
    for (int s = 0; s < nsrc; ++s) {
        double c[3];
        pps_rhs_gauss_source(seed, s, wmin, wmax, c, &w[s], &a[s]);
        cx[s] = c[0]; cy[s] = c[1]; cz[s] = c[2];
    }
    OG_FOR_SOLVER(g, c)
    {
        const int k = (int)(c / g->sk), j = (int)((c % g->sk) / g->sj), i = (int)(c % g->sj);
        b[c] = -pps_rhs_gauss_eval(nsrc, cx, cy, cz, w, a, og_coord(g, 0, i), og_coord(g, 1, j), og_coord(g, 2, k));
    }
    free(cx); free(cy); free(cz); free(w); free(a);
}

/* separable evaluation of sum_k c_k sin(kx pi gi/(np-1)) sin(...) sin(...) */
void og_rhs_spectrum(const og_grid* g, uint64_t seed, int K, double* b)
{
    const int* L = g->lim_data;
    const int n0 = g->np[0], n1 = g->np[1], n2 = g->np[2];
    double* C = malloc(sizeof(double) * (size_t)K * K * K);
    for (int kz = 1; kz <= K; ++kz)
        for (int ky = 1; ky <= K; ++ky)
            for (int kx = 1; kx <= K; ++kx) C[(kx - 1) + K * ((ky - 1) + K * (kz - 1))] = pps_rhs_spec_coef(seed, K, kx, ky, kz);
    double* Sx = malloc(sizeof(double) * (size_t)K * n0);
    double* Sy = malloc(sizeof(double) * (size_t)K * n1);
    double* Sz = malloc(sizeof(double) * (size_t)K * n2);
    for (int k = 1; k <= K; ++k) {
        for (int i = 0; i < n0; ++i) Sx[(k - 1) + (size_t)K * i] = sin(k * OG_PI * (double)i / (n0 - 1));
        for (int i = 0; i < n1; ++i) Sy[(k - 1) + (size_t)K * i] = sin(k * OG_PI * (double)i / (n1 - 1));
        for (int i = 0; i < n2; ++i) Sz[(k - 1) + (size_t)K * i] = sin(k * OG_PI * (double)i / (n2 - 1));
    }
    /* T1[kx,ky,gk] = sum_kz C Sz ; T2[kx,gj,gk] = sum_ky T1 Sy ; f = sum_kx T2 Sx */
    double* T1 = calloc((size_t)K * K * n2, sizeof(double));
    for (int gk = 0; gk < n2; ++gk)
        for (int ky = 0; ky < K; ++ky)
            for (int kx = 0; kx < K; ++kx) {
                double acc = 0;
                for (int kz = 0; kz < K; ++kz) acc += C[kx + K * (ky + K * kz)] * Sz[kz + (size_t)K * gk];
                T1[kx + K * (ky + (size_t)K * gk)] = acc;
            }
    double* T2 = calloc((size_t)K * n1 * n2, sizeof(double));
    for (int gk = 0; gk < n2; ++gk)
        for (int gj = 0; gj < n1; ++gj)
            for (int kx = 0; kx < K; ++kx) {
                double acc = 0;
                for (int ky = 0; ky < K; ++ky) acc += T1[kx + K * (ky + (size_t)K * gk)] * Sy[ky + (size_t)K * gj];
                T2[kx + K * (gj + (size_t)n1 * gk)] = acc;
            }
    OG_FOR_SOLVER(g, c)
    {
        const int k = (int)(c / g->sk), j = (int)((c % g->sk) / g->sj), i = (int)(c % g->sj);
        const int gi = i - L[0], gj = j - L[2], gk = k - L[4];
        double acc = 0;
        for (int kx = 0; kx < K; ++kx) acc += T2[kx + K * (gj + (size_t)n1 * gk)] * Sx[kx + (size_t)K * gi];
        b[c] = -acc;
    }
    free(C); free(Sx); free(Sy); free(Sz); free(T1); free(T2);
}

/* The recipes' random parameters, exported for the tests that check their laws */
uint64_t og_mix64(uint64_t z) { return pps_mix64(z); }
double og_rng_normal(uint64_t seed, uint64_t stream, uint64_t ctr) { return pps_rng_normal(seed, stream, ctr); }
void og_rhs_gauss_source(uint64_t seed, int s, double wmin, double wmax, double c[3], double* w, double* a)
{
    pps_rhs_gauss_source(seed, s, wmin, wmax, c, w, a);
}
double og_rhs_spec_coef(uint64_t seed, int K, int kx, int ky, int kz) { return pps_rhs_spec_coef(seed, K, kx, ky, kz); }
