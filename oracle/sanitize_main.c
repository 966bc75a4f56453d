/*
 * sanitize_main.c -- runs the oracle's solver paths on small grids so that a
 * build with -fsanitize=address,undefined checks their memory accesses and
 * arithmetic (tests/test_oracle_sanitize.py).  Test infrastructure only.
 *
 * Cases: CG, BiCGSTAB with Chebyshev (FP64 and the FP32 alpaka form), nested
 * BiCGSTAB, Chebyshev as the main solver, FCG, FP32-inner refinement, one
 * preconditioner application; Dirichlet and mixed faces; ragged grids.
 */
#include "pps_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int run(const int np[3], const int bc[6], const og_stage* chain, int nst, int precondition_only)
{
    og_grid g;
    const double ds[3] = {1.0 / (np[0] - 1), 1.0 / (np[1] - 1), 1.0 / (np[2] - 1)};
    const double origin[3] = {0.5, 0.25, 1.0};
    if (og_grid_init(&g, np, ds, origin, bc) != 0) return 1;
    double* x = calloc((size_t)g.ntot, sizeof(double));
    double* b = calloc((size_t)g.ntot, sizeof(double));
    double* flux[6] = {0};
    for (int f = 0; f < 6; ++f) flux[f] = calloc((size_t)og_face_size(&g, f), sizeof(double));
    og_set_problem(&g, 0, x, b, flux);
    double hist[4096];
    og_result res;
    memset(&res, 0, sizeof res);
    res.history = hist;
    res.history_cap = 4096;
    int rc;
    if (precondition_only) {
        double* out = calloc((size_t)g.ntot, sizeof(double));
        rc = og_precondition(&g, chain, nst, out, b, &res);
        free(out);
    } else {
        rc = og_solve(&g, chain, nst, x, b, (const double* const*)flux, &res);
    }
    printf("np %d %d %d kind %d: rc %d iterations %d true residual %.3e\n", np[0], np[1], np[2], chain[0].kind, rc,
           res.iterations, res.error_compute_operator);
    for (int f = 0; f < 6; ++f) free(flux[f]);
    free(x);
    free(b);
    return rc != 0 || !(res.error_compute_operator == res.error_compute_operator);
}

int main(void)
{
    const int dir[6] = {0, 0, 0, 0, 0, 0}, mixed[6] = {0, 1, 1, 0, 1, 0};
    const int cube[3] = {14, 14, 14}, ragged[3] = {13, 9, 17};
    og_stage none = {0}, cg = {0}, bicg = {0}, cheb = {0}, cheb32 = {0}, fcg = {0}, ir = {0}, cg32 = {0}, inner = {0};
    none.kind = OG_NONE;
    cg.kind = OG_CG; cg.tol = 1e-10; cg.maxit = 2000;
    bicg.kind = OG_BICGSTAB; bicg.tol = 1e-10; bicg.maxit = 2000;
    cheb.kind = OG_CHEB; cheb.maxit = 11; cheb.cheb_eps = 1e-4; cheb.cheb_rmin = 500; cheb.cheb_rmax = 1 - 1e-4;
    cheb32 = cheb; cheb32.maxit = 24; cheb32.cheb_eps = 0; cheb32.cheb_rmin = 100; cheb32.precision = 32;
    fcg.kind = OG_FCG; fcg.tol = 1e-10; fcg.maxit = 2000;
    ir.kind = OG_IR; ir.tol = 1e-10; ir.maxit = 50; ir.ir_first_tol = 1e-2;
    cg32.kind = OG_CG; cg32.tol = 1e-5; cg32.maxit = 2000;
    inner.kind = OG_BICGSTAB; inner.tol = 1e-6; inner.maxit = 150;
    int bad = 0;
    {
        og_stage c[] = {cg, none};
        bad |= run(cube, dir, c, 2, 0);
        bad |= run(ragged, dir, c, 2, 0);
    }
    {
        og_stage c[] = {bicg, cheb, none};
        bad |= run(cube, mixed, c, 3, 0);
        bad |= run(ragged, mixed, c, 3, 0);
    }
    {
        og_stage c[] = {bicg, cheb32, none};
        bad |= run(ragged, mixed, c, 3, 0);
        bad |= run(ragged, mixed, c + 1, 2, 1);
    }
    {
        og_stage c[] = {bicg, inner, none};
        bad |= run(cube, mixed, c, 3, 0);
    }
    {
        og_stage c[] = {cheb, none};
        c[0].maxit = 24;
        c[0].cheb_eps = 0;
        c[0].cheb_rmin = 1;
        c[0].cheb_rmax = 1;
        bad |= run(ragged, dir, c, 2, 0);
    }
    {
        og_stage c[] = {fcg, none};
        bad |= run(ragged, dir, c, 2, 0);
    }
    {
        og_stage c[] = {ir, cg32, none};
        bad |= run(ragged, dir, c, 3, 0);
    }
    return bad;
}
