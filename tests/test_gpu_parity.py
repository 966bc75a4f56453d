"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars: FP64 solves match the oracle (itself bit-exact with the reference CPU
solver on these cases) to 1e-12 relative max-norm with iteration counts within
2%; kernel-level FP64 operations agree to a few ulps of the data scale; every
converging form reaches the FP64 true relative residual it was asked for.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.caseargs import data_region, decode, oracle_problem, relmax

pytestmark = pytest.mark.gpu

pps = pytest.importorskip("parallelpoissonsolver_amd")
L = pps._lib


def ctx_for(G, bcs):
    g = G.g
    return L.Context(list(g.np), list(g.ds), list(g.origin), bcs)


def upload_problem(C, G, x, b, flux, bcs):
    C.upload(L.FIELD_X, x)
    C.upload(L.FIELD_B, b)
    for f in range(6):
        if bcs[f] == 1:
            C.set_flux(f, flux[f])


def gpu_solve_case(case):
    d = decode(case)
    G, x, b, flux = oracle_problem(d)
    C = ctx_for(G, d["bc"])
    upload_problem(C, G, x, b, flux, d["bc"])
    res = C.solve(d["chain"])
    xg = C.download(L.FIELD_X)
    C.close()
    xo, bo = x.copy(), b.copy()
    ro = G.solve(d["chain"], xo, bo, flux)
    return G, xg, res, xo, ro


# ------------------------------------------------------------ kernel level

@pytest.mark.parametrize("bcs", [(0, 0, 0, 0, 0, 0), (0, 1, 1, 0, 1, 0), (1, 1, 1, 1, 1, 1)])
def test_apply_A_matches_oracle(bcs):
    n = (21, 18, 35)  # ragged, x not a multiple of the 64-wide tile, z not of kz
    ds = (0.07, 0.05, 0.03)
    G = O.OracleGrid(n, ds, (0.1, 0.2, 0.3), bcs)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(G.shape)
    C = ctx_for(G, bcs)
    C.upload(L.FIELD_T0, x)
    C.fill(L.FIELD_T1, 0.0)
    C.apply_A(L.FIELD_T1, L.FIELD_T0)
    yg = C.download(L.FIELD_T1)
    yo = G.apply_A(x)
    sl = G.solver_slice()
    scale = np.abs(yo[sl]).max()
    assert np.abs(yg[sl] - yo[sl]).max() / scale < 1e-14
    # nothing outside the solver region is written
    mask = np.ones(G.shape, bool)
    mask[sl] = False
    assert np.all(yg[mask] == 0)


def test_residual_dot_and_bnorm():
    n = (34, 34, 34)
    ds = (1 / 33,) * 3
    bcs = (0, 1, 1, 0, 1, 0)
    G = O.OracleGrid(n, ds, (3, 2.5, 10), bcs)
    x, b, flux = G.set_problem(0)
    C = ctx_for(G, bcs)
    upload_problem(C, G, x, b, flux, bcs)
    # ||b_adj|| (iterativeSolverBase.hpp:171-201)
    bt = b.copy()
    G.adjust_b(x, bt, flux)
    sl = G.solver_slice()
    nb_o = float(np.sqrt((bt[sl] ** 2).sum()))
    assert abs(C.norm_b_adjusted() - nb_o) / nb_o < 1e-14
    # r = b - A x and ||r||
    rn_g = C.residual(L.FIELD_T0, L.FIELD_X, L.FIELD_B)
    rn_o, r_o = G.residual(x, b)
    assert abs(rn_g - rn_o) / rn_o < 1e-13
    rg = C.download(L.FIELD_T0)
    assert np.abs(rg[sl] - r_o[sl]).max() / np.abs(r_o[sl]).max() < 1e-13
    # deterministic dot: two calls give identical bits
    d1, d2 = C.dot(L.FIELD_T0, L.FIELD_B), C.dot(L.FIELD_T0, L.FIELD_B)
    assert d1 == d2
    assert abs(d1 - float((rg[sl] * b[sl]).sum())) / abs(d1) < 1e-13


@pytest.mark.parametrize("with_flux,norm", [(False, 1.0), (True, 1.0), (True, 3.7)])
def test_ghost_fill_matches_oracle(with_flux, norm):
    n = (18, 20, 22)
    ds = (0.1, 0.1, 0.1)
    bcs = (0, 1, 1, 0, 1, 0)
    G = O.OracleGrid(n, ds, (3, 2.5, 10), bcs)
    x, b, flux = G.set_problem(0)
    rng = np.random.default_rng(3)
    f = rng.standard_normal(G.shape)
    C = ctx_for(G, bcs)
    for q in range(6):
        if bcs[q] == 1:
            C.set_flux(q, flux[q])
    C.upload(L.FIELD_T0, f)
    C.ghost_fill(L.FIELD_T0, with_flux, norm)
    fg = C.download(L.FIELD_T0)
    fo = f.copy()
    G.reset_neumann(fo, flux if with_flux else None, norm)
    assert np.array_equal(fg, fo) or np.abs(fg - fo).max() < 1e-14 * np.abs(fo).max()


@pytest.mark.parametrize("kind", ["normal", "gauss", "spectrum"])
def test_rhs_generators_match_oracle(kind):
    n = (34, 30, 26)
    ds = tuple(1.0 / (v - 1) for v in n)
    G = O.OracleGrid(n, ds, (0, 0, 0), (0,) * 6)
    C = ctx_for(G, (0,) * 6)
    C.fill(L.FIELD_B, 0.0)
    if kind == "normal":
        C.rhs_normal(2024)
        bo = G.rhs("normal", 2024)
    elif kind == "gauss":
        C.rhs_gauss(3, 64, 0.01, 0.1)
        bo = G.rhs("gauss", 3)
    else:
        C.rhs_spectrum(9, 12)
        bo = G.rhs("spectrum", 9, K=12)
    bg = C.download(L.FIELD_B)
    assert relmax(bg, bo) < 1e-13
    if kind == "normal":
        sl = G.solver_slice()
        assert abs(bg[sl].mean()) < 1e-14


# ----------------------------------------------------------------- solves

SOLVE_CASES = ["cg_none_dir_p0_18", "bicgstab_none_dir_p0_18", "cg_cheb11_dir_p0_18", "bicgstab_cheb24_alp_18",
               "bicgstab_cheb11_cpu_18", "bicgstab_cheb4_alp_18", "cg_none_alp_18", "bicgstab_none_alp_18",
               "bicgstab_bicgstab_alp_18", "cg_cg_dir_18", "cheb24_main_dir_18", "cg_none_sine_18",
               "cg_none_normal_18", "bicgstab_none_normal_18", "cg_none_gauss_18", "cg_none_spectrum_18",
               "cg_none_sine_66", "cg_none_p0_unit_66", "cg_none_normal_66", "bicgstab_none_normal_66",
               "cg_none_gauss_66", "bicgstab_cheb24_alp_64", "cg_cheb11_dir_34", "cg_none_p0_34"]


@pytest.mark.parametrize("name", SOLVE_CASES)
def test_solve_matches_oracle(golden, name):
    cases, _ = golden
    c = cases[name]
    G, xg, res, xo, ro = gpu_solve_case(c)
    it_g, it_o = res["iterations"], ro["iterations"]
    assert it_o == c["iterations"]  # the oracle is the reference here
    sl = G.solver_slice()
    diff = relmax(xg[sl], xo[sl])
    info = (it_g, it_o, diff, res["error_compute_operator"], ro["error_compute_operator"])
    print(f"[solve] {name}: GPU {it_g} it, oracle {it_o} it, max rel dx {diff:.3e}, true residual "
          f"{res['error_compute_operator']:.3e} / {ro['error_compute_operator']:.3e}", flush=True)
    if ro["error_compute_operator"] < 1e-10:
        assert res["error_compute_operator"] < 1e-10, info
    else:
        # cases whose true residual stays large in the reference too (CG with
        # Neumann faces never refreshes p's ghosts; main-loop Chebyshev is a
        # fixed-degree smoother): same value as the reference
        assert abs(res["error_compute_operator"] - ro["error_compute_operator"]) <= 1e-8 * ro["error_compute_operator"]
    if c["args"][c["args"].index("--solver") + 1] == "bicgstab":
        # BiCGSTAB amplifies reduction-order rounding (the reference's own 1-rank
        # and 4-rank runs of one problem take 103 and 108 iterations,
        # bicgstab_none_normal_34 / _r4), so the two solves stop at different
        # iterates, and a change of the GPU's summation order alone moves a
        # count by one or two (bicgstab_none_normal_18: 54, then 53, against 55).
        # Bar: iterations within the reference's own 5 % (at least 2), and the
        # solutions equal to the accuracy both solves reached:
        # ||dx||/||x|| <= kappa (r1 + r2), the perturbation bound of two
        # relative residuals r1, r2 (measured: up to 3.5e-9 at 66^3, where the
        # bound is 3.2e-7; cases that reach 1e-12 agree to 1e-12 and below).
        assert abs(it_g - it_o) <= max(2, round(0.05 * it_o)), info
        e = G.g.eig_global
        kappa = e[1] / e[0]
        assert diff <= kappa * (res["error_compute_operator"] + ro["error_compute_operator"]), info
    else:
        assert abs(it_g - it_o) <= max(0, round(0.02 * it_o)), info
        assert diff < 1e-12, info
        if len(ro["history"]) and len(res["history"]):
            # residual histories follow the reference: the first ten entries to
            # 1e-9, the rest to 2 %.  A right-hand side made of a few spectral
            # components (the k <= 8 spectrum at 18^3) is the exception: once CG
            # has resolved them its residual is a few components of rounding
            # origin, and the two summation orders steer those apart by more
            # before the solutions meet again at the tolerance (equal counts,
            # solutions within 1e-12 as asserted above): 10 % there.
            h = min(len(ro["history"]), len(res["history"]))
            a, b = res["history"][:h], ro["history"][:h]
            dev = np.max((np.abs(a - b) - 1e-12) / np.maximum(b, 1e-300))  # allclose's rtol with atol 1e-12
            print(f"[solve] {name}: largest relative history deviation {dev:.3e}", flush=True)
            assert dev <= (0.1 if "spectrum" in name else 2e-2), info
            assert np.allclose(res["history"][:min(h, 10)], ro["history"][:min(h, 10)], rtol=1e-9), info


def test_config1_manufactured_sine_one_iteration(golden):
    """64^3 interior, b = -3 pi^2 sin sin sin: b is one discrete eigenvector, so CG
    converges in one iteration to u * (pi h/2)^2 / sin^2(pi h/2)."""
    cases, _ = golden
    c = cases["cg_none_sine_66"]
    G, xg, res, xo, ro = gpu_solve_case(c)
    assert res["iterations"] == 1 == c["iterations"]
    h = 1.0 / 65
    ax = [np.arange(68) - 1] * 3
    Z, Y, X = np.meshgrid(ax[2] * h, ax[1] * h, ax[0] * h, indexing="ij")
    u = np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z)
    uh = u * (np.pi * h / 2) ** 2 / np.sin(np.pi * h / 2) ** 2
    sl = G.solver_slice()
    assert np.abs(xg[sl] - uh[sl]).max() < 1e-12


def test_jacobi_equals_unpreconditioned():
    """Jacobi for the constant-coefficient operator is a scaled identity: CG and
    BiCGSTAB with Jacobi give the unpreconditioned iterates up to rounding."""
    n = 34
    G = O.OracleGrid((n,) * 3, (1 / (n - 1),) * 3, (0, 0, 0), (0,) * 6)
    b = G.rhs("normal", 7)
    x = G.zeros()
    for kind in (L.CG, L.BICGSTAB):
        out = []
        for pre in (L.NONE, L.JACOBI):
            C = ctx_for(G, (0,) * 6)
            C.upload(L.FIELD_X, x)
            C.upload(L.FIELD_B, b)
            res = C.solve([dict(kind=kind, tol=1e-10, maxit=10000), dict(kind=pre)])
            out.append((res, C.download(L.FIELD_X)))
            C.close()
        (r0, x0), (r1, x1) = out
        tolr = 0.02 if kind == L.CG else 0.05  # BiCGSTAB: reduction-order sensitivity
        assert abs(r0["iterations"] - r1["iterations"]) <= max(1, round(tolr * r0["iterations"]))
        assert r1["error_compute_operator"] < 1e-10
        e = G.g.eig_global
        bound = 1e-9 if kind == L.CG else 10 * e[1] / e[0] * (r0["error_compute_operator"] + r1["error_compute_operator"])
        assert relmax(x1, x0) < bound


@pytest.mark.parametrize("kind", [L.CG, L.BICGSTAB, L.FCG])
@pytest.mark.parametrize("tol", [1e-10, 1e-14])
def test_true_residual_finish(kind, tol):
    """pps_stage.true_finish: without it the solve stops on the recursive
    residual (the reference's rule); with it the true FP64 residual of the
    returned x, recomputed by the oracle, is below tol.  When the recursive
    stop already meets tol on the true residual nothing changes, bit for bit.
    At tol = 1e-14 the recursive residual has drifted below the true one (the
    oracle's CG and FCG stop at a true 1.05e-14 at 66^3), so the finish has to
    replace it and resume."""
    n = 66
    G = O.OracleGrid((n,) * 3, (1 / (n - 1),) * 3, (0, 0, 0), (0,) * 6)
    b = G.rhs("normal", 13)
    x = G.zeros()
    out = {}
    for fin in (0, 10):
        C = ctx_for(G, (0,) * 6)
        C.upload(L.FIELD_X, x)
        C.upload(L.FIELD_B, b)
        res = C.solve([dict(kind=kind, tol=tol, maxit=20000, true_finish=fin), dict(kind=L.NONE)])
        out[fin] = (res, C.download(L.FIELD_X))
        C.close()
    (r0, x0), (r1, x1) = out[0], out[10]
    sl = G.solver_slice()
    bt = b.copy()
    G.adjust_b(x, bt, None)
    nb = np.sqrt((bt[sl] ** 2).sum())
    true0 = G.residual(x0, b)[0] / nb
    true1 = G.residual(x1, b)[0] / nb
    print(f"[finish] kind {kind} tol {tol:.0e}: recursive stop {r0['iterations']} it true {true0:.3e} "
          f"(lib {r0['error_compute_operator']:.3e}); finish {r1['iterations']} it true {true1:.3e} "
          f"(lib {r1['error_compute_operator']:.3e})", flush=True)
    assert abs(true1 - r1["error_compute_operator"]) <= 1e-3 * true1 + 1e-16
    assert true1 < tol
    if true0 < tol:
        assert r1["iterations"] == r0["iterations"] and np.array_equal(x1, x0)
    else:
        assert r1["iterations"] > r0["iterations"]


def test_fcg_with_fixed_preconditioner_equals_cg():
    n = 34
    G = O.OracleGrid((n,) * 3, (1 / (n - 1),) * 3, (0, 0, 0), (0,) * 6)
    b = G.rhs("gauss", 5)
    x = G.zeros()
    out = {}
    for kind in (L.CG, L.FCG):
        C = ctx_for(G, (0,) * 6)
        C.upload(L.FIELD_X, x)
        C.upload(L.FIELD_B, b)
        out[kind] = (C.solve([dict(kind=kind, tol=1e-10, maxit=10000), dict(kind=L.NONE)]), C.download(L.FIELD_X))
        C.close()
    (rc, xc), (rf, xf) = out[L.CG], out[L.FCG]
    assert abs(rc["iterations"] - rf["iterations"]) <= max(1, round(0.02 * rc["iterations"]))
    assert rf["error_compute_operator"] < 1e-10
    assert relmax(xf, xc) < 1e-9
    # and the oracle's FCG agrees
    xo, bo = x.copy(), b.copy()
    ro = G.solve([dict(kind=O.FCG, tol=1e-10), dict(kind=O.NONE)], xo, bo)
    assert abs(ro["iterations"] - rf["iterations"]) <= max(1, round(0.02 * ro["iterations"]))
    assert relmax(xf, xo) < 1e-9


@pytest.mark.parametrize("inner", [L.CG, L.FCG])
def test_mixed_precision_ir_reaches_fp64_tolerance(inner):
    n = 66
    G = O.OracleGrid((n,) * 3, (1 / (n - 1),) * 3, (0, 0, 0), (0,) * 6)
    b = G.rhs("gauss", 3)
    x = G.zeros()
    C = ctx_for(G, (0,) * 6)
    C.upload(L.FIELD_X, x)
    C.upload(L.FIELD_B, b)
    res = C.solve([dict(kind=L.IR, tol=1e-10, maxit=50), dict(kind=inner, precision=32, tol=1e-5, maxit=5000),
                   dict(kind=L.NONE, precision=32)])
    xg = C.download(L.FIELD_X)
    assert res["error_compute_operator"] < 1e-10
    assert res["outer_iterations"] >= 2
    # the oracle's FP64 CG solution of the same system
    xo, bo = x.copy(), b.copy()
    G.solve([dict(kind=O.CG, tol=1e-12), dict(kind=O.NONE)], xo, bo)
    assert relmax(xg, xo) < 1e-8
    # true residual recomputed by the oracle on the GPU solution
    sl = G.solver_slice()
    rn, _ = G.residual(xg, b)
    bt = b.copy()
    G.adjust_b(x, bt, None)
    assert rn / np.sqrt((bt[sl] ** 2).sum()) < 1e-10


def test_ir_inner_fp32_cg_matches_oracle_structure():
    """The FP32 inner CG of the refinement follows the oracle's FP32 CG: same
    outer step count, inner counts within 2%."""
    n = 34
    G = O.OracleGrid((n,) * 3, (1 / (n - 1),) * 3, (0, 0, 0), (0,) * 6)
    b = G.rhs("normal", 11)
    x = G.zeros()
    C = ctx_for(G, (0,) * 6)
    C.upload(L.FIELD_X, x)
    C.upload(L.FIELD_B, b)
    res = C.solve([dict(kind=L.IR, tol=1e-10, maxit=50), dict(kind=L.CG, precision=32, tol=1e-5, maxit=5000),
                   dict(kind=L.NONE, precision=32)])
    xo, bo = x.copy(), b.copy()
    ro = G.solve([dict(kind=O.IR, tol=1e-10, maxit=50), dict(kind=O.CG, tol=1e-5, maxit=5000), dict(kind=O.NONE)],
                 xo, bo)
    assert res["outer_iterations"] == ro["outer_iterations"]
    assert abs(res["iterations"] - ro["iterations"]) <= max(2, round(0.02 * ro["iterations"]))
    assert res["error_compute_operator"] < 1e-10 and ro["error_compute_operator"] < 1e-10


def test_ir_first_step_tolerance_matches_oracle():
    """ir_first_tol (a looser inner tolerance for the first outer step) on the
    GPU follows the oracle: same outer step count, inner counts within 2 %,
    both first outer residuals near the looser tolerance, and the FP64 true
    residual at the outer tolerance."""
    n = 34
    G = O.OracleGrid((n,) * 3, (1 / (n - 1),) * 3, (0, 0, 0), (0,) * 6)
    b = G.rhs("gauss", 3)
    x = G.zeros()
    C = ctx_for(G, (0,) * 6)
    C.upload(L.FIELD_X, x)
    C.upload(L.FIELD_B, b)
    res = C.solve([dict(kind=L.IR, tol=1e-10, maxit=50, first_tol=1e-2),
                   dict(kind=L.CG, precision=32, tol=1e-5, maxit=5000), dict(kind=L.NONE, precision=32)])
    xo, bo = x.copy(), b.copy()
    ro = G.solve([dict(kind=O.IR, tol=1e-10, maxit=50, first_tol=1e-2), dict(kind=O.CG, tol=1e-5, maxit=5000),
                  dict(kind=O.NONE)], xo, bo)
    print(f"[ir] GPU {res['iterations']} inner ({res['outer_iterations']} outer), history {list(res['history'])}; "
          f"oracle {ro['iterations']} ({ro['outer_iterations']}), {list(ro['history'])}", flush=True)
    assert res["outer_iterations"] == ro["outer_iterations"]
    assert abs(res["iterations"] - ro["iterations"]) <= max(2, round(0.02 * ro["iterations"]))
    assert 1e-3 < res["history"][1] < 2e-2 and 1e-3 < ro["history"][1] < 2e-2
    assert res["error_compute_operator"] < 1e-10 and ro["error_compute_operator"] < 1e-10


def test_ir_adaptive_inner_tolerance_matches_oracle():
    """ir_adapt (an outer step asks the inner solve for no more than
    adapt * tol / e) on the GPU follows the oracle: same outer step count,
    inner counts within 2 %, the same residual before the last step to 1 %,
    and the FP64 true residual at the outer tolerance."""
    n = 34
    G = O.OracleGrid((n,) * 3, (1 / (n - 1),) * 3, (0, 0, 0), (0,) * 6)
    b = G.rhs("gauss", 3)
    x = G.zeros()
    C = ctx_for(G, (0,) * 6)
    C.upload(L.FIELD_X, x)
    C.upload(L.FIELD_B, b)
    res = C.solve([dict(kind=L.IR, tol=1e-10, maxit=50, adapt=0.5),
                   dict(kind=L.CG, precision=32, tol=1e-4, maxit=5000), dict(kind=L.NONE, precision=32)])
    xo, bo = x.copy(), b.copy()
    ro = G.solve([dict(kind=O.IR, tol=1e-10, maxit=50, adapt=0.5), dict(kind=O.CG, tol=1e-4, maxit=5000),
                  dict(kind=O.NONE)], xo, bo)
    print(f"[ir adapt] GPU {res['iterations']} inner ({res['outer_iterations']} outer), history {list(res['history'])}; "
          f"oracle {ro['iterations']} ({ro['outer_iterations']}), {list(ro['history'])}", flush=True)
    assert res["outer_iterations"] == ro["outer_iterations"]
    assert abs(res["iterations"] - ro["iterations"]) <= max(2, round(0.02 * ro["iterations"]))
    k = ro["outer_iterations"] - 1
    assert abs(res["history"][k] - ro["history"][k]) <= 0.01 * ro["history"][k]
    assert res["error_compute_operator"] < 1e-10 and ro["error_compute_operator"] < 1e-10


ALP_CHEB = dict(cheb_eps=0.0, cheb_rmin=100.0, cheb_rmax=1 - 1e-4)  # solverSetup.hpp (alpaka):59-62


@pytest.mark.parametrize("n", [34, 50])
@pytest.mark.parametrize("m", [7, 24])
def test_ir_fp32_cg_with_fp32_chebyshev_matches_oracle(n, m):
    """Refinement with an FP32 inner CG preconditioned by FP32 Chebyshev(m)
    (the fused pairs in FP32 throughout, the last pair writing an FP32 field)
    against the oracle's refinement with its FP32 Chebyshev (og_cheb_f), on the
    reference test problem with Dirichlet faces (with the reference's Neumann
    CG, whose p ghosts are never refreshed, the inner CG does not converge in
    either): same outer step count, inner counts within 5 % (at least 2: FP32
    sweeps in another summation order), both at the FP64 true residual 1e-10."""
    bcs = (0,) * 6
    G = O.OracleGrid((n,) * 3, (1 / (n - 1),) * 3, (3, 2.5, 10), bcs)
    x, b, flux = G.set_problem(0)
    C = ctx_for(G, bcs)
    upload_problem(C, G, x, b, flux, bcs)
    res = C.solve([dict(kind=L.IR, tol=1e-10, maxit=50), dict(kind=L.CG, precision=32, tol=1e-5, maxit=5000),
                   dict(kind=L.CHEBYSHEV, precision=32, maxit=m)])
    C.close()
    xo, bo = x.copy(), b.copy()
    ro = G.solve([dict(kind=O.IR, tol=1e-10, maxit=50), dict(kind=O.CG, tol=1e-5, maxit=5000),
                  dict(kind=O.CHEB, maxit=m)], xo, bo, flux)
    print(f"[ir-cheb32] n {n} m {m}: GPU {res['iterations']} inner ({res['outer_iterations']} outer), "
          f"oracle {ro['iterations']} ({ro['outer_iterations']}); true residual "
          f"{res['error_compute_operator']:.2e} / {ro['error_compute_operator']:.2e}", flush=True)
    assert res["outer_iterations"] == ro["outer_iterations"]
    assert abs(res["iterations"] - ro["iterations"]) <= max(2, round(0.05 * ro["iterations"]))
    assert res["error_compute_operator"] < 1e-10 and ro["error_compute_operator"] < 1e-10


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("bcs", [(0, 0, 0, 0, 0, 0), (0, 1, 1, 0, 1, 0)])
@pytest.mark.parametrize("m", [4, 11, 24])
def test_chebyshev_preconditioner_matches_oracle(prec, bcs, m):
    """One Chebyshev preconditioner application (pps_precondition) against the
    oracle's (og_precondition).  precision 32 is the alpaka T_data_chebyshev =
    float form (chebyshevIterationAlpaka.hpp:149-196): FP64 input with its
    Neumann ghosts reset, cast to FP32, FP32 sweeps, widened output.  Bars:
    FP64 1e-12 relative max-norm; FP32 5e-6 (the oracle's FP32 form differs
    from its FP64 form by 1e-7 .. 6e-7 on these inputs)."""
    n = (34, 30, 26)
    ds = tuple(1.0 / (v - 1) for v in n)
    G = O.OracleGrid(n, ds, (3, 2.5, 10), bcs)
    inp = np.random.default_rng(5).standard_normal(G.shape)
    chain = [dict(kind=L.CHEBYSHEV, precision=prec, maxit=m, **ALP_CHEB)]
    C = ctx_for(G, bcs)
    C.upload(L.FIELD_T0, inp)
    C.fill(L.FIELD_T1, 0.0)
    its = C.precondition(chain, L.FIELD_T1, L.FIELD_T0)
    yg = C.download(L.FIELD_T1)
    ing = C.download(L.FIELD_T0)
    C.close()
    ino = inp.copy()
    yo, info = G.precondition([dict(kind=O.CHEB, precision=prec, maxit=m, **ALP_CHEB)], ino)
    sl = G.solver_slice()
    diff = relmax(yg[sl], yo[sl])
    print(f"[cheb-pre] prec {prec} bcs {bcs} m {m}: max rel diff {diff:.3e}", flush=True)
    assert its == info["precond_iterations"] == m
    assert diff < (1e-12 if prec == 64 else 5e-6), diff
    # the input's Neumann ghosts were reset in place, as the oracle's
    assert np.array_equal(ing, ino)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("bcs", [(0, 0, 0, 0, 0, 0), (1, 1, 1, 1, 1, 1), (1, 0, 0, 1, 1, 0), (0, 1, 1, 0, 0, 1)])
@pytest.mark.parametrize("m", [5, 7, 24])
def test_cheb_pairs_match_single_sweeps(prec, bcs, m):
    """The fused two-sweep Chebyshev kernel (k_cheb_pair, the default) against
    one sweep per launch, on a grid of partial tiles in x, y and z (131 x 35 x
    43 interior: 64-wide tiles, 16-row tiles, 16-plane chunks), every Neumann
    face set (the pair resets its intermediate's ghosts itself).  m = 5 is one
    pair ending in the output, 7 two pairs, 24 one single sweep then ten pairs.
    Both paths evaluate the same expressions, so the bar is rounding level:
    1e-14 (FP64), 1e-6 (FP32) relative max-norm."""
    n = (133, 37, 45)
    ds = tuple(1.0 / (v - 1) for v in n)
    G = O.OracleGrid(n, ds, (0, 0, 0), bcs)
    inp = np.random.default_rng(11).standard_normal(G.shape)
    chain = [dict(kind=L.CHEBYSHEV, precision=prec, maxit=m, **ALP_CHEB)]
    out = {}
    for pairs in (1, 0):
        C = ctx_for(G, bcs)
        C.set_cheb_pairs(pairs)
        C.upload(L.FIELD_T0, inp)
        C.fill(L.FIELD_T1, 0.0)
        its = C.precondition(chain, L.FIELD_T1, L.FIELD_T0)
        out[pairs] = C.download(L.FIELD_T1)
        C.close()
        assert its == m
    diff = relmax(out[1], out[0])
    print(f"[cheb-pairs] prec {prec} bcs {bcs} m {m}: max rel diff {diff:.3e}", flush=True)
    assert diff <= (1e-14 if prec == 64 else 1e-6), diff


@pytest.mark.parametrize("n", [(4, 4, 4), (5, 7, 6), (3, 66, 9)])
@pytest.mark.parametrize("bcs", [(1, 1, 1, 1, 1, 1), (0, 1, 1, 0, 1, 0)])
def test_cheb_pairs_on_tiny_grids(n, bcs):
    """Fused pairs against single sweeps where one tile holds both Neumann
    ghosts of a direction and the mirror points sit next to them: 2 to 5
    interior points per direction (plus one direction of two tiles), FP64,
    m = 24.  Bar: 1e-14 relative max-norm."""
    ds = tuple(1.0 / (v - 1) for v in n)
    G = O.OracleGrid(n, ds, (0, 0, 0), bcs)
    inp = np.random.default_rng(13).standard_normal(G.shape)
    chain = [dict(kind=L.CHEBYSHEV, maxit=24, **ALP_CHEB)]
    out = {}
    for pairs in (1, 0):
        C = ctx_for(G, bcs)
        C.set_cheb_pairs(pairs)
        C.upload(L.FIELD_T0, inp)
        C.fill(L.FIELD_T1, 0.0)
        assert C.precondition(chain, L.FIELD_T1, L.FIELD_T0) == 24
        out[pairs] = C.download(L.FIELD_T1)
        C.close()
    diff = relmax(out[1], out[0])
    print(f"[cheb-pairs-tiny] n {n} bcs {bcs}: max rel diff {diff:.3e}", flush=True)
    assert diff <= 1e-14, diff


@pytest.mark.parametrize("n,bcs", [(34, (0, 1, 1, 0, 1, 0)), (66, (0, 0, 0, 0, 0, 0))])
def test_bicgstab_fp32_chebyshev_solve_matches_oracle(n, bcs):
    """The reference's default GPU composition with T_data_chebyshev = float:
    BiCGSTAB (FP64) preconditioned by Chebyshev(24, global, epsilon 0,
    rescaleEigMin 100) in FP32, GPU against the oracle on the same problem.
    Bars: iterations within 2 % (at least 1), solutions within
    10 kappa (r_gpu + r_oracle) as for every BiCGSTAB case, both at 1e-10."""
    G = O.OracleGrid((n,) * 3, (1 / (n - 1),) * 3, (3, 2.5, 10), bcs)
    x, b, flux = G.set_problem(0)
    chain = [dict(kind=L.BICGSTAB, tol=1e-10, maxit=1000),
             dict(kind=L.CHEBYSHEV, precision=32, maxit=24, **ALP_CHEB)]
    C = ctx_for(G, bcs)
    upload_problem(C, G, x, b, flux, bcs)
    res = C.solve(chain)
    xg = C.download(L.FIELD_X)
    C.close()
    xo, bo = x.copy(), b.copy()
    ro = G.solve([dict(kind=O.BICGSTAB, tol=1e-10, maxit=1000),
                  dict(kind=O.CHEB, precision=32, maxit=24, **ALP_CHEB)], xo, bo, flux)
    sl = G.solver_slice()
    diff = relmax(xg[sl], xo[sl])
    it_g, it_o = res["iterations"], ro["iterations"]
    e = G.g.eig_global
    kappa = e[1] / e[0]
    print(f"[cheb32-solve] {n}^3 bcs {bcs}: GPU {it_g} it, oracle {it_o} it, max rel dx {diff:.3e}, "
          f"true residual {res['error_compute_operator']:.3e} / {ro['error_compute_operator']:.3e}", flush=True)
    assert res["error_compute_operator"] < 1e-10 and ro["error_compute_operator"] < 1e-10
    assert abs(it_g - it_o) <= max(1, round(0.02 * it_o)), (it_g, it_o)
    # two preconditioner applications of maxIteration each per iteration, as
    # the alpaka main loop counts them (BiCGstabAlpaka.hpp:214,290), whatever
    # iterations were queued past convergence
    assert res["precond_iterations"] == 2 * 24 * it_g and ro["precond_iterations"] == 2 * 24 * it_o
    assert diff <= 10 * kappa * (res["error_compute_operator"] + ro["error_compute_operator"]), diff


def test_large_bicgstab_chebyshev_matches_reference(golden):
    """The reference's default GPU composition (BiCGSTAB + Chebyshev(24),
    global eigenvalues, epsilon 0, rescale 100) on its test problem (origin
    {3, 2.5, 10}, ds 0.1, bcsType {0,1,1,0,1,0}) at 130^3, one rank, against
    the compiled reference's own run (bicgstab_cheb24_alp_130): the fused
    Chebyshev pairs at a size where every block marches full z-chunks.  Bars
    as for every BiCGSTAB case: iterations within 5 % (at least 2), the
    solution sample within kappa (r_gpu + r_ref), both true residuals below
    the tolerance when the reference's is."""
    cases, _ = golden
    name = "bicgstab_cheb24_alp_130"
    if name not in cases:
        pytest.skip("large reference cases not generated")
    c = cases[name]
    d = decode(c)
    G, x, b, flux = oracle_problem(d)
    C = ctx_for(G, d["bc"])
    upload_problem(C, G, x, b, flux, d["bc"])
    res = C.solve(d["chain"])
    xg = C.download(L.FIELD_X)
    C.close()
    it_g, it_r = res["iterations"], c["iterations"]
    samp = data_region(G, xg)[::c["x_sample_step"][0], ::c["x_sample_step"][1], ::c["x_sample_step"][2]].ravel()
    diff = relmax(samp, np.array(c["x_sample"]))
    e = G.g.eig_global
    kappa = e[1] / e[0]
    print(f"[{name}] GPU {it_g} it, reference {it_r} it, sample max rel dx {diff:.3e}, true residual "
          f"{res['error_compute_operator']:.3e} / {c['error_compute_operator']:.3e}", flush=True)
    assert abs(it_g - it_r) <= max(2, round(0.05 * it_r)), (it_g, it_r)
    if c["error_compute_operator"] < 1e-10:
        assert res["error_compute_operator"] < 1e-10
    assert diff <= kappa * (res["error_compute_operator"] + c["error_compute_operator"]), diff


@pytest.mark.parametrize("name", ["cg_none_normal_258_r8", "bicgstab_none_normal_258_r8", "cg_none_gauss_130_r8"])
def test_large_iteration_counts_match_reference(golden, name):
    """config-2 grid (256^3 interior, seeded N(0,1) rhs): iteration counts of the
    reference's own multi-rank runs."""
    cases, _ = golden
    if name not in cases:
        pytest.skip("large reference cases not generated")
    c = cases[name]
    d = decode(c)
    G = O.OracleGrid(d["np"], d["ds"], d["origin"], d["bc"])
    C = ctx_for(G, d["bc"])
    C.fill(L.FIELD_X, 0.0)
    C.fill(L.FIELD_B, 0.0)
    if d["rhs"] == "normal":
        C.rhs_normal(d["seed"])
    else:
        C.rhs_gauss(d["seed"])
    res = C.solve(d["chain"])
    xg = C.download(L.FIELD_X)
    C.close()
    it_r = c["iterations"]
    if d["chain"][0]["kind"] == O.CG:
        assert abs(res["iterations"] - it_r) <= max(1, round(0.02 * it_r)), (res["iterations"], it_r)
    else:
        # Unpreconditioned BiCGSTAB on this case converges erratically: the
        # reference's own count moves from 664 to 897 between rank layouts,
        # which differ only in the order of the global sums
        # (bicgstab_258_rank_spread.json).  The GPU count must fall inside
        # that spread, widened by 5 %.
        with open(os.path.join(os.path.dirname(__file__), "golden", "bicgstab_258_rank_spread.json")) as f:
            spread = [r["iterations"] for r in json.load(f)["runs"]]
        print(f"[{name}] GPU {res['iterations']} it, reference layouts {spread}", flush=True)
        assert it_r in spread
        assert 0.95 * min(spread) <= res["iterations"] <= 1.05 * max(spread), (res["iterations"], spread)
    assert res["error_compute_operator"] < 1e-10
    samp = data_region(G, xg)[::c["x_sample_step"][0], ::c["x_sample_step"][1], ::c["x_sample_step"][2]].ravel()
    e = G.g.eig_global
    bound = 1e-12 if d["chain"][0]["kind"] == O.CG else 10 * e[1] / e[0] * (res["error_compute_operator"] + c["error_compute_operator"])
    assert relmax(samp, np.array(c["x_sample"])) < bound
