"""The oracle's FP64/FP32 iterative refinement with a looser inner tolerance
for the first outer step (ir_first_tol).  The refinement has no reference
counterpart; what fixes the behaviour is the algorithm itself: the first inner
solve stops at the looser tolerance (so the first outer residual is no
smaller than about that factor), the later ones at the inner stage's own, and
the FP64 true residual still reaches the outer tolerance."""
import numpy as np

from oracle import oracle as O


def _solve(first_tol):
    n = 34
    G = O.OracleGrid((n,) * 3, (1 / (n - 1),) * 3, (0, 0, 0), (0,) * 6)
    b = G.rhs("gauss", 3)
    x = G.zeros()
    ir = dict(kind=O.IR, tol=1e-10, maxit=50)
    if first_tol:
        ir["first_tol"] = first_tol
    res = G.solve([ir, dict(kind=O.CG, tol=1e-5, maxit=5000), dict(kind=O.NONE)], x, b)
    return res


def test_first_outer_step_uses_its_own_tolerance():
    plain, loose = _solve(0.0), _solve(1e-2)
    for r in (plain, loose):
        assert r["error_compute_operator"] < 1e-10
    hp, hl = np.asarray(plain["history"]), np.asarray(loose["history"])
    # history[k]: true residual after k outer steps (history[0] = 1 after normalisation)
    assert abs(hp[0] - 1.0) < 1e-12 and abs(hl[0] - 1.0) < 1e-12
    assert hp[1] < 1e-4 and 1e-3 < hl[1] < 2e-2
    assert loose["iterations"] < plain["iterations"]


def _solve_adapt(adapt):
    n = 34
    G = O.OracleGrid((n,) * 3, (1 / (n - 1),) * 3, (0, 0, 0), (0,) * 6)
    b = G.rhs("gauss", 3)
    x = G.zeros()
    ir = dict(kind=O.IR, tol=1e-10, maxit=50, adapt=adapt)
    return G.solve([ir, dict(kind=O.CG, tol=1e-4, maxit=5000), dict(kind=O.NONE)], x, b)


def test_adaptive_inner_tolerance_loosens_only_the_last_step():
    """ir_adapt = 0.5: an outer step asks the inner solve for no more than
    0.5 * tol / e, e its outer residual.  With an inner tolerance of 1e-4 the
    first two steps (e = 1, then ~1e-4) need more than 1e-4 and are unchanged,
    so the outer residuals before the last step are identical; the last step
    (e ~ 1e-8) asks for ~5e-3 instead of 1e-4 and takes fewer inner
    iterations, and the true residual still reaches 1e-10."""
    plain, adapt = _solve_adapt(0.0), _solve_adapt(0.5)
    hp, ha = np.asarray(plain["history"]), np.asarray(adapt["history"])
    assert plain["outer_iterations"] == adapt["outer_iterations"] == 3
    np.testing.assert_array_equal(hp[:3], ha[:3])
    assert 1e-9 < hp[2] < 1e-7  # the residual before the last step
    assert adapt["iterations"] < plain["iterations"]
    assert hp[3] < ha[3] < 1e-10
    # the last step's inner solve stopped near 0.5 * tol / e (its own relative
    # residual, not the outer one: the FP32 solve's error moves the outer
    # residual by less than a factor 10 from it)
    assert 0.05 * 0.5e-10 / hp[2] < ha[3] / hp[2] < 10 * 0.5e-10 / hp[2]
