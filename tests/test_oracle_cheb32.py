"""Pins the oracle's FP32 Chebyshev preconditioner under an FP64 caller (the
alpaka T_data_chebyshev = float form, chebyshevIterationAlpaka.hpp:119-196,
kernelsAlpakaChebyshev.hpp:7-18,115-273).

The reference's alpaka build cannot be compiled here, so this form is pinned by:
  * the FP64 Chebyshev of the oracle, which the compiled CPU reference pins bit
    for bit (tests/test_oracle_golden.py): the two agree to FP32 rounding and
    differ by more than FP64 rounding (the sweeps did run in FP32);
  * the closed form in the DST eigenbasis: on a Dirichlet grid a discrete sine
    mode v is an eigenvector of A (A v = mu v), so the preconditioner returns
    p(mu) v with p the scalar recurrence;
  * a BiCGSTAB solve preconditioned by it converging like the FP64 one.
"""
import numpy as np
import pytest

from oracle import oracle as O

ALP = dict(cheb_eps=0.0, cheb_rmin=100.0, cheb_rmax=1 - 1e-4)  # solverSetup.hpp (alpaka):59-62


def cheb(m, prec, **kw):
    d = dict(kind=O.CHEB, maxit=m, precision=prec, **ALP)
    d.update(kw)
    return d


@pytest.mark.parametrize("bcs", [(0, 0, 0, 0, 0, 0), (0, 1, 1, 0, 1, 0)])
@pytest.mark.parametrize("m", [4, 11, 24])
def test_fp32_chebyshev_matches_fp64_to_fp32_rounding(bcs, m):
    n = (34, 30, 26)
    ds = tuple(1.0 / (v - 1) for v in n)
    G = O.OracleGrid(n, ds, (3, 2.5, 10), bcs)
    rng = np.random.default_rng(5)
    inp = rng.standard_normal(G.shape)
    y64, _ = G.precondition([cheb(m, 64)], inp.copy())
    y32, info = G.precondition([cheb(m, 32)], inp.copy())
    assert info["precond_iterations"] == m
    sl = G.solver_slice()
    scale = np.abs(y64[sl]).max()
    diff = np.abs(y32[sl] - y64[sl]).max() / scale
    assert diff < 5e-6, diff          # FP32 rounding over m sweeps (measured 1e-7 .. 6e-7)
    assert diff > 1e-9, diff          # and not FP64 arithmetic
    # only the solver region is written
    mask = np.ones(G.shape, bool)
    mask[sl] = False
    assert np.all(y32[mask] == 0)


def scalar_cheb(mu, m, theta, delta):
    """The preconditioner on one eigenvector: the recurrence on scalars, with
    the reference's buffer rotation (returns the y two sweeps before the last)."""
    sigma = theta / delta
    rho_old = 1 / sigma
    rho = 1 / (2 * sigma - rho_old)
    b = 1.0
    Z = b / theta
    Y = 2 * rho / delta * (2 * b + mu * b / theta)
    W = 0.0
    for _ in range(2, m + 1):
        rho_old, rho = rho, 1 / (2 * sigma - rho)
        W = rho * (2 * sigma * Y + 2 / delta * (b + mu * Y) - rho_old * Z)
        Z, Y = Y, Z
        W, Y = Y, W
    return -W


@pytest.mark.parametrize("modes", [(1, 1, 1), (2, 3, 1), (7, 5, 9)])
def test_fp32_chebyshev_closed_form_on_sine_modes(modes):
    n = (20, 18, 16)  # data points incl. the Dirichlet layers
    ds = (0.05, 0.07, 0.06)
    G = O.OracleGrid(n, ds, (0, 0, 0), (0,) * 6)
    sl = G.solver_slice()
    ni = [v - 2 for v in n]  # solver points per direction
    ax = [np.sin(modes[d] * np.pi * np.arange(1, ni[d] + 1) / (ni[d] + 1)) for d in range(3)]
    v = G.zeros()
    v[sl] = ax[2][:, None, None] * ax[1][None, :, None] * ax[0][None, None, :]
    mu = -sum(4 / ds[d] ** 2 * np.sin(modes[d] * np.pi / (2 * (ni[d] + 1))) ** 2 for d in range(3))
    e = G.g.eig_global
    for m in (4, 11, 24):
        theta = (e[0] * ALP["cheb_rmin"] + e[1] * ALP["cheb_rmax"]) * 0.5
        delta = (e[1] * ALP["cheb_rmax"] - e[0] * ALP["cheb_rmin"]) * 0.5
        p = scalar_cheb(mu, m, theta, delta)
        y32, _ = G.precondition([cheb(m, 32)], v.copy())
        y64, _ = G.precondition([cheb(m, 64)], v.copy())
        assert np.abs(y64[sl] - p * v[sl]).max() < 1e-12 * abs(p)
        assert np.abs(y32[sl] - p * v[sl]).max() < 1e-5 * abs(p), m


def test_bicgstab_with_fp32_chebyshev_converges_like_fp64():
    n = 34
    G = O.OracleGrid((n,) * 3, (1 / (n - 1),) * 3, (3, 2.5, 10), (0, 1, 1, 0, 1, 0))
    x0, b0, flux = G.set_problem(0)
    its = {}
    for prec in (64, 32):
        x, b = x0.copy(), b0.copy()
        r = G.solve([dict(kind=O.BICGSTAB, tol=1e-10, maxit=1000), cheb(24, prec)], x, b, flux)
        assert r["error_compute_operator"] < 1e-10
        its[prec] = r["iterations"]
    assert abs(its[32] - its[64]) <= 2, its
