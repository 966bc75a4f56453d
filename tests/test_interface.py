"""The host-side mirror of the reference interface (no GPU needed): BlockGrid
limits and eigenvalues, setProblem fields and Neumann flux planes against the
oracle (which is pinned to the compiled reference)."""
import numpy as np
import pytest

import parallelpoissonsolver_amd as pps
from oracle import oracle as O


@pytest.mark.parametrize("bcs", [(0,) * 6, (0, 1, 1, 0, 1, 0), (0, 1, 0, 1, 0, 1), (1,) * 6])
def test_blockgrid_matches_oracle(bcs):
    n, ds, org = (18, 20, 22), (0.1, 0.05, 0.2), (3, 2.5, 10)
    g = pps.BlockGrid((1, 1, 1), 0, n, ds, org, (1, 1, 1), bcs)
    G = O.OracleGrid(n, ds, org, bcs)
    assert list(g.getIndexLimitsData()) == G.lim_data
    assert list(g.getIndexLimitsSolver()) == G.lim_solver
    assert np.allclose(g.getEigenValuesGlobal(), list(G.g.eig_global), rtol=1e-15)
    assert np.allclose(g.getEigenValuesLocal(), list(G.g.eig_local), rtol=1e-15)
    assert g.getNtotLocalGuards() == G.g.ntot


def test_blockgrid_rejects_multi_rank():
    with pytest.raises(ValueError):
        pps.BlockGrid((2, 1, 1), 0, (18, 18, 18), (0.1,) * 3, (0,) * 3, (1, 1, 1), (0,) * 6)


@pytest.mark.parametrize("bcs", [(0,) * 6, (0, 1, 1, 0, 1, 0), (1, 1, 0, 1, 1, 1)])
def test_set_problem_and_flux_match_oracle(bcs):
    n, ds, org = (18, 20, 22), (0.1, 0.05, 0.2), (3, 2.5, 10)
    g = pps.BlockGrid((1, 1, 1), 0, n, ds, org, (1, 1, 1), bcs)
    spec = pps.BaseCG(3, np.float64, 1, 100, True, True, pps.NoneSolver())
    s = spec(g, pps.ExactSolutionAndBCs(), pps.CommunicatorMPI(g))
    x = np.zeros(g.getNtotLocalGuards())
    b = np.zeros(g.getNtotLocalGuards())
    s.setProblem(x, b)
    G = O.OracleGrid(n, ds, org, bcs)
    xo, bo, flux = G.set_problem(0)
    assert np.allclose(x.reshape(G.shape), xo, rtol=1e-15, atol=0)
    assert np.allclose(b.reshape(G.shape), bo, rtol=1e-15, atol=1e-15)
    planes = s._flux_planes()
    for f in range(6):
        if bcs[f] == 1:
            assert np.allclose(planes[f], flux[f], rtol=1e-15, atol=1e-15)
        else:
            assert f not in planes


def test_spec_chain_mirrors_template_composition():
    inner = pps.ChebyshevIteration(3, np.float64, 10000, 24, False, False, pps.NoneSolver(), T_data_chebyshev=np.float32,
                                   eps=0.0, rescale_min=100.0, rescale_max=1 - 1e-4)
    spec = pps.BiCGSTAB(3, np.float64, 1, 1000, True, True, inner)
    ch = spec.chain()
    assert [c["kind"] for c in ch] == [pps._lib.BICGSTAB, pps._lib.CHEBYSHEV, pps._lib.NONE]
    assert ch[0]["tol"] == 1e-10 and ch[0]["maxit"] == 1000
    assert ch[1]["maxit"] == 24 and ch[1]["precision"] == 32 and ch[1]["cheb_rmin"] == 100.0
    with pytest.raises(ValueError):
        inner(None, None)
