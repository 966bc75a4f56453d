"""The reference's main-program flow compiled against the C++ facade
(include/pps/pps_solvers.hpp, driver tests/cpp/facade_main.cpp) drives the GPU
library exactly as the Python mirror of the same interface does."""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
EXE = os.path.join(ROOT, "tests", "cpp", "facade_main")


def test_facade_header_compiles_standalone():
    """The header is self-contained C++17 over the C ABI (no HIP headers)."""
    src = os.path.join(ROOT, "tests", "cpp", "facade_main.cpp")
    p = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-I",
                        os.path.join(ROOT, "include"), src], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr


def reference_problem():
    """The ExactSolutionAndBCs of tests/cpp/facade_main.cpp: the reference's shipped
    test problem (solverPoissonMPI_CPU/include/solverSetup.hpp:44-113), which the
    Python mirror's ExactSolutionAndBCs also is."""
    import parallelpoissonsolver_amd as pps
    return pps.ExactSolutionAndBCs()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["cg", "bicgstab"])
def test_cpp_facade_matches_python_mirror(kind):
    import parallelpoissonsolver_amd as pps
    n = 18
    # BaseCG never refreshes p's Neumann ghosts (baseCG.hpp:124 resets them with
    # isMainLoop false), so its true residual stays large on Neumann faces, as in
    # the reference: CG runs on Dirichlet faces, BiCGSTAB on mixed ones
    bcs = (0, 0, 0, 0, 0, 0) if kind == "cg" else (0, 1, 1, 0, 1, 0)
    print(f"[facade] {kind} {n}^3 bcs {bcs}", flush=True)
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "x.bin")
        p = subprocess.run([EXE, str(n), kind, *map(str, bcs), out], capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stderr
        res = json.loads(p.stdout.strip().splitlines()[-1])
        xc = np.fromfile(out, dtype=np.float64)
    ds = 1.0 / (n - 1)
    grid = pps.BlockGrid((1, 1, 1), 0, (n, n, n), (ds, ds, ds), (0.5, 0.25, 1.0), (1, 1, 1), bcs)
    none = pps.NoneSolver(3, np.float64, 1, 1)
    if kind == "cg":
        spec = pps.BaseCG(3, np.float64, 1, 5000, True, True, none)
    else:
        spec = pps.BiCGSTAB(3, np.float64, 1, 1000, True, True,
                            pps.ChebyshevIteration(3, np.float64, 1, 11, False, False, none))
    solver = spec(grid, reference_problem(), pps.CommunicatorMPI(grid))
    x = np.zeros(grid.getNtotLocalGuards())
    b = np.zeros(grid.getNtotLocalGuards())
    solver.setProblem(x, b)
    solver(x, b)
    assert res["iterations"] == solver.getNumIterationFinal()
    assert res["error_compute_operator"] < 1e-10
    assert abs(res["error_compute_operator"] - solver.getErrorComputeOperator()) <= 1e-3 * solver.getErrorComputeOperator()
    assert np.max(np.abs(xc - x)) <= 1e-12 * np.max(np.abs(x))


@pytest.mark.gpu
def test_cpp_facade_mixed_precision_ir():
    """MixedPrecisionIR with a first-step inner tolerance, driven through the
    C++ facade and through the Python mirror: the same chain, so the same
    inner iteration count and solution; FP64 refinement around FP32 CG
    reaches the FP64 tolerance."""
    import parallelpoissonsolver_amd as pps
    n = 18
    bcs = (0,) * 6
    print(f"[facade] ir {n}^3", flush=True)
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "x.bin")
        p = subprocess.run([EXE, str(n), "ir", *map(str, bcs), out], capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stderr
        res = json.loads(p.stdout.strip().splitlines()[-1])
        xc = np.fromfile(out, dtype=np.float64)
    ds = 1.0 / (n - 1)
    grid = pps.BlockGrid((1, 1, 1), 0, (n, n, n), (ds, ds, ds), (0.5, 0.25, 1.0), (1, 1, 1), bcs)
    inner = pps.BaseCG(3, np.float32, 100000, 5000, False, False, pps.NoneSolver(3, np.float32, 1, 1))
    solver = pps.MixedPrecisionIR(1, 50, inner, first_inner_tol=1e-2)(grid, reference_problem(), pps.CommunicatorMPI(grid))
    x = np.zeros(grid.getNtotLocalGuards())
    b = np.zeros(grid.getNtotLocalGuards())
    solver.setProblem(x, b)
    solver(x, b)
    assert res["iterations"] > 0 and res["iterations"] == solver.getNumIterationFinal()
    assert res["error_compute_operator"] < 1e-10
    assert np.max(np.abs(xc - x)) <= 1e-12 * np.max(np.abs(x))


def _run_facade(kind, n, bcs, outdir=None):
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "x.bin")
        args = [EXE, str(n), kind, *map(str, bcs), out] + ([outdir] if outdir else [])
        p = subprocess.run(args, capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stderr
        res = json.loads(p.stdout.strip().splitlines()[-1])
        return res, p.stdout, np.fromfile(out, dtype=np.float64)


def _printed(stdout, pattern):
    import re
    m = re.search(pattern, stdout)
    assert m, pattern
    return float(m.group(1))


def _oracle_problem(n, bcs, chain):
    """The oracle's solve of facade_main's problem and its errors against u."""
    from oracle import oracle as O
    ds = 1.0 / (n - 1)
    G = O.OracleGrid((n, n, n), (ds, ds, ds), (0.5, 0.25, 1.0), bcs)
    x, b, flux = G.set_problem(0)
    r = G.solve(chain, x, b, flux)
    L = G.lim_data
    sl = (slice(L[4], L[5]), slice(L[2], L[3]), slice(L[0], L[1]))
    X = 0.5 + (np.arange(L[0], L[1]) - L[0]) * ds
    Y = 0.25 + (np.arange(L[2], L[3]) - L[2]) * ds
    Z = 1.0 + (np.arange(L[4], L[5]) - L[4]) * ds
    Z, Y, X = np.meshgrid(Z, Y, X, indexing="ij")
    u = np.sin(X) + np.cos(Y) + 3 * np.sin(Z) + X * X * Y * Z + X * X + 10
    err = np.abs(x[sl] - u)
    return G, x, r, err


CASES = {"cg": ("cg_dir_18", (0, 0, 0, 0, 0, 0)), "bicgstab": ("bicgstab_cheb11_18", (0, 1, 1, 0, 1, 0))}


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["cg", "bicgstab"])
def test_cpp_facade_reports_and_output_files(kind):
    """The reference main program's calls after the solve (main.cpp:103-127):
    checkSolutionLocalGlobal prints the errors the oracle's solution has,
    getDurationSolver().count() is the device time, and writeResidualHistory /
    solution.dat (the alpaka build's writers) produce the files the reference
    solver writes for the same case (tests/golden/files, made by
    tests/golden/make_file_formats.py): the same layout and length, the same
    iteration and preconditioner-iteration counts, the same residual history
    and solution to the parity bars."""
    from oracle import oracle as O
    n = 18
    fixture, bcs = CASES[kind]
    chain = ([dict(kind=O.CG, tol=1e-10), dict(kind=O.NONE)] if kind == "cg" else
             [dict(kind=O.BICGSTAB, tol=1e-10), dict(kind=O.CHEB, maxit=11), dict(kind=O.NONE)])
    print(f"[facade-files] {kind}", flush=True)
    with tempfile.TemporaryDirectory() as outdir:
        res, stdout, xc = _run_facade(kind, n, bcs, outdir)
        hist = open(os.path.join(outdir, "residualHistory.txt")).read().split("\n")
        sol = np.fromfile(os.path.join(outdir, "solution.dat"), dtype=np.float64)
    G, xo, ro, err = _oracle_problem(n, bcs, chain)
    # checkSolutionLocalGlobal's printout against the oracle's solution (6 printed digits)
    avg = _printed(stdout, r"Max error local block avg (\S+) in rank 0")
    pt = _printed(stdout, r"Max error local point (\S+) in rank 0")
    assert abs(avg - err.mean()) <= 1e-5 * err.mean()
    assert abs(pt - err.max()) <= 1e-5 * err.max()
    # the sum of |x - u| (the return value): CG matches the oracle to 1e-12, BiCGSTAB to ~1e-9 per point
    assert abs(res["error_local"] - err.sum()) <= (1e-8 if kind == "cg" else 1e-5) * err.sum()
    assert _printed(stdout, r"SolverInFunction time: (\S+) seconds") > 0
    assert "timeTotal " in stdout
    # the files of the reference solver for the same case
    ref_hist = open(os.path.join(ROOT, "tests", "golden", "files", fixture + ".residualHistory.txt")).read().split("\n")
    ref_sol = np.fromfile(os.path.join(ROOT, "tests", "golden", "files", fixture + ".solution.dat"), dtype=np.float64)
    # header: duration, iterations (within the 2 % parity bar of the
    # reference's), preconditioner iterations (Chebyshev(11) twice per BiCGSTAB
    # iteration, BiCGstabAlpaka.hpp:214,290); then one residual per iteration 0..iters
    it, it_ref = int(hist[1]), int(ref_hist[1])
    print(f"[facade-files] {kind}: GPU {it} it, reference {it_ref} it", flush=True)
    assert float(hist[0]) > 0 and it == res["iterations"]
    assert abs(it - it_ref) <= max(1, round(0.02 * it_ref))
    assert int(hist[2]) == res["precond_iterations"] == (0 if kind == "cg" else 2 * 11 * it)
    assert int(ref_hist[2]) == (0 if kind == "cg" else 2 * 11 * it_ref)
    assert hist[-1] == ref_hist[-1] == ""
    h, hr = np.array([float(v) for v in hist[3:-1]]), np.array([float(v) for v in ref_hist[3:-1]])
    assert h.size == it + 1 and hr.size == it_ref + 1
    m = min(h.size, hr.size)
    assert np.allclose(h[:10], hr[:10], rtol=1e-5) and np.allclose(h[:m], hr[:m], rtol=2e-2)
    assert sol.size == ref_sol.size == G.shape[0] * G.shape[1] * G.shape[2]
    assert np.array_equal(sol, xc)
    # the Python mirror of the same interface writes the same history file
    import parallelpoissonsolver_amd as pps
    ds = 1.0 / (n - 1)
    grid = pps.BlockGrid((1, 1, 1), 0, (n, n, n), (ds, ds, ds), (0.5, 0.25, 1.0), (1, 1, 1), bcs)
    none = pps.NoneSolver(3, np.float64, 1, 1)
    spec = (pps.BaseCG(3, np.float64, 1, 5000, True, True, none) if kind == "cg" else
            pps.BiCGSTAB(3, np.float64, 1, 1000, True, True, pps.ChebyshevIteration(3, np.float64, 1, 11, False, False, none)))
    mirror = spec(grid, reference_problem(), pps.CommunicatorMPI(grid))
    xm, bm = np.zeros(grid.getNtotLocalGuards()), np.zeros(grid.getNtotLocalGuards())
    mirror.setProblem(xm, bm)
    mirror(xm, bm)
    assert abs(mirror.checkSolutionLocalGlobal(xm, verbose=False) - res["error_local"]) <= 1e-10 * res["error_local"]
    with tempfile.TemporaryDirectory() as tmp:
        mirror.writeResidualHistory(os.path.join(tmp, "h.txt"))
        mh = open(os.path.join(tmp, "h.txt")).read().split("\n")
    assert mh[1:] == hist[1:]
    e = G.g.eig_global
    bar = 1e-12 if kind == "cg" else e[1] / e[0] * 2e-10
    assert np.abs(sol - ref_sol).max() <= bar * np.abs(ref_sol).max()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["alp", "alp32"])
def test_cpp_facade_alpaka_types(kind):
    """The alpaka build's default composition through its own type names
    (BiCGstabAlpaka + ChebyshevIterationAlpaka<..., T_data_chebyshev, global,
    ...>, NoneSolverAlpaka, AlpakaHelper; the solver sets the problem itself,
    BiCGstabAlpaka.hpp:136-138) matches the oracle with the alpaka Chebyshev
    parameters (epsilon 0, rescaleEigMin 100), in FP64 and with
    T_data_chebyshev = float."""
    from oracle import oracle as O
    n = 18
    bcs = (0, 1, 1, 0, 1, 0)
    print(f"[facade-alp] {kind}", flush=True)
    res, stdout, xc = _run_facade(kind, n, bcs)
    prec = 32 if kind == "alp32" else 64
    chain = [dict(kind=O.BICGSTAB, tol=1e-10),
             dict(kind=O.CHEB, maxit=24, precision=prec, cheb_eps=0.0, cheb_rmin=100.0, cheb_rmax=1 - 1e-4),
             dict(kind=O.NONE)]
    G, xo, ro, err = _oracle_problem(n, bcs, chain)
    assert abs(res["iterations"] - ro["iterations"]) <= max(1, round(0.02 * ro["iterations"]))
    assert res["precond_iterations"] == 2 * 24 * res["iterations"]
    assert res["error_compute_operator"] < 1e-10
    L = G.lim_data
    sl = (slice(L[4], L[5]), slice(L[2], L[3]), slice(L[0], L[1]))
    e = G.g.eig_global
    d = np.abs(xc.reshape(G.shape)[sl] - xo[sl]).max() / np.abs(xo[sl]).max()
    assert d <= e[1] / e[0] * (res["error_compute_operator"] + ro["error_compute_operator"]), d


def test_reference_main_compiles_against_dropin():
    """The reference's own main program and configuration headers
    (solverPoissonMPI_CPU/src/main.cpp, include/inputParam.hpp,
    include/solverSetup.hpp), copied unchanged to a scratch directory, compile
    against include/pps/dropin/ in place of the reference's solver headers."""
    ref = "/root/reference/solverPoissonMPI_CPU"
    mpi_inc = "/opt/conda/include"
    if not os.path.isdir(ref) or not os.path.exists(os.path.join(mpi_inc, "mpi.h")):
        pytest.skip("reference sources or MPI headers not present")
    import shutil
    with tempfile.TemporaryDirectory() as tmp:
        for f in ("src/main.cpp", "include/inputParam.hpp", "include/solverSetup.hpp"):
            shutil.copy(os.path.join(ref, f), tmp)
        p = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", mpi_inc, "-I", tmp, "-I",
                            os.path.join(ROOT, "include", "pps", "dropin"), os.path.join(tmp, "main.cpp")],
                           capture_output=True, text=True)
    assert p.returncode == 0, p.stderr


@pytest.mark.gpu
def test_reference_main_runs_on_gpu():
    """The reference's main program built against the drop-in headers
    (tests/cpp/ref_main_dropin, made by the build from the reference's own
    main.cpp and configuration) runs its shipped configuration, the one of
    solverPoissonMPI_CPU/run/solverScoreP.o (128 x 128 x 256, BiCGSTAB +
    Chebyshev(11), 1e-8), on the GPU with one rank, and agrees with the
    reference solver's one-rank run of it (tests/golden/recorded_run.json)."""
    import re
    exe = os.path.join(ROOT, "tests", "cpp", "ref_main_dropin")
    if not os.path.exists(exe):
        pytest.skip("ref_main_dropin not built (the build makes it where the reference sources are present)")
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "recorded_run.json")))["rerun_111"]
    print("[ref-main] 128x128x256 BiCGSTAB + Chebyshev(11) on one GPU", flush=True)
    p = subprocess.run([exe, "1", "1", "1"], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    out = p.stdout
    m = re.search(r"Iterative solver finished with iter: (\d+) error from algo (\S+) error r=b-Ax (\S+)", out)
    assert m, out[-2000:]
    it, err_op = int(m.group(1)), float(m.group(3))
    avg = _printed(out, r"Max error local block avg (\S+) in rank 0")
    pt = _printed(out, r"Max error local point (\S+) in rank 0")
    print(f"[ref-main] GPU {it} it, residual {err_op:.4e}, errors {avg:.6g} / {pt:.6g}; reference one rank "
          f"{fx['iterations']} it, {fx['error_compute_operator']:.4e}, {fx['max_block_avg']:.6g} / {fx['max_point']:.6g}",
          flush=True)
    assert "End program." in out
    assert err_op < 1e-8
    # BiCGSTAB at 4.2M unknowns: the reduction order alone moves the count by a
    # few per cent (the reference's own 1-rank and 4-rank runs of one problem
    # take 103 and 108 iterations; the recorded 64-rank run 155, its rerun here
    # 152), so the bar is the 5 % of the other BiCGSTAB runs of this size
    # (measured 101 against 104)
    assert abs(it - fx["iterations"]) <= round(0.05 * fx["iterations"])
    # errors against the exact solution: at a 1e-8 residual the solver's own
    # error is a few per cent of them and depends on the iterate the solve
    # stops at (the recorded 64-rank run and its rerun here differ by 2.5 % and
    # 0.8 %; the GPU's two summation orders gave 0.3 % / 0.1 % and 2.0 % / 2.6 %)
    assert abs(avg - fx["max_block_avg"]) <= 0.05 * fx["max_block_avg"]
    assert abs(pt - fx["max_point"]) <= 0.05 * fx["max_point"]
