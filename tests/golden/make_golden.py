#!/usr/bin/env python3
"""Generate the golden fixtures from the compiled reference CPU solver.

Runs oracle/_ref/ref_solver (the reference's own solver classes, built from
/root/reference/solverPoissonMPI_CPU/include by oracle/Makefile) on a set of
grids, boundary sets, right-hand sides and solver chains, and stores what the
reference produced:

  ref_cases.json   per case: options, iterations, recursive and true residual,
                   residual history, solution checksums and samples
  ref_vectors.npz  full solution arrays of the small (18^3-point) cases

Run in the build container only (it needs /root/reference to build the
reference and mpirun for the multi-rank cases):

    make -C oracle ref && python3 tests/golden/make_golden.py [--large]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref", "ref_solver")
MPIRUN = "/opt/conda/bin/mpirun"

CPU_CHEB = ["--cheb-eps", "1e-4", "--cheb-rescale", "500", "0.9999"]       # CPU solverSetup.hpp:35-38
ALP_CHEB = ["--cheb-eps", "0", "--cheb-rescale", "100", "0.9999"]          # alpaka solverSetup.hpp:70-73
OFF = ["--origin", "0.5", "0.25", "1.0"]


def unit(n):
    ds = 1.0 / (n - 1)
    return ["--np", str(n), str(n), str(n), "--ds", repr(ds), repr(ds), repr(ds)]


def grid(n, ds):
    return ["--np", str(n), str(n), str(n), "--ds", repr(ds), repr(ds), repr(ds)]


BC_DIR = ["--bc", "0", "0", "0", "0", "0", "0"]
BC_ALP = ["--bc", "0", "1", "1", "0", "1", "0"]   # alpaka inputParam.hpp:56
BC_CPU = ["--bc", "0", "1", "0", "1", "0", "1"]   # CPU inputParam.hpp:46

# (name, ranks, args, store_full_vector)
SMALL = [
    ("cg_none_dir_p0_18", (1, 1, 1), unit(18) + OFF + BC_DIR + ["--solver", "cg", "--precond", "none"], True),
    ("bicgstab_none_dir_p0_18", (1, 1, 1), unit(18) + OFF + BC_DIR + ["--solver", "bicgstab", "--precond", "none"], True),
    ("cg_cheb11_dir_p0_18", (1, 1, 1), unit(18) + OFF + BC_DIR + ["--solver", "cg", "--precond", "cheb", "--cheb-m", "11"] + CPU_CHEB, True),
    ("bicgstab_cheb24_alp_18", (1, 1, 1), grid(18, 0.1) + ["--origin", "3", "2.5", "10"] + BC_ALP
     + ["--solver", "bicgstab", "--precond", "cheb", "--cheb-m", "24"] + ALP_CHEB, True),
    ("bicgstab_cheb11_cpu_18", (1, 1, 1), grid(18, 0.1) + BC_CPU + ["--solver", "bicgstab", "--precond", "cheb", "--cheb-m", "11"] + CPU_CHEB, True),
    ("bicgstab_cheb4_alp_18", (1, 1, 1), grid(18, 0.1) + ["--origin", "3", "2.5", "10"] + BC_ALP
     + ["--solver", "bicgstab", "--precond", "cheb", "--cheb-m", "4"] + ALP_CHEB, True),
    ("cg_none_alp_18", (1, 1, 1), grid(18, 0.1) + ["--origin", "3", "2.5", "10"] + BC_ALP + ["--solver", "cg", "--precond", "none"], True),
    ("bicgstab_none_alp_18", (1, 1, 1), grid(18, 0.1) + ["--origin", "3", "2.5", "10"] + BC_ALP + ["--solver", "bicgstab", "--precond", "none"], True),
    ("bicgstab_bicgstab_alp_18", (1, 1, 1), grid(18, 0.1) + ["--origin", "3", "2.5", "10"] + BC_ALP + ["--solver", "bicgstab", "--precond", "bicgstab"], True),
    ("cg_cg_dir_18", (1, 1, 1), unit(18) + OFF + BC_DIR + ["--solver", "cg", "--precond", "cg"], True),
    ("cheb24_main_dir_18", (1, 1, 1), unit(18) + OFF + BC_DIR + ["--solver", "cheb", "--cheb-m", "24", "--cheb-eps", "0", "--cheb-rescale", "1", "1"], True),
    ("cg_none_sine_18", (1, 1, 1), unit(18) + BC_DIR + ["--problem", "1", "--solver", "cg", "--precond", "none"], True),
    ("cg_none_normal_18", (1, 1, 1), unit(18) + BC_DIR + ["--problem", "2", "--rhs", "normal", "--seed", "7", "--solver", "cg", "--precond", "none"], True),
    ("bicgstab_none_normal_18", (1, 1, 1), unit(18) + BC_DIR + ["--problem", "2", "--rhs", "normal", "--seed", "7", "--solver", "bicgstab", "--precond", "none"], True),
    ("cg_none_gauss_18", (1, 1, 1), unit(18) + BC_DIR + ["--problem", "2", "--rhs", "gauss", "--seed", "11", "--solver", "cg", "--precond", "none"], True),
    ("cg_none_spectrum_18", (1, 1, 1), unit(18) + BC_DIR + ["--problem", "2", "--rhs", "spectrum", "--spec-k", "8", "--seed", "5", "--solver", "cg", "--precond", "none"], True),
]

MEDIUM = [
    # config 1: 64^3 interior unknowns, manufactured sine solution, plain CG, 1e-10
    ("cg_none_sine_66", (1, 1, 1), unit(66) + BC_DIR + ["--problem", "1", "--solver", "cg", "--precond", "none"], False),
    ("cg_none_p0_unit_66", (1, 1, 1), unit(66) + BC_DIR + ["--problem", "0", "--solver", "cg", "--precond", "none"], False),
    ("cg_none_normal_66", (1, 1, 1), unit(66) + BC_DIR + ["--problem", "2", "--rhs", "normal", "--seed", "2024", "--solver", "cg", "--precond", "none"], False),
    ("bicgstab_none_normal_66", (1, 1, 1), unit(66) + BC_DIR + ["--problem", "2", "--rhs", "normal", "--seed", "2024", "--solver", "bicgstab", "--precond", "none"], False),
    ("cg_none_gauss_66", (1, 1, 1), unit(66) + BC_DIR + ["--problem", "2", "--rhs", "gauss", "--seed", "3", "--solver", "cg", "--precond", "none"], False),
    ("cg_none_multiscale_66", (1, 1, 1), unit(66) + BC_DIR + ["--problem", "2", "--rhs", "multiscale", "--spec-k", "16", "--seed", "9", "--solver", "cg", "--precond", "none"], False),
    # the alpaka build's default test problem (alpaka inputParam.hpp:53-57) and composition (:36)
    ("bicgstab_cheb24_alp_64", (1, 1, 1), grid(64, 0.1) + ["--origin", "3", "2.5", "10"] + BC_ALP
     + ["--solver", "bicgstab", "--precond", "cheb", "--cheb-m", "24"] + ALP_CHEB, False),
    ("cg_cheb11_dir_34", (1, 1, 1), unit(34) + OFF + BC_DIR + ["--solver", "cg", "--precond", "cheb", "--cheb-m", "11"] + CPU_CHEB, False),
    # multi-rank runs of the reference: the iterates are decomposition independent
    ("cg_none_p0_34_r8", (2, 2, 2), unit(34) + OFF + BC_DIR + ["--solver", "cg", "--precond", "none"], False),
    ("cg_none_p0_34", (1, 1, 1), unit(34) + OFF + BC_DIR + ["--solver", "cg", "--precond", "none"], False),
    ("bicgstab_none_normal_34_r4", (2, 2, 1), unit(34) + BC_DIR + ["--problem", "2", "--rhs", "normal", "--seed", "7", "--solver", "bicgstab", "--precond", "none"], False),
    ("bicgstab_none_normal_34", (1, 1, 1), unit(34) + BC_DIR + ["--problem", "2", "--rhs", "normal", "--seed", "7", "--solver", "bicgstab", "--precond", "none"], False),
]

LARGE = [
    # config 2 grid (256^3 interior) with its seeded N(0,1) right-hand side
    ("cg_none_normal_258_r8", (2, 2, 2), unit(258) + BC_DIR + ["--problem", "2", "--rhs", "normal", "--seed", "2024", "--solver", "cg", "--precond", "none"], False),
    ("bicgstab_none_normal_258_r8", (2, 2, 2), unit(258) + BC_DIR + ["--problem", "2", "--rhs", "normal", "--seed", "2024", "--solver", "bicgstab", "--precond", "none"], False),
    ("cg_none_gauss_130_r8", (2, 2, 2), unit(130) + BC_DIR + ["--problem", "2", "--rhs", "gauss", "--seed", "3", "--solver", "cg", "--precond", "none"], False),
    # the alpaka default composition on its test problem at 130^3, one rank (the
    # preconditioner runs without halo exchange, communicationOFF, so a
    # multi-rank run would precondition block by block)
    ("bicgstab_cheb24_alp_130", (1, 1, 1), grid(130, 0.1) + ["--origin", "3", "2.5", "10"] + BC_ALP
     + ["--solver", "bicgstab", "--precond", "cheb", "--cheb-m", "24"] + ALP_CHEB, False),
]


def assemble(prefix, ranks, npg):
    """Reassemble per-rank local guarded arrays into the global data region (k, j, i)."""
    px, py, pz = ranks
    nl = [npg[0] // px, npg[1] // py, npg[2] // pz]
    out = np.zeros((npg[2], npg[1], npg[0]))
    for r in range(px * py * pz):
        gx, gy, gz = r % px, (r // px) % py, r // (px * py)
        loc = np.fromfile(f"{prefix}.x.r{r}.bin").reshape(nl[2] + 2, nl[1] + 2, nl[0] + 2)
        out[gz * nl[2]:(gz + 1) * nl[2], gy * nl[1]:(gy + 1) * nl[1], gx * nl[0]:(gx + 1) * nl[0]] = loc[1:-1, 1:-1, 1:-1]
    return out


def run_case(name, ranks, args, full, tmp):
    prefix = os.path.join(tmp, name)
    nproc = ranks[0] * ranks[1] * ranks[2]
    cmd = [REF, "--ranks", *map(str, ranks), *args, "--tol", "1e-10", "--out", prefix]
    if nproc > 1:
        cmd = [MPIRUN, "-n", str(nproc)] + cmd
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"{name}: {p.stderr[-2000:]}")
    meta = json.load(open(prefix + ".json"))
    npg = [int(v) for v in args[args.index("--np") + 1: args.index("--np") + 4]]
    x = assemble(prefix, ranks, npg)  # data region, k j i
    step = [max(1, s // 8) for s in x.shape]
    case = {
        "name": name,
        "ranks": list(ranks),
        "args": args,
        "tol": 1e-10,
        "iterations": meta["iterations"],
        "error_from_iteration": meta["error_from_iteration"],
        "error_compute_operator": meta["error_compute_operator"],
        "solver_time": meta["solver_time"],
        "history": meta["history"],
        "x_sum": float(x.sum()),
        "x_sumsq": float((x * x).sum()),
        "x_maxabs": float(np.abs(x).max()),
        "x_sample_step": step,
        "x_sample": x[::step[0], ::step[1], ::step[2]].ravel().tolist(),
    }
    vec = x if full else None
    print(f"{name}: iters {meta['iterations']} true residual {meta['error_compute_operator']:.3e}", flush=True)
    return case, vec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--large", action="store_true", help="also run the 130^3/258^3 multi-rank cases")
    ap.add_argument("--only", nargs="*", default=None, help="run only these cases and merge them into the fixtures")
    a = ap.parse_args()
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: run `make -C oracle ref` first")
    cases, vecs = [], {}
    with tempfile.TemporaryDirectory() as tmp:
        todo = SMALL + MEDIUM + (LARGE if a.large else [])
        if a.only:
            todo = [t for t in SMALL + MEDIUM + LARGE if t[0] in a.only]
        for name, ranks, args, full in todo:
            c, v = run_case(name, ranks, args, full, tmp)
            cases.append(c)
            if v is not None:
                vecs[name] = v
    out_json = os.path.join(HERE, "ref_cases.json")
    if (a.only or not a.large) and os.path.exists(out_json):
        # keep the previously generated cases not rerun
        old = json.load(open(out_json))
        have = {c["name"] for c in cases}
        cases += [c for c in old["cases"] if c["name"] not in have]
    json.dump({"generator": "tests/golden/make_golden.py", "reference": "solverPoissonMPI_CPU (compiled by oracle/Makefile)",
               "cases": cases}, open(out_json, "w"), indent=1)
    if a.only:
        # keep the stored vectors of the cases not rerun
        old_vecs = dict(np.load(os.path.join(HERE, "ref_vectors.npz")))
        old_vecs.update(vecs)
        vecs = old_vecs
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **vecs)


if __name__ == "__main__":
    main()
