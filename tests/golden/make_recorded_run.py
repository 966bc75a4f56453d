#!/usr/bin/env python3
"""Fixture of the one run the reference records: solverPoissonMPI_CPU/run/solverScoreP.o.

That file holds the output of the reference CPU solver with its shipped
configuration (inputParam.hpp:25-46, solverSetup.hpp:26-40: 128 x 128 x 256
points, ds 0.1, origin 0, bcs {0,1,0,1,0,1}, BiCGSTAB to 1e2 * 1e-10 with a
communication-free Chebyshev(11) preconditioner, epsilon 1e-4, rescale 500 /
1 - 1e-4) on 4 x 4 x 4 MPI ranks.  This script runs the same configuration
with the reference solver compiled here (oracle/_ref/ref_solver, the
reference's own headers) on 4 x 4 x 4 ranks and on one rank, and stores:

  record     the lines of solverScoreP.o the checks use (data of the reference)
  rerun_444  the same lines from the 4 x 4 x 4 rerun
  rerun_111  the one-rank run: the configuration as one GPU sees it

into recorded_run.json.  Run in the build container (it needs /root/reference
and mpirun; the 64-rank run takes a few minutes on 8 cores):

    make -C oracle ref && python3 tests/golden/make_recorded_run.py
"""
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref", "ref_solver")
MPIRUN = "/opt/conda/bin/mpirun"
RECORD = "/root/reference/solverPoissonMPI_CPU/run/solverScoreP.o"
ARGS = ["--np", "128", "128", "256", "--ds", "0.1", "0.1", "0.1", "--origin", "0", "0", "0",
        "--bc", "0", "1", "0", "1", "0", "1", "--problem", "0", "--solver", "bicgstab", "--precond", "cheb",
        "--cheb-m", "11", "--cheb-eps", "1e-4", "--cheb-rescale", "500", "0.9999", "--tol", "1e-8"]


def parse(text):
    """The solver's progress and result lines (BiCGSTAB.hpp debug print, main.cpp:103-113)."""
    out = {"debug": {}}
    for line in text.splitlines():
        line = line.strip()
        m = re.match(r"Debug in BiCGSTAB iter (\d+) (.*)", line)
        if m:
            out["debug"][m.group(1)] = m.group(2)
        m = re.search(r"norm fieldB (\S+)", line)
        if m:
            out["norm_fieldB"] = m.group(1)
        m = re.match(r"Iterative solver finished with iter: (\d+) error from algo (\S+) error r=b-Ax (\S+)", line)
        if m:
            out["iterations"] = int(m.group(1))
            out["error_from_iteration"] = float(m.group(2))
            out["error_compute_operator"] = float(m.group(3))
        m = re.match(r"Max error local block avg (\S+) in rank (\d+)", line)
        if m:
            out["max_block_avg"], out["max_block_avg_rank"] = float(m.group(1)), int(m.group(2))
        m = re.match(r"Max error local point (\S+) in rank (\d+)", line)
        if m:
            out["max_point"], out["max_point_rank"] = float(m.group(1)), int(m.group(2))
    return out


def rerun(ranks):
    n = ranks[0] * ranks[1] * ranks[2]
    cmd = [MPIRUN, "-n", str(n), REF, "--ranks", *map(str, ranks), *ARGS]
    print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(p.stderr[-2000:])
    r = parse(p.stdout)
    r["ranks"] = list(ranks)
    print({k: v for k, v in r.items() if k != "debug"}, flush=True)
    return r


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: run `make -C oracle ref` first")
    record = parse(open(RECORD).read())
    record["ranks"] = [4, 4, 4]
    out = {"generator": "tests/golden/make_recorded_run.py",
           "source": "solverPoissonMPI_CPU/run/solverScoreP.o",
           "args": ARGS, "record": record, "rerun_444": rerun((4, 4, 4)), "rerun_111": rerun((1, 1, 1))}
    json.dump(out, open(os.path.join(HERE, "recorded_run.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
