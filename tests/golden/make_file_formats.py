#!/usr/bin/env python3
"""Fixtures of the reference's two output files for the C++ facade's cases.

The alpaka build writes residualHistory.txt (iterativeSolverBaseAlpaka.hpp:
620-638) and solution.dat (solverPoissonMPI_alpaka/src/main.cpp:135-146).
oracle/_ref/ref_solver (the reference's CPU solver classes, compiled here)
writes both for the cases tests/cpp/facade_main.cpp runs on the GPU (18^3
points, ds 1/17, origin (0.5, 0.25, 1.0), the reference's own test problem
u = sin x + cos y + 3 sin z + x^2 y z + x^2 + 10 of its solverSetup.hpp):

  files/cg_dir_18.*             BaseCG, no preconditioner, Dirichlet faces
  files/bicgstab_cheb11_18.*    BiCGSTAB + Chebyshev(11), bcs {0,1,1,0,1,0}

Run in the build container:  make -C oracle ref && python3 tests/golden/make_file_formats.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref", "ref_solver")
DS = repr(1.0 / 17)
COMMON = ["--ranks", "1", "1", "1", "--np", "18", "18", "18", "--ds", DS, DS, DS, "--origin", "0.5", "0.25", "1.0",
          "--problem", "0", "--tol", "1e-10"]
CASES = {
    "cg_dir_18": ["--bc", "0", "0", "0", "0", "0", "0", "--solver", "cg", "--precond", "none"],
    "bicgstab_cheb11_18": ["--bc", "0", "1", "1", "0", "1", "0", "--solver", "bicgstab", "--precond", "cheb",
                           "--cheb-m", "11", "--cheb-eps", "1e-4", "--cheb-rescale", "500", "0.9999"],
}


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: run `make -C oracle ref` first")
    out = os.path.join(HERE, "files")
    os.makedirs(out, exist_ok=True)
    for name, args in CASES.items():
        sol = os.path.join(out, name + ".solution.dat")
        if os.path.exists(sol):
            os.remove(sol)  # solution.dat is opened without truncation
        cmd = [REF, *COMMON, *args, "--solution-dat", sol, "--residual-history",
               os.path.join(out, name + ".residualHistory.txt")]
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError(p.stderr[-2000:])
        print(name, p.stdout.strip().splitlines()[-3:], flush=True)


if __name__ == "__main__":
    main()
