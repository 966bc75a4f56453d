"""Iteration counts of the reference's own unpreconditioned BiCGSTAB on the
258^3 N(0,1) case (bicgstab_none_normal_258_r8 of ref_cases.json) over several
MPI rank layouts.  The layouts differ only in the order of the global sums, and
the counts show how far that order alone moves this solve: the GPU count is
judged against this spread.

    make -C oracle ref && python3 tests/golden/make_bicgstab_spread.py

Writes tests/golden/bicgstab_258_rank_spread.json.  Test infrastructure only.
"""
import json
import os
import subprocess
import sys
import tempfile

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import make_golden as M  # noqa: E402

LAYOUTS = [(2, 2, 2), (2, 2, 1), (4, 2, 1), (1, 1, 1)]
ARGS = M.unit(258) + M.BC_DIR + ["--problem", "2", "--rhs", "normal", "--seed", "2024", "--solver", "bicgstab",
                                  "--precond", "none"]


def main():
    out = {"case": "bicgstab_none_normal_258_r8", "args": ARGS, "tol": 1e-10, "runs": []}
    with tempfile.TemporaryDirectory() as tmp:
        for r in LAYOUTS:
            n = r[0] * r[1] * r[2]
            prefix = os.path.join(tmp, "b%d%d%d" % r)
            cmd = [M.REF, "--ranks", *map(str, r), *ARGS, "--tol", "1e-10", "--out", prefix]
            if n > 1:
                cmd = [M.MPIRUN, "-n", str(n)] + cmd
            subprocess.run(cmd, check=True, capture_output=True)
            m = json.load(open(prefix + ".json"))
            out["runs"].append({"ranks": list(r), "iterations": m["iterations"],
                                "error_compute_operator": m["error_compute_operator"]})
            print(r, m["iterations"], m["error_compute_operator"], flush=True)
            for f in os.listdir(tmp):
                os.remove(os.path.join(tmp, f))
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "bicgstab_258_rank_spread.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
