"""The oracle's solver paths under AddressSanitizer, LeakSanitizer and
UndefinedBehaviorSanitizer: oracle/sanitize_main.c drives CG, BiCGSTAB with
FP64 and FP32 Chebyshev, nested BiCGSTAB, main-loop Chebyshev, FCG and FP32
refinement on Dirichlet and mixed faces and ragged grids, built with gcc on
the CPU (the oracle is plain C99)."""
import os
import shutil
import subprocess
import tempfile

import pytest

from tests.conftest import ROOT


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not present")
def test_oracle_runs_clean_under_sanitizers():
    src = os.path.join(ROOT, "oracle")
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "og_sanitize")
        p = subprocess.run(["gcc", "-std=c99", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                            "-fno-sanitize-recover=undefined", "-I", src, os.path.join(src, "sanitize_main.c"),
                            os.path.join(src, "pps_oracle.c"), "-lm", "-o", exe], capture_output=True, text=True)
        assert p.returncode == 0, p.stderr
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
        r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
    assert r.stdout.count("rc 0") == 10
