"""The C ABI library loads and exports every entry point include/pps_abi.h declares
(no compute: these run without a GPU)."""
import os
import re
import subprocess

import pytest

from tests.conftest import ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "pps_abi.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pps_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared()
    for must in ("pps_create", "pps_destroy", "pps_solve", "pps_upload", "pps_download", "pps_apply_A",
                 "pps_ghost_fill", "pps_residual", "pps_dot", "pps_time_kernel"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from parallelpoissonsolver_amd import _lib
    L = _lib.lib()
    for name in declared():
        assert hasattr(L, name), name
    assert set(declared()) <= set(_lib.SIGNATURES)


def test_exported_symbols_are_c_linkage():
    lib = os.path.join(ROOT, "parallelpoissonsolver_amd", "libpps.so")
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    for name in declared():
        assert name in exported, name


def test_abi_version():
    from parallelpoissonsolver_amd import _lib
    assert _lib.lib().pps_abi_version() == 1


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_create_without_gpu_fails_loudly():
    from parallelpoissonsolver_amd import _lib
    with pytest.raises(_lib.PPSError) as e:
        _lib.Context((18, 18, 18), (0.1, 0.1, 0.1))
    assert "PPS_ERR_NOGPU" in str(e.value)


def test_create_rejects_bad_grids():
    """Argument checks run before any device call (solverPoissonMPI_alpaka/src/main.cpp:48-59
    aborts on a bad DIM or rank layout; the C ABI returns an error instead)."""
    from parallelpoissonsolver_amd import _lib
    with pytest.raises(_lib.PPSError) as e:
        _lib.Context((18, 18, 18), (0.1, 0.1, 0.1), nranks=(2, 1, 1))
    assert "PPS_ERR_UNSUPPORTED" in str(e.value)
    with pytest.raises(_lib.PPSError):
        _lib.Context((18, 18, 18), (0.1, 0.1, 0.1), guards=(2, 1, 1))
    with pytest.raises(_lib.PPSError):
        _lib.Context((18, 2, 18), (0.1, 0.1, 0.1))
    with pytest.raises(_lib.PPSError):
        _lib.Context((18, 18, 18), (0.1, 0.1, 0.1), bcs=(0, 2, 0, 0, 0, 0))
