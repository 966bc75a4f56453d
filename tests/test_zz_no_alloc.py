"""pps_residual_noalloc, the allocation-free variant of pps_residual: the same
r = b - A x and ||r|| on caller-owned dense device arrays, with its reduction
partials in the library's static device scratch (no hipMalloc, no context).
Checked against pps_residual on the same inputs, bit for bit, on two calls in
a row (the second on a smaller grid), and its size limit on the host."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import gpu_available

pps = pytest.importorskip("parallelpoissonsolver_amd")
L = pps._lib
ERR_ARG, ERR_UNSUPPORTED = -1, -4


def _fn():
    return L.lib().pps_residual_noalloc


def _grid(n, ds, bcs):
    g = L.CGrid()
    g.dim = 3
    g.nranks[:] = [1, 1, 1]
    g.my_rank = 0
    g.npglobal[:] = list(n)
    g.ds[:] = list(ds)
    g.origin[:] = [0.0, 0.0, 0.0]
    g.guards[:] = [1, 1, 1]
    g.bcs_type[:] = list(bcs)
    return g


def test_no_alloc_rejects_bad_arguments_on_the_host():
    """Null pointers and grids beyond the scratch (an x-y plane of more than
    131072 stencil tiles: 65536 x 8400 points is 1024 x 132 tiles of 64 x 64
    points and more with any smaller tile height) are refused before any
    device call."""
    f = _fn()
    g = _grid((18, 18, 18), (0.1,) * 3, (0,) * 6)
    assert f(ctypes.byref(g), None, 8, 8, None, None, None) == ERR_ARG
    big = _grid((65536, 8400, 3), (0.1,) * 3, (1,) * 6)
    rc = f(ctypes.byref(big), 8, 8, 8, None, None, None)
    assert rc == ERR_UNSUPPORTED
    assert b"131072 tiles" in L.lib().pps_last_error(None)


class Hip:
    """hipMalloc'ed test buffers from the HIP runtime libpps.so runs on (the
    caller owns its inputs and outputs; only the entry point allocates nothing)."""

    def __init__(self):
        L.lib()
        self.h = ctypes.CDLL("libamdhip64.so.7")  # the runtime libpps.so already loaded
        for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize", "hipStreamCreate",
                     "hipStreamDestroy", "hipStreamSynchronize"):
            getattr(self.h, name).restype = ctypes.c_int
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]

    def check(self, rc, what):
        assert rc == 0, f"{what}: hip error {rc}"

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        self.check(self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)), "hipMalloc")
        return p.value

    def h2d(self, dst, a):
        a = np.ascontiguousarray(a)
        self.check(self.h.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")

    def d2h(self, a, src):
        self.check(self.h.hipMemcpy(a.ctypes.data, src, a.nbytes, 2), "hipMemcpy D2H")


@pytest.mark.gpu
def test_no_alloc_residual_matches_pps_residual():
    assert gpu_available()
    hip = Hip()
    f = _fn()
    cases = [((66, 58, 50), (0.02, 0.03, 0.025), (0, 1, 1, 0, 1, 0), 1),
             ((34, 30, 26), (0.05, 0.04, 0.06), (0, 0, 0, 0, 0, 0), 2)]  # second call: a smaller grid
    for call, (n, ds, bcs, seed) in enumerate(cases):
        print(f"[no-alloc] call {call}: grid {n} bcs {bcs}", flush=True)
        G = O.OracleGrid(n, ds, (0, 0, 0), bcs)
        rng = np.random.default_rng(seed)
        x = rng.standard_normal(G.shape)
        b = rng.standard_normal(G.shape)
        C = L.Context(n, ds, (0, 0, 0), bcs)
        C.upload(L.FIELD_X, x)
        C.upload(L.FIELD_B, b)
        C.fill(L.FIELD_T0, 0.0)
        nrm_ref = C.residual(L.FIELD_T0, L.FIELD_X, L.FIELD_B)
        r_ref = C.download(L.FIELD_T0)
        C.close()
        xd, bd, rd, nd = hip.alloc(x.nbytes), hip.alloc(b.nbytes), hip.alloc(x.nbytes), hip.alloc(8)
        hip.h2d(xd, x)
        hip.h2d(bd, b)
        hip.h2d(rd, np.zeros_like(x))
        stream = ctypes.c_void_p()
        if call == 1:  # the second call on a stream of the caller's
            hip.check(hip.h.hipStreamCreate(ctypes.byref(stream)), "hipStreamCreate")
        nh = ctypes.c_double(-1.0)
        rc = f(ctypes.byref(_grid(n, ds, bcs)), xd, bd, rd, nd, ctypes.byref(nh), stream)
        assert rc == 0, L.lib().pps_last_error(None)
        hip.check(hip.h.hipDeviceSynchronize(), "sync")
        r = np.empty_like(x)
        hip.d2h(r, rd)
        nrm_dev = np.empty(1)
        hip.d2h(nrm_dev, nd)
        if stream.value:
            hip.h.hipStreamDestroy(stream)
        for p in (xd, bd, rd, nd):
            hip.h.hipFree(ctypes.c_void_p(p))
        sl = G.solver_slice()
        assert np.array_equal(r[sl], r_ref[sl])
        mask = np.ones(G.shape, bool)
        mask[sl] = False
        assert np.all(r[mask] == 0)  # only the solver region is written
        assert nh.value == nrm_ref
        assert nrm_dev[0] == nrm_ref
        # and the oracle's residual norm
        no, _ = G.residual(x, b)
        assert abs(nh.value - no) <= 1e-13 * no
