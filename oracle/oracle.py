"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py, always as the checker.  The product path
(parallelpoissonsolver_amd) never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

NONE, CG, BICGSTAB, CHEB, JACOBI, FCG, IR, USER = range(8)


class Grid(ctypes.Structure):
    _fields_ = [
        ("np", ctypes.c_int * 3),
        ("ng", ctypes.c_int * 3),
        ("sj", ctypes.c_long),
        ("sk", ctypes.c_long),
        ("ntot", ctypes.c_long),
        ("ds", ctypes.c_double * 3),
        ("origin", ctypes.c_double * 3),
        ("bc", ctypes.c_int * 6),
        ("lim_data", ctypes.c_int * 6),
        ("lim_solver", ctypes.c_int * 6),
        ("eig_local", ctypes.c_double * 2),
        ("eig_global", ctypes.c_double * 2),
    ]


USER_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(Grid),
                           ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float))


class Stage(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("maxit", ctypes.c_int),
        ("cheb_eps", ctypes.c_double),
        ("cheb_rmin", ctypes.c_double),
        ("cheb_rmax", ctypes.c_double),
        ("cheb_local", ctypes.c_int),
        ("user_fn", USER_FN),
        ("user", ctypes.c_void_p),
        ("ir_first_tol", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("ir_adapt", ctypes.c_double),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("precond_iterations", ctypes.c_long),
        ("outer_iterations", ctypes.c_int),
        ("error_from_iteration", ctypes.c_double),
        ("error_compute_operator", ctypes.c_double),
        ("norm_b", ctypes.c_double),
        ("history", ctypes.POINTER(ctypes.c_double)),
        ("history_cap", ctypes.c_int),
        ("history_len", ctypes.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        gp = ctypes.POINTER(Grid)
        L.og_grid_init.argtypes = [gp, ctypes.c_int * 3, ctypes.c_double * 3, ctypes.c_double * 3, ctypes.c_int * 6]
        L.og_stencil_point.argtypes = [gp, dp, ctypes.c_long]
        L.og_stencil_point.restype = ctypes.c_double
        L.og_apply_A.argtypes = [gp, dp, dp]
        L.og_apply_A_f32.argtypes = [gp, fp, fp]
        L.og_set_problem.argtypes = [gp, ctypes.c_int, dp, dp, dp * 6]
        L.og_face_size.argtypes = [gp, ctypes.c_int]
        L.og_face_size.restype = ctypes.c_long
        L.og_reset_neumann.argtypes = [gp, dp, ctypes.c_void_p, ctypes.c_double]
        L.og_adjust_b.argtypes = [gp, dp, dp, ctypes.c_void_p]
        L.og_residual.argtypes = [gp, dp, dp, dp]
        L.og_residual.restype = ctypes.c_double
        L.og_solve.argtypes = [gp, ctypes.POINTER(Stage), ctypes.c_int, dp, dp, ctypes.c_void_p, ctypes.POINTER(Result)]
        L.og_precondition.argtypes = [gp, ctypes.POINTER(Stage), ctypes.c_int, dp, dp, ctypes.POINTER(Result)]
        L.og_rhs_normal.argtypes = [gp, ctypes.c_uint64, dp]
        L.og_rhs_gauss.argtypes = [gp, ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_double, dp]
        L.og_rhs_spectrum.argtypes = [gp, ctypes.c_uint64, ctypes.c_int, dp]
        _lib = L
    return _lib


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class OracleGrid:
    """One-rank BlockGrid (blockGrid.hpp:11-367) with guarded arrays in (k, j, i) numpy order."""

    def __init__(self, npglobal, ds, origin=(0.0, 0.0, 0.0), bcs=(0, 0, 0, 0, 0, 0)):
        self.g = Grid()
        rc = lib().og_grid_init(ctypes.byref(self.g), (ctypes.c_int * 3)(*npglobal), (ctypes.c_double * 3)(*ds),
                                (ctypes.c_double * 3)(*origin), (ctypes.c_int * 6)(*bcs))
        if rc != 0:
            raise ValueError("bad grid")
        self.shape = (self.g.ng[2], self.g.ng[1], self.g.ng[0])

    @property
    def lim_solver(self):
        return list(self.g.lim_solver)

    @property
    def lim_data(self):
        return list(self.g.lim_data)

    def solver_slice(self):
        L = self.g.lim_solver
        return (slice(L[4], L[5]), slice(L[2], L[3]), slice(L[0], L[1]))

    def zeros(self, dtype=np.float64):
        return np.zeros(self.shape, dtype=dtype)

    def face_shape(self, f):
        d = f // 2
        ng = list(self.g.ng)
        if d == 0:
            return (ng[2], ng[1])
        if d == 1:
            return (ng[2], ng[0])
        return (ng[1], ng[0])

    def new_flux(self):
        return [np.zeros(self.face_shape(f)) for f in range(6)]

    def _flux_arg(self, flux):
        if flux is None:
            return None
        arr = (ctypes.POINTER(ctypes.c_double) * 6)()
        for f in range(6):
            arr[f] = _dp(flux[f]) if flux[f] is not None else ctypes.POINTER(ctypes.c_double)()
        return arr

    def set_problem(self, kind):
        x, b = self.zeros(), self.zeros()
        flux = self.new_flux()
        lib().og_set_problem(ctypes.byref(self.g), kind, _dp(x), _dp(b), self._flux_arg(flux))
        return x, b, flux

    def apply_A(self, x, out=None):
        y = self.zeros(x.dtype) if out is None else out
        if x.dtype == np.float32:
            lib().og_apply_A_f32(ctypes.byref(self.g), _fp(x), _fp(y))
        else:
            lib().og_apply_A(ctypes.byref(self.g), _dp(x), _dp(y))
        return y

    def reset_neumann(self, field, flux=None, norm=1.0):
        lib().og_reset_neumann(ctypes.byref(self.g), _dp(field), self._flux_arg(flux), norm)

    def adjust_b(self, x, b, flux=None):
        lib().og_adjust_b(ctypes.byref(self.g), _dp(x), _dp(b), self._flux_arg(flux))

    def residual(self, x, b):
        r = self.zeros()
        n = lib().og_residual(ctypes.byref(self.g), _dp(x), _dp(b), _dp(r))
        return n, r

    def rhs(self, kind, seed, **kw):
        b = self.zeros()
        if kind == "normal":
            lib().og_rhs_normal(ctypes.byref(self.g), seed, _dp(b))
        elif kind == "gauss":
            lib().og_rhs_gauss(ctypes.byref(self.g), seed, kw.get("nsrc", 64), kw.get("wmin", 0.01), kw.get("wmax", 0.1), _dp(b))
        elif kind == "spectrum":
            lib().og_rhs_spectrum(ctypes.byref(self.g), seed, kw.get("K", 16), _dp(b))
        else:
            raise ValueError(kind)
        return b

    @staticmethod
    def _stages(chain, keep):
        st = (Stage * len(chain))()
        for i, c in enumerate(chain):
            st[i].kind = c["kind"]
            st[i].tol = c.get("tol", 1e-10)
            st[i].maxit = c.get("maxit", 100000)
            st[i].cheb_eps = c.get("cheb_eps", 1e-4)
            st[i].cheb_rmin = c.get("cheb_rmin", 500.0)
            st[i].cheb_rmax = c.get("cheb_rmax", 1 - 1e-4)
            st[i].cheb_local = c.get("cheb_local", 0)
            st[i].ir_first_tol = c.get("first_tol", 0.0)
            st[i].precision = c.get("precision", 64)
            st[i].ir_adapt = c.get("adapt", 0.0)
            if "user_fn" in c:
                fn = USER_FN(c["user_fn"])
                keep.append(fn)
                st[i].user_fn = fn
        return st

    def precondition(self, chain, inp):
        """out = M(inp) for the preconditioner chain (chain[0] applied, chain[1..]
        nested); `inp` gets its Neumann ghosts reset in place."""
        keep = []
        st = self._stages(chain, keep)
        out = self.zeros()
        res = Result()
        rc = lib().og_precondition(ctypes.byref(self.g), st, len(chain), _dp(out), _dp(inp), ctypes.byref(res))
        if rc != 0:
            raise ValueError(f"og_precondition failed {rc}")
        return out, {"iterations": res.iterations, "precond_iterations": res.precond_iterations}

    def solve(self, chain, x, b, flux=None, history_cap=200000):
        """chain: list of dicts {kind, tol, maxit, cheb_eps, cheb_rmin, cheb_rmax, cheb_local, precision, user_fn}."""
        keep = []
        st = self._stages(chain, keep)
        hist = np.zeros(history_cap)
        res = Result()
        res.history = _dp(hist)
        res.history_cap = history_cap
        rc = lib().og_solve(ctypes.byref(self.g), st, len(chain), _dp(x), _dp(b), self._flux_arg(flux), ctypes.byref(res))
        if rc != 0:
            raise ValueError(f"og_solve failed {rc}")
        return {
            "iterations": res.iterations,
            "precond_iterations": res.precond_iterations,
            "outer_iterations": res.outer_iterations,
            "error_from_iteration": res.error_from_iteration,
            "error_compute_operator": res.error_compute_operator,
            "history": hist[: res.history_len].copy(),
        }
