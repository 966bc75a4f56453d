"""ctypes binding of include/pps_abi.h (libpps.so, built in-tree for gfx950).

There is no fallback: if the HIP library is missing or no GPU is present the
calls raise, so a silent CPU path can never stand in for the device kernels.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpps.so")

PPS_OK = 0
ERRORS = {-1: "PPS_ERR_ARG", -2: "PPS_ERR_HIP", -3: "PPS_ERR_NOMEM", -4: "PPS_ERR_UNSUPPORTED", -5: "PPS_ERR_STATE",
          -6: "PPS_ERR_NOGPU"}

FIELD_X, FIELD_B, FIELD_T0, FIELD_T1, FIELD_T2, FIELD_T3 = range(6)
NONE, CG, BICGSTAB, CHEBYSHEV, JACOBI, FCG, IR, NN = range(8)
(K_CG_STENCIL, K_CG_UPDATE, K_RESIDUAL, K_COPY, K_CHEB_SWEEP, K_CG_STENCIL_F32, K_CG_UPDATE_F32, K_FCG_STENCIL,
 K_FCG_STENCIL_F32, K_COPY_FLAT, K_CHEB_PAIR) = range(11)


class PPSError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class CGrid(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("nranks", ctypes.c_int * 3),
        ("my_rank", ctypes.c_int),
        ("npglobal", ctypes.c_int * 3),
        ("ds", ctypes.c_double * 3),
        ("origin", ctypes.c_double * 3),
        ("guards", ctypes.c_int * 3),
        ("bcs_type", ctypes.c_int * 6),
        ("bcs_value", ctypes.c_double * 6),
    ]


class CStage(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("max_iter", ctypes.c_int),
        ("cheb_eps", ctypes.c_double),
        ("cheb_rescale_min", ctypes.c_double),
        ("cheb_rescale_max", ctypes.c_double),
        ("cheb_local", ctypes.c_int),
        ("stall_window", ctypes.c_int),
        ("stall_ratio", ctypes.c_double),
        ("ir_first_tol", ctypes.c_double),
        ("true_finish", ctypes.c_int),
        ("ir_adapt", ctypes.c_double),
    ]


class CResult(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("outer_iterations", ctypes.c_int),
        ("precond_iterations", ctypes.c_longlong),
        ("error_from_iteration", ctypes.c_double),
        ("error_compute_operator", ctypes.c_double),
        ("norm_b", ctypes.c_double),
        ("seconds", ctypes.c_double),
        ("fallback_switches", ctypes.c_int),
        ("history_len", ctypes.c_int),
    ]


class CNNTrainCfg(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("nrhs", ctypes.c_int),
        ("snapshots", ctypes.c_int),
        ("kmax", ctypes.c_int),
        ("seed", ctypes.c_ulonglong),
        ("batch", ctypes.c_int),
        ("steps", ctypes.c_int),
        ("lr", ctypes.c_double),
        ("lr_final", ctypes.c_double),
        ("white_fraction", ctypes.c_double),
    ]


class CNNTrainStats(ctypes.Structure):
    _fields_ = [
        ("pairs", ctypes.c_longlong),
        ("steps", ctypes.c_int),
        ("seconds", ctypes.c_double),
        ("loss_first", ctypes.c_double),
        ("loss_last", ctypes.c_double),
    ]


# every symbol declared in include/pps_abi.h, with its ctypes signature
_VP = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_FP = ctypes.POINTER(ctypes.c_float)
_LLP = ctypes.POINTER(ctypes.c_longlong)
SIGNATURES = {
    "pps_abi_version": ([], ctypes.c_int),
    "pps_create": ([ctypes.POINTER(CGrid), ctypes.POINTER(_VP)], ctypes.c_int),
    "pps_destroy": ([_VP], None),
    "pps_last_error": ([_VP], ctypes.c_char_p),
    "pps_get_extent": ([_VP, _IP], ctypes.c_int),
    "pps_upload": ([_VP, ctypes.c_int, _DP], ctypes.c_int),
    "pps_download": ([_VP, ctypes.c_int, _DP], ctypes.c_int),
    "pps_fill": ([_VP, ctypes.c_int, ctypes.c_double], ctypes.c_int),
    "pps_copy_field": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "pps_set_neumann_flux": ([_VP, ctypes.c_int, _DP], ctypes.c_int),
    "pps_solve": ([_VP, ctypes.POINTER(CStage), ctypes.c_int, ctypes.POINTER(CResult)], ctypes.c_int),
    "pps_precondition": ([_VP, ctypes.POINTER(CStage), ctypes.c_int, ctypes.c_int, ctypes.c_int, _LLP], ctypes.c_int),
    "pps_get_history": ([_VP, _DP, ctypes.c_int], ctypes.c_int),
    "pps_synchronize": ([_VP], ctypes.c_int),
    "pps_apply_A": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "pps_ghost_fill": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_double], ctypes.c_int),
    "pps_residual": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _DP], ctypes.c_int),
    "pps_residual_noalloc": ([ctypes.POINTER(CGrid), _VP, _VP, _VP, _VP, _DP, _VP], ctypes.c_int),
    "pps_dot": ([_VP, ctypes.c_int, ctypes.c_int, _DP], ctypes.c_int),
    "pps_norm_b_adjusted": ([_VP, _DP], ctypes.c_int),
    "pps_rhs_normal": ([_VP, ctypes.c_uint64], ctypes.c_int),
    "pps_rhs_gauss": ([_VP, ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_double], ctypes.c_int),
    "pps_rhs_spectrum": ([_VP, ctypes.c_uint64, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "pps_time_kernel": ([_VP, ctypes.c_int, ctypes.c_int, _DP, _DP], ctypes.c_int),
    "pps_set_tiling": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "pps_set_cheb_pairs": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "pps_set_deferred_x": ([_VP, ctypes.c_int], ctypes.c_int),
    "pps_nn_create": ([ctypes.c_uint64, ctypes.POINTER(_VP)], ctypes.c_int),
    "pps_nn_destroy": ([_VP], None),
    "pps_nn_last_error": ([_VP], ctypes.c_char_p),
    "pps_nn_num_params": ([_VP, _LLP, _LLP], ctypes.c_int),
    "pps_nn_get_params": ([_VP, _FP, ctypes.c_longlong], ctypes.c_int),
    "pps_nn_set_params": ([_VP, _FP, ctypes.c_longlong], ctypes.c_int),
    "pps_nn_forward": ([_VP, _FP, ctypes.c_int, _IP, _FP], ctypes.c_int),
    "pps_nn_activation": ([_VP, ctypes.c_int, _FP], ctypes.c_int),
    "pps_nn_loss_grad": ([_VP, _FP, _DP, ctypes.c_int, _IP, _DP, _FP], ctypes.c_int),
    "pps_nn_generate": ([_VP, ctypes.POINTER(CNNTrainCfg), _LLP], ctypes.c_int),
    "pps_nn_dataset": ([_VP, ctypes.c_longlong, _FP, _DP], ctypes.c_int),
    "pps_nn_train": ([_VP, ctypes.POINTER(CNNTrainCfg), ctypes.POINTER(CNNTrainStats)], ctypes.c_int),
    "pps_nn_apply_field": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "pps_nn_apply_field_prec": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _DP], ctypes.c_int),
    "pps_nn_apply_update": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                             ctypes.c_double, ctypes.c_int], ctypes.c_int),
    "pps_nn_time_layer": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _DP, _DP, _DP], ctypes.c_int),
    "pps_nn_attach": ([_VP, _VP], ctypes.c_int),
}

_lib = None


def lib():
    """Load libpps.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libpps.so not built at {LIB_PATH}: run `python3 -m parallelpoissonsolver_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_DP)


class Context:
    """One device context (pps_create): fields X, B (+ scratch T0..T3) on the GPU."""

    def __init__(self, npglobal, ds, origin=(0.0, 0.0, 0.0), bcs=(0, 0, 0, 0, 0, 0), nranks=(1, 1, 1), my_rank=0,
                 guards=(1, 1, 1)):
        g = CGrid()
        g.dim = 3
        g.nranks[:] = list(nranks)
        g.my_rank = my_rank
        g.npglobal[:] = [int(v) for v in npglobal]
        g.ds[:] = [float(v) for v in ds]
        g.origin[:] = [float(v) for v in origin]
        g.guards[:] = list(guards)
        g.bcs_type[:] = [int(v) for v in bcs]
        self.grid = g
        self.h = _VP()
        rc = lib().pps_create(ctypes.byref(g), ctypes.byref(self.h))
        if rc != 0:
            raise PPSError(rc, lib().pps_last_error(None).decode())
        self.npglobal = tuple(int(v) for v in npglobal)
        self.ng = tuple(n + 2 for n in self.npglobal)
        self.shape = (self.ng[2], self.ng[1], self.ng[0])
        self.bcs = tuple(int(v) for v in bcs)

    def close(self):
        if self.h:
            lib().pps_destroy(self.h)
            self.h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc < 0:
            raise PPSError(rc, lib().pps_last_error(self.h).decode())
        return rc

    def _arr(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != self.shape[0] * self.shape[1] * self.shape[2]:
            raise ValueError(f"field must have {self.shape} guarded points")
        return a

    def upload(self, field, a):
        a = self._arr(a)
        self._chk(lib().pps_upload(self.h, field, _dp(a)))

    def download(self, field):
        a = np.empty(self.shape)
        self._chk(lib().pps_download(self.h, field, _dp(a)))
        return a

    def fill(self, field, v):
        self._chk(lib().pps_fill(self.h, field, float(v)))

    def copy_field(self, dst, src):
        self._chk(lib().pps_copy_field(self.h, dst, src))

    def set_flux(self, face, plane):
        plane = np.ascontiguousarray(plane, dtype=np.float64)
        self._chk(lib().pps_set_neumann_flux(self.h, face, _dp(plane)))

    @staticmethod
    def _stages(chain):
        st = (CStage * len(chain))()
        for i, c in enumerate(chain):
            st[i].kind = c["kind"]
            st[i].precision = c.get("precision", 64)
            st[i].tol = c.get("tol", 1e-10)
            st[i].max_iter = c.get("maxit", 100000)
            st[i].cheb_eps = c.get("cheb_eps", 1e-4)
            st[i].cheb_rescale_min = c.get("cheb_rmin", 500.0)
            st[i].cheb_rescale_max = c.get("cheb_rmax", 1 - 1e-4)
            st[i].cheb_local = c.get("cheb_local", 0)
            st[i].stall_window = c.get("stall_window", 0)
            st[i].stall_ratio = c.get("stall_ratio", 0.0)
            st[i].ir_first_tol = c.get("first_tol", 0.0)
            st[i].true_finish = c.get("true_finish", 0)
            st[i].ir_adapt = c.get("adapt", 0.0)
        return st

    def precondition(self, chain, out, inp):
        """out = M(inp) for the preconditioner chain (pps_precondition); returns
        the iterations spent."""
        it = ctypes.c_longlong()
        self._chk(lib().pps_precondition(self.h, self._stages(chain), len(chain), out, inp, ctypes.byref(it)))
        return it.value

    def solve(self, chain):
        st = self._stages(chain)
        res = CResult()
        self._chk(lib().pps_solve(self.h, st, len(chain), ctypes.byref(res)))
        out = {k: getattr(res, k) for k, _ in CResult._fields_}
        n = res.history_len
        hist = np.zeros(max(n, 1))
        if n > 0:
            self._chk(lib().pps_get_history(self.h, _dp(hist), n))
        out["history"] = hist[:n]
        return out

    def apply_A(self, y, x):
        self._chk(lib().pps_apply_A(self.h, y, x))

    def ghost_fill(self, field, with_flux=False, norm=1.0):
        self._chk(lib().pps_ghost_fill(self.h, field, int(with_flux), float(norm)))

    def residual(self, r, x, b):
        v = ctypes.c_double()
        self._chk(lib().pps_residual(self.h, r, x, b, ctypes.byref(v)))
        return v.value

    def dot(self, a, b):
        v = ctypes.c_double()
        self._chk(lib().pps_dot(self.h, a, b, ctypes.byref(v)))
        return v.value

    def norm_b_adjusted(self):
        v = ctypes.c_double()
        self._chk(lib().pps_norm_b_adjusted(self.h, ctypes.byref(v)))
        return v.value

    def rhs_normal(self, seed):
        self._chk(lib().pps_rhs_normal(self.h, seed))

    def rhs_gauss(self, seed, nsrc=64, wmin=0.01, wmax=0.1):
        self._chk(lib().pps_rhs_gauss(self.h, seed, nsrc, wmin, wmax))

    def rhs_spectrum(self, seed, kmax=16, accumulate=False):
        self._chk(lib().pps_rhs_spectrum(self.h, seed, kmax, int(accumulate)))

    def time_kernel(self, kernel, reps=20):
        ms, nbytes = ctypes.c_double(), ctypes.c_double()
        self._chk(lib().pps_time_kernel(self.h, kernel, reps, ctypes.byref(ms), ctypes.byref(nbytes)))
        return ms.value, nbytes.value

    def nn_apply(self, out, inp, prec=None, dot=None):
        """out = M(inp) with the attached network.  prec None: FP64 fields
        (pps_nn_apply_field); 64 or 32: the network on FP64 fields or on FP32
        copies (pps_nn_apply_field_prec).  With a dot field, the output layer
        also sums out . dot in its epilogue, as FCG's gamma; returned."""
        if prec is None and dot is None:
            self._chk(lib().pps_nn_apply_field(self.h, out, inp))
            return None
        g = ctypes.c_double(0.0)
        self._chk(lib().pps_nn_apply_field_prec(self.h, out, inp, 64 if prec is None else prec,
                                                -1 if dot is None else dot, ctypes.byref(g)))
        return None if dot is None else g.value

    def nn_apply_update(self, out, r_prev, ap_prev, r_cur, alpha, rr, prec=64):
        """out = M(r_cur), r_cur = r_prev - alpha * ap_prev formed inside the
        network's first layer (FCG's fused residual update); rr = ||r_cur||^2."""
        self._chk(lib().pps_nn_apply_update(self.h, out, r_prev, ap_prev, r_cur, float(alpha), float(rr), prec))

    def nn_time_layer(self, layer, reps=10, prec=32):
        """(ms per launch, algorithmic bytes, flops) of one layer of the attached network,
        on FP32 (the mixed-precision solve's) or FP64 fields."""
        ms, nb, fl = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._chk(lib().pps_nn_time_layer(self.h, layer, prec, reps, ctypes.byref(ms), ctypes.byref(nb),
                                          ctypes.byref(fl)))
        return ms.value, nb.value, fl.value

    def set_deferred_x(self, on=True):
        """CG and learned-FCG x updates deferred over a ring of 16 directions (True, the default, where the
        device holds it) or applied every iteration (False)."""
        self._chk(lib().pps_set_deferred_x(self.h, int(bool(on))))

    def set_cheb_pairs(self, pairs=1, kz=0):
        """Chebyshev sweeps two per pass (1, the default) or one per pass (0); kz: z-chunk of the fused kernel."""
        self._chk(lib().pps_set_cheb_pairs(self.h, int(pairs), int(kz)))

    def set_tiling(self, kz_stencil=0, nontemporal=0):
        """Stencil z-chunk (0 keeps it); output stores: 0 automatic, 1 non-temporal, -1 plain."""
        self._chk(lib().pps_set_tiling(self.h, kz_stencil, nontemporal))

    def synchronize(self):
        self._chk(lib().pps_synchronize(self.h))


def _fp(a):
    return a.ctypes.data_as(_FP)


class Network:
    """Learned multilevel preconditioner (pps_nn_*): 14 BF16 MFMA convolutions
    over 3 levels, trained on the GPU; attach it to a Context to use PPS_NN
    stages (below an FCG stage)."""

    N_ACTS = 13
    LEVEL = (0, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 1, 1)  # level of activation i

    def __init__(self, seed=0):
        self.h = _VP()
        rc = lib().pps_nn_create(seed, ctypes.byref(self.h))
        if rc != 0:
            raise PPSError(rc, lib().pps_nn_last_error(None).decode())
        st, eff = ctypes.c_longlong(), ctypes.c_longlong()
        lib().pps_nn_num_params(self.h, ctypes.byref(st), ctypes.byref(eff))
        self.n_params, self.n_effective = st.value, eff.value

    def close(self):
        if self.h:
            lib().pps_nn_destroy(self.h)
            self.h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc < 0:
            raise PPSError(rc, lib().pps_nn_last_error(self.h).decode())
        return rc

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        self._chk(lib().pps_nn_get_params(self.h, _fp(p), p.size))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._chk(lib().pps_nn_set_params(self.h, _fp(p), p.size))

    @staticmethod
    def _dims(x):
        if x.ndim != 4:
            raise ValueError("batch must be [B][n2][n1][n0]")
        return (ctypes.c_int * 3)(x.shape[3], x.shape[2], x.shape[1])

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty_like(x)
        self._chk(lib().pps_nn_forward(self.h, _fp(x), x.shape[0], self._dims(x), _fp(out)))
        self._last = x.shape
        return out

    def activation(self, i):
        B, n2, n1, n0 = self._last
        d = [n2, n1, n0]
        for _ in range(self.LEVEL[i]):
            d = [(v + 1) // 2 for v in d]
        out = np.empty((B, *d, 16), np.float32)
        self._chk(lib().pps_nn_activation(self.h, i, _fp(out)))
        return out

    def loss_grad(self, x, g):
        x = np.ascontiguousarray(x, np.float32)
        g = np.ascontiguousarray(g, np.float64)
        loss = ctypes.c_double()
        grad = np.empty(self.n_params, np.float32)
        self._chk(lib().pps_nn_loss_grad(self.h, _fp(x), _dp(g), x.shape[0], self._dims(x), ctypes.byref(loss),
                                         _fp(grad)))
        self._last = x.shape
        return loss.value, grad

    @staticmethod
    def _cfg(n=64, nrhs=256, snapshots=16, kmax=32, seed=1, batch=16, steps=1000, lr=1e-3, lr_final=1e-5,
             white_fraction=0.0):
        c = CNNTrainCfg()
        c.n, c.nrhs, c.snapshots, c.kmax, c.seed = n, nrhs, snapshots, kmax, seed
        c.batch, c.steps, c.lr, c.lr_final = batch, steps, lr, lr_final
        c.white_fraction = white_fraction
        return c

    def generate(self, **kw):
        c = self._cfg(**kw)
        pairs = ctypes.c_longlong()
        self._chk(lib().pps_nn_generate(self.h, ctypes.byref(c), ctypes.byref(pairs)))
        self._gen_n = c.n
        return pairs.value

    def dataset(self, index):
        n = self._gen_n
        r = np.empty((n, n, n), np.float32)
        g = ctypes.c_double()
        self._chk(lib().pps_nn_dataset(self.h, index, _fp(r), ctypes.byref(g)))
        return r, g.value

    def train(self, **kw):
        c = self._cfg(**kw)
        st = CNNTrainStats()
        self._chk(lib().pps_nn_train(self.h, ctypes.byref(c), ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in CNNTrainStats._fields_}

    def attach(self, ctx):
        self._chk(lib().pps_nn_attach(ctx.h, self.h))
        ctx._nn = self  # keep the network alive while the context uses it
