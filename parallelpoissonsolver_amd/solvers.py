"""Host-side mirror of the reference's solver interface, driving libpps.

The reference composes its solver at compile time (inputParam.hpp:22-37):

    using T_Solver = BiCGSTAB<DIM, T_data, tollMainSolver, iterMaxMainSolver,
                              isbiCGMainLoop1, communicationON, T_Preconditioner2>;
    T_Solver solver(blockGrid, exactSolutionAndBCs, communicator);
    solver.setProblem(fieldX, fieldB);
    solver(fieldX, fieldB, operatorA);
    solver.getNumIterationFinal(); ...

Here the same composition is written with spec objects that carry the template
parameters, and calling a spec with the grid objects gives the solver:

    T_Solver = BiCGSTAB(DIM, T_data, tollMainSolver, iterMaxMainSolver, True, True,
                        ChebyshevIteration(DIM, T_data, tollPreconditionerSolver, chebyshevMax,
                                           False, False, NoneSolver(DIM, T_data, 1, 1)))
    solver = T_Solver(blockGrid, exactSolutionAndBCs, communicator)
    solver.setProblem(fieldX, fieldB)
    solver(fieldX, fieldB)

Names, argument meaning and result getters follow solverPoissonMPI_CPU/include
and solverPoissonMPI_alpaka/include.  Every solve runs on the GPU through the C
ABI (include/pps_abi.h); there is no host fallback.
"""
import math

import numpy as np

from . import _lib

# ---- configuration constants of solverSetup.hpp (CPU: :27-40, alpaka: :55-73)
DIM = 3
PI = 3.141592653589793
tollScalingFactor = 1e-10
orderNeumanBcs = 2
trackErrorFromIterationHistory = True
epsilon = 1e-4
rescaleEigMin = 500.0
rescaleEigMax = 1 - 1e-4


class BlockGrid:
    """BlockGrid<DIM,T_data> (blockGrid.hpp:11-367) for one GPU (rank grid 1x1x1)."""

    def __init__(self, nranks, my_rank, npglobal, ds, origin, guards, bcsType, bcsValue=(0, 0, 0, 0, 0, 0)):
        if int(nranks[0]) * int(nranks[1]) * int(nranks[2]) != 1 or my_rank != 0:
            raise ValueError("one GPU: the rank grid must be 1x1x1 (communicationMPI halos become device ghost fills)")
        if tuple(guards) != (1, 1, 1):
            raise ValueError("guards must be {1,1,1} for the 7-point stencil")
        self.nranks_ = tuple(int(v) for v in nranks)
        self.my_rank_ = int(my_rank)
        self.npglobal_ = tuple(int(v) for v in npglobal)
        self.ds_ = tuple(float(v) for v in ds)
        self.origin_ = tuple(float(v) for v in origin)
        self.guards_ = tuple(int(v) for v in guards)
        self.bcsType_ = tuple(int(v) for v in bcsType)
        self.bcsValue_ = tuple(float(v) for v in bcsValue)
        n = self.npglobal_
        self.indexLimitsData_ = (1, n[0] + 1, 1, n[1] + 1, 1, n[2] + 1)
        ls = list(self.indexLimitsData_)
        for d in range(3):
            if self.bcsType_[2 * d] == 0:
                ls[2 * d] += 1
            if self.bcsType_[2 * d + 1] == 0:
                ls[2 * d + 1] -= 1
        self.indexLimitsSolver_ = tuple(ls)
        emin = emax = gmin = gmax = 0.0
        for d in range(3):
            cnt = ls[2 * d + 1] - ls[2 * d]
            dd = self.ds_[d] * self.ds_[d]
            emin += 4 * math.sin(1 * PI / 2 / (cnt + 1)) ** 2 / dd
            emax += 4 * math.sin(cnt * PI / 2 / (cnt + 1)) ** 2 / dd
            ni = n[d] - (self.bcsType_[2 * d] == 0) - (self.bcsType_[2 * d + 1] == 0)
            gmin += 4 * math.sin(1 * PI / 2 / (ni + 1)) ** 2 / dd
            gmax += 4 * math.sin(ni * PI / 2 / (ni + 1)) ** 2 / dd
        self.eigenValuesLocal_ = (emin, emax)
        self.eigenValuesGlobal_ = (gmin, gmax)

    def getNranks(self): return self.nranks_
    def getMyrank(self): return self.my_rank_
    def getNpglobal(self): return self.npglobal_
    def getDs(self): return self.ds_
    def getOrigin(self): return self.origin_
    def getGuards(self): return self.guards_
    def getGlobalLocation(self): return (0, 0, 0)
    def getNlocalNoGuards(self): return self.npglobal_
    def getNlocalGuards(self): return tuple(v + 2 for v in self.npglobal_)
    def getIndexLimitsData(self): return self.indexLimitsData_
    def getIndexLimitsSolver(self): return self.indexLimitsSolver_
    def getNumCommunication(self): return 0
    def getNtotLocalGuards(self): return int(np.prod(self.getNlocalGuards()))
    def getNtotLocalNoGuards(self): return int(np.prod(self.npglobal_))
    def getHasBoundary(self): return (True,) * 6
    def getHasCommunication(self): return (False,) * 6
    def getBcsType(self): return self.bcsType_
    def getBcsValue(self): return self.bcsValue_
    def getEigenValuesLocal(self): return self.eigenValuesLocal_
    def getEigenValuesGlobal(self): return self.eigenValuesGlobal_
    def getNtotNpglobal(self): return int(np.prod(self.npglobal_))

    # guarded coordinate axes (iterativeSolverBase.hpp:547-549)
    def axes(self):
        return [self.origin_[d] + (np.arange(self.npglobal_[d] + 2) - 1) * self.ds_[d] for d in range(3)]

    def shape(self):
        ng = self.getNlocalGuards()
        return (ng[2], ng[1], ng[0])


class ExactSolutionAndBCs:
    """Test problem of solverSetup.hpp (CPU :48-113, alpaka :95-202); subclass to change it.
    The methods take numpy arrays and evaluate point-wise."""

    def setFieldB(self, x, y, z):
        return -np.sin(x) - np.cos(y) - 3 * np.sin(z) + 2 * y * z + 2

    def trueSolutionFxyz(self, x, y, z):
        return np.sin(x) + np.cos(y) + 3 * np.sin(z) + x * x * y * z + x * x + 10

    def trueSolutionDdir(self, x, y, z, dir):
        if dir == 0:
            return np.cos(x) + 2 * x * y * z + 2 * x
        if dir == 1:
            return -np.sin(y) + x * x * z
        return 3 * np.cos(z) + x * x * y


class CommunicatorMPI:
    """CommunicatorMPI (communicationMPI.hpp): with one GPU there are no
    neighbours; the ghost layers are filled on the device by the solver."""

    def __init__(self, blockGrid):
        self.blockGrid = blockGrid

    def __call__(self, data):
        return None

    def waitAllandCheckRcv(self):
        return None


class MatrixFreeOperatorA:
    """MatrixFreeOperatorA (matrixFreeOperatorA.hpp): the 7-point +Laplacian; the
    solvers apply it on the device."""

    def __init__(self, blockGrid):
        self.blockGrid = blockGrid


# --------------------------------------------------------------- solver specs

def _dtype_bits(t):
    if t in (np.float32, "float", "float32", 32):
        return 32
    return 64


class SolverSpec:
    """The template arguments of one solver type of the chain."""

    def __init__(self, kind, params, pre=None):
        self.kind = kind
        self.params = params
        self.pre = pre

    def chain(self):
        stages, s = [], self
        while s is not None:
            stages.append(s.stage())
            s = s.pre
        return stages

    def stage(self):
        p = self.params
        st = {"kind": self.kind, "precision": p.get("precision", 64), "tol": p.get("tol", 1e-10),
              "maxit": p.get("maxit", 1)}
        for k in ("cheb_eps", "cheb_rmin", "cheb_rmax", "cheb_local", "stall_window", "stall_ratio", "first_tol",
                  "adapt"):
            if k in p:
                st[k] = p[k]
        return st

    def __call__(self, blockGrid, exactSolutionAndBCs, communicatorMPI=None, alpakaHelper=None):
        if not self.params.get("isMainLoop", True):
            raise ValueError("the outer solver of a chain must be the main loop (isMainLoop = true)")
        return Solver(self, blockGrid, exactSolutionAndBCs)


def NoneSolver(DIM=3, T_data=np.float64, tolerance=1, maxIteration=1):
    """NoneSolver<DIM,T_data,tolerance,maxIteration> (noneSolver.hpp:14-28)."""
    return SolverSpec(_lib.NONE, {"precision": _dtype_bits(T_data)})


def Jacobi(DIM=3, T_data=np.float64):
    """Diagonal preconditioner (no reference counterpart)."""
    return SolverSpec(_lib.JACOBI, {"precision": _dtype_bits(T_data)})


def BaseCG(DIM, T_data, tolerance, maxIteration, isMainLoop, communicationON, T_Preconditioner):
    """BaseCG<...> (baseCG.hpp:14-270)."""
    return SolverSpec(_lib.CG, {"precision": _dtype_bits(T_data), "tol": tolerance * tollScalingFactor,
                                "maxit": int(maxIteration), "isMainLoop": isMainLoop}, T_Preconditioner)


def BiCGSTAB(DIM, T_data, tolerance, maxIteration, isMainLoop, communicationON, T_Preconditioner):
    """BiCGSTAB<...> (BiCGSTAB.hpp:14-335; alpaka BiCGstabAlpaka.hpp:19-888)."""
    return SolverSpec(_lib.BICGSTAB, {"precision": _dtype_bits(T_data), "tol": tolerance * tollScalingFactor,
                                      "maxit": int(maxIteration), "isMainLoop": isMainLoop}, T_Preconditioner)


def ChebyshevIteration(DIM, T_data, tolerance, maxIteration, isMainLoop, communicationON, T_Preconditioner,
                       T_data_chebyshev=None, global_=True, eps=None, rescale_min=None, rescale_max=None):
    """ChebyshevIteration<...> (chebyshevIteration.hpp:15-157); the alpaka form adds
    T_data_chebyshev and global (chebyshevIterationAlpaka.hpp:16)."""
    prec = _dtype_bits(T_data_chebyshev if T_data_chebyshev is not None else T_data)
    return SolverSpec(_lib.CHEBYSHEV, {
        "precision": prec, "tol": tolerance * tollScalingFactor, "maxit": int(maxIteration), "isMainLoop": isMainLoop,
        "cheb_eps": epsilon if eps is None else eps,
        "cheb_rmin": rescaleEigMin if rescale_min is None else rescale_min,
        "cheb_rmax": rescaleEigMax if rescale_max is None else rescale_max,
        "cheb_local": 0 if global_ else 1}, T_Preconditioner)


def FCG(DIM, T_data, tolerance, maxIteration, isMainLoop, communicationON, T_Preconditioner):
    """Flexible CG, truncation 1 (no reference counterpart)."""
    return SolverSpec(_lib.FCG, {"precision": _dtype_bits(T_data), "tol": tolerance * tollScalingFactor,
                                 "maxit": int(maxIteration), "isMainLoop": isMainLoop}, T_Preconditioner)


def MixedPrecisionIR(tolerance, maxOuter, inner, first_inner_tol=0.0, adapt=0.0):
    """FP64 outer refinement r = b - A u, u += d around an FP32 inner solver
    spec (BaseCG or FCG with T_data float32).  first_inner_tol > 0: the inner
    tolerance of the first outer step.  adapt in (0, 1): an outer step asks the
    inner solver for no more reduction than the outer tolerance still needs,
    times adapt (pps_stage::ir_adapt).  No reference counterpart."""
    return SolverSpec(_lib.IR, {"precision": 64, "tol": tolerance * tollScalingFactor, "maxit": int(maxOuter),
                                "isMainLoop": True, "first_tol": float(first_inner_tol), "adapt": float(adapt)},
                      inner)


def LearnedPreconditioner(T_data=np.float32, stall_window=0, stall_ratio=0.0):
    """Learned multilevel preconditioner held by the solver's device context."""
    return SolverSpec(_lib.NN, {"precision": _dtype_bits(T_data), "stall_window": stall_window,
                                "stall_ratio": stall_ratio})


# -------------------------------------------------------------------- solver

class Solver:
    """An instantiated T_Solver bound to a grid and a device context."""

    def __init__(self, spec, blockGrid, exact):
        self.spec = spec
        self.blockGrid = blockGrid
        self.exact = exact
        self._ctx = None
        self.numIterationFinal_ = 0
        self.numIterationPreconditionerFinal_ = 0
        self.errorFromIteration_ = -1.0
        self.errorComputeOperator_ = -1.0
        self.durationSolver_ = 0.0
        self.normFieldB_ = 1.0
        self.errorFromIterationHistory_ = np.zeros(0)
        self.outerIterations_ = 0

    @property
    def ctx(self):
        """Device context, created on first use (the solve needs the GPU; setProblem does not)."""
        if self._ctx is None:
            g = self.blockGrid
            self._ctx = _lib.Context(g.getNpglobal(), g.getDs(), g.getOrigin(), g.getBcsType())
        return self._ctx

    # iterativeSolverBase.hpp:51-55,537-603
    def setProblem(self, fieldX, fieldB):
        g = self.blockGrid
        X, B = self._view(fieldX), self._view(fieldB)
        ax = g.axes()
        L = g.getIndexLimitsData()
        Z, Y, Xc = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        sl = (slice(L[4], L[5]), slice(L[2], L[3]), slice(L[0], L[1]))
        B[sl] = self.exact.setFieldB(Xc[sl], Y[sl], Z[sl])
        bc = g.getBcsType()
        for d in range(3):
            for side in range(2):
                if bc[2 * d + side] != 0:
                    continue
                s = list(sl)
                plane = L[2 * d] if side == 0 else L[2 * d + 1] - 1
                s[2 - d] = slice(plane, plane + 1)
                s = tuple(s)
                X[s] = self.exact.trueSolutionFxyz(Xc[s], Y[s], Z[s])

    def _view(self, a):
        shp = self.blockGrid.shape()
        if a.dtype != np.float64 or not a.flags.c_contiguous:
            raise TypeError("fields must be C-contiguous float64 arrays")
        return a.reshape(shp)

    def _flux_planes(self):
        g = self.blockGrid
        ax = g.axes()
        L = g.getIndexLimitsData()
        planes = {}
        for f in range(6):
            if g.getBcsType()[f] != 1:
                continue
            d = f // 2
            pos = L[2 * d] if f % 2 == 0 else L[2 * d + 1] - 1
            others = [e for e in range(3) if e != d]  # lower direction fastest
            ng = g.getNlocalGuards()
            plane = np.zeros((ng[others[1]], ng[others[0]]))
            c = [None, None, None]
            c[d] = np.full((1, 1), ax[d][pos])
            B_, A_ = np.meshgrid(ax[others[1]], ax[others[0]], indexing="ij")
            c[others[0]] = A_
            c[others[1]] = B_
            vals = self.exact.trueSolutionDdir(c[0], c[1], c[2], d)
            vals = np.broadcast_to(vals, plane.shape)
            s1 = slice(L[2 * others[1]], L[2 * others[1] + 1])
            s0 = slice(L[2 * others[0]], L[2 * others[0] + 1])
            plane[s1, s0] = vals[s1, s0]
            planes[f] = plane
        return planes

    # BiCGstabAlpaka::operator()(T_data*, T_data*) (BiCGstabAlpaka.hpp:68-475): copy in,
    # solve on the device, copy x back
    def __call__(self, fieldX, fieldB, operatorA=None):
        X = self._view(fieldX)
        self.ctx.upload(_lib.FIELD_X, X)
        self.ctx.upload(_lib.FIELD_B, self._view(fieldB))
        for f, plane in self._flux_planes().items():
            self.ctx.set_flux(f, plane)
        res = self.solve_device()
        X[...] = self.ctx.download(_lib.FIELD_X)
        return res

    def solve_device(self):
        """Solve with X and B already resident on the device."""
        res = self.ctx.solve(self.spec.chain())
        self.numIterationFinal_ = res["iterations"]
        self.numIterationPreconditionerFinal_ = res["precond_iterations"]
        self.errorFromIteration_ = res["error_from_iteration"]
        self.errorComputeOperator_ = res["error_compute_operator"]
        self.durationSolver_ = res["seconds"]
        self.normFieldB_ = res["norm_b"]
        self.outerIterations_ = res["outer_iterations"]
        self.errorFromIterationHistory_ = res["history"]
        self.last = res
        return res

    def getNumIterationFinal(self): return self.numIterationFinal_
    def getNumIterationPreconditionerFinal(self): return self.numIterationPreconditionerFinal_
    def getErrorFromIteration(self): return self.errorFromIteration_
    def getErrorComputeOperator(self): return self.errorComputeOperator_
    def getDurationSolver(self): return self.durationSolver_
    def getErrorFromIterationHistory(self): return self.errorFromIterationHistory_

    # iterativeSolverBase.hpp:283-408 (one rank): prints the largest block-average
    # and point errors against trueSolutionFxyz on the data region and returns
    # the sum of |x - u| (errorLocal)
    def checkSolutionLocalGlobal(self, fieldX, verbose=True):
        g = self.blockGrid
        X = self._view(fieldX)
        L = g.getIndexLimitsData()
        ax = g.axes()
        sl = (slice(L[4], L[5]), slice(L[2], L[3]), slice(L[0], L[1]))
        Z, Y, Xc = np.meshgrid(ax[2][sl[0]], ax[1][sl[1]], ax[0][sl[2]], indexing="ij")
        err = np.abs(X[sl] - self.exact.trueSolutionFxyz(Xc, Y, Z))
        error_local = float(err.sum())
        if verbose:
            print(f"Max error local block avg {error_local / g.getNtotLocalNoGuards():g} in rank 0")
            print(f"Max error local point {float(err.max()):g} in rank 0")
        return error_local

    # iterativeSolverBaseAlpaka.hpp:620-638: duration, iterations, preconditioner
    # iterations, then the history up to the first non-positive entry and at most
    # maxIteration entries
    def writeResidualHistory(self, path="residualHistory.txt"):
        maxit = self.spec.chain()[0].get("maxit", 0)
        with open(path, "w") as f:
            f.write(f"{self.durationSolver_:g}\n{self.numIterationFinal_}\n{self.numIterationPreconditionerFinal_}\n")
            for i, v in enumerate(self.errorFromIterationHistory_):
                if i >= maxit or not v > 0:
                    break
                f.write(f"{v:g}\n")
