"""MI355X-native Krylov path of the ParallelPoissonSolver finite-difference
Poisson solve: hand-written gfx950 HIP kernels behind the C ABI of
include/pps_abi.h (libpps.so), driven from here through ctypes."""
from ._lib import Context, PPSError, lib  # noqa: F401
from .solvers import (  # noqa: F401
    DIM, BaseCG, BiCGSTAB, BlockGrid, ChebyshevIteration, CommunicatorMPI, ExactSolutionAndBCs, FCG, Jacobi,
    LearnedPreconditioner, MatrixFreeOperatorA, MixedPrecisionIR, NoneSolver, Solver, SolverSpec)
