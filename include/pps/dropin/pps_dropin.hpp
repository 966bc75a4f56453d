// pps_dropin.hpp -- shared part of the drop-in headers in this directory.
//
// The reference's main program (solverPoissonMPI_CPU/src/main.cpp) and its
// user configuration (inputParam.hpp, solverSetup.hpp) compile unchanged
// against this directory in place of the reference's solver headers:
//
//     mpicxx -std=c++17 -I<dir with the user's inputParam.hpp + solverSetup.hpp>
//            -I<repo>/include/pps/dropin main.cpp -L<repo>/parallelpoissonsolver_amd -lpps
//
// The user's solverSetup.hpp (tolerance scale, Chebyshev bounds,
// ExactSolutionAndBCs, T_data) is read first; the solver types, BlockGrid,
// CommunicatorMPI and MatrixFreeOperatorA then come from the C++ facade
// include/pps/pps_solvers.hpp, which runs the solve on the GPU through the C ABI.
#pragma once

#include "solverSetup.hpp"

#ifndef PPS_USE_SOLVER_SETUP
#define PPS_USE_SOLVER_SETUP 1
#endif
#include "../pps_solvers.hpp"

using pps::BaseCG;
using pps::BiCGSTAB;
using pps::BlockGrid;
using pps::ChebyshevIteration;
using pps::CommunicatorMPI;
using pps::FCG;
using pps::Jacobi;
using pps::LearnedPreconditioner;
using pps::MatrixFreeOperatorA;
using pps::MixedPrecisionIR;
using pps::NoneSolver;
