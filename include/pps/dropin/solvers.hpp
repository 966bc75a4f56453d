// Drop-in for the reference's solvers.hpp (noneSolver.hpp, baseCG.hpp, BiCGSTAB.hpp, chebyshevIteration.hpp), see pps_dropin.hpp.
#pragma once
#include "pps_dropin.hpp"
