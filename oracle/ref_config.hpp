/*
 * ref_config.hpp -- run-time configuration for building the reference CPU
 * solver (solverPoissonMPI_CPU) as the oracle's cross-check binary.
 *
 * The reference is configured by editing its solverSetup.hpp / inputParam.hpp
 * and recompiling (README.md:36).  This header plays the role of solverSetup.hpp
 * for the oracle build: it is included first, defines the same include guard
 * (solverPoissonMPI_CPU/include/solverSetup.hpp:1-2) so the reference's copy is
 * skipped, and supplies every name the reference's solver headers read:
 *   T_data, PI, tollScalingFactor, orderNeumanBcs, trackErrorFromIterationHistory,
 *   epsilon, rescaleEigMin, rescaleEigMax, getMPIType<>, ExactSolutionAndBCs.
 * Values the reference only multiplies into run-time arithmetic are run-time
 * variables here, so one binary serves every grid and tolerance; names the
 * reference uses in `if constexpr` stay constexpr.
 */
#ifndef SOLVERSETUP_HPP
#define SOLVERSETUP_HPP

#include <mpi.h>
#include <cmath>
#include <array>

using T_data = double;

template <typename T>
inline MPI_Datatype getMPIType();
template <>
inline MPI_Datatype getMPIType<float>() { return MPI_FLOAT; }
template <>
inline MPI_Datatype getMPIType<double>() { return MPI_DOUBLE; }

constexpr T_data PI = 3.141592653589793;

/* effective tolerance = template tolerance * tollScalingFactor (baseCG.hpp:22) */
inline T_data tollScalingFactor = 1e-10;
constexpr int orderNeumanBcs = 2;
constexpr bool trackErrorFromIterationHistory = true;

/* Chebyshev spectrum bounds (chebyshevIteration.hpp:22-26) */
inline T_data epsilon = 1e-4;
inline T_data rescaleEigMin = 500;
inline T_data rescaleEigMax = 1 - 1e-4;

/* Which manufactured problem ExactSolutionAndBCs describes.
 *  0: the reference's own test problem (solverSetup.hpp:48-66 of the CPU code)
 *  1: u = sin(pi x) sin(pi y) sin(pi z)
 *  2: u = 0 (array-fed right-hand sides; boundary values and fluxes vanish) */
inline int refProblemKind = 0;

template <int DIM, typename T>
class ExactSolutionAndBCs {
public:
    inline T setFieldB(const T x, const T y, const T z) const
    {
        if (refProblemKind == 0) return -sin(x) - cos(y) - 3 * sin(z) + 2 * y * z + 2;
        if (refProblemKind == 1) return -3 * PI * PI * sin(PI * x) * sin(PI * y) * sin(PI * z);
        return 0;
    }
    inline T trueSolutionFxyz(const T x, const T y, const T z) const
    {
        if (refProblemKind == 0) return sin(x) + cos(y) + 3 * sin(z) + x * x * y * z + x * x + 10;
        if (refProblemKind == 1) return sin(PI * x) * sin(PI * y) * sin(PI * z);
        return 0;
    }
    inline T trueSolutionDdir(const T x, const T y, const T z, const int dir) const
    {
        if (refProblemKind == 0) {
            if (dir == 0) return cos(x) + 2 * x * y * z + 2 * x;
            if (dir == 1) return -sin(y) + x * x * z;
            if (dir == 2) return 3 * cos(z) + x * x * y;
            return -100;
        }
        if (refProblemKind == 1) {
            if (dir == 0) return PI * cos(PI * x) * sin(PI * y) * sin(PI * z);
            if (dir == 1) return PI * sin(PI * x) * cos(PI * y) * sin(PI * z);
            if (dir == 2) return PI * sin(PI * x) * sin(PI * y) * cos(PI * z);
            return -100;
        }
        return 0;
    }
};

#endif
