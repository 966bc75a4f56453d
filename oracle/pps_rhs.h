/*
 * pps_rhs.h -- seeded right-hand-side recipes shared by the CPU oracle and the
 * reference driver (test infrastructure only; the GPU library carries its own
 * implementation of the same recipes and is checked against this one).
 *
 * The reference has no array-fed right-hand sides: it evaluates
 * ExactSolutionAndBCs::setFieldB at every data point
 * (solverPoissonMPI_CPU/include/iterativeSolverBase.hpp:537-555).  The Poisson
 * workloads solved here need seeded random sources instead, so the recipes are
 * defined by this file:
 *
 *  - every random number is a pure function of (seed, stream, counter) through
 *    a splitmix64 cascade, so CPU, GPU and any domain decomposition agree;
 *  - uniform doubles lie strictly inside (0,1); normals use Box-Muller;
 *  - coordinates are those of the reference's vertex-centred grid,
 *    x = origin + i_data * ds (iterativeSolverBase.hpp:547-549);
 *  - the field f is written as b = -f on the solver region (A is the +Laplacian,
 *    matrixFreeOperatorA.hpp:33-38) and 0 on the rest of the data region.
 */
#ifndef PPS_RHS_H
#define PPS_RHS_H

#include <math.h>
#include <stdint.h>

#define PPS_RHS_STREAM_POINT 1u   /* i.i.d. normal field, counter = global data index */
#define PPS_RHS_STREAM_SRC   2u   /* Gaussian-source centres and widths               */
#define PPS_RHS_STREAM_AMP   4u   /* Gaussian-source amplitudes                       */
#define PPS_RHS_STREAM_SPEC  3u   /* k^-2 spectrum coefficients                       */

static inline uint64_t pps_mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t pps_rng_u64(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    return pps_mix64(pps_mix64(pps_mix64(seed) ^ stream) ^ ctr);
}

/* uniform in (0,1): 53 random bits, offset by half an ulp so 0 never occurs */
static inline double pps_rng_u01(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    return ((double)(pps_rng_u64(seed, stream, ctr) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

static inline double pps_rng_normal(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    const double u1 = pps_rng_u01(seed, stream, 2u * ctr);
    const double u2 = pps_rng_u01(seed, stream, 2u * ctr + 1u);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* Gaussian source s of the "sum of Gaussian sources" recipe:
 * centre uniform in [0,1]^3, width uniform in [wmin,wmax], amplitude N(0,1). */
static inline void pps_rhs_gauss_source(uint64_t seed, int s, double wmin, double wmax,
                                        double c[3], double* w, double* a)
{
    c[0] = pps_rng_u01(seed, PPS_RHS_STREAM_SRC, 4u * (uint64_t)s + 0u);
    c[1] = pps_rng_u01(seed, PPS_RHS_STREAM_SRC, 4u * (uint64_t)s + 1u);
    c[2] = pps_rng_u01(seed, PPS_RHS_STREAM_SRC, 4u * (uint64_t)s + 2u);
    *w = wmin + (wmax - wmin) * pps_rng_u01(seed, PPS_RHS_STREAM_SRC, 4u * (uint64_t)s + 3u);
    *a = pps_rng_normal(seed, PPS_RHS_STREAM_AMP, (uint64_t)s);
}

/* f(x,y,z) of the Gaussian-source recipe, given the source table */
static inline double pps_rhs_gauss_eval(int nsrc, const double* cx, const double* cy,
                                        const double* cz, const double* w, const double* a,
                                        double x, double y, double z)
{
    double f = 0.0;
    for (int s = 0; s < nsrc; ++s) {
        const double dx = x - cx[s], dy = y - cy[s], dz = z - cz[s];
        const double r2 = dx * dx + dy * dy + dz * dz;
        f += a[s] * exp(-r2 / (2.0 * w[s] * w[s]));
    }
    return f;
}

/* coefficient of the DST mode (kx,ky,kz), 1 <= k <= K, of the k^-2 spectrum recipe */
static inline double pps_rhs_spec_coef(uint64_t seed, int K, int kx, int ky, int kz)
{
    const uint64_t idx = (uint64_t)(kx - 1) + (uint64_t)K * ((uint64_t)(ky - 1) + (uint64_t)K * (uint64_t)(kz - 1));
    const double k2 = (double)kx * kx + (double)ky * ky + (double)kz * kz;
    return pps_rng_normal(seed, PPS_RHS_STREAM_SPEC, idx) / k2;
}

#endif
