// pps_rhs.hip -- seeded right-hand-side generators (device side).
//
// The reference fills b from ExactSolutionAndBCs::setFieldB at every data
// point (iterativeSolverBase.hpp:537-555).  The Poisson workloads here are
// seeded random sources instead; each generator writes b = -f on the solver
// region (A is the +Laplacian, matrixFreeOperatorA.hpp:33-38).  The random
// numbers are a pure function of (seed, stream, counter) -- a splitmix64
// cascade -- so any decomposition and the CPU checker produce the same field.
#include "pps_kernels.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace pps {

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double u01(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    const uint64_t r = mix64(mix64(mix64(seed) ^ stream) ^ ctr);
    return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ double normal01(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    const double a = u01(seed, stream, 2u * ctr), b = u01(seed, stream, 2u * ctr + 1u);
    return sqrt(-2.0 * log(a)) * cos(6.283185307179586 * b);
}

// i.i.d. N(0,1) at every solver point, counter = global data index; partial
// sums of f for the mean subtraction
__global__ __launch_bounds__(256) void k_rhs_normal(Layout L, Region R, int n0, int n1, uint64_t seed, double* b,
                                                   double* partial)
{
    __shared__ double sm[4];
    const int nx = R.hi[0] - R.lo[0], ny = R.hi[1] - R.lo[1], nz = R.hi[2] - R.lo[2];
    const long long n = (long long)nx * ny * nz;
    double s = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int i = R.lo[0] + (int)(t % nx);
        const int j = R.lo[1] + (int)((t / nx) % ny);
        const int k = R.lo[2] + (int)(t / ((long long)nx * ny));
        const uint64_t idx = (uint64_t)(i - 1) + (uint64_t)n0 * ((uint64_t)(j - 1) + (uint64_t)n1 * (uint64_t)(k - 1));
        const double f = normal01(seed, 1u, idx);
        b[L.at(i, j, k)] = -f;
        s += f;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
}

__global__ void k_sum_partials(const double* partial, int n, double* out)
{
    __shared__ double sm[4];
    double s = 0.0;
    for (int t = threadIdx.x; t < n; t += blockDim.x) s += partial[t];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = sm[0] + sm[1] + sm[2] + sm[3];
}

__global__ __launch_bounds__(256) void k_add_mean(Layout L, Region R, double* b, const double* sum, double count)
{
    const int nx = R.hi[0] - R.lo[0], ny = R.hi[1] - R.lo[1], nz = R.hi[2] - R.lo[2];
    const long long n = (long long)nx * ny * nz;
    const double mean = *sum / count;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int i = R.lo[0] + (int)(t % nx);
        const int j = R.lo[1] + (int)((t / nx) % ny);
        const int k = R.lo[2] + (int)(t / ((long long)nx * ny));
        b[L.at(i, j, k)] += mean;
    }
}

void launch_rhs_normal(hipStream_t s, const Layout& L, const Region& R, const int np[3], uint64_t seed, double* b,
                       double* scratch /* >= 2049 doubles */)
{
    const int nblk = 2048;
    hipLaunchKernelGGL(k_rhs_normal, dim3(nblk), dim3(256), 0, s, L, R, np[0], np[1], seed, b, scratch);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, scratch, nblk, scratch + nblk);
    hipLaunchKernelGGL(k_add_mean, dim3(nblk), dim3(256), 0, s, L, R, b, scratch + nblk, (double)R.count());
}

// sum of Gaussian sources; src = [cx, cy, cz, w, a] x nsrc
__global__ __launch_bounds__(256) void k_rhs_gauss(Layout L, Region R, const double* src, int nsrc, double ox,
                                                  double oy, double oz, double dx, double dy, double dz, double* b)
{
    __shared__ double tab[5 * 256];
    for (int t = threadIdx.x; t < 5 * nsrc && t < 5 * 256; t += blockDim.x) tab[t] = src[t];
    __syncthreads();
    const int nx = R.hi[0] - R.lo[0], ny = R.hi[1] - R.lo[1], nz = R.hi[2] - R.lo[2];
    const long long n = (long long)nx * ny * nz;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int i = R.lo[0] + (int)(t % nx);
        const int j = R.lo[1] + (int)((t / nx) % ny);
        const int k = R.lo[2] + (int)(t / ((long long)nx * ny));
        const double x = ox + (i - 1) * dx, y = oy + (j - 1) * dy, z = oz + (k - 1) * dz;
        double f = 0.0;
        for (int q = 0; q < nsrc; ++q) {
            const double ddx = x - tab[q], ddy = y - tab[nsrc + q], ddz = z - tab[2 * nsrc + q];
            const double w = tab[3 * nsrc + q];
            f += tab[4 * nsrc + q] * exp(-(ddx * ddx + ddy * ddy + ddz * ddz) / (2.0 * w * w));
        }
        b[L.at(i, j, k)] = -f;
    }
}

void launch_rhs_gauss(hipStream_t s, const Layout& L, const Region& R, const double* src_dev, int nsrc,
                      const double origin[3], const double ds[3], double* b)
{
    hipLaunchKernelGGL(k_rhs_gauss, dim3(4096), dim3(256), 0, s, L, R, src_dev, nsrc, origin[0], origin[1], origin[2],
                       ds[0], ds[1], ds[2], b);
}

// separable k^-2 spectrum: T1[kx,ky,gk] = sum_kz C Sz ; T2[kx,gj,gk] = sum_ky T1 Sy ;
// b = -(sum_kx T2 Sx) (+ b when accumulating)
__global__ void k_spec1(const double* C, const double* Sz, int K, int n2, double* T1)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)K * K * n2) return;
    const int kx = (int)(t % K), ky = (int)((t / K) % K), gk = (int)(t / ((long long)K * K));
    double acc = 0.0;
    for (int kz = 0; kz < K; ++kz) acc += C[kx + K * (ky + (long long)K * kz)] * Sz[kz + (long long)K * gk];
    T1[kx + K * (ky + (long long)K * gk)] = acc;
}

__global__ void k_spec2(const double* T1, const double* Sy, int K, int n1, int n2, double* T2)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)K * n1 * n2) return;
    const int kx = (int)(t % K);
    const int gj = (int)((t / K) % n1);
    const int gk = (int)(t / ((long long)K * n1));
    double acc = 0.0;
    for (int ky = 0; ky < K; ++ky) acc += T1[kx + K * (ky + (long long)K * gk)] * Sy[ky + (long long)K * gj];
    T2[kx + K * (gj + (long long)n1 * gk)] = acc;
}

__global__ __launch_bounds__(256) void k_spec3(Layout L, Region R, const double* T2, const double* Sx, int K, int n1,
                                              int accumulate, double* b)
{
    const int nx = R.hi[0] - R.lo[0], ny = R.hi[1] - R.lo[1], nz = R.hi[2] - R.lo[2];
    const long long n = (long long)nx * ny * nz;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int i = R.lo[0] + (int)(t % nx);
        const int j = R.lo[1] + (int)((t / nx) % ny);
        const int k = R.lo[2] + (int)(t / ((long long)nx * ny));
        const int gi = i - 1, gj = j - 1, gk = k - 1;
        double acc = 0.0;
        for (int kx = 0; kx < K; ++kx) acc += T2[kx + K * (gj + (long long)n1 * gk)] * Sx[kx + (long long)K * gi];
        const long long c = L.at(i, j, k);
        b[c] = accumulate ? b[c] - acc : -acc;
    }
}

void launch_rhs_spectrum(hipStream_t s, const Layout& L, const Region& R, const int np[3], int K, const double* C,
                         const double* Sx, const double* Sy, const double* Sz, double* T1, double* T2, int accumulate,
                         double* b)
{
    const long long n1 = (long long)K * K * np[2];
    hipLaunchKernelGGL(k_spec1, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, s, C, Sz, K, np[2], T1);
    const long long n2 = (long long)K * np[1] * np[2];
    hipLaunchKernelGGL(k_spec2, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s, T1, Sy, K, np[1], np[2], T2);
    hipLaunchKernelGGL(k_spec3, dim3(4096), dim3(256), 0, s, L, R, T2, Sx, K, np[1], accumulate, b);
}

}  // namespace pps
