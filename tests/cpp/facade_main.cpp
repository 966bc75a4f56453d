// facade_main.cpp -- the reference's main-program flow
// (solverPoissonMPI_CPU/src/main.cpp:77-110) compiled against the C++ facade
// include/pps/pps_solvers.hpp: grid, exact solution, solver type chain,
// setProblem, solve on the GPU, getters.  Run by tests/test_cpp_facade.py.
//
//   facade_main <n> <cg|bicgstab> <bc x- x+ y- y+ z- z+> <x out file>
#include "pps/pps_solvers.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace pps;
constexpr int DIM = 3;
using T_data = double;
constexpr T_data PI = 3.141592653589793;

// u = sin(pi x) sin(pi y) sin(pi z); the right-hand side is its Laplacian
template <int D, typename T>
class ExactSolutionAndBCs {
public:
    T setFieldB(T x, T y, T z) const { return -3 * PI * PI * trueSolutionFxyz(x, y, z); }
    T trueSolutionFxyz(T x, T y, T z) const { return std::sin(PI * x) * std::sin(PI * y) * std::sin(PI * z); }
    T trueSolutionDdir(T x, T y, T z, int dir) const
    {
        if (dir == 0) return PI * std::cos(PI * x) * std::sin(PI * y) * std::sin(PI * z);
        if (dir == 1) return PI * std::sin(PI * x) * std::cos(PI * y) * std::sin(PI * z);
        return PI * std::sin(PI * x) * std::sin(PI * y) * std::cos(PI * z);
    }
};

using T_NoneSolver = NoneSolver<DIM, T_data, 1, 1>;
using T_Preconditioner2 = ChebyshevIteration<DIM, T_data, 1, 11, false, false, T_NoneSolver>;
using T_SolverBiCG = BiCGSTAB<DIM, T_data, 1, 1000, true, true, T_Preconditioner2>;
using T_SolverCG = BaseCG<DIM, T_data, 1, 5000, true, true, T_NoneSolver>;
// FP64 refinement around FP32 CG to 1e-5, the first inner solve to 1e-2
using T_InnerIR = BaseCG<DIM, float, 100000, 5000, false, false, NoneSolver<DIM, float, 1, 1>>;
using T_SolverIR = MixedPrecisionIR<DIM, 1, 50, T_InnerIR, 10000>;

template <class T_Solver>
static int run(const BlockGrid<DIM, T_data>& blockGrid, const char* out)
{
    CommunicatorMPI<DIM, T_data> communicator(blockGrid);
    ExactSolutionAndBCs<DIM, T_data> exactSolutionAndBCs;
    MatrixFreeOperatorA<DIM, T_data> operatorA(blockGrid);
    T_Solver solver(blockGrid, exactSolutionAndBCs, communicator);
    std::vector<T_data> fieldX(blockGrid.getNtotLocalGuards(), 0.0), fieldB(blockGrid.getNtotLocalGuards(), 0.0);
    solver.setProblem(fieldX.data(), fieldB.data());
    solver(fieldX.data(), fieldB.data(), operatorA);
    std::printf("{\"iterations\": %d, \"precond_iterations\": %d, \"error_from_iteration\": %.17g, "
                "\"error_compute_operator\": %.17g}\n",
                solver.getNumIterationFinal(), solver.getNumIterationPreconditionerFinal(),
                solver.getErrorFromIteration(), solver.getErrorComputeOperator());
    FILE* f = std::fopen(out, "wb");
    if (!f) return 1;
    std::fwrite(fieldX.data(), sizeof(T_data), fieldX.size(), f);
    std::fclose(f);
    return 0;
}

int main(int argc, char** argv)
{
    if (argc != 10) {
        std::fprintf(stderr, "usage: %s n cg|bicgstab|ir bc*6 xout\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[1]);
    std::array<int, 6> bcs{};
    for (int f = 0; f < 6; ++f) bcs[f] = std::atoi(argv[3 + f]);
    const T_data ds = 1.0 / (n - 1);
    BlockGrid<DIM, T_data> blockGrid({1, 1, 1}, 0, {n, n, n}, {ds, ds, ds}, {0.5, 0.25, 1.0}, {1, 1, 1}, bcs,
                                     {0, 0, 0, 0, 0, 0});
    if (std::strcmp(argv[2], "cg") == 0) return run<T_SolverCG>(blockGrid, argv[9]);
    if (std::strcmp(argv[2], "ir") == 0) return run<T_SolverIR>(blockGrid, argv[9]);
    return run<T_SolverBiCG>(blockGrid, argv[9]);
}
