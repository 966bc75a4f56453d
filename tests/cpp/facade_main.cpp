// facade_main.cpp -- the reference's main-program flow
// (solverPoissonMPI_CPU/src/main.cpp:77-127; alpaka form
// solverPoissonMPI_alpaka/src/main.cpp:83-160) compiled against the C++ facade
// include/pps/pps_solvers.hpp: grid, exact solution, solver type chain,
// setProblem, solve on the GPU, the getters, checkSolutionLocalGlobal,
// getDurationSolver().count(), writeResidualHistory and solution.dat.  Run by
// tests/test_cpp_facade.py.
//
//   facade_main <n> <cg|bicgstab|ir|alp|alp32> <bc x- x+ y- y+ z- z+> <x out file> [output dir]
//
// The optional output directory receives residualHistory.txt and solution.dat.
#include "pps/pps_solvers.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace pps;
constexpr int DIM = 3;
using T_data = double;

// the reference's shipped test problem (solverPoissonMPI_CPU/include/solverSetup.hpp:44-113):
// u = sin x + cos y + 3 sin z + x^2 y z + x^2 + 10, f = its Laplacian
template <int D, typename T>
class ExactSolutionAndBCs {
public:
    T setFieldB(T x, T y, T z) const { return -std::sin(x) - std::cos(y) - 3 * std::sin(z) + 2 * y * z + 2; }
    T trueSolutionFxyz(T x, T y, T z) const { return std::sin(x) + std::cos(y) + 3 * std::sin(z) + x * x * y * z + x * x + 10; }
    T trueSolutionDdir(T x, T y, T z, int dir) const
    {
        if (dir == 0) return std::cos(x) + 2 * x * y * z + 2 * x;
        if (dir == 1) return -std::sin(y) + x * x * z;
        return 3 * std::cos(z) + x * x * y;
    }
};

using T_NoneSolver = NoneSolver<DIM, T_data, 1, 1>;
using T_Preconditioner2 = ChebyshevIteration<DIM, T_data, 1, 11, false, false, T_NoneSolver>;
using T_SolverBiCG = BiCGSTAB<DIM, T_data, 1, 1000, true, true, T_Preconditioner2>;
using T_SolverCG = BaseCG<DIM, T_data, 1, 5000, true, true, T_NoneSolver>;
// FP64 refinement around FP32 CG to 1e-5, the first inner solve to 1e-2
using T_InnerIR = BaseCG<DIM, float, 100000, 5000, false, false, NoneSolver<DIM, float, 1, 1>>;
using T_SolverIR = MixedPrecisionIR<DIM, 1, 50, T_InnerIR, 10000>;
// the alpaka build's default composition (solverPoissonMPI_alpaka/include/inputParam.hpp:25-37,
// solverSetup.hpp:59-63): BiCGSTAB + global Chebyshev(24), in FP64 or with T_data_chebyshev = float
using T_NoneSolverAlpaka = NoneSolverAlpaka<DIM, T_data, T_data, 10000, 500>;
template <typename T_cheb>
using T_PreconditionerChebGlobal =
    ChebyshevIterationAlpaka<DIM, T_data, T_cheb, true, 10000, 24, false, false, T_NoneSolverAlpaka>;
template <typename T_cheb>
using T_SolverAlp = BiCGstabAlpaka<DIM, T_data, T_data, 1, 1000, true, true, T_PreconditionerChebGlobal<T_cheb>>;

// CPU form: setProblem then solver(x, b, operatorA) (main.cpp:94-99)
template <class T_Solver>
static void solve_cpu_form(T_Solver& solver, const BlockGrid<DIM, T_data>& blockGrid, std::vector<T_data>& fieldX,
                           std::vector<T_data>& fieldB)
{
    MatrixFreeOperatorA<DIM, T_data> operatorA(blockGrid);
    solver.setProblem(fieldX.data(), fieldB.data());
    solver(fieldX.data(), fieldB.data(), operatorA);
}

// the reference's reporting after the solve (main.cpp:103-127)
template <class T_Solver>
static int report(T_Solver& solver, const BlockGrid<DIM, T_data>& blockGrid, std::vector<T_data>& fieldX,
                  const char* out, const std::string& dir, std::chrono::high_resolution_clock::time_point startSolver)
{
    auto endSolver = std::chrono::high_resolution_clock::now();
    std::cout << "Iterative solver finished with iter: " << solver.getNumIterationFinal()
              << " error from algo " << solver.getErrorFromIteration() << " error r=b-Ax "
              << solver.getErrorComputeOperator() << " errorAvgtot "
              << solver.getErrorComputeOperator() / static_cast<T_data>(blockGrid.getNtotNpglobal()) << std::endl;
    const T_data errorLocal = solver.checkSolutionLocalGlobal(fieldX.data());
    std::chrono::duration<double> durationSolver = endSolver - startSolver;
    std::cout << "Solver time: " << durationSolver.count() << " seconds" << std::endl;
    std::cout << "SolverInFunction time: " << solver.getDurationSolver().count() << " seconds" << std::endl;
    solver.timeCounter.printAverageTime(solver.getNumIterationFinal());
    if (!dir.empty()) {
        solver.writeResidualHistory(dir + "/residualHistory.txt");
        if (!writeSolutionDat(blockGrid, fieldX.data(), dir + "/solution.dat")) return 3;
    }
    std::printf("{\"iterations\": %d, \"precond_iterations\": %d, \"error_from_iteration\": %.17g, "
                "\"error_compute_operator\": %.17g, \"error_local\": %.17g}\n",
                solver.getNumIterationFinal(), solver.getNumIterationPreconditionerFinal(),
                solver.getErrorFromIteration(), solver.getErrorComputeOperator(), errorLocal);
    std::fflush(stdout);
    FILE* f = std::fopen(out, "wb");
    if (!f) return 1;
    std::fwrite(fieldX.data(), sizeof(T_data), fieldX.size(), f);
    std::fclose(f);
    return 0;
}

template <class T_Solver, bool alpaka>
static int run(const BlockGrid<DIM, T_data>& blockGrid, const char* out, const std::string& dir)
{
    CommunicatorMPI<DIM, T_data> communicator(blockGrid);
    ExactSolutionAndBCs<DIM, T_data> exactSolutionAndBCs;
    std::vector<T_data> fieldX(blockGrid.getNtotLocalGuards(), 0.0), fieldB(blockGrid.getNtotLocalGuards(), 0.0);
    auto startSolver = std::chrono::high_resolution_clock::now();
    if constexpr (alpaka) {
        // alpaka form: the solver sets the problem itself (main.cpp:90-104)
        AlpakaHelper<DIM, T_data> alpakaHelper(blockGrid);
        T_Solver solver(blockGrid, exactSolutionAndBCs, communicator, alpakaHelper);
        solver(fieldX.data(), fieldB.data());
        return report(solver, blockGrid, fieldX, out, dir, startSolver);
    } else {
        T_Solver solver(blockGrid, exactSolutionAndBCs, communicator);
        solve_cpu_form(solver, blockGrid, fieldX, fieldB);
        return report(solver, blockGrid, fieldX, out, dir, startSolver);
    }
}

int main(int argc, char** argv)
{
    if (argc != 10 && argc != 11) {
        std::fprintf(stderr, "usage: %s n cg|bicgstab|ir|alp|alp32 bc*6 xout [outdir]\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[1]);
    std::array<int, 6> bcs{};
    for (int f = 0; f < 6; ++f) bcs[f] = std::atoi(argv[3 + f]);
    const std::string dir = argc == 11 ? argv[10] : "";
    const T_data ds = 1.0 / (n - 1);
    BlockGrid<DIM, T_data> blockGrid({1, 1, 1}, 0, {n, n, n}, {ds, ds, ds}, {0.5, 0.25, 1.0}, {1, 1, 1}, bcs,
                                     {0, 0, 0, 0, 0, 0});
    const std::string kind = argv[2];
    if (kind == "cg") return run<T_SolverCG, false>(blockGrid, argv[9], dir);
    if (kind == "ir") return run<T_SolverIR, false>(blockGrid, argv[9], dir);
    if (kind == "alp") return run<T_SolverAlp<double>, true>(blockGrid, argv[9], dir);
    if (kind == "alp32") return run<T_SolverAlp<float>, true>(blockGrid, argv[9], dir);
    return run<T_SolverBiCG, false>(blockGrid, argv[9], dir);
}
