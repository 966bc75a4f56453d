// ref_driver.cpp -- driver that runs the reference CPU solver classes
// (solverPoissonMPI_CPU/include/*.hpp, compiled where they lie) on a grid,
// boundary set, right-hand side and solver chain chosen at run time, and dumps
// what the oracle and the golden fixtures need: iterations, recursive and true
// residual, residual history and the solution array.
//
// It follows the reference driver solverPoissonMPI_CPU/src/main.cpp:19-135
// (rank-grid check, BlockGrid, CommunicatorMPI, T_Solver, setProblem, solve,
// checkSolutionLocalGlobal) with the compile-time choices of inputParam.hpp
// turned into command-line options.  Test infrastructure only.
#include "ref_config.hpp"

#include <chrono>
#include <cstdio>
#include <fstream>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "solvers.hpp"
#include "blockGrid.hpp"
#include "communicationMPI.hpp"
#include "matrixFreeOperatorA.hpp"

extern "C" {
#include "pps_rhs.h"
}

#ifndef REF_MAXIT
#define REF_MAXIT 100000
#endif

constexpr int DIM = 3;
using Grid = BlockGrid<DIM, T_data>;
using Comm = CommunicatorMPI<DIM, T_data>;
using Exact = ExactSolutionAndBCs<DIM, T_data>;
using OpA = MatrixFreeOperatorA<DIM, T_data>;

// inner solvers, as the aliases of solverPoissonMPI_CPU/include/inputParam.hpp:22-31
using NoneT = NoneSolver<DIM, T_data, 10000, 150>;
template <int M>
using ChebT = ChebyshevIteration<DIM, T_data, 10000, M, false, false, NoneT>;
using InnerBiCG = BiCGSTAB<DIM, T_data, 10000, 150, false, false, NoneT>;
using InnerCG = BaseCG<DIM, T_data, 10000, 150, false, false, NoneT>;

struct Opts {
    std::array<int, 3> nranks{1, 1, 1};
    std::array<int, 3> np{18, 18, 18};
    std::array<T_data, 3> ds{1.0 / 17, 1.0 / 17, 1.0 / 17};
    std::array<T_data, 3> origin{0, 0, 0};
    std::array<int, 6> bc{0, 0, 0, 0, 0, 0};
    std::string solver = "cg", precond = "none", rhs = "function", out;
    std::string solutionDat, residualHistory;  // the alpaka build's output files
    int problem = 0, chebM = 11, nsrc = 64, specK = 16;
    unsigned long long seed = 1234;
    double tol = 1e-10, wmin = 0.01, wmax = 0.1;
};

// Exposes the protected residual history of a reference solver (iterativeSolverBase.hpp:637).
template <class S>
struct Exposed : S {
    using S::S;
    const T_data* history() const { return this->errorFromIterationHistory_; }
};

static double coord(const Grid& g, int dir, int idx)
{
    // iterativeSolverBase.hpp:547-549
    return g.getOrigin()[dir] + (idx - g.getIndexLimitsData()[2 * dir]) * g.getDs()[dir] +
           g.getGlobalLocation()[dir] * g.getNlocalNoGuards()[dir] * g.getDs()[dir];
}

// Overwrites b on the solver region with -f of the requested seeded recipe.
static void fillRhs(const Opts& o, const Grid& g, T_data* b)
{
    if (o.rhs == "function") return;
    const auto ls = g.getIndexLimitsSolver();
    const auto ld = g.getIndexLimitsData();
    const auto ng = g.getNlocalGuards();
    const auto nl = g.getNlocalNoGuards();
    const auto gl = g.getGlobalLocation();
    const long sj = ng[0], sk = (long)ng[0] * ng[1];
    std::vector<double> cx, cy, cz, w, a, coef;
    if (o.rhs == "gauss" || o.rhs == "multiscale") {
        for (int s = 0; s < o.nsrc; ++s) {
            double c[3], ws, as;
            pps_rhs_gauss_source(o.seed, s, o.wmin, o.wmax, c, &ws, &as);
            cx.push_back(c[0]); cy.push_back(c[1]); cz.push_back(c[2]); w.push_back(ws); a.push_back(as);
        }
    }
    if (o.rhs == "spectrum" || o.rhs == "multiscale") {
        const int K = o.specK;
        coef.resize((size_t)K * K * K);
        for (int kz = 1; kz <= K; ++kz)
            for (int ky = 1; ky <= K; ++ky)
                for (int kx = 1; kx <= K; ++kx)
                    coef[(kx - 1) + K * ((ky - 1) + K * (kz - 1))] = pps_rhs_spec_coef(o.seed, K, kx, ky, kz);
    }
    double localSum = 0;
    for (int k = ls[4]; k < ls[5]; ++k)
        for (int j = ls[2]; j < ls[3]; ++j)
            for (int i = ls[0]; i < ls[1]; ++i) {
                const long gi = (i - ld[0]) + (long)gl[0] * nl[0];
                const long gj = (j - ld[2]) + (long)gl[1] * nl[1];
                const long gk = (k - ld[4]) + (long)gl[2] * nl[2];
                double f = 0;
                if (o.rhs == "normal") {
                    const unsigned long long idx = gi + (unsigned long long)o.np[0] * (gj + (unsigned long long)o.np[1] * gk);
                    f = pps_rng_normal(o.seed, PPS_RHS_STREAM_POINT, idx);
                }
                if (!cx.empty())
                    f += pps_rhs_gauss_eval(o.nsrc, cx.data(), cy.data(), cz.data(), w.data(), a.data(),
                                            coord(g, 0, i), coord(g, 1, j), coord(g, 2, k));
                if (!coef.empty()) {
                    const int K = o.specK;
                    double acc = 0;
                    for (int kz = 1; kz <= K; ++kz) {
                        const double sz = sin(kz * PI * (double)gk / (o.np[2] - 1));
                        for (int ky = 1; ky <= K; ++ky) {
                            const double syz = sz * sin(ky * PI * (double)gj / (o.np[1] - 1));
                            for (int kx = 1; kx <= K; ++kx)
                                acc += coef[(kx - 1) + K * ((ky - 1) + K * (kz - 1))] * syz *
                                       sin(kx * PI * (double)gi / (o.np[0] - 1));
                        }
                    }
                    f += acc;
                }
                b[i + sj * j + sk * k] = -f;
                localSum += f;
            }
    if (o.rhs == "normal") {
        // mean-subtract over the global solver region
        double tot = 0;
        MPI_Allreduce(&localSum, &tot, 1, MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);
        long cnt = (long)(ls[1] - ls[0]) * (ls[3] - ls[2]) * (ls[5] - ls[4]), totCnt = 0;
        MPI_Allreduce(&cnt, &totCnt, 1, MPI_LONG, MPI_SUM, MPI_COMM_WORLD);
        const double mean = tot / totCnt;
        for (int k = ls[4]; k < ls[5]; ++k)
            for (int j = ls[2]; j < ls[3]; ++j)
                for (int i = ls[0]; i < ls[1]; ++i) b[i + sj * j + sk * k] += mean;
    }
}

template <class Solver>
static int run(const Opts& o, Grid& grid, Comm& comm, const Exact& exact, int rank)
{
    OpA opA(grid);
    Exposed<Solver> solver(grid, exact, comm);
    const int n = grid.getNtotLocalGuards();
    std::vector<T_data> x(n, 0.0), b(n, 0.0);
    solver.setProblem(x.data(), b.data());
    fillRhs(o, grid, b.data());

    MPI_Barrier(MPI_COMM_WORLD);
    auto t0 = std::chrono::high_resolution_clock::now();
    solver(x.data(), b.data(), opA);
    auto t1 = std::chrono::high_resolution_clock::now();
    MPI_Barrier(MPI_COMM_WORLD);

    const int iters = solver.getNumIterationFinal();
    if (rank == 0)
        std::cout << "Iterative solver finished with iter: " << iters << " error from algo "
                  << solver.getErrorFromIteration() << " error r=b-Ax " << solver.getErrorComputeOperator()
                  << std::endl;
    solver.checkSolutionLocalGlobal(x.data());
    const double tSolve = std::chrono::duration<double>(t1 - t0).count();
    if (rank == 0) std::cout << "Solver time: " << tSolve << " seconds" << std::endl;

    if (!o.residualHistory.empty() && rank == 0) {
        // writeResidualHistory (solverPoissonMPI_alpaka/include/iterativeSolverBaseAlpaka.hpp:620-638)
        // applied to this solver's history.  The alpaka history array is zero
        // filled past the last iteration (:75-76); the CPU one is not
        // (iterativeSolverBase.hpp:43), so only the entries the solve wrote,
        // 0..iters, are read.  Preconditioner iterations as the alpaka main
        // loop counts them: Chebyshev adds maxIteration per application
        // (BiCGstabAlpaka.hpp:214,290), two applications per BiCGSTAB iteration.
        long precIters = 0;
        if (o.precond == "cheb") precIters = (long)o.chebM * (o.solver == "bicgstab" ? 2L * iters : iters + 1L);
        std::ofstream outFile(o.residualHistory);
        outFile << solver.getDurationSolver().count() << "\n";
        outFile << iters << "\n";
        outFile << precIters << "\n";
        for (int i = 0; i < REF_MAXIT && i <= iters && solver.history()[i] > 0; i++) outFile << solver.history()[i] << "\n";
    }
    if (!o.solutionDat.empty()) {
        // solution.dat (solverPoissonMPI_alpaka/src/main.cpp:135-146): each rank's
        // guarded array at offset sizeof(T_data) * ntotLocalGuards * rank
        MPI_File fileSolution;
        MPI_Status status;
        MPI_Offset offset = sizeof(T_data) * (MPI_Offset)grid.getNtotLocalGuards() * rank;
        MPI_File_open(MPI_COMM_WORLD, o.solutionDat.c_str(), MPI_MODE_CREATE | MPI_MODE_WRONLY, MPI_INFO_NULL,
                      &fileSolution);
        MPI_File_seek(fileSolution, offset, MPI_SEEK_SET);
        MPI_File_write(fileSolution, x.data(), grid.getNtotLocalGuards(), getMPIType<T_data>(), &status);
        MPI_File_close(&fileSolution);
    }

    if (!o.out.empty()) {
        char path[4096];
        std::snprintf(path, sizeof path, "%s.x.r%d.bin", o.out.c_str(), rank);
        FILE* f = std::fopen(path, "wb");
        if (!f) { std::perror(path); return 2; }
        std::fwrite(x.data(), sizeof(T_data), n, f);
        std::fclose(f);
        if (rank == 0) {
            std::snprintf(path, sizeof path, "%s.json", o.out.c_str());
            f = std::fopen(path, "w");
            if (!f) { std::perror(path); return 2; }
            std::fprintf(f, "{\"iterations\": %d, \"error_from_iteration\": %.17g, \"error_compute_operator\": %.17g,"
                            " \"solver_time\": %.9g, \"solver_in_function_time\": %.9g, \"nranks\": [%d, %d, %d],"
                            " \"history\": [",
                         iters, solver.getErrorFromIteration(), solver.getErrorComputeOperator(), tSolve,
                         solver.getDurationSolver().count(), o.nranks[0], o.nranks[1], o.nranks[2]);
            const bool tracked = o.solver != "cheb";
            const int nh = tracked ? std::min(iters, REF_MAXIT - 1) + 1 : 0;
            for (int h = 0; h < nh; ++h) std::fprintf(f, "%s%.17g", h ? ", " : "", solver.history()[h]);
            std::fprintf(f, "]}\n");
            std::fclose(f);
        }
    }
    return 0;
}

template <int M>
static int dispatchCheb(const Opts& o, Grid& g, Comm& c, const Exact& e, int r)
{
    if (o.solver == "cg") return run<BaseCG<DIM, T_data, 1, REF_MAXIT, true, true, ChebT<M>>>(o, g, c, e, r);
    if (o.solver == "bicgstab") return run<BiCGSTAB<DIM, T_data, 1, REF_MAXIT, true, true, ChebT<M>>>(o, g, c, e, r);
    return -1;
}

static int dispatch(const Opts& o, Grid& g, Comm& c, const Exact& e, int r)
{
    if (o.solver == "cheb") {
        switch (o.chebM) {
        case 4: return run<ChebyshevIteration<DIM, T_data, 1, 4, true, true, NoneT>>(o, g, c, e, r);
        case 11: return run<ChebyshevIteration<DIM, T_data, 1, 11, true, true, NoneT>>(o, g, c, e, r);
        case 24: return run<ChebyshevIteration<DIM, T_data, 1, 24, true, true, NoneT>>(o, g, c, e, r);
        }
        return -1;
    }
    if (o.precond == "none") {
        if (o.solver == "cg") return run<BaseCG<DIM, T_data, 1, REF_MAXIT, true, true, NoneT>>(o, g, c, e, r);
        if (o.solver == "bicgstab") return run<BiCGSTAB<DIM, T_data, 1, REF_MAXIT, true, true, NoneT>>(o, g, c, e, r);
    }
    if (o.precond == "bicgstab") {
        if (o.solver == "cg") return run<BaseCG<DIM, T_data, 1, REF_MAXIT, true, true, InnerBiCG>>(o, g, c, e, r);
        if (o.solver == "bicgstab") return run<BiCGSTAB<DIM, T_data, 1, REF_MAXIT, true, true, InnerBiCG>>(o, g, c, e, r);
    }
    if (o.precond == "cg") {
        if (o.solver == "cg") return run<BaseCG<DIM, T_data, 1, REF_MAXIT, true, true, InnerCG>>(o, g, c, e, r);
        if (o.solver == "bicgstab") return run<BiCGSTAB<DIM, T_data, 1, REF_MAXIT, true, true, InnerCG>>(o, g, c, e, r);
    }
    if (o.precond == "cheb") {
        switch (o.chebM) {
        case 4: return dispatchCheb<4>(o, g, c, e, r);
        case 11: return dispatchCheb<11>(o, g, c, e, r);
        case 24: return dispatchCheb<24>(o, g, c, e, r);
        }
    }
    return -1;
}

int main(int argc, char** argv)
{
    MPI_Init(&argc, &argv);
    int world = 0, rank = 0;
    MPI_Comm_size(MPI_COMM_WORLD, &world);
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);

    Opts o;
    for (int a = 1; a < argc; ++a) {
        std::string k = argv[a];
        auto need = [&](int cnt) {
            if (a + cnt >= argc) { std::cerr << "missing value for " << k << std::endl; MPI_Abort(MPI_COMM_WORLD, 2); }
        };
        if (k == "--ranks") { need(3); for (int d = 0; d < 3; ++d) o.nranks[d] = std::atoi(argv[++a]); }
        else if (k == "--np") { need(3); for (int d = 0; d < 3; ++d) o.np[d] = std::atoi(argv[++a]); }
        else if (k == "--ds") { need(3); for (int d = 0; d < 3; ++d) o.ds[d] = std::atof(argv[++a]); }
        else if (k == "--origin") { need(3); for (int d = 0; d < 3; ++d) o.origin[d] = std::atof(argv[++a]); }
        else if (k == "--bc") { need(6); for (int d = 0; d < 6; ++d) o.bc[d] = std::atoi(argv[++a]); }
        else if (k == "--problem") { need(1); o.problem = std::atoi(argv[++a]); }
        else if (k == "--rhs") { need(1); o.rhs = argv[++a]; }
        else if (k == "--seed") { need(1); o.seed = std::strtoull(argv[++a], nullptr, 10); }
        else if (k == "--nsrc") { need(1); o.nsrc = std::atoi(argv[++a]); }
        else if (k == "--width") { need(2); o.wmin = std::atof(argv[++a]); o.wmax = std::atof(argv[++a]); }
        else if (k == "--spec-k") { need(1); o.specK = std::atoi(argv[++a]); }
        else if (k == "--solver") { need(1); o.solver = argv[++a]; }
        else if (k == "--precond") { need(1); o.precond = argv[++a]; }
        else if (k == "--cheb-m") { need(1); o.chebM = std::atoi(argv[++a]); }
        else if (k == "--cheb-eps") { need(1); epsilon = std::atof(argv[++a]); }
        else if (k == "--cheb-rescale") { need(2); rescaleEigMin = std::atof(argv[++a]); rescaleEigMax = std::atof(argv[++a]); }
        else if (k == "--tol") { need(1); o.tol = std::atof(argv[++a]); }
        else if (k == "--out") { need(1); o.out = argv[++a]; }
        else if (k == "--solution-dat") { need(1); o.solutionDat = argv[++a]; }
        else if (k == "--residual-history") { need(1); o.residualHistory = argv[++a]; }
        else { std::cerr << "unknown option " << k << std::endl; MPI_Abort(MPI_COMM_WORLD, 2); }
    }
    tollScalingFactor = o.tol;
    refProblemKind = o.problem;

    // solverPoissonMPI_CPU/src/main.cpp:51-55
    if (o.nranks[0] * o.nranks[1] * o.nranks[2] != world) {
        std::cerr << "Error: configuration of MPI ranks not coherent! DIM = " << DIM << " worldSize " << world
                  << " ranks " << o.nranks[0] << " " << o.nranks[1] << " " << o.nranks[2] << " " << std::endl;
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    const std::array<int, 3> guards{1, 1, 1};
    const std::array<T_data, 6> bcsValue{0, 0, 0, 0, 0, 0};
    Grid grid(o.nranks, rank, o.np, o.ds, o.origin, guards, o.bc, bcsValue);
    Comm comm(grid);
    Exact exact;
    const int rc = dispatch(o, grid, comm, exact, rank);
    if (rc < 0 && rank == 0) std::cerr << "unsupported solver chain " << o.solver << ":" << o.precond << std::endl;
    MPI_Finalize();
    return rc < 0 ? 3 : rc;
}
