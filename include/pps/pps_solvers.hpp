// pps_solvers.hpp -- C++ drop-in for the reference's solver types, driving
// libpps.so through the C ABI of pps_abi.h.
//
// The reference composes its solver at compile time and calls it on host
// arrays (solverPoissonMPI_CPU/include/inputParam.hpp:22-37, src/main.cpp:77-110):
//
//     using T_Solver = BiCGSTAB<DIM, T_data, tollMainSolver, iterMaxMainSolver,
//                               isbiCGMainLoop1, communicationON, T_Preconditioner2>;
//     T_Solver solver(blockGrid, exactSolutionAndBCs, communicator);
//     solver.setProblem(fieldX, fieldB);
//     solver(fieldX, fieldB, operatorA);
//     solver.getNumIterationFinal(); solver.getErrorComputeOperator(); ...
//
// The same lines compile against this header: the template chain becomes the
// stage array of pps_solve, the arrays are copied to the GPU and x is copied
// back, and every getter reports the device solve.  One GPU: the rank grid is
// 1x1x1 and CommunicatorMPI's halo exchange is the device ghost fill.  Like the
// reference, a configuration error prints to std::cerr and exits with -1.
#pragma once

#include "../pps_abi.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <iostream>
#include <string>
#include <type_traits>
#include <vector>

namespace pps {

// Tolerance scale and Chebyshev spectrum bounds.  With PPS_USE_SOLVER_SETUP
// defined, the caller's own solverSetup.hpp (included before this header, as
// the drop-in headers of include/pps/dropin/ do) sets them, as it does for the
// reference; otherwise the defaults of the reference's CPU solverSetup.hpp:22,
// 37-39 hold, and the alpaka types take those of the alpaka
// solverSetup.hpp:59-62 (epsilon 0, rescaleEigMin 100).
#ifdef PPS_USE_SOLVER_SETUP
constexpr double tollScalingFactor = ::tollScalingFactor;
constexpr double epsilon = ::epsilon;
constexpr double rescaleEigMin = ::rescaleEigMin;
constexpr double rescaleEigMax = ::rescaleEigMax;
constexpr double alpakaEpsilon = ::epsilon;
constexpr double alpakaRescaleEigMin = ::rescaleEigMin;
constexpr double alpakaRescaleEigMax = ::rescaleEigMax;
#else
constexpr double tollScalingFactor = 1e-10;
constexpr double epsilon = 1e-4;
constexpr double rescaleEigMin = 500;
constexpr double rescaleEigMax = 1 - 1e-4;
constexpr double alpakaEpsilon = 0;
constexpr double alpakaRescaleEigMin = 100;
constexpr double alpakaRescaleEigMax = 1 - 1e-4;
#endif

namespace detail {
[[noreturn]] inline void config_error(const std::string& msg)
{
    std::cerr << "pps: " << msg << std::endl;
    std::exit(-1);
}
template <typename T> constexpr int precision() { return std::is_same<T, float>::value ? 32 : 64; }
inline pps_stage stage(int kind, int prec, double tol, int maxit)
{
    pps_stage s{};
    s.kind = kind;
    s.precision = prec;
    s.tol = tol;
    s.max_iter = maxit;
    s.cheb_eps = epsilon;
    s.cheb_rescale_min = rescaleEigMin;
    s.cheb_rescale_max = rescaleEigMax;
    return s;
}
// a solver type whose operator()(x, b) sets the problem itself, as the alpaka
// main solvers do (BiCGstabAlpaka.hpp:136-138)
template <class T, class = void> struct SetsProblemInside : std::false_type {};
template <class T> struct SetsProblemInside<T, std::void_t<decltype(T::kSetsProblemInside)>>
    : std::integral_constant<bool, T::kSetsProblemInside> {};
}  // namespace detail

// TimeCounter<DIM,T_data> (solverPoissonMPI_alpaka/include/solverSetup.hpp:206-286):
// the per-phase times of the alpaka main loop.  The device path runs a whole
// iteration as one queued sequence of fused kernels with no host-timed phases,
// so only the total (the device time of the solve) is filled; the phase
// fields stay zero.
template <int DIM, typename T_data>
class TimeCounter {
public:
    using dur = std::chrono::duration<double>;
    dur timeTotPreconditioner1{0}, timeTotPreconditioner2{0}, timeTotCommunication1{0}, timeTotCommunication2{0};
    dur timeTotKernel1{0}, timeTotKernel2{0}, timeTotKernel3{0}, timeTotKernel4{0}, timeTotKernel5{0}, timeTotKernel6{0};
    dur timeTotResetNeumanBCs{0}, timeTotAllReduction1{0}, timeTotAllReduction2{0};
    dur timeDevice{0};  // device time of the whole solve
    dur timeTotal{0};
    dur getTimeTotal()
    {
        timeTotal = timeTotPreconditioner1 + timeTotPreconditioner2 + timeTotCommunication1 + timeTotCommunication2 +
                    timeTotKernel1 + timeTotKernel2 + timeTotKernel3 + timeTotKernel4 + timeTotKernel5 + timeTotKernel6 +
                    timeTotResetNeumanBCs + timeTotAllReduction1 + timeTotAllReduction2 + timeDevice;
        return timeTotal;
    }
    // the reference's printout, milliseconds per cycle
    void printAverageTime(const int numCycles)
    {
        const double n = numCycles > 0 ? numCycles : 1;
        auto ms = [n](dur d) { return d.count() * 1000 / n; };
        std::cout << "Time in milliseconds " << std::endl;
        std::cout << "timeTotPreconditioner1 " << ms(timeTotPreconditioner1) << std::endl;
        std::cout << "timeTotPreconditioner2 " << ms(timeTotPreconditioner2) << std::endl;
        std::cout << "timeTotCommunication1 " << ms(timeTotCommunication1) << std::endl;
        std::cout << "timeTotCommunication2 " << ms(timeTotCommunication2) << std::endl;
        std::cout << "timeTotAllReduction1 " << ms(timeTotAllReduction1) << std::endl;
        std::cout << "timeTotAllReduction2 " << ms(timeTotAllReduction2) << std::endl;
        std::cout << "timeTotKernel1 " << ms(timeTotKernel1) << std::endl;
        std::cout << "timeTotKernel2 " << ms(timeTotKernel2) << std::endl;
        std::cout << "timeTotKernel3 " << ms(timeTotKernel3) << std::endl;
        std::cout << "timeTotKernel4 " << ms(timeTotKernel4) << std::endl;
        std::cout << "timeTotKernel5 " << ms(timeTotKernel5) << std::endl;
        std::cout << "timeTotKernel6 " << ms(timeTotKernel6) << std::endl;
        std::cout << "timeTotPreconditionerTot " << ms(timeTotPreconditioner1 + timeTotPreconditioner2) << std::endl;
        std::cout << "timeTotCommunicationTot " << ms(timeTotCommunication1 + timeTotCommunication2) << std::endl;
        std::cout << "timeTotAllReductionTot " << ms(timeTotAllReduction1 + timeTotAllReduction2) << std::endl;
        std::cout << "timeTotKernelsBBiCGstabTot "
                  << ms(timeTotKernel1 + timeTotKernel2 + timeTotKernel3 + timeTotKernel4 + timeTotKernel5 + timeTotKernel6)
                  << std::endl;
        std::cout << "timeTotResetNeumanBCs " << ms(timeTotResetNeumanBCs) << std::endl;
        std::cout << "timeTotDevice " << ms(timeDevice) << std::endl;
        std::cout << "timeTotal " << ms(getTimeTotal()) << std::endl;
    }
};

// BlockGrid<DIM,T_data> (blockGrid.hpp:11-367) on one GPU
template <int DIM, typename T_data>
class BlockGrid {
public:
    BlockGrid(const std::array<int, 3>& nranks, const int my_rank, const std::array<int, 3>& npglobal,
              const std::array<T_data, 3>& ds, const std::array<T_data, 3>& origin, const std::array<int, 3>& guards,
              const std::array<int, 6>& bcsType, const std::array<T_data, 6>& bcsValue)
        : nranks_(nranks), my_rank_(my_rank), npglobal_(npglobal), ds_(ds), origin_(origin), guards_(guards),
          bcsType_(bcsType), bcsValue_(bcsValue)
    {
        static_assert(DIM == 3, "this build implements DIM = 3");
        if (nranks[0] * nranks[1] * nranks[2] != 1 || my_rank != 0)
            detail::config_error("one GPU: the rank grid must be 1x1x1");
        if (guards[0] != 1 || guards[1] != 1 || guards[2] != 1) detail::config_error("guards must be {1,1,1}");
        for (int d = 0; d < 3; ++d) {
            indexLimitsData_[2 * d] = 1;
            indexLimitsData_[2 * d + 1] = npglobal[d] + 1;
            indexLimitsSolver_[2 * d] = 1 + (bcsType[2 * d] == 0);
            indexLimitsSolver_[2 * d + 1] = npglobal[d] + 1 - (bcsType[2 * d + 1] == 0);
        }
    }
    const std::array<int, 3>& getNranks() const { return nranks_; }
    int getMyrank() const { return my_rank_; }
    const std::array<int, 3>& getNpglobal() const { return npglobal_; }
    const std::array<T_data, 3>& getDs() const { return ds_; }
    const std::array<T_data, 3>& getOrigin() const { return origin_; }
    const std::array<int, 3>& getGuards() const { return guards_; }
    const std::array<int, 6>& getBcsType() const { return bcsType_; }
    const std::array<T_data, 6>& getBcsValue() const { return bcsValue_; }
    std::array<int, 3> getNlocalGuards() const { return {npglobal_[0] + 2, npglobal_[1] + 2, npglobal_[2] + 2}; }
    const std::array<int, 3>& getNlocalNoGuards() const { return npglobal_; }
    int getNtotLocalGuards() const { return (npglobal_[0] + 2) * (npglobal_[1] + 2) * (npglobal_[2] + 2); }
    int getNtotLocalNoGuards() const { return npglobal_[0] * npglobal_[1] * npglobal_[2]; }
    int getNtotNpglobal() const { return getNtotLocalNoGuards(); }
    const std::array<int, 6>& getIndexLimitsData() const { return indexLimitsData_; }
    const std::array<int, 6>& getIndexLimitsSolver() const { return indexLimitsSolver_; }
    // coordinate of guarded index i along d (iterativeSolverBase.hpp:547-549)
    T_data coord(int d, int i) const { return origin_[d] + (i - 1) * ds_[d]; }

private:
    std::array<int, 3> nranks_;
    int my_rank_;
    std::array<int, 3> npglobal_;
    std::array<T_data, 3> ds_, origin_;
    std::array<int, 3> guards_;
    std::array<int, 6> bcsType_;
    std::array<T_data, 6> bcsValue_;
    std::array<int, 6> indexLimitsData_{}, indexLimitsSolver_{};
};

// CommunicatorMPI / MatrixFreeOperatorA: on one GPU the halo exchange and the
// operator live in the library; the types keep the call sites unchanged
template <int DIM, typename T_data>
class CommunicatorMPI {
public:
    explicit CommunicatorMPI(const BlockGrid<DIM, T_data>&) {}
};
template <int DIM, typename T_data>
class MatrixFreeOperatorA {
public:
    explicit MatrixFreeOperatorA(const BlockGrid<DIM, T_data>&) {}
};
// AlpakaHelper (solverPoissonMPI_alpaka/include/alpakaHelper.hpp): device,
// queues and work divisions of the alpaka build; the library holds its own
template <int DIM, typename T_data>
class AlpakaHelper {
public:
    explicit AlpakaHelper(const BlockGrid<DIM, T_data>&) {}
};

// solution.dat as the alpaka main program writes it (main.cpp:135-146): the
// rank's whole guarded array fieldX, raw T_data in host byte order, at byte
// offset sizeof(T_data) * ntotLocalGuards * rank of a file opened for writing
// without truncation (MPI_MODE_CREATE | MPI_MODE_WRONLY).  Returns false if
// the file cannot be written.
template <int DIM, typename T_data>
bool writeSolutionDat(const BlockGrid<DIM, T_data>& blockGrid, const T_data fieldX[],
                      const std::string& path = "solution.dat")
{
    std::fstream f(path, std::ios::in | std::ios::out | std::ios::binary);
    if (!f.is_open()) {
        f.clear();
        f.open(path, std::ios::out | std::ios::binary);  // create
    }
    if (!f.is_open()) return false;
    const std::streamoff off =
        (std::streamoff)sizeof(T_data) * (std::streamoff)blockGrid.getNtotLocalGuards() * blockGrid.getMyrank();
    f.seekp(off);
    f.write(reinterpret_cast<const char*>(fieldX), (std::streamsize)sizeof(T_data) * blockGrid.getNtotLocalGuards());
    return (bool)f;
}

// ------------------------------------------------------------- solver base

template <class Derived, int DIM, typename T_data>
class SolverFacade {
public:
    // the alpaka constructor (BiCGstabAlpaka.hpp:22-23) takes the AlpakaHelper too
    template <class T_Exact, class T_Comm>
    SolverFacade(const BlockGrid<DIM, T_data>& g, const T_Exact& exact, T_Comm& comm, const AlpakaHelper<DIM, T_data>&)
        : SolverFacade(g, exact, comm)
    {
    }
    template <class T_Exact, class T_Comm>
    SolverFacade(const BlockGrid<DIM, T_data>& g, const T_Exact& exact, T_Comm&) : g_(g)
    {
        static_assert(std::is_same<T_data, double>::value, "host fields are double (T_data of solverSetup.hpp)");
        setFieldB_ = [&exact](T_data x, T_data y, T_data z) { return exact.setFieldB(x, y, z); };
        trueSolution_ = [&exact](T_data x, T_data y, T_data z) { return exact.trueSolutionFxyz(x, y, z); };
        trueDdir_ = [&exact](T_data x, T_data y, T_data z, int d) { return exact.trueSolutionDdir(x, y, z, d); };
        Derived::stages(chain_);
        maxIteration_ = chain_[0].max_iter;
        pps_grid pg{};
        pg.dim = DIM;
        for (int d = 0; d < 3; ++d) {
            pg.nranks[d] = g.getNranks()[d];
            pg.npglobal[d] = g.getNpglobal()[d];
            pg.ds[d] = g.getDs()[d];
            pg.origin[d] = g.getOrigin()[d];
            pg.guards[d] = g.getGuards()[d];
        }
        pg.my_rank = g.getMyrank();
        for (int f = 0; f < 6; ++f) {
            pg.bcs_type[f] = g.getBcsType()[f];
            pg.bcs_value[f] = g.getBcsValue()[f];
        }
        if (pps_create(&pg, &ctx_) != PPS_OK) detail::config_error(pps_last_error(nullptr));
    }
    ~SolverFacade()
    {
        if (ctx_) pps_destroy(ctx_);
    }
    SolverFacade(const SolverFacade&) = delete;
    SolverFacade& operator=(const SolverFacade&) = delete;

    // setProblem (iterativeSolverBase.hpp:51-55,537-603): b = setFieldB on the
    // data points, x = u_exact on the Dirichlet boundary layers
    void setProblem(T_data fieldX[], T_data fieldB[])
    {
        const auto L = g_.getIndexLimitsData();
        for (int k = L[4]; k < L[5]; ++k)
            for (int j = L[2]; j < L[3]; ++j)
                for (int i = L[0]; i < L[1]; ++i)
                    fieldB[idx(i, j, k)] = setFieldB_(g_.coord(0, i), g_.coord(1, j), g_.coord(2, k));
        for (int f = 0; f < 6; ++f) {
            if (g_.getBcsType()[f] != 0) continue;
            const int d = f / 2, pos = (f % 2) ? L[2 * d + 1] - 1 : L[2 * d];
            int lo[3] = {L[0], L[2], L[4]}, hi[3] = {L[1], L[3], L[5]};
            lo[d] = pos;
            hi[d] = pos + 1;
            for (int k = lo[2]; k < hi[2]; ++k)
                for (int j = lo[1]; j < hi[1]; ++j)
                    for (int i = lo[0]; i < hi[0]; ++i)
                        fieldX[idx(i, j, k)] = trueSolution_(g_.coord(0, i), g_.coord(1, j), g_.coord(2, k));
        }
    }

    // BiCGstabAlpaka::operator()(T_data*, T_data*) (BiCGstabAlpaka.hpp:68-475):
    // copy in, solve on the device, copy x back.  The alpaka main solvers set
    // the problem themselves first (BiCGstabAlpaka.hpp:136-138); the CPU ones
    // leave it to the caller's setProblem (main.cpp:94)
    void operator()(T_data fieldX[], T_data fieldB[])
    {
        if constexpr (detail::SetsProblemInside<Derived>::value) setProblem(fieldX, fieldB);
        check(pps_upload(ctx_, PPS_FIELD_X, fieldX));
        check(pps_upload(ctx_, PPS_FIELD_B, fieldB));
        set_flux();
        pps_result r{};
        check(pps_solve(ctx_, chain_.data(), (int)chain_.size(), &r));
        check(pps_download(ctx_, PPS_FIELD_X, fieldX));
        numIterationFinal_ = r.iterations;
        numIterationPreconditionerFinal_ = (int)r.precond_iterations;
        errorFromIteration_ = r.error_from_iteration;
        errorComputeOperator_ = r.error_compute_operator;
        durationSolver_ = std::chrono::duration<double>(r.seconds);
        timeCounter.timeDevice += durationSolver_;
        outerIterations_ = r.outer_iterations;
        fallbackSwitches_ = r.fallback_switches;
        history_.assign(r.history_len, 0.0);
        if (r.history_len > 0 && pps_get_history(ctx_, history_.data(), r.history_len) < 0)
            detail::config_error(pps_last_error(ctx_));
    }
    // CPU form solver(x, b, operatorA) (baseCG.hpp:44)
    template <class T_Op>
    void operator()(T_data fieldX[], T_data fieldB[], T_Op&)
    {
        (*this)(fieldX, fieldB);
    }

    int getNumIterationFinal() const { return numIterationFinal_; }
    int getNumIterationPreconditionerFinal() const { return numIterationPreconditionerFinal_; }
    T_data getErrorFromIteration() const { return errorFromIteration_; }
    T_data getErrorComputeOperator() const { return errorComputeOperator_; }
    // iterativeSolverBase.hpp:410 (the device time of the solve, HIP events)
    std::chrono::duration<double> getDurationSolver() const { return durationSolver_; }
    const std::vector<double>& getErrorFromIterationHistory() const { return history_; }

    // checkSolutionLocalGlobal (iterativeSolverBase.hpp:283-408): resets x's
    // Neumann ghosts with the exact flux, compares x with trueSolutionFxyz on
    // the data region, prints the largest block-average and point errors as
    // the reference's rank 0 does, and returns the sum of |x - u| (errorLocal).
    // One GPU: the rank grid is one block, so both maxima are this block's.
    T_data checkSolutionLocalGlobal(T_data fieldX[])
    {
        reset_neumann_host(fieldX);
        const auto L = g_.getIndexLimitsData();
        T_data errorLocal = 0, errorLocalMax = -1;
        for (int k = L[4]; k < L[5]; ++k)
            for (int j = L[2]; j < L[3]; ++j)
                for (int i = L[0]; i < L[1]; ++i) {
                    const T_data u = trueSolution_(g_.coord(0, i), g_.coord(1, j), g_.coord(2, k));
                    const T_data err = std::abs(fieldX[idx(i, j, k)] - u);
                    errorLocal += err;
                    if (err > errorLocalMax) errorLocalMax = err;
                }
        std::cout << "Max error local block avg " << errorLocal / g_.getNtotLocalNoGuards() << " in rank " << 0
                  << std::endl;
        std::cout << "Max error local point " << errorLocalMax << " in rank " << 0 << std::endl;
        return errorLocal;
    }

    // writeResidualHistory (iterativeSolverBaseAlpaka.hpp:620-638): duration,
    // main iterations, preconditioner iterations, then the recursive residual
    // of iterations 0, 1, ... up to the first non-positive entry or
    // maxIteration entries, one value per line in default stream format
    void writeResidualHistory(const std::string& path = "residualHistory.txt") const
    {
        std::ofstream outFile(path);
        if (!outFile) std::cerr << "Error: Could not open the file!" << std::endl;
        outFile << getDurationSolver().count() << "\n";
        outFile << getNumIterationFinal() << "\n";
        outFile << getNumIterationPreconditionerFinal() << "\n";
        const int n = std::min<int>(maxIteration_, (int)history_.size());
        for (int i = 0; i < n && history_[i] > 0; i++) outFile << history_[i] << "\n";
        outFile.close();
        std::cout << "residualHistory has been written to " << path << std::endl;
    }

    // pps_stage::true_finish of the main solver (CG, BiCGSTAB, FCG): after the
    // reference's recursive stop, resume from r = b - A x until the true FP64
    // residual is below the tolerance, at most n times (0, the default: the
    // reference's stop rule)
    void setTrueResidualFinish(int n) { chain_[0].true_finish = n; }

    TimeCounter<DIM, T_data> timeCounter;
    int getOuterIterations() const { return outerIterations_; }
    int getFallbackSwitches() const { return fallbackSwitches_; }
    // the device context (kernel-level entry points, pps_nn_attach)
    void* context() const { return ctx_; }

private:
    int idx(int i, int j, int k) const
    {
        const auto ng = g_.getNlocalGuards();
        return i + ng[0] * (j + ng[1] * k);
    }
    // resetNeumanBCs<true,true> on a host array (iterativeSolverBase.hpp:62-169,
    // orderNeumanBcs 2, normFieldB_ 1 after the solve): guard = mirrored data
    // point -+ 2 ds du/dn
    void reset_neumann_host(T_data* f) const
    {
        const auto L = g_.getIndexLimitsData();
        const auto ng = g_.getNlocalGuards();
        const int stride[3] = {1, ng[0], ng[0] * ng[1]};
        for (int face = 0; face < 6; ++face) {
            if (g_.getBcsType()[face] != 1) continue;
            const int d = face / 2, side = face % 2;
            int lo[3] = {L[0], L[2], L[4]}, hi[3] = {L[1], L[3], L[5]};
            const int pos = side ? L[2 * d + 1] - 1 : L[2 * d];
            lo[d] = pos;
            hi[d] = pos + 1;
            const T_data ds = g_.getDs()[d];
            for (int k = lo[2]; k < hi[2]; ++k)
                for (int j = lo[1]; j < hi[1]; ++j)
                    for (int i = lo[0]; i < hi[0]; ++i) {
                        const int p = idx(i, j, k);
                        const T_data q = trueDdir_(g_.coord(0, i), g_.coord(1, j), g_.coord(2, k), d);
                        if (side == 0) f[p - stride[d]] = f[p + stride[d]] - 2 * ds * q;
                        else f[p + stride[d]] = f[p - stride[d]] + 2 * ds * q;
                    }
        }
    }
    void check(int rc)
    {
        if (rc != PPS_OK) detail::config_error(pps_last_error(ctx_));
    }
    // du/dn on the Neumann faces: trueSolutionDdir on the face's data points,
    // plane of the other two guarded extents, lower direction fastest
    void set_flux()
    {
        const auto L = g_.getIndexLimitsData();
        const auto ng = g_.getNlocalGuards();
        for (int f = 0; f < 6; ++f) {
            if (g_.getBcsType()[f] != 1) continue;
            const int d = f / 2, pos = (f % 2) ? L[2 * d + 1] - 1 : L[2 * d];
            const int o0 = d == 0 ? 1 : 0, o1 = d == 2 ? 1 : 2;
            std::vector<double> plane((size_t)ng[o0] * ng[o1], 0.0);
            for (int i1 = L[2 * o1]; i1 < L[2 * o1 + 1]; ++i1)
                for (int i0 = L[2 * o0]; i0 < L[2 * o0 + 1]; ++i0) {
                    T_data c[3];
                    c[d] = g_.coord(d, pos);
                    c[o0] = g_.coord(o0, i0);
                    c[o1] = g_.coord(o1, i1);
                    plane[(size_t)i1 * ng[o0] + i0] = trueDdir_(c[0], c[1], c[2], d);
                }
            check(pps_set_neumann_flux(ctx_, f, plane.data()));
        }
    }

    const BlockGrid<DIM, T_data>& g_;
    std::function<T_data(T_data, T_data, T_data)> setFieldB_, trueSolution_;
    std::function<T_data(T_data, T_data, T_data, int)> trueDdir_;
    std::vector<pps_stage> chain_;
    void* ctx_ = nullptr;
    int numIterationFinal_ = 0, numIterationPreconditionerFinal_ = 0, outerIterations_ = 0, fallbackSwitches_ = 0;
    T_data errorFromIteration_ = -1, errorComputeOperator_ = -1;
    std::chrono::duration<double> durationSolver_{0};
    std::vector<double> history_;
    int maxIteration_ = 0;
};

// ------------------------------------------------------------- solver types

// NoneSolver<DIM,T_data,tolerance,maxIteration> (noneSolver.hpp:14-28)
template <int DIM, typename T_data, int tolerance, int maxIteration>
struct NoneSolver {
    static void stages(std::vector<pps_stage>& v) { v.push_back(detail::stage(PPS_NONE, detail::precision<T_data>(), 0, 1)); }
};

// diagonal preconditioner (no reference counterpart)
template <int DIM, typename T_data>
struct Jacobi {
    static void stages(std::vector<pps_stage>& v) { v.push_back(detail::stage(PPS_JACOBI, detail::precision<T_data>(), 0, 1)); }
};

#define PPS_KRYLOV_TYPE(NAME, KIND)                                                                                  \
    template <int DIM, typename T_data, int tolerance, int maxIteration, bool isMainLoop, bool communicationON,      \
              typename T_Preconditioner>                                                                             \
    class NAME : public SolverFacade<NAME<DIM, T_data, tolerance, maxIteration, isMainLoop, communicationON,         \
                                          T_Preconditioner>,                                                         \
                                     DIM, double> {                                                                  \
    public:                                                                                                          \
        using SolverFacade<NAME, DIM, double>::SolverFacade;                                                         \
        static void stages(std::vector<pps_stage>& v)                                                                \
        {                                                                                                            \
            v.push_back(detail::stage(KIND, detail::precision<T_data>(), tolerance * tollScalingFactor, maxIteration)); \
            T_Preconditioner::stages(v);                                                                             \
        }                                                                                                            \
    };

// BaseCG<...> (baseCG.hpp:14-270)
PPS_KRYLOV_TYPE(BaseCG, PPS_CG)
// BiCGSTAB<...> (BiCGSTAB.hpp:14-335; BiCGstabAlpaka.hpp:19-888)
PPS_KRYLOV_TYPE(BiCGSTAB, PPS_BICGSTAB)
// flexible CG, truncation 1 (no reference counterpart)
PPS_KRYLOV_TYPE(FCG, PPS_FCG)
#undef PPS_KRYLOV_TYPE

// ChebyshevIteration<...> (chebyshevIteration.hpp:15-157); maxIteration is the
// degree chebyshevMax (solverSetup.hpp:40)
template <int DIM, typename T_data, int tolerance, int maxIteration, bool isMainLoop, bool communicationON,
          typename T_Preconditioner>
class ChebyshevIteration
    : public SolverFacade<ChebyshevIteration<DIM, T_data, tolerance, maxIteration, isMainLoop, communicationON,
                                             T_Preconditioner>,
                          DIM, double> {
public:
    using SolverFacade<ChebyshevIteration, DIM, double>::SolverFacade;
    static void stages(std::vector<pps_stage>& v)
    {
        v.push_back(detail::stage(PPS_CHEBYSHEV, detail::precision<T_data>(), tolerance * tollScalingFactor, maxIteration));
        T_Preconditioner::stages(v);
    }
};

// FP64 outer refinement r = b - A u, u += d around an FP32 inner solver type
// (BaseCG or FCG with T_data = float); no reference counterpart.
// firstInnerTolMicro > 0: the inner tolerance of the first outer step, in
// units of 1e-6 (pps_stage::ir_first_tol); adaptPercent > 0: an outer step
// asks the inner solver for no more reduction than the outer tolerance still
// needs, times adaptPercent / 100 (pps_stage::ir_adapt)
template <int DIM, int tolerance, int maxOuter, typename T_Inner, int firstInnerTolMicro = 0, int adaptPercent = 0>
class MixedPrecisionIR
    : public SolverFacade<MixedPrecisionIR<DIM, tolerance, maxOuter, T_Inner, firstInnerTolMicro, adaptPercent>, DIM,
                          double> {
public:
    using SolverFacade<MixedPrecisionIR, DIM, double>::SolverFacade;
    static void stages(std::vector<pps_stage>& v)
    {
        pps_stage s = detail::stage(PPS_IR, 64, tolerance * tollScalingFactor, maxOuter);
        s.ir_first_tol = firstInnerTolMicro * 1e-6;
        s.ir_adapt = adaptPercent * 0.01;
        v.push_back(s);
        T_Inner::stages(v);
    }
};

// learned multilevel preconditioner (attach a network with pps_nn_attach on
// solver.context()); stallWindow/stallRatio: Jacobi fallback when FCG stalls
template <typename T_data, int stallWindow = 0, int stallRatioPercent = 0>
struct LearnedPreconditioner {
    static void stages(std::vector<pps_stage>& v)
    {
        pps_stage s = detail::stage(PPS_NN, detail::precision<T_data>(), 0, 1);
        s.stall_window = stallWindow;
        s.stall_ratio = stallRatioPercent / 100.0;
        v.push_back(s);
    }
};

// ------------------------------------------------------- alpaka solver types
//
// The aliases of solverPoissonMPI_alpaka/include/inputParam.hpp:25-37 with
// the alpaka template parameters.  Their constructors take the AlpakaHelper,
// and a main solver's operator()(x, b) sets the problem itself.

// NoneSolverAlpaka<DIM,T_data,T_data_noneSolver,tolerance,maxIteration> (noneSolverAlpaka.hpp:17)
template <int DIM, typename T_data, typename T_data_noneSolver, int tolerance, int maxIteration>
struct NoneSolverAlpaka {
    static void stages(std::vector<pps_stage>& v)
    {
        v.push_back(detail::stage(PPS_NONE, detail::precision<T_data_noneSolver>(), 0, 1));
    }
};

// BiCGstabAlpaka<DIM,T_data,T_data_BiCGStab,tolerance,maxIteration,isMainLoop,communicationON,T_Preconditioner>
// (BiCGstabAlpaka.hpp:19)
template <int DIM, typename T_data, typename T_data_BiCGStab, int tolerance, int maxIteration, bool isMainLoop,
          bool communicationON, typename T_Preconditioner>
class BiCGstabAlpaka
    : public SolverFacade<BiCGstabAlpaka<DIM, T_data, T_data_BiCGStab, tolerance, maxIteration, isMainLoop,
                                         communicationON, T_Preconditioner>,
                          DIM, double> {
public:
    static constexpr bool kSetsProblemInside = isMainLoop;
    using SolverFacade<BiCGstabAlpaka, DIM, double>::SolverFacade;
    static void stages(std::vector<pps_stage>& v)
    {
        static_assert(std::is_same<T_data_BiCGStab, double>::value, "BiCGSTAB vectors are FP64");
        v.push_back(detail::stage(PPS_BICGSTAB, 64, tolerance * tollScalingFactor, maxIteration));
        T_Preconditioner::stages(v);
    }
};

// ChebyshevIterationAlpaka<DIM,T_data,T_data_chebyshev,global,tolerance,maxIteration,isMainLoop,
// communicationON,T_Preconditioner> (chebyshevIterationAlpaka.hpp:16).  global: the rescaled global
// eigenvalue bounds (epsilon, rescaleEigMin/Max), local: the block's own bounds, not rescaled
// (:26-30); T_data_chebyshev = float runs the sweeps in FP32 on a cast copy of the input (:149-196)
template <int DIM, typename T_data, typename T_data_chebyshev, bool global, int tolerance, int maxIteration,
          bool isMainLoop, bool communicationON, typename T_Preconditioner>
class ChebyshevIterationAlpaka
    : public SolverFacade<ChebyshevIterationAlpaka<DIM, T_data, T_data_chebyshev, global, tolerance, maxIteration,
                                                   isMainLoop, communicationON, T_Preconditioner>,
                          DIM, double> {
public:
    static constexpr bool kSetsProblemInside = isMainLoop;
    using SolverFacade<ChebyshevIterationAlpaka, DIM, double>::SolverFacade;
    static void stages(std::vector<pps_stage>& v)
    {
        pps_stage s = detail::stage(PPS_CHEBYSHEV, detail::precision<T_data_chebyshev>(), tolerance * tollScalingFactor,
                                    maxIteration);
        s.cheb_eps = alpakaEpsilon;
        s.cheb_rescale_min = alpakaRescaleEigMin;
        s.cheb_rescale_max = alpakaRescaleEigMax;
        s.cheb_local = global ? 0 : 1;
        v.push_back(s);
        T_Preconditioner::stages(v);
    }
};

}  // namespace pps
