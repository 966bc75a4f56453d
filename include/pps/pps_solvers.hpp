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

#include <array>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <iostream>
#include <string>
#include <type_traits>
#include <vector>

namespace pps {

// solverSetup.hpp:22,37-39 (CPU form)
constexpr double tollScalingFactor = 1e-10;
constexpr double epsilon = 1e-4;
constexpr double rescaleEigMin = 500;
constexpr double rescaleEigMax = 1 - 1e-4;

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
}  // namespace detail

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

// ------------------------------------------------------------- solver base

template <class Derived, int DIM, typename T_data>
class SolverFacade {
public:
    template <class T_Exact, class T_Comm>
    SolverFacade(const BlockGrid<DIM, T_data>& g, const T_Exact& exact, T_Comm&) : g_(g)
    {
        static_assert(std::is_same<T_data, double>::value, "host fields are double (T_data of solverSetup.hpp)");
        setFieldB_ = [&exact](T_data x, T_data y, T_data z) { return exact.setFieldB(x, y, z); };
        trueSolution_ = [&exact](T_data x, T_data y, T_data z) { return exact.trueSolutionFxyz(x, y, z); };
        trueDdir_ = [&exact](T_data x, T_data y, T_data z, int d) { return exact.trueSolutionDdir(x, y, z, d); };
        Derived::stages(chain_);
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
    // copy in, solve on the device, copy x back
    void operator()(T_data fieldX[], T_data fieldB[])
    {
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
        durationSolver_ = r.seconds;
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
    double getDurationSolver() const { return durationSolver_; }
    const std::vector<double>& getErrorFromIterationHistory() const { return history_; }
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
    double durationSolver_ = 0;
    std::vector<double> history_;
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
// units of 1e-6 (pps_stage::ir_first_tol)
template <int DIM, int tolerance, int maxOuter, typename T_Inner, int firstInnerTolMicro = 0>
class MixedPrecisionIR
    : public SolverFacade<MixedPrecisionIR<DIM, tolerance, maxOuter, T_Inner, firstInnerTolMicro>, DIM, double> {
public:
    using SolverFacade<MixedPrecisionIR, DIM, double>::SolverFacade;
    static void stages(std::vector<pps_stage>& v)
    {
        pps_stage s = detail::stage(PPS_IR, 64, tolerance * tollScalingFactor, maxOuter);
        s.ir_first_tol = firstInnerTolMicro * 1e-6;
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

}  // namespace pps
