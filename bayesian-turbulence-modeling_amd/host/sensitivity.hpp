// sensitivity.hpp -- the warm-started finite-difference true-model gradient
// on the MI355X engine.
//
// Same contract as the reference ParameterSensitivity for the calls on the
// forward path (incompressible/include/ParameterSensitivity.hpp:45-110,
// incompressible/src/ParameterSensitivity.cpp:9-36, :73-83, :692-745):
//   solveState(θ11)         cold solve, the converged state kept as the fixed point
//   observe / logLik(θ11)   observations / Gaussian log-likelihood at that state
//   etaJacobianWarmFD(θ11)  ∂η/∂θ and ∂logL/∂θ by central differences, every
//                           perturbed solve warm-started from the fixed point
// The 22 perturbed solves of one gradient are submitted together and run
// concurrently on the device (one workgroup each), where the reference runs
// them one after the other.  The adjoint groundwork of the reference class
// (residual, dResidualDTheta, dObsDTheta, the tangent solves) is not on the
// forward path and is not provided.
#pragma once
#include <map>
#include <utility>
#include <vector>

#include "forward_model.hpp"

namespace rsh {

// TangentGradientResult (ParameterSensitivity.hpp) as the warm-FD path fills it.
struct TangentGradientResult {
    std::vector<std::vector<double>> dObsDTheta;   // nObs x 11
    std::vector<double> logLikGradient;            // 11
    std::vector<int> krylovIters;                  // 11 (0: no Krylov solve on this path)
    std::vector<double> krylovRelRes;              // 11
    std::vector<int> krylovConverged;              // 11: both perturbed solves did not diverge
    int nResidualEvals = 0;                        // total outer iterations of the perturbed solves
    int nColors = 0;
};

class ParameterSensitivity {
public:
    ParameterSensitivity(const Mesh& mesh, const ObservationOperator& obs, const FlowBoundaryConditions& bcs,
                         double nu, const SolverSettings& settings = {}, const Vec3& Uinit = Vec3(1, 0, 0),
                         double pInit = 0.0, double kInit = 1e-4, double omegaInit = 1.0);
    ~ParameterSensitivity();
    ParameterSensitivity(const ParameterSensitivity&) = delete;
    ParameterSensitivity& operator=(const ParameterSensitivity&) = delete;

    EvaluationStatus solveState(const std::vector<double>& theta11);
    bool hasState() const { return hasState_; }
    int nState() const { return 4 * mesh_.nCells(); }
    int nCells() const { return mesh_.nCells(); }
    int nObs() const { return obs_.nObs(); }
    std::vector<double> observe(const std::vector<double>& theta11);
    double logLik(const std::vector<double>& theta11);
    TangentGradientResult etaJacobianWarmFD(const std::vector<double>& theta11, double hRel = 5e-4,
                                            double hFloor = 1e-7, int warmMaxIter = 0, double warmTol = 0.0);
    const FlowFields& state() const { return state_; }

private:
    const Mesh& mesh_;
    ObservationOperator obs_;
    FlowBoundaryConditions bcs_;
    double nu_;
    SolverSettings settings_;
    Vec3 Uinit_;
    double pInit_, kInit_, omegaInit_;
    FlowFields state_;
    bool hasState_ = false;
    int meshId_ = -1;
    uint64_t ctxGen_ = 0;                               // engine generation the ids belong to
    std::map<std::pair<int, double>, int> problems_;   // (maxIterations, convergenceTol) -> problem id

    int problemFor(const SolverSettings& s);
    int64_t submit(int problem, const std::vector<double>& theta11, const FlowFields* warm);
    // pred / logLik (optional): the observables the device evaluated at the end of the solve
    void fetch(int64_t ticket, FlowFields& f, ConvergenceInfo& conv, std::vector<double>* pred = nullptr,
               double* logLik = nullptr);
};

}  // namespace rsh
