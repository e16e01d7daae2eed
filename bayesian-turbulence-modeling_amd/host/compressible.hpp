// compressible.hpp -- the compressible forward model on the MI355X engine.
//
// Same contract as the reference's compressible layer: IdealGasEOS
// (compressible/include/IdealGasEOS.hpp:6-38), CompressibleBoundaryConditions
// (compressible/include/CompressibleBCs.hpp:10-72) and
// CompressibleForwardModel (compressible/include/CompressibleForwardModel.hpp,
// compressible/src/CompressibleForwardModel.cpp:6-110): bounds check, unpack
// to the 11 SST coefficients, a cold uniform start, the compressible SIMPLE
// solve, status classification, observables on the thermodynamic pressure
// with nu = mu(T_0) / rho_0, Gaussian log-likelihood.  The solve runs in
// librsg.so (rsg_comp_problem_create / rsg_comp_submit / rsg_comp_fetch).
#pragma once
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "forward_model.hpp"
#include "mesh.hpp"
#include "model.hpp"
#include "rsg.h"

namespace rsh {

struct IdealGasEOS {
    double gamma = 1.4, R = 287.0, Pr = 0.72, Pr_T = 0.9;
    double mu_ref = 1.716e-5, T_ref = 273.15, S_suth = 110.4;
    double Cp() const { return gamma * R / (gamma - 1.0); }
    double Cv() const { return R / (gamma - 1.0); }
    double viscosity(double T) const {
        const double ratio = T / T_ref;
        return mu_ref * ratio * std::sqrt(ratio) * (T_ref + S_suth) / (T + S_suth);
    }
    double conductivity(double mu, double muT) const { return mu * Cp() / Pr + muT * Cp() / Pr_T; }
    double density(double p, double T) const { return p / (R * T); }
    double temperature(double p, double rho) const { return p / (rho * R); }
    double pressure(double rho, double T) const { return rho * R * T; }
    double soundSpeed(double T) const { return std::sqrt(gamma * R * T); }
    double machNumber(double Umag, double T) const { return Umag / soundSpeed(T); }
    double totalEnthalpy(double T, double Umag2) const { return Cp() * T + 0.5 * Umag2; }
};

enum class TBCType { Dirichlet, Neumann, AdiabaticWall };

struct PatchTBC {
    TBCType type = TBCType::Neumann;
    double value = 300.0;
    std::string patchName;
};

struct CompressibleBoundaryConditions {
    std::vector<PatchBC> velocityBC, pressureBC, kBC, omegaBC;
    std::vector<PatchTBC> temperatureBC;
    // wall: no-slip, zero-gradient p, adiabatic, k-omega wall values; inlet:
    // velocity (Uin, 0, 0), T_in, k, omega; outlet: thermodynamic p_out,
    // zero-gradient otherwise; any other patch zero-gradient
    static CompressibleBoundaryConditions channelDefaults(const Mesh& mesh, double Uin, double T_in, double p_out,
                                                          double kIn, double omIn);
    FlowBoundaryConditions flow() const;
    // per-boundary-face RSG_TBC_* types and values (RSG_TBC_KEEP: face in no patch)
    void resolveT(const Mesh& mesh, std::vector<int32_t>& type, std::vector<double>& value) const;
};

// Host copy of a compressible solve's fields: p thermodynamic, pMech the
// solver's mechanical working pressure.
struct CompressibleFields {
    FlowFields flow;
    std::vector<double> pMech, T, rho, Tb;
};

class CompressibleForwardModel {
public:
    CompressibleForwardModel(const Mesh& mesh, const InferenceParameterSet& paramSet, const ObservationOperator& obsOp,
                             const CompressibleBoundaryConditions& bcs, const IdealGasEOS& eos,
                             const SolverSettings& settings = {}, const Vec3& uInit = Vec3(1, 0, 0),
                             double pInit = 101325.0, double TInit = 300.0, double kInit = 1e-4,
                             double omegaInit = 100.0);
    ~CompressibleForwardModel();
    CompressibleForwardModel(const CompressibleForwardModel&) = delete;
    CompressibleForwardModel& operator=(const CompressibleForwardModel&) = delete;

    EvaluationResult evaluate(const std::vector<double>& theta);
    // Many thetas solved concurrently (kernel-per-phase, up to four at once);
    // each result equals what evaluate() returns for that theta alone.
    std::vector<EvaluationResult> evaluateBatch(const std::vector<std::vector<double>>& thetas);
    double penalizedLogLikelihood(const std::vector<double>& theta);
    const InferenceParameterSet& paramSet() const { return paramSet_; }
    bool hasLastFields() const { return hasLastFields_; }
    const CompressibleFields& lastFields() const { return lastFields_; }
    const ConvergenceInfo& lastConvergence() const { return lastConv_; }
    double lastEnergyIterations() const { return lastEnergyIters_; }

    int64_t submit(const std::vector<double>& theta);
    EvaluationResult complete(int64_t ticket);

private:
    const Mesh& mesh_;
    InferenceParameterSet paramSet_;
    ObservationOperator obsOp_;
    CompressibleBoundaryConditions bcs_;
    IdealGasEOS eos_;
    SolverSettings settings_;
    Vec3 uInit_;
    double pInit_, TInit_, kInit_, omegaInit_;
    CompressibleFields lastFields_;
    bool hasLastFields_ = false;
    ConvergenceInfo lastConv_;
    double lastEnergyIters_ = 0.0;
    int meshId_ = -1, problemId_ = -1;
    uint64_t ctxGen_ = 0;
    std::mutex mu_;
    void ensureProblem();
};

}  // namespace rsh
