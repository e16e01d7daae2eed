// compressible.cpp -- CompressibleForwardModel over the librsg.so C-ABI.
#include "compressible.hpp"

#include <algorithm>
#include <stdexcept>

namespace rsh {

// CompressibleBCs.hpp:27-72
CompressibleBoundaryConditions CompressibleBoundaryConditions::channelDefaults(const Mesh& mesh, double Uin,
                                                                               double T_in, double p_out,
                                                                               double kIn, double omIn) {
    CompressibleBoundaryConditions bc;
    const int np = mesh.nPatches();
    bc.velocityBC.resize(np); bc.pressureBC.resize(np); bc.temperatureBC.resize(np);
    bc.kBC.resize(np); bc.omegaBC.resize(np);
    for (int p = 0; p < np; ++p) {
        const Patch& pat = mesh.patch(p);
        bc.velocityBC[p].patchName = bc.pressureBC[p].patchName = bc.temperatureBC[p].patchName = pat.name;
        bc.kBC[p].patchName = bc.omegaBC[p].patchName = pat.name;
        if (pat.type == "wall") {
            bc.velocityBC[p].type = BCType::WallNoSlip;
            bc.pressureBC[p].type = BCType::Neumann;
            bc.temperatureBC[p].type = TBCType::AdiabaticWall;
            bc.kBC[p].type = BCType::WallKOmega;
            bc.omegaBC[p].type = BCType::WallKOmega;
        } else if (pat.type == "inlet") {
            bc.velocityBC[p].type = BCType::InletVelocity; bc.velocityBC[p].vecValue = Vec3(Uin, 0, 0);
            bc.pressureBC[p].type = BCType::Neumann;
            bc.temperatureBC[p].type = TBCType::Dirichlet; bc.temperatureBC[p].value = T_in;
            bc.kBC[p].type = BCType::Dirichlet; bc.kBC[p].value = kIn;
            bc.omegaBC[p].type = BCType::Dirichlet; bc.omegaBC[p].value = omIn;
        } else if (pat.type == "outlet") {
            bc.velocityBC[p].type = BCType::Neumann;
            bc.pressureBC[p].type = BCType::OutletPressure; bc.pressureBC[p].value = p_out;
            bc.temperatureBC[p].type = TBCType::Neumann;
            bc.kBC[p].type = BCType::Neumann;
            bc.omegaBC[p].type = BCType::Neumann;
        } else {
            bc.temperatureBC[p].type = TBCType::Neumann;
        }
    }
    return bc;
}

FlowBoundaryConditions CompressibleBoundaryConditions::flow() const {
    FlowBoundaryConditions f;
    f.velocityBC = velocityBC; f.pressureBC = pressureBC; f.kBC = kBC; f.omegaBC = omegaBC;
    return f;
}

// applyTemperatureBC (CompressibleBCs.hpp:75-92): Dirichlet fixes the face
// value, Neumann and AdiabaticWall copy the owner cell's
void CompressibleBoundaryConditions::resolveT(const Mesh& mesh, std::vector<int32_t>& type,
                                              std::vector<double>& value) const {
    const int np = mesh.nPatches(), nIF = mesh.nInternalFaces(), nB = mesh.nFaces() - nIF;
    if ((int)temperatureBC.size() < np)
        throw std::invalid_argument("CompressibleBoundaryConditions: every mesh patch needs a temperature BC");
    type.assign(nB, RSG_TBC_KEEP);
    value.assign(nB, 0.0);
    for (int p = 0; p < np; ++p) {
        const PatchTBC& t = temperatureBC[p];
        for (int f : mesh.patch(p).faces) {
            const int b = f - nIF;
            if (b < 0 || b >= nB) throw std::invalid_argument("patch face is not a boundary face");
            type[b] = t.type == TBCType::Dirichlet ? RSG_TBC_DIRICHLET : RSG_TBC_ZERO_GRADIENT;
            value[b] = t.value;
        }
    }
}

CompressibleForwardModel::CompressibleForwardModel(const Mesh& mesh, const InferenceParameterSet& paramSet,
                                                   const ObservationOperator& obsOp,
                                                   const CompressibleBoundaryConditions& bcs, const IdealGasEOS& eos,
                                                   const SolverSettings& settings, const Vec3& uInit, double pInit,
                                                   double TInit, double kInit, double omegaInit)
    : mesh_(mesh), paramSet_(paramSet), obsOp_(obsOp), bcs_(bcs), eos_(eos), settings_(settings), uInit_(uInit),
      pInit_(pInit), TInit_(TInit), kInit_(kInit), omegaInit_(omegaInit) {
    // CompressibleForwardModel.cpp:21-31: a drag observable would integrate a
    // kinematic-pressure formula over a dimensional pressure
    for (const Observable& o : obsOp_.observables())
        if (o.type == ObsType::Drag)
            throw std::invalid_argument(
                "CompressibleForwardModel does not support generic Drag observations; use the DBNS "
                "wall-observation path");
}

CompressibleForwardModel::~CompressibleForwardModel() {
    if (problemId_ >= 0 || meshId_ >= 0) {
        if (rsg_context* c = Engine::get().existing(ctxGen_)) {
            if (problemId_ >= 0) rsg_problem_release(c, problemId_);
            if (meshId_ >= 0) rsg_mesh_release(c, meshId_);
        }
    }
}

void CompressibleForwardModel::ensureProblem() {
    if (problemId_ >= 0) return;
    rsg_context* c = Engine::get().ctx();
    if (!mesh_.hasWallDistance()) computeWallDistanceOnDevice(const_cast<Mesh&>(mesh_));
    int32_t mid = -1, pid = -1;
    int rc = rsg_mesh_upload(c, &mesh_.desc(), &mid);
    if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
    ResolvedBC R;
    bcs_.flow().resolve(mesh_, R);
    std::vector<int32_t> tT;
    std::vector<double> tV;
    bcs_.resolveT(mesh_, tT, tV);
    rsg_comp_desc d{};
    d.t_type = tT.data(); d.t_value = tV.data();
    d.gamma = eos_.gamma; d.R = eos_.R; d.Pr = eos_.Pr; d.Pr_T = eos_.Pr_T;
    d.mu_ref = eos_.mu_ref; d.T_ref = eos_.T_ref; d.S_suth = eos_.S_suth;
    d.alpha_t = settings_.alphaT;
    // the reference's compressible model builds SSTModel(coeffs) with the
    // default (standard SST) variant whatever the settings name
    SolverSettings s = settings_;
    s.sstVariant = 0;
    rsg_settings cs = s.toC();
    rc = rsg_comp_problem_create(c, mid, &R.desc, &d, &cs, &pid);
    if (rc != RSG_OK) { rsg_mesh_release(c, mid); throw std::runtime_error(engineError(c, rc)); }
    try {
        setProblemObservations(c, pid, mesh_, obsOp_);
    } catch (...) {
        rsg_problem_release(c, pid);
        rsg_mesh_release(c, mid);
        throw;
    }
    meshId_ = mid; problemId_ = pid;
    ctxGen_ = Engine::get().generation();
}

int64_t CompressibleForwardModel::submit(const std::vector<double>& theta) {
    std::lock_guard<std::mutex> l(mu_);
    ensureProblem();
    rsg_context* c = Engine::get().ctx();
    std::vector<double> coeffs = InferenceParameterSet::toVector(paramSet_.unpack(theta));
    rsg_comp_init init{};
    init.U[0] = uInit_.x; init.U[1] = uInit_.y; init.U[2] = uInit_.z;
    init.p = pInit_; init.T = TInit_; init.k = kInit_; init.omega = omegaInit_;
    int64_t t = -1;
    const int rc = rsg_comp_submit(c, problemId_, coeffs.data(), &init, &t);
    if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
    return t;
}

EvaluationResult CompressibleForwardModel::complete(int64_t ticket) {
    rsg_context* c = Engine::get().ctx();
    int rc = rsg_wait(c, &ticket, 1, -1.0);
    if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
    const int n = mesh_.nCells(), nB = mesh_.nFaces() - mesh_.nInternalFaces();
    CompressibleFields F;
    FlowFields& f = F.flow;
    f.resize(n, nB);
    F.pMech.assign(n, 0.0); F.T.assign(n, 0.0); F.rho.assign(n, 0.0); F.Tb.assign(std::max(nB, 1), 0.0);
    ConvergenceInfo conv;
    if (settings_.recordHistory) conv.history.assign(6 * (size_t)std::max(settings_.maxIterations, 1), 0.0);
    rsg_result r{};
    r.U = f.U.data(); r.p = f.p.data(); r.k = f.k.data(); r.omega = f.omega.data(); r.nuT = f.nuT.data();
    r.F1 = f.F1.data(); r.F2 = f.F2.data(); r.Pk = f.Pk.data(); r.CDkw = f.CDkw.data();
    r.U_bnd = f.Ub.data(); r.p_bnd = f.pb.data(); r.k_bnd = f.kb.data(); r.om_bnd = f.omb.data();
    r.history = conv.history.empty() ? nullptr : conv.history.data();
    r.history_cap = (int32_t)(conv.history.size() / 6);
    std::vector<double> pred(obsOp_.nObs());
    r.predictions = pred.empty() ? nullptr : pred.data();
    r.predictions_cap = (int32_t)pred.size();
    rc = rsg_fetch(c, ticket, &r);
    double energyIters = 0.0;
    if (rc == RSG_OK && r.status == RSG_TICKET_DONE)
        rc = rsg_comp_fetch(c, ticket, F.T.data(), F.rho.data(), F.pMech.data(), F.Tb.data(), &energyIters);
    rsg_release(c, ticket);
    if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
    if (r.status != RSG_TICKET_DONE) throw std::runtime_error("device solve failed: " + std::string(rsg_last_error(c)));
    conv.history.resize(6 * (size_t)r.history_len);
    conv.finalIter = r.final_iter; conv.converged = r.converged; conv.diverged = r.diverged;
    for (int i = 0; i < 6; ++i) conv.innerIters[i] = r.inner_iters[i];
    f.turbEstablished = r.turb_established != 0;

    EvaluationResult res;
    res.simpleIters = r.final_iter;
    res.latency = r.t_end - r.t_submit;
    res.deviceTime = r.t_end - r.t_start;
    std::lock_guard<std::mutex> l(mu_);
    lastConv_ = conv;
    lastEnergyIters_ = energyIters;
    // CompressibleForwardModel.cpp:55-68: the fields are kept even on divergence
    lastFields_ = std::move(F);
    hasLastFields_ = true;
    if (conv.diverged) {
        res.status = EvaluationStatus::Diverged;
        res.loglik = -1e30;
        return res;
    }
    if (r.n_predictions != obsOp_.nObs()) throw std::runtime_error("device observables missing");
    res.predictions = std::move(pred);
    res.loglik = obsOp_.nObs() > 0 ? r.log_lik : 0.0;
    res.status = conv.converged ? EvaluationStatus::Converged : EvaluationStatus::Unconverged;
    return res;
}

// CompressibleForwardModel.cpp:36-97
EvaluationResult CompressibleForwardModel::evaluate(const std::vector<double>& theta) {
    EvaluationResult result;
    if (!paramSet_.inBounds(theta)) {
        result.status = EvaluationStatus::InvalidParameters;
        result.loglik = -1e30;
        return result;
    }
    return complete(submit(theta));
}

std::vector<EvaluationResult> CompressibleForwardModel::evaluateBatch(const std::vector<std::vector<double>>& thetas) {
    std::vector<EvaluationResult> out(thetas.size());
    std::vector<int64_t> tickets(thetas.size(), -1);
    {
        AdmissionHold hold;
        for (size_t i = 0; i < thetas.size(); ++i) {
            if (!paramSet_.inBounds(thetas[i])) {
                out[i].status = EvaluationStatus::InvalidParameters;
                out[i].loglik = -1e30;
                continue;
            }
            tickets[i] = submit(thetas[i]);
        }
    }
    for (size_t i = 0; i < thetas.size(); ++i)
        if (tickets[i] >= 0) out[i] = complete(tickets[i]);
    return out;
}

// CompressibleForwardModel.cpp:104-110
double CompressibleForwardModel::penalizedLogLikelihood(const std::vector<double>& theta) {
    const EvaluationResult r = evaluate(theta);
    if (!std::isfinite(r.loglik) || r.status == EvaluationStatus::Diverged) return -1e30;
    return r.loglik;
}

}  // namespace rsh
