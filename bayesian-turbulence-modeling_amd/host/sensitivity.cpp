// sensitivity.cpp -- ParameterSensitivity (warm-FD gradient) on the engine.
#include "sensitivity.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace rsh {

namespace {

void check11(const std::vector<double>& theta11) {
    if (theta11.size() != 11)
        throw std::invalid_argument("theta must be the full 11-coefficient vector (InferenceParameters order)");
}

}  // namespace

// ParameterSensitivity.cpp:9-18
ParameterSensitivity::ParameterSensitivity(const Mesh& mesh, const ObservationOperator& obs,
                                           const FlowBoundaryConditions& bcs, double nu,
                                           const SolverSettings& settings, const Vec3& Uinit, double pInit,
                                           double kInit, double omegaInit)
    : mesh_(mesh), obs_(obs), bcs_(bcs), nu_(nu), settings_(settings), Uinit_(Uinit), pInit_(pInit),
      kInit_(kInit), omegaInit_(omegaInit) {
    state_.resize(mesh_.nCells(), mesh_.nFaces() - mesh_.nInternalFaces());
}

ParameterSensitivity::~ParameterSensitivity() {
    // never starts an engine (see ForwardModel::~ForwardModel)
    if (meshId_ < 0) return;
    if (rsg_context* c = Engine::get().existing(ctxGen_)) {
        for (const auto& kv : problems_) rsg_problem_release(c, kv.second);
        rsg_mesh_release(c, meshId_);
    }
}

int ParameterSensitivity::problemFor(const SolverSettings& s) {
    const std::pair<int, double> key(s.maxIterations, s.convergenceTol);
    auto it = problems_.find(key);
    if (it != problems_.end()) return it->second;
    rsg_context* c = Engine::get().ctx();
    if (meshId_ < 0) {
        if (!mesh_.hasWallDistance()) computeWallDistanceOnDevice(const_cast<Mesh&>(mesh_));
        int32_t mid = -1;
        const int rc = rsg_mesh_upload(c, &mesh_.desc(), &mid);
        if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
        meshId_ = mid;
        ctxGen_ = Engine::get().generation();
    }
    ResolvedBC R;
    bcs_.resolve(mesh_, R);
    const rsg_settings cs = s.toC();
    int32_t pid = -1;
    const int rc = rsg_problem_create(c, meshId_, &R.desc, nu_, &cs, &pid);
    if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
    try {
        setProblemObservations(c, pid, mesh_, obs_);
    } catch (...) {
        rsg_problem_release(c, pid);
        throw;
    }
    problems_[key] = pid;
    return pid;
}

// warm == nullptr: SIMPLESolver::initUniform(f, Uinit, pInit, kInit, omegaInit);
// otherwise the solve starts from *warm (fields, boundary values, turbEstablished).
int64_t ParameterSensitivity::submit(int problem, const std::vector<double>& theta11, const FlowFields* warm) {
    rsg_context* c = Engine::get().ctx();
    const std::vector<double> coeffs = InferenceParameterSet::toVector(InferenceParameterSet::fromVector(theta11));
    rsg_init init{};
    init.U[0] = Uinit_.x; init.U[1] = Uinit_.y; init.U[2] = Uinit_.z;
    init.p = pInit_; init.k = kInit_; init.omega = omegaInit_;
    if (warm) {
        const FlowFields& f = *warm;
        init.kind = 1;
        init.U_cells = f.U.data(); init.p_cells = f.p.data(); init.k_cells = f.k.data(); init.om_cells = f.omega.data();
        init.nuT = f.nuT.data(); init.F1 = f.F1.data(); init.F2 = f.F2.data(); init.Pk = f.Pk.data();
        init.CDkw = f.CDkw.data();
        init.U_bnd = f.Ub.data(); init.p_bnd = f.pb.data(); init.k_bnd = f.kb.data(); init.om_bnd = f.omb.data();
        init.turb_established = f.turbEstablished ? 1 : 0;
    }
    int64_t t = -1;
    const int rc = rsg_submit(c, problem, coeffs.data(), &init, RSG_PREC_FP64, &t);   // copies the warm state
    if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
    return t;
}

void ParameterSensitivity::fetch(int64_t ticket, FlowFields& f, ConvergenceInfo& conv, std::vector<double>* pred,
                                 double* logLik) {
    rsg_context* c = Engine::get().ctx();
    int rc = rsg_wait(c, &ticket, 1, -1.0);
    if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
    f.resize(mesh_.nCells(), mesh_.nFaces() - mesh_.nInternalFaces());
    rsg_result r{};
    r.U = f.U.data(); r.p = f.p.data(); r.k = f.k.data(); r.omega = f.omega.data(); r.nuT = f.nuT.data();
    r.F1 = f.F1.data(); r.F2 = f.F2.data(); r.Pk = f.Pk.data(); r.CDkw = f.CDkw.data();
    r.U_bnd = f.Ub.data(); r.p_bnd = f.pb.data(); r.k_bnd = f.kb.data(); r.om_bnd = f.omb.data();
    if (pred) {
        pred->assign(obs_.nObs(), 0.0);
        r.predictions = pred->empty() ? nullptr : pred->data();
        r.predictions_cap = (int32_t)pred->size();
    }
    rc = rsg_fetch(c, ticket, &r);
    rsg_release(c, ticket);
    if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
    if (r.status != RSG_TICKET_DONE) throw std::runtime_error("device solve failed: " + std::string(rsg_last_error(c)));
    if (pred && r.n_predictions != obs_.nObs()) throw std::runtime_error("device observables missing");
    if (logLik) *logLik = obs_.nObs() > 0 ? r.log_lik : 0.0;
    conv.finalIter = r.final_iter; conv.converged = r.converged; conv.diverged = r.diverged;
    for (int i = 0; i < 6; ++i) conv.innerIters[i] = r.inner_iters[i];
    f.turbEstablished = r.turb_established != 0;
}

// ParameterSensitivity.cpp:26-36
EvaluationStatus ParameterSensitivity::solveState(const std::vector<double>& theta11) {
    check11(theta11);
    FlowFields f;
    ConvergenceInfo conv;
    fetch(submit(problemFor(settings_), theta11, nullptr), f, conv);
    state_ = std::move(f);
    hasState_ = true;
    if (conv.diverged) return EvaluationStatus::Diverged;
    if (conv.converged) return EvaluationStatus::Converged;
    return EvaluationStatus::Unconverged;
}

// ParameterSensitivity.cpp:73-83.  The reference recomputes the closure
// (nuT, F1, F2, Pk, CDkw) from θ at the stored state first; the observation
// operator reads only U and p (ObservationOperator.hpp:159-378), so the
// observations and the log-likelihood are those of the stored state.
std::vector<double> ParameterSensitivity::observe(const std::vector<double>& theta11) {
    check11(theta11);
    return obs_.evaluate(mesh_, state_, nu_);
}

double ParameterSensitivity::logLik(const std::vector<double>& theta11) {
    return obs_.logLikelihood(observe(theta11));
}

// ParameterSensitivity.cpp:692-745: central differences of η and logL, each
// perturbed solve warm-started from the fixed point with the looser warm cap
// and tolerance; β* and a1 (the coefficients nonlinear in the closure) get
// the smaller relative step.  All 22 solves are in flight at once.
TangentGradientResult ParameterSensitivity::etaJacobianWarmFD(const std::vector<double>& theta11, double hRel,
                                                              double hFloor, int warmMaxIter, double warmTol) {
    check11(theta11);
    const int nObsv = obs_.nObs();
    TangentGradientResult out;
    out.dObsDTheta.assign(nObsv, std::vector<double>(11, 0.0));
    out.logLikGradient.assign(11, 0.0);
    out.krylovIters.assign(11, 0);
    out.krylovRelRes.assign(11, 0.0);
    out.krylovConverged.assign(11, 0);
    if (!hasState_) return out;

    SolverSettings warm = settings_;
    if (warmMaxIter > 0) warm.maxIterations = warmMaxIter;
    if (warmTol > 0.0) warm.convergenceTol = warmTol;
    const int pid = problemFor(warm);
    auto stepFor = [&](int j, double tj) {
        const double hr = (j == 8 || j == 9) ? std::min(hRel, 1e-5) : hRel;
        return std::max(hr * std::abs(tj), hFloor);
    };
    std::vector<double> h(11);
    std::vector<int64_t> tickets(22);
    {
        AdmissionHold hold;
        for (int j = 0; j < 11; ++j) {
            h[j] = stepFor(j, theta11[j]);
            std::vector<double> tp = theta11, tm = theta11;
            tp[j] += h[j]; tm[j] -= h[j];
            tickets[2 * j] = submit(pid, tp, &state_);
            tickets[2 * j + 1] = submit(pid, tm, &state_);
        }
    }
    for (int j = 0; j < 11; ++j) {
        std::vector<double> eta[2];
        double ll[2] = {0.0, 0.0};
        bool ok[2] = {true, true};
        for (int s = 0; s < 2; ++s) {
            FlowFields f;
            ConvergenceInfo conv;
            fetch(tickets[2 * j + s], f, conv, &eta[s], &ll[s]);     // observables evaluated on the device
            out.nResidualEvals += conv.finalIter;
            ok[s] = !conv.diverged;
        }
        out.krylovConverged[j] = (ok[0] && ok[1]) ? 1 : 0;
        if (!(ok[0] && ok[1])) continue;                      // the column stays 0 on divergence
        for (int i = 0; i < nObsv; ++i) out.dObsDTheta[i][j] = (eta[0][i] - eta[1][i]) / (2.0 * h[j]);
        out.logLikGradient[j] = (ll[0] - ll[1]) / (2.0 * h[j]);
    }
    return out;
}

}  // namespace rsh
