// forward_model.cpp -- ForwardModel over the librsg.so C-ABI.
#include "forward_model.hpp"

#include <cmath>
#include <cstdlib>
#include <stdexcept>

namespace rsh {

Engine& Engine::get() {
    static Engine* e = new Engine();   // never destroyed: shutdown() is explicit
    return *e;
}

std::string engineError(rsg_context* ctx, int rc) {
    std::string msg = ctx ? rsg_last_error(ctx) : "no engine";
    return "rsg error " + std::to_string(rc) + ": " + msg;
}

AdmissionHold::AdmissionHold() : ctx_(Engine::get().ctx()), gen_(Engine::get().generation()) {
    const int rc = rsg_hold_admission(ctx_, 1);
    if (rc != RSG_OK) throw std::runtime_error(engineError(ctx_, rc));
}

AdmissionHold::~AdmissionHold() {
    if (rsg_context* c = Engine::get().existing(gen_)) rsg_hold_admission(c, 0);
}

rsg_context* Engine::ctx() {
    std::lock_guard<std::mutex> l(mu_);
    if (ctx_) return ctx_;
    uint64_t budget = 16ull << 30;
    if (const char* s = std::getenv("RSG_HBM_BUDGET_GB")) budget = (uint64_t)(std::atof(s) * (1ull << 30));
    int dev = 0;
    if (const char* s = std::getenv("RSG_DEVICE")) dev = std::atoi(s);
    rsg_context* c = nullptr;
    int rc = rsg_context_create(dev, budget, &c);
    if (rc != RSG_OK)
        throw std::runtime_error("cannot start the MI355X engine (rsg_context_create returned " +
                                 std::to_string(rc) + "): no usable HIP device");
    ctx_ = c;
    ++gen_;
    return ctx_;
}

rsg_context* Engine::existing(uint64_t gen) {
    std::lock_guard<std::mutex> l(mu_);
    return (ctx_ && gen == gen_) ? ctx_ : nullptr;
}

uint64_t Engine::generation() {
    std::lock_guard<std::mutex> l(mu_);
    return gen_;
}

void Engine::shutdown() {
    std::lock_guard<std::mutex> l(mu_);
    if (ctx_) { rsg_context_destroy(ctx_); ctx_ = nullptr; }
}

void Engine::setLaunch(int itersPerLaunch, int blockSize) {
    rsg_context* c = ctx();
    int rc = rsg_set_launch(c, itersPerLaunch, blockSize);
    if (rc != RSG_OK) throw std::invalid_argument(engineError(c, rc));
}

void Engine::setSchedule(int policy) {
    rsg_context* c = ctx();
    int rc = rsg_set_schedule(c, policy);
    if (rc != RSG_OK) throw std::invalid_argument(engineError(c, rc));
}

void Engine::setTimeSlice(int microseconds) {
    rsg_context* c = ctx();
    int rc = rsg_set_time_slice(c, microseconds);
    if (rc != RSG_OK) throw std::invalid_argument(engineError(c, rc));
}

void Engine::setGrid(int minCells, int itersPerLaunch) {
    rsg_context* c = ctx();
    int rc = rsg_set_grid(c, minCells, itersPerLaunch);
    if (rc != RSG_OK) throw std::invalid_argument(engineError(c, rc));
}

void Engine::setGridLevels(int maxLevels) {
    rsg_context* c = ctx();
    int rc = rsg_set_grid_levels(c, maxLevels);
    if (rc != RSG_OK) throw std::invalid_argument(engineError(c, rc));
}

void Engine::setGraphs(bool enable) {
    rsg_context* c = ctx();
    int rc = rsg_set_graphs(c, enable ? 1 : 0);
    if (rc != RSG_OK) throw std::invalid_argument(engineError(c, rc));
}

void Engine::setTeam(int minCells, int blocks, int itersPerLaunch) {
    rsg_context* c = ctx();
    int rc = rsg_set_team(c, minCells, blocks, itersPerLaunch);
    if (rc != RSG_OK) throw std::invalid_argument(engineError(c, rc));
}

std::vector<double> Engine::stats() {
    std::vector<double> s(5, 0.0);
    rsg_context* c = ctx();
    rsg_stats(c, s.data());
    return s;
}

bool WarmStartCache::findNearest(const std::vector<double>& theta, Entry& out) const {
    std::lock_guard<std::mutex> l(mu_);
    const Entry* best = nullptr;
    double bestDist = threshold_;
    for (const Entry& e : cache_) {
        double s = 0.0;
        for (size_t i = 0; i < theta.size() && i < e.theta.size(); ++i) { double d = theta[i] - e.theta[i]; s += d * d; }
        double d = std::sqrt(s);
        if (d < bestDist) { bestDist = d; best = &e; }
    }
    if (!best) return false;
    out = *best;
    return true;
}

void WarmStartCache::store(const std::vector<double>& theta, const FlowFields& f) {
    std::lock_guard<std::mutex> l(mu_);
    cache_.push_back({theta, f});
    while ((int)cache_.size() > maxSize_) cache_.pop_front();
}

ForwardModel::ForwardModel(const Mesh& mesh, const InferenceParameterSet& paramSet,
                           const ObservationOperator& obsOp, const FlowBoundaryConditions& bcs, double nu,
                           const SolverSettings& settings, const Vec3& Uinit, double pInit, double kInit,
                           double omegaInit)
    : mesh_(mesh), paramSet_(paramSet), obsOp_(obsOp), bcs_(bcs), nu_(nu), settings_(settings),
      Uinit_(Uinit), pInit_(pInit), kInit_(kInit), omegaInit_(omegaInit), cache_(50, 0.5) {}

ForwardModel::~ForwardModel() {
    // never starts an engine: at interpreter exit the context may already be
    // shut down, and its ids went with it
    if (problemId_ >= 0 || meshId_ >= 0) {
        if (rsg_context* c = Engine::get().existing(ctxGen_)) {
            if (problemId_ >= 0) rsg_problem_release(c, problemId_);
            if (meshId_ >= 0) rsg_mesh_release(c, meshId_);
        }
    }
}

double computeWallDistanceOnDevice(Mesh& mesh) {
    rsg_context* c = Engine::get().ctx();
    const std::vector<double> cc = mesh.cellCenters(), wp = mesh.wallFaceCenters();
    std::vector<double> d(mesh.nCells());
    double secs = 0.0;
    const int rc = rsg_wall_distance(c, mesh.nCells(), cc.data(), (int32_t)(wp.size() / 3), wp.data(), d.data(), &secs);
    if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
    mesh.setWallDistance(std::move(d));
    return secs;
}

void setProblemObservations(rsg_context* c, int problem, const Mesh& mesh, const ObservationOperator& op) {
    if (op.nObs() == 0) return;
    ResolvedObservations R;
    op.resolve(mesh, R);
    const int rc = rsg_problem_set_observations(c, problem, (int32_t)R.obs.size(), R.obs.data(),
                                                (int32_t)R.faces.size(), R.faces.data(), R.faceX.data());
    if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
}

void ForwardModel::ensureProblem() {
    if (problemId_ >= 0) return;
    rsg_context* c = Engine::get().ctx();
    if (!mesh_.hasWallDistance()) computeWallDistanceOnDevice(const_cast<Mesh&>(mesh_));
    int32_t mid = -1, pid = -1;
    int rc = rsg_mesh_upload(c, &mesh_.desc(), &mid);
    if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
    ResolvedBC R;
    bcs_.resolve(mesh_, R);
    rsg_settings s = settings_.toC();
    rc = rsg_problem_create(c, mid, &R.desc, nu_, &s, &pid);
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

int64_t ForwardModel::submit(const std::vector<double>& theta, bool useCache) {
    std::lock_guard<std::mutex> l(mu_);
    ensureProblem();
    rsg_context* c = Engine::get().ctx();
    SSTCoefficients co = paramSet_.unpack(theta);
    std::vector<double> coeffs = InferenceParameterSet::toVector(co);
    rsg_init init{};
    init.U[0] = Uinit_.x; init.U[1] = Uinit_.y; init.U[2] = Uinit_.z;
    init.p = pInit_; init.k = kInit_; init.omega = omegaInit_;
    WarmStartCache::Entry hit;
    if (useCache && cache_.findNearest(theta, hit)) {
        const FlowFields& f = hit.fields;
        init.kind = 1;
        init.U_cells = f.U.data(); init.p_cells = f.p.data(); init.k_cells = f.k.data(); init.om_cells = f.omega.data();
        init.nuT = f.nuT.data(); init.F1 = f.F1.data(); init.F2 = f.F2.data(); init.Pk = f.Pk.data(); init.CDkw = f.CDkw.data();
        init.U_bnd = f.Ub.data(); init.p_bnd = f.pb.data(); init.k_bnd = f.kb.data(); init.om_bnd = f.omb.data();
        init.turb_established = f.turbEstablished ? 1 : 0;
    }
    if (!collectFields_) init.kind |= RSG_INIT_NO_FIELDS;
    int64_t t = -1;
    int rc = rsg_submit(c, problemId_, coeffs.data(), &init, precision_, &t);   // copies the warm state
    if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
    if (!collectFields_) noFieldTickets_.insert(t);
    return t;
}

EvaluationResult ForwardModel::complete(int64_t ticket, const std::vector<double>& theta, bool storeCache) {
    rsg_context* c = Engine::get().ctx();
    int rc = rsg_wait(c, &ticket, 1, -1.0);
    if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
    bool fields;
    {
        std::lock_guard<std::mutex> l(mu_);
        fields = noFieldTickets_.erase(ticket) == 0;
    }
    FlowFields f;
    ConvergenceInfo conv;
    if (settings_.recordHistory) conv.history.assign(6 * (size_t)std::max(settings_.maxIterations, 1), 0.0);
    rsg_result r{};
    if (fields) {
        f.resize(mesh_.nCells(), mesh_.nFaces() - mesh_.nInternalFaces());
        r.U = f.U.data(); r.p = f.p.data(); r.k = f.k.data(); r.omega = f.omega.data(); r.nuT = f.nuT.data();
        r.F1 = f.F1.data(); r.F2 = f.F2.data(); r.Pk = f.Pk.data(); r.CDkw = f.CDkw.data();
        r.U_bnd = f.Ub.data(); r.p_bnd = f.pb.data(); r.k_bnd = f.kb.data(); r.om_bnd = f.omb.data();
    }
    r.history = conv.history.empty() ? nullptr : conv.history.data();
    r.history_cap = (int32_t)(conv.history.size() / 6);
    std::vector<double> pred(obsOp_.nObs());
    r.predictions = pred.empty() ? nullptr : pred.data();
    r.predictions_cap = (int32_t)pred.size();
    rc = rsg_fetch(c, ticket, &r);
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
    // ForwardModel.cpp:70-97
    if (conv.diverged) {
        res.status = EvaluationStatus::Diverged;
        res.loglik = -1e6;
        hasLastFields_ = false;
        return res;
    } else if (conv.converged) {
        res.status = EvaluationStatus::Converged;
        if (storeCache && fields) cache_.store(theta, f);
    } else {
        res.status = EvaluationStatus::Unconverged;
    }
    // a solve whose fields stayed on the device leaves no last fields (an
    // earlier solve's would not be this theta's)
    if (fields) lastFields_ = std::move(f);
    hasLastFields_ = fields;
    // ForwardModel.cpp:95-97, evaluated on the device at the end of the solve
    if (r.n_predictions != obsOp_.nObs()) throw std::runtime_error("device observables missing");
    res.predictions = std::move(pred);
    res.loglik = obsOp_.nObs() > 0 ? r.log_lik : 0.0;
    return res;
}

EvaluationResult ForwardModel::evaluate(const std::vector<double>& theta) {
    EvaluationResult result;
    if (!paramSet_.inBounds(theta)) {
        result.status = EvaluationStatus::InvalidParameters;
        result.loglik = -1e30;
        return result;
    }
    try {
        int64_t t = submit(theta, true);
        return complete(t, theta, true);
    } catch (const std::exception& e) {
        result.status = EvaluationStatus::Unknown;
        result.loglik = -1e6;
        if (settings_.verbose) fprintf(stderr, "ForwardModel exception: %s\n", e.what());
        // a missing GPU is not a solver outcome: surface it to the caller
        if (std::string(e.what()).find("cannot start the MI355X engine") != std::string::npos) throw;
    }
    return result;
}

std::vector<EvaluationResult> ForwardModel::evaluateBatch(const std::vector<std::vector<double>>& thetas) {
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
            tickets[i] = submit(thetas[i], true);
        }
    }
    for (size_t i = 0; i < thetas.size(); ++i) {
        if (tickets[i] < 0) continue;
        try {
            out[i] = complete(tickets[i], thetas[i], true);
        } catch (const std::exception& e) {
            out[i].status = EvaluationStatus::Unknown;
            out[i].loglik = -1e6;
        }
    }
    return out;
}

std::vector<EvaluationResult> evaluateMany(const std::vector<ForwardModel*>& fms,
                                           const std::vector<std::vector<double>>& thetas) {
    if (fms.size() != thetas.size()) throw std::invalid_argument("evaluate_many: models and thetas differ in length");
    std::vector<EvaluationResult> out(fms.size());
    std::vector<int64_t> tickets(fms.size(), -1);
    {
        AdmissionHold hold;
        for (size_t i = 0; i < fms.size(); ++i) {
            if (!fms[i]->paramSet().inBounds(thetas[i])) {
                out[i].status = EvaluationStatus::InvalidParameters;
                out[i].loglik = -1e30;
                continue;
            }
            tickets[i] = fms[i]->submit(thetas[i], true);
        }
    }
    for (size_t i = 0; i < fms.size(); ++i) {
        if (tickets[i] < 0) continue;
        try {
            out[i] = fms[i]->complete(tickets[i], thetas[i], true);
        } catch (const std::exception&) {
            out[i].status = EvaluationStatus::Unknown;
            out[i].loglik = -1e6;
        }
    }
    return out;
}

// ForwardModel.cpp:109-122
double ForwardModel::penalizedLogLikelihood(const std::vector<double>& theta) {
    EvaluationResult r = evaluate(theta);
    switch (r.status) {
        case EvaluationStatus::Converged:
        case EvaluationStatus::Unconverged: return r.loglik;
        case EvaluationStatus::Diverged:
        case EvaluationStatus::Unknown: return -1e6;
        case EvaluationStatus::InvalidParameters: return -1e30;
    }
    return -1e30;
}

}  // namespace rsh
