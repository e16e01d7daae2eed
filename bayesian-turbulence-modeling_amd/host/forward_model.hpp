// forward_model.hpp -- theta -> EvaluationResult on the MI355X engine.
//
// Same contract as the reference ForwardModel (incompressible/include/
// ForwardModel.hpp:75-139, incompressible/src/ForwardModel.cpp:29-122):
// bounds check, unpack to the 11 SST coefficients, warm start from the
// nearest cached converged state or a uniform initial state, SIMPLE solve,
// status classification, predictions and Gaussian log-likelihood.  The solve
// itself runs in librsg.so (include/rsg.h); this class only marshals.
#pragma once
#include <deque>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "mesh.hpp"
#include "model.hpp"
#include "rsg.h"

namespace rsh {

enum class EvaluationStatus { Converged, Unconverged, Diverged, InvalidParameters, Unknown };

struct EvaluationResult {
    EvaluationStatus status = EvaluationStatus::Unknown;
    double loglik = -1e30;
    std::vector<double> predictions;
    int simpleIters = 0;
    // engine timing of the solve (seconds): submit -> result, device residency
    double latency = 0.0, deviceTime = 0.0;
};

// Process-wide engine context on one GPU (created on first use).
class Engine {
public:
    static Engine& get();
    rsg_context* ctx();           // throws if the GPU engine cannot start
    // The running context if it is still the one of generation `gen` (the
    // value generation() had when the caller's ids were created), else null;
    // never starts an engine.  For destructors: ids of a context that has
    // been shut down are gone with it.
    rsg_context* existing(uint64_t gen);
    uint64_t generation();
    void shutdown();
    void setLaunch(int itersPerLaunch, int blockSize);
    void setTimeSlice(int microseconds);
    void setSchedule(int policy);
    // members on meshes of at least minCells cells run as teams of `blocks`
    // workgroups (0: sized by the cell count); minCells 0 turns team solves off
    void setTeam(int minCells, int blocks, int itersPerLaunch);
    // members on meshes of at least minCells cells run kernel-per-phase on the
    // whole GPU (0 turns it off)
    void setGrid(int minCells, int itersPerLaunch);
    // p' AMG levels of kernel-per-phase members (0: the reference's cap)
    void setGridLevels(int maxLevels);
    // hipGraph replay of AMG V-cycles on the whole-GPU paths
    void setGraphs(bool enable);
    std::vector<double> stats();
private:
    std::mutex mu_;
    rsg_context* ctx_ = nullptr;
    uint64_t gen_ = 0;            // bumped every time a context is created
};

std::string engineError(rsg_context* ctx, int rc);

// Holds the engine's admission while it lives (rsg_hold_admission), so that
// the solves a batch submits under it start together in one launch.  Nothing
// that waits for a solve may run while it is held.
class AdmissionHold {
public:
    AdmissionHold();
    ~AdmissionHold();
    AdmissionHold(const AdmissionHold&) = delete;
    AdmissionHold& operator=(const AdmissionHold&) = delete;
private:
    rsg_context* ctx_;
    uint64_t gen_;
};

// Mesh::computeWallDistance on the device (rsg_wall_distance): the same
// distances, bit for bit; returns the kernel's seconds.
double computeWallDistanceOnDevice(Mesh& mesh);

// Hands a problem its observables (resolved against the mesh) so every solve
// evaluates them on the device when it finishes.
void setProblemObservations(rsg_context* ctx, int problem, const Mesh& mesh, const ObservationOperator& op);

// WarmStartCache, ForwardModel.hpp:18-72 (value copies under the lock)
class WarmStartCache {
public:
    struct Entry { std::vector<double> theta; FlowFields fields; };
    explicit WarmStartCache(int maxSize = 50, double threshold = 0.5) : maxSize_(maxSize), threshold_(threshold) {}
    bool findNearest(const std::vector<double>& theta, Entry& out) const;
    void store(const std::vector<double>& theta, const FlowFields& f);
    int size() const { std::lock_guard<std::mutex> l(mu_); return (int)cache_.size(); }
    void clear() { std::lock_guard<std::mutex> l(mu_); cache_.clear(); }
private:
    int maxSize_;
    double threshold_;
    mutable std::mutex mu_;
    std::deque<Entry> cache_;
};

class ForwardModel {
public:
    ForwardModel(const Mesh& mesh, const InferenceParameterSet& paramSet, const ObservationOperator& obsOp,
                 const FlowBoundaryConditions& bcs, double nu, const SolverSettings& settings = {},
                 const Vec3& Uinit = Vec3(1, 0, 0), double pInit = 0.0, double kInit = 1e-4,
                 double omegaInit = 1.0);
    ~ForwardModel();
    ForwardModel(const ForwardModel&) = delete;
    ForwardModel& operator=(const ForwardModel&) = delete;

    EvaluationResult evaluate(const std::vector<double>& theta);
    // Many thetas solved concurrently on the device (one workgroup each);
    // each result equals what evaluate() returns for that theta alone, given
    // the same warm-start cache contents.
    std::vector<EvaluationResult> evaluateBatch(const std::vector<std::vector<double>>& thetas);
    double penalizedLogLikelihood(const std::vector<double>& theta);

    WarmStartCache& cache() { return cache_; }
    const InferenceParameterSet& paramSet() const { return paramSet_; }
    const FlowFields& lastFields() const { return lastFields_; }
    bool hasLastFields() const { return hasLastFields_; }
    const ConvergenceInfo& lastConvergence() const { return lastConv_; }
    const Mesh& mesh() const { return mesh_; }
    void setPrecision(int p) { precision_ = p; }
    int precision() const { return precision_; }
    // false: solves submitted afterwards leave their fields on the device
    // (RSG_INIT_NO_FIELDS); evaluate returns predictions, log-likelihood and
    // status only, and neither last_fields nor the warm-start cache is updated
    void setCollectFields(bool on) { std::lock_guard<std::mutex> l(mu_); collectFields_ = on; }
    bool collectFields() const { return collectFields_; }

    // Submission split for servers: returns an engine ticket.
    int64_t submit(const std::vector<double>& theta, bool useCache);
    // Finishes a ticket into a result (and the retained state).
    EvaluationResult complete(int64_t ticket, const std::vector<double>& theta, bool storeCache);

private:
    const Mesh& mesh_;
    InferenceParameterSet paramSet_;
    ObservationOperator obsOp_;
    FlowBoundaryConditions bcs_;
    double nu_;
    SolverSettings settings_;
    Vec3 Uinit_;
    double pInit_, kInit_, omegaInit_;
    WarmStartCache cache_;
    FlowFields lastFields_;
    bool hasLastFields_ = false;
    ConvergenceInfo lastConv_;
    int precision_ = RSG_PREC_FP64;
    bool collectFields_ = true;
    std::set<int64_t> noFieldTickets_;   // submitted with RSG_INIT_NO_FIELDS, not yet completed
    int meshId_ = -1, problemId_ = -1;
    uint64_t ctxGen_ = 0;         // engine generation the ids belong to
    std::mutex mu_;
    void ensureProblem();
};

// Ensemble evaluation: every (model, theta) pair is submitted before any is
// awaited, so all members share the device (one workgroup each).
std::vector<EvaluationResult> evaluateMany(const std::vector<ForwardModel*>& fms,
                                           const std::vector<std::vector<double>>& thetas);

}  // namespace rsh
