// model.cpp -- boundary conditions, settings, parameter sets, observations.
#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <deque>
#include <stdexcept>

namespace rsh {

// ------------------------------------------------------------------- BCs
// channelDefaults BoundaryCondition.hpp:44-98
FlowBoundaryConditions FlowBoundaryConditions::channelDefaults(const Mesh& mesh, double Uin, double kIn,
                                                               double omIn) {
    FlowBoundaryConditions bc;
    const int np = mesh.nPatches();
    bc.velocityBC.resize(np); bc.pressureBC.resize(np); bc.kBC.resize(np); bc.omegaBC.resize(np);
    for (int p = 0; p < np; ++p) {
        const Patch& pat = mesh.patch(p);
        for (auto* v : {&bc.velocityBC, &bc.pressureBC, &bc.kBC, &bc.omegaBC}) (*v)[p].patchName = pat.name;
        if (pat.type == "wall") {
            bc.velocityBC[p].type = BCType::WallNoSlip;
            bc.pressureBC[p].type = BCType::Neumann;
            bc.kBC[p].type = BCType::WallKOmega;
            bc.omegaBC[p].type = BCType::WallKOmega;
        } else if (pat.type == "inlet") {
            bc.velocityBC[p].type = BCType::InletVelocity; bc.velocityBC[p].vecValue = Vec3(Uin, 0, 0);
            bc.kBC[p].type = BCType::Dirichlet; bc.kBC[p].value = kIn;
            bc.omegaBC[p].type = BCType::Dirichlet; bc.omegaBC[p].value = omIn;
        } else if (pat.type == "outlet") {
            bc.pressureBC[p].type = BCType::OutletPressure;
        } else if (pat.type == "symmetry") {
            bc.velocityBC[p].type = BCType::Symmetry;
        }
    }
    return bc;
}

// flatPlateDefaults :108-122
FlowBoundaryConditions FlowBoundaryConditions::flatPlateDefaults(const Mesh& mesh, double Uinf, double kIn,
                                                                 double omIn) {
    FlowBoundaryConditions bc = channelDefaults(mesh, Uinf, kIn, omIn);
    for (int p = 0; p < mesh.nPatches(); ++p) {
        if (mesh.patch(p).name != "top") continue;
        bc.velocityBC[p].type = BCType::Dirichlet; bc.velocityBC[p].vecValue = Vec3(Uinf, 0, 0);
        bc.velocityBC[p].vecProfile.clear();
        bc.pressureBC[p].type = BCType::Neumann;
        bc.kBC[p].type = BCType::Dirichlet; bc.kBC[p].value = kIn;
        bc.omegaBC[p].type = BCType::Dirichlet; bc.omegaBC[p].value = omIn;
    }
    return bc;
}

// couetteDefaults :134-152
FlowBoundaryConditions FlowBoundaryConditions::couetteDefaults(const Mesh& mesh, double Uwall, double kIn,
                                                               double omIn) {
    FlowBoundaryConditions bc = channelDefaults(mesh, 0.5 * Uwall, kIn, omIn);
    for (int p = 0; p < mesh.nPatches(); ++p) {
        const Patch& pat = mesh.patch(p);
        if (pat.name == "top") {
            bc.velocityBC[p].type = BCType::WallMoving; bc.velocityBC[p].vecValue = Vec3(Uwall, 0, 0);
        } else if (pat.name == "inlet") {
            for (auto* v : {&bc.velocityBC, &bc.pressureBC, &bc.kBC, &bc.omegaBC}) {
                (*v)[p].type = BCType::Neumann; (*v)[p].value = 0.0; (*v)[p].vecValue = Vec3();
            }
        }
    }
    return bc;
}

int FlowBoundaryConditions::patchIndex(const Mesh& mesh, const std::string& name) {
    for (int p = 0; p < mesh.nPatches(); ++p)
        if (mesh.patch(p).name == name) return p;
    throw std::runtime_error("FlowBoundaryConditions: unknown patch '" + name + "'");
}

void FlowBoundaryConditions::setVelocityProfile(const Mesh& mesh, const std::string& name,
                                                const std::vector<Vec3>& vals) {
    int p = patchIndex(mesh, name);
    if (vals.size() != mesh.patch(p).faces.size()) throw std::runtime_error("velocity profile size != patch face count");
    if (p >= (int)velocityBC.size()) throw std::runtime_error("velocity BC list does not cover patch '" + name + "'");
    velocityBC[p].vecProfile = vals;
}
void FlowBoundaryConditions::setKProfile(const Mesh& mesh, const std::string& name, const std::vector<double>& vals) {
    int p = patchIndex(mesh, name);
    if (vals.size() != mesh.patch(p).faces.size()) throw std::runtime_error("k profile size != patch face count");
    if (p >= (int)kBC.size()) throw std::runtime_error("k BC list does not cover patch '" + name + "'");
    kBC[p].scalarProfile = vals;
}
void FlowBoundaryConditions::setOmegaProfile(const Mesh& mesh, const std::string& name,
                                             const std::vector<double>& vals) {
    int p = patchIndex(mesh, name);
    if (vals.size() != mesh.patch(p).faces.size()) throw std::runtime_error("omega profile size != patch face count");
    if (p >= (int)omegaBC.size()) throw std::runtime_error("omega BC list does not cover patch '" + name + "'");
    omegaBC[p].scalarProfile = vals;
}

// applyVelocityBC / applyPressureBC / applyKBC / applyOmegaBC (:196-300)
// decide one value per boundary face; the wall omega value depends on beta1
// and is computed on the device.
void FlowBoundaryConditions::resolve(const Mesh& mesh, ResolvedBC& R) const {
    const int np = mesh.nPatches(), nIF = mesh.nInternalFaces(), nB = mesh.nFaces() - nIF;
    if ((int)velocityBC.size() < np || (int)pressureBC.size() < np || (int)kBC.size() < np ||
        (int)omegaBC.size() < np)
        throw std::invalid_argument("FlowBoundaryConditions: every mesh patch needs a velocity, pressure, k and omega BC");
    R.velT.assign(nB, RSG_BC_KEEP); R.pT.assign(nB, RSG_BC_KEEP); R.kT.assign(nB, RSG_BC_KEEP); R.omT.assign(nB, RSG_BC_KEEP);
    R.velV.assign(3 * (size_t)nB, 0.0); R.pV.assign(nB, 0.0); R.kV.assign(nB, 0.0); R.omV.assign(nB, 0.0);
    for (int p = 0; p < np; ++p) {
        const Patch& pat = mesh.patch(p);
        for (size_t j = 0; j < pat.faces.size(); ++j) {
            const int b = pat.faces[j] - nIF;
            if (b < 0 || b >= nB) throw std::invalid_argument("patch face is not a boundary face");
            const PatchBC& v = velocityBC[p];
            R.velT[b] = (int)v.type;
            Vec3 u = v.vecValue;
            if ((v.type == BCType::InletVelocity || v.type == BCType::Dirichlet) && !v.vecProfile.empty()) u = v.vecProfile[j];
            R.velV[3 * b] = u.x; R.velV[3 * b + 1] = u.y; R.velV[3 * b + 2] = u.z;
            R.pT[b] = (int)pressureBC[p].type; R.pV[b] = pressureBC[p].value;
            const PatchBC& kb = kBC[p];
            R.kT[b] = (int)kb.type;
            R.kV[b] = (kb.type == BCType::Dirichlet && !kb.scalarProfile.empty()) ? kb.scalarProfile[j] : kb.value;
            const PatchBC& ob = omegaBC[p];
            R.omT[b] = (int)ob.type;
            R.omV[b] = (ob.type == BCType::Dirichlet && !ob.scalarProfile.empty()) ? ob.scalarProfile[j] : ob.value;
        }
    }
    rsg_bc_desc& d = R.desc;
    d.vel_type = R.velT.data(); d.vel_value = R.velV.data(); d.p_type = R.pT.data(); d.p_value = R.pV.data();
    d.k_type = R.kT.data(); d.k_value = R.kV.data(); d.om_type = R.omT.data(); d.om_value = R.omV.data();
}

// -------------------------------------------------------------- settings
static int32_t solverCode(const std::string& s) {
    if (s == "AMG") return RSG_SOLVER_AMG;
    if (s == "PCG") return RSG_SOLVER_PCG;
    if (s == "BiCGSTAB") return RSG_SOLVER_BICGSTAB;
    if (s == "GaussSeidel") return RSG_SOLVER_GAUSS_SEIDEL;
    return RSG_SOLVER_BICGSTAB;   // makeSolver fallback (LinearSolver.hpp:664)
}

rsg_settings SolverSettings::toC() const {
    rsg_settings s;
    s.max_iterations = maxIterations; s.convergence_tol = convergenceTol; s.divergence_limit = divergenceLimit;
    s.alpha_u = alphaU; s.alpha_p = alphaP; s.alpha_k = alphaK; s.alpha_omega = alphaOmega;
    s.inner_iterations = innerIterations; s.inner_tolerance = innerTolerance;
    s.turb_start_iter = turbStartIter; s.turb_update_interval = turbUpdateInterval;
    s.nut_floor_iters = nuTFloorIters; s.rhie_chow_all_meshes = rhieChowAllMeshes ? 1 : 0;
    s.body_force[0] = bodyForce.x; s.body_force[1] = bodyForce.y; s.body_force[2] = bodyForce.z;
    s.k_min = kMin; s.k_max = kMax; s.omega_min = omegaMin;
    s.use_wall_functions = useWallFunctions ? 1 : 0; s.von_karman = vonKarman; s.sst_variant = sstVariant;
    s.pressure_solver = solverCode(pressureSolver);
    s.momentum_solver = solverCode(momentumSolver);
    s.turbulence_solver = solverCode(turbulenceSolver);
    s.amg_refresh = amgRefresh; s.coarse_refresh = coarseRefresh; s.record_history = recordHistory ? 1 : 0;
    return s;
}

// ------------------------------------------------------------- parameters
std::vector<double> InferenceParameterSet::toVector(const SSTCoefficients& c) {
    return {c.sigma_k1, c.sigma_w1, c.beta1, c.alpha1, c.sigma_k2, c.sigma_w2,
            c.beta2, c.alpha2, c.betaStar, c.a1, c.kappa};
}
SSTCoefficients InferenceParameterSet::fromVector(const std::vector<double>& v) {
    if (v.size() != 11) throw std::runtime_error("Expected 11 SST coefficients");
    SSTCoefficients c;
    c.sigma_k1 = v[0]; c.sigma_w1 = v[1]; c.beta1 = v[2]; c.alpha1 = v[3];
    c.sigma_k2 = v[4]; c.sigma_w2 = v[5]; c.beta2 = v[6]; c.alpha2 = v[7];
    c.betaStar = v[8]; c.a1 = v[9]; c.kappa = v[10];
    return c;
}
SSTCoefficients InferenceParameterSet::unpack(const std::vector<double>& theta) const {
    if ((int)theta.size() != nActive())
        throw std::runtime_error("theta size mismatch: expected " + std::to_string(nActive()) + ", got " +
                                 std::to_string(theta.size()));
    std::vector<double> full = toVector(base);
    for (int i = 0; i < nActive(); ++i) full[activeIndices[i]] = theta[i];
    return fromVector(full);
}
std::vector<double> InferenceParameterSet::pack(const SSTCoefficients& c) const {
    std::vector<double> full = toVector(c), th(nActive());
    for (int i = 0; i < nActive(); ++i) th[i] = full[activeIndices[i]];
    return th;
}
std::vector<std::string> InferenceParameterSet::activeNames() const {
    static const char* all[] = {"sigma_k1", "sigma_w1", "beta1", "alpha1", "sigma_k2", "sigma_w2",
                                "beta2", "alpha2", "betaStar", "a1", "kappa"};
    std::vector<std::string> o;
    for (int i : activeIndices) o.push_back(all[i]);
    return o;
}
// InferenceParameters.hpp:78-97
std::vector<double> InferenceParameterSet::lowerBounds() const {
    static const double lo[] = {0.1, 0.1, 0.01, 0.1, 0.1, 0.1, 0.01, 0.1, 0.01, 0.1, 0.35};
    std::vector<double> o;
    for (int i : activeIndices) o.push_back(lo[i]);
    return o;
}
std::vector<double> InferenceParameterSet::upperBounds() const {
    static const double hi[] = {2.0, 2.0, 0.2, 1.5, 2.0, 2.0, 0.2, 1.5, 0.2, 0.8, 0.50};
    std::vector<double> o;
    for (int i : activeIndices) o.push_back(hi[i]);
    return o;
}
// :102-111
bool InferenceParameterSet::inBounds(const std::vector<double>& theta) const {
    if ((int)theta.size() != nActive()) return false;
    auto lo = lowerBounds(), hi = upperBounds();
    for (int i = 0; i < nActive(); ++i) {
        if (theta[i] < lo[i] || theta[i] > hi[i]) return false;
        if (std::isnan(theta[i]) || std::isinf(theta[i])) return false;
    }
    return true;
}
InferenceParameterSet InferenceParameterSet::fromIndices(const std::string& name, const std::vector<int>& idx) {
    if (idx.empty()) throw std::runtime_error("fromIndices: need at least one active index");
    for (size_t i = 0; i < idx.size(); ++i) {
        if (idx[i] < 0 || idx[i] > 10) throw std::runtime_error("fromIndices: index out of range [0,10]");
        for (size_t j = i + 1; j < idx.size(); ++j)
            if (idx[i] == idx[j]) throw std::runtime_error("fromIndices: duplicate index");
    }
    return {name, idx, SSTCoefficients{}};
}

void FlowFields::resize(int nc, int nb) {
    n = nc; nB = nb;
    U.assign(3 * (size_t)nc, 0.0);
    for (auto* v : {&p, &k, &omega, &nuT, &F1, &F2, &Pk, &CDkw}) v->assign(nc, 0.0);
    Ub.assign(3 * (size_t)nb, 0.0);
    for (auto* v : {&pb, &kb, &omb}) v->assign(nb, 0.0);
}

// ----------------------------------------------------------- observations
void ObservationOperator::addDrag(const std::string& wallPatch, double Cd_obs, double sigma, double refArea,
                                  double refVel, double sigma_model) {
    Observable o; o.type = ObsType::Drag; o.meas = {Cd_obs, sigma, sigma_model};
    o.patchName = wallPatch; o.referenceArea = refArea; o.refVelocity = refVel;
    obs_.push_back(o);
}
void ObservationOperator::addPressureTap(const Vec3& loc, double Cp_obs, double sigma, double refVel,
                                         double sigma_model, double refPressure, double refDensity) {
    Observable o; o.type = ObsType::PressureTap; o.meas = {Cp_obs, sigma, sigma_model};
    o.location = loc; o.refVelocity = refVel; o.refPressure = refPressure; o.refDensity = refDensity;
    obs_.push_back(o);
}
void ObservationOperator::addVelocityProfile(const Vec3& loc, int comp, double Uobs, double sigma, double sigma_model) {
    Observable o; o.type = ObsType::VelocityProfile; o.meas = {Uobs, sigma, sigma_model};
    o.location = loc; o.component = comp;
    obs_.push_back(o);
}
void ObservationOperator::addSkinFriction(const std::string& wallPatch, const Vec3& loc, double Cf_obs,
                                          double sigma, double refVel, double sigma_model) {
    Observable o; o.type = ObsType::SkinFriction; o.meas = {Cf_obs, sigma, sigma_model};
    o.patchName = wallPatch; o.location = loc; o.refVelocity = refVel;
    obs_.push_back(o);
}
void ObservationOperator::addSeparationPoint(const std::string& wallPatch, double xSep_obs, double sigma,
                                             double sigma_model) {
    Observable o; o.type = ObsType::SeparationPoint; o.meas = {xSep_obs, sigma, sigma_model}; o.patchName = wallPatch;
    obs_.push_back(o);
}
void ObservationOperator::addReattachmentLength(const std::string& wallPatch, double xr_obs, double sigma,
                                                double sigma_model) {
    Observable o; o.type = ObsType::ReattachmentLength; o.meas = {xr_obs, sigma, sigma_model}; o.patchName = wallPatch;
    obs_.push_back(o);
}

namespace {
Vec3 cellU(const FlowFields& f, int c) { return Vec3(f.U[3 * c], f.U[3 * c + 1], f.U[3 * c + 2]); }
int nearestCell(const Mesh& m, const Vec3& loc) {
    int best = 0; double bd = 1e30;
    for (int c = 0; c < m.nCells(); ++c) {
        double d = (m.cell(c).center - loc).norm2();
        if (d < bd) { bd = d; best = c; }
    }
    return best;
}
}  // namespace

// ObservationOperator::evaluate and helpers, ObservationOperator.hpp:159-364
std::vector<double> ObservationOperator::evaluate(const Mesh& m, const FlowFields& f, double nu) const {
    std::vector<double> pred(obs_.size(), 0.0);
    const int nIF = m.nInternalFaces();
    for (size_t i = 0; i < obs_.size(); ++i) {
        const Observable& o = obs_[i];
        switch (o.type) {
            case ObsType::Drag: {
                const Patch& pat = m.patch(m.patchByName(o.patchName));
                double force = 0.0;
                for (int fi : pat.faces) {
                    const Face& face = m.face(fi);
                    const int ow = face.owner;
                    const double Sf = face.area;
                    force += f.p[ow] * face.normal.x * Sf;
                    const double delta = std::max(face.delta, 1e-20);
                    const double Uw_x = f.Ub[3 * (fi - nIF)];
                    const double tau_x = nu * (f.U[3 * ow] - Uw_x) / delta;
                    force += tau_x * Sf;
                }
                const double dyn = 0.5 * o.refVelocity * o.refVelocity;
                pred[i] = force / (dyn * o.referenceArea);
                break;
            }
            case ObsType::PressureTap: {
                int c = nearestCell(m, o.location);
                const double dyn = 0.5 * o.refDensity * o.refVelocity * o.refVelocity;
                pred[i] = (f.p[c] - o.refPressure) / std::max(dyn, 1e-20);
                break;
            }
            case ObsType::VelocityProfile: {
                int c = nearestCell(m, o.location);
                pred[i] = f.U[3 * c + (o.component == 0 ? 0 : o.component == 1 ? 1 : 2)];
                break;
            }
            case ObsType::SkinFriction: {
                const Patch& pat = m.patch(m.patchByName(o.patchName));
                int best = pat.faces.empty() ? -1 : pat.faces[0];
                double bd = 1e30;
                for (int fi : pat.faces) {
                    double d = (m.face(fi).center - o.location).norm2();
                    if (d < bd) { bd = d; best = fi; }
                }
                if (best < 0) { pred[i] = 0.0; break; }
                const Face& face = m.face(best);
                const double delta = std::max(face.delta, 1e-20);
                Vec3 Uc = cellU(f, face.owner);
                Vec3 Un = face.normal * Uc.dot(face.normal);
                Vec3 Ut = Uc - Un;
                const double tau = nu * Ut.norm() / delta;
                const double dynP = 0.5 * o.refVelocity * o.refVelocity;
                pred[i] = tau / std::max(dynP, 1e-20);
                break;
            }
            case ObsType::SeparationPoint:
            case ObsType::ReattachmentLength: {
                const bool sep = o.type == ObsType::SeparationPoint;
                const Patch& pat = m.patch(m.patchByName(o.patchName));
                double out = sep ? 0.0 : (pat.faces.empty() ? 0.0 : m.face(pat.faces.back()).center.x);
                bool havePrev = false;
                double prevTau = sep ? 1.0 : -1.0, prevX = 0.0;
                for (int fi : pat.faces) {
                    const Face& face = m.face(fi);
                    const double delta = std::max(face.delta, 1e-20);
                    Vec3 Uc = cellU(f, face.owner);
                    Vec3 Ut = Uc - face.normal * Uc.dot(face.normal);
                    const double tau = nu * Ut.x / delta;
                    const double x = face.center.x;
                    if (sep && havePrev && prevTau > 0 && tau <= 0) {
                        double denom = prevTau - tau;
                        double frac = (std::abs(denom) > 1e-30) ? prevTau / denom : 0.0;
                        out = prevX + frac * (x - prevX); break;
                    }
                    if (!sep && havePrev && prevTau < 0 && tau >= 0) {
                        double denom = tau - prevTau;
                        double frac = (std::abs(denom) > 1e-30) ? (-prevTau) / denom : 0.0;
                        out = prevX + frac * (x - prevX); break;
                    }
                    prevTau = tau; prevX = x; havePrev = true;
                }
                pred[i] = out;
                break;
            }
        }
    }
    return pred;
}

// Gaussian log-likelihood, ObservationOperator.hpp:367-378
// The searches of findNearestCell / findNearestWallFace (ObservationOperator.hpp:
// 160-183) and the patch face lists that computeDrag / computeSeparationPoint /
// computeReattachmentLength walk; the device evaluates the rest.
void ObservationOperator::resolve(const Mesh& m, ResolvedObservations& out) const {
    out.obs.clear(); out.faces.clear(); out.faceX.clear();
    for (const Observable& o : obs_) {
        rsg_observation r{};
        r.ref_velocity = o.refVelocity; r.ref_area = o.referenceArea;
        r.ref_density = o.refDensity; r.ref_pressure = o.refPressure;
        r.observed = o.meas.observed; r.sigma = o.meas.sigma_obs; r.sigma_model = o.meas.sigma_model;
        r.index = -1;
        switch (o.type) {
            case ObsType::Drag:
            case ObsType::SeparationPoint:
            case ObsType::ReattachmentLength: {
                r.type = o.type == ObsType::Drag ? RSG_OBS_DRAG
                       : o.type == ObsType::SeparationPoint ? RSG_OBS_SEPARATION : RSG_OBS_REATTACHMENT;
                const Patch& pat = m.patch(m.patchByName(o.patchName));
                r.face_begin = (int32_t)out.faces.size();
                for (int fi : pat.faces) { out.faces.push_back(fi); out.faceX.push_back(m.face(fi).center.x); }
                r.face_end = (int32_t)out.faces.size();
                break;
            }
            case ObsType::PressureTap:
            case ObsType::VelocityProfile:
                r.type = o.type == ObsType::PressureTap ? RSG_OBS_PRESSURE_TAP : RSG_OBS_VELOCITY;
                r.index = nearestCell(m, o.location);
                r.component = o.component == 0 ? 0 : o.component == 1 ? 1 : 2;
                break;
            case ObsType::SkinFriction: {
                r.type = RSG_OBS_SKIN_FRICTION;
                const Patch& pat = m.patch(m.patchByName(o.patchName));
                int best = pat.faces.empty() ? -1 : pat.faces[0];
                double bd = 1e30;
                for (int fi : pat.faces) {
                    double d = (m.face(fi).center - o.location).norm2();
                    if (d < bd) { bd = d; best = fi; }
                }
                r.index = best;
                break;
            }
        }
        out.obs.push_back(r);
    }
}

double ObservationOperator::logLikelihood(const std::vector<double>& pred) const {
    double ll = 0.0;
    for (size_t i = 0; i < obs_.size(); ++i) {
        double var = obs_[i].meas.totalVariance();
        if (var < 1e-30) var = 1e-30;
        double diff = pred[i] - obs_[i].meas.observed;
        ll += -0.5 * diff * diff / var;
    }
    return ll;
}

// Field.hpp:174-230
double oddEvenEnergyRatio(const Mesh& mesh, const std::vector<double>& phi) {
    const int n = mesh.nCells();
    if (n == 0) return 0.0;
    std::vector<std::vector<int>> adj(n);
    for (int fi = 0; fi < mesh.nInternalFaces(); ++fi) {
        const Face& f = mesh.face(fi);
        adj[f.owner].push_back(f.neighbor);
        adj[f.neighbor].push_back(f.owner);
    }
    std::vector<int> colour(n, -1);
    std::deque<int> q;
    for (int seed = 0; seed < n; ++seed) {
        if (colour[seed] != -1) continue;
        colour[seed] = 0; q.push_back(seed);
        while (!q.empty()) {
            int c = q.front(); q.pop_front();
            for (int nb : adj[c]) {
                if (colour[nb] == -1) { colour[nb] = 1 - colour[c]; q.push_back(nb); }
                else if (colour[nb] == colour[c]) throw std::invalid_argument("oddEvenEnergyRatio requires a bipartite cell graph");
            }
        }
    }
    double weight = 0.0, mean = 0.0, modeMean = 0.0;
    for (int c = 0; c < n; ++c) {
        double w = std::max(mesh.cell(c).volume, 0.0), q2 = colour[c] == 0 ? 1.0 : -1.0;
        weight += w; mean += w * phi[c]; modeMean += w * q2;
    }
    if (weight <= 1e-300) return 0.0;
    mean /= weight; modeMean /= weight;
    double proj = 0.0, fe = 0.0, me = 0.0;
    for (int c = 0; c < n; ++c) {
        double w = std::max(mesh.cell(c).volume, 0.0);
        double v = phi[c] - mean, q2 = (colour[c] == 0 ? 1.0 : -1.0) - modeMean;
        proj += w * v * q2; fe += w * v * v; me += w * q2 * q2;
    }
    if (fe <= 1e-300 || me <= 1e-300) return 0.0;
    return std::abs(proj) / std::sqrt(fe * me);
}

}  // namespace rsh
