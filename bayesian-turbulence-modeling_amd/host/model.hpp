// model.hpp -- host-side model description: boundary conditions, solver
// settings, SST coefficients and inference parameter sets, flow state and the
// observation operator.
//
// Reference counterparts (same names, same argument meaning and defaults):
//   FlowBoundaryConditions   core/include/BoundaryCondition.hpp:12-190
//   SolverSettings           core/include/FlowFields.hpp:53-132
//   SSTCoefficients          core/include/SSTModel.hpp:9-40
//   InferenceParameterSet    core/include/InferenceParameters.hpp:18-161
//   FlowFields               core/include/FlowFields.hpp:9-33
//   ObservationOperator      core/include/ObservationOperator.hpp:12-378
#pragma once
#include <string>
#include <vector>

#include "mesh.hpp"
#include "rsg.h"

namespace rsh {

enum class BCType { Dirichlet, Neumann, InletVelocity, OutletPressure, WallNoSlip, WallKOmega, WallMoving, Symmetry, Cyclic };

struct PatchBC {
    BCType type = BCType::Neumann;
    double value = 0.0;
    Vec3 vecValue;
    std::string patchName;
    std::vector<Vec3> vecProfile;
    std::vector<double> scalarProfile;
};

// Per-boundary-face arrays handed to rsg_problem_create.
struct ResolvedBC {
    std::vector<int32_t> velT, pT, kT, omT;
    std::vector<double> velV, pV, kV, omV;
    rsg_bc_desc desc{};
};

struct FlowBoundaryConditions {
    std::vector<PatchBC> velocityBC, pressureBC, kBC, omegaBC;

    static FlowBoundaryConditions channelDefaults(const Mesh& mesh, double Uin, double kIn, double omIn);
    static FlowBoundaryConditions bfsDefaults(const Mesh& mesh, double Uin, double kIn, double omIn) {
        return channelDefaults(mesh, Uin, kIn, omIn);
    }
    static FlowBoundaryConditions flatPlateDefaults(const Mesh& mesh, double Uinf, double kIn, double omIn);
    static FlowBoundaryConditions couetteDefaults(const Mesh& mesh, double Uwall, double kIn, double omIn);

    static int patchIndex(const Mesh& mesh, const std::string& name);
    void setVelocityProfile(const Mesh& mesh, const std::string& name, const std::vector<Vec3>& vals);
    void setKProfile(const Mesh& mesh, const std::string& name, const std::vector<double>& vals);
    void setOmegaProfile(const Mesh& mesh, const std::string& name, const std::vector<double>& vals);

    // face-by-face values of every BC (profiles applied), in boundary-face order
    void resolve(const Mesh& mesh, ResolvedBC& out) const;
};

struct SolverSettings {
    int maxIterations = 500;
    double convergenceTol = 1e-5;
    double divergenceLimit = 1e6;
    double alphaU = 0.7, alphaP = 0.3, alphaK = 0.5, alphaOmega = 0.5, alphaT = 0.7;
    int innerIterations = 200;
    double innerTolerance = 1e-3;
    int turbStartIter = 5, turbUpdateInterval = 1;
    int nuTFloorIters = 500;
    bool rhieChowAllMeshes = false;
    double alphaInjection = 0.3;
    Vec3 bodyForce = Vec3(0.0, 0.0, 0.0);
    double kMin = 1e-10, kMax = 1e10, omegaMin = 1e-6;
    int reportInterval = 10;
    bool verbose = true;
    std::string pressureSolver = "AMG", momentumSolver = "BiCGSTAB", turbulenceSolver = "BiCGSTAB";
    bool useWallFunctions = false;
    double vonKarman = 0.41, wallFnE = 9.0;
    int sstVariant = 0;
    // engine knobs (no reference counterpart)
    int amgRefresh = 5000;
    int coarseRefresh = 10;
    bool recordHistory = false;

    rsg_settings toC() const;
};

struct SSTCoefficients {
    double sigma_k1 = 0.85, sigma_w1 = 0.5, beta1 = 0.075, alpha1 = 5.0 / 9.0;
    double sigma_k2 = 1.0, sigma_w2 = 0.856, beta2 = 0.0828, alpha2 = 0.44;
    double betaStar = 0.09, a1 = 0.31, kappa = 0.41;
};

struct InferenceParameterSet {
    std::string name;
    std::vector<int> activeIndices;
    SSTCoefficients base;

    int nActive() const { return (int)activeIndices.size(); }
    static std::vector<double> toVector(const SSTCoefficients& c);
    static SSTCoefficients fromVector(const std::vector<double>& v);
    SSTCoefficients unpack(const std::vector<double>& theta) const;
    std::vector<double> pack(const SSTCoefficients& c) const;
    std::vector<std::string> activeNames() const;
    std::vector<double> lowerBounds() const;
    std::vector<double> upperBounds() const;
    bool inBounds(const std::vector<double>& theta) const;
    static InferenceParameterSet a1_betaStar() { return {"a1_betaStar", {9, 8}, SSTCoefficients{}}; }
    static InferenceParameterSet inletTurb() { return {"inletTurb", {2, 1}, SSTCoefficients{}}; }
    static InferenceParameterSet nearWall4() { return {"nearWall4", {9, 8, 2, 0}, SSTCoefficients{}}; }
    static InferenceParameterSet all11() { return {"all11", {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10}, SSTCoefficients{}}; }
    static InferenceParameterSet live10() { return {"live10", {0, 1, 2, 3, 4, 5, 6, 7, 8, 9}, SSTCoefficients{}}; }
    static InferenceParameterSet fromIndices(const std::string& name, const std::vector<int>& idx);
};

// Host copy of a solved state (cell values + boundary-face values).
struct FlowFields {
    int n = 0, nB = 0;
    std::vector<double> U;            // [3n] interleaved
    std::vector<double> p, k, omega, nuT, F1, F2, Pk, CDkw;
    std::vector<double> Ub, pb, kb, omb;   // [3nB], [nB] ...
    bool turbEstablished = false;
    void resize(int nc, int nb);
};

struct ConvergenceInfo {
    int finalIter = 0;
    bool converged = false, diverged = false;
    std::vector<double> history;      // 6 per outer iteration
    double innerIters[6] = {0, 0, 0, 0, 0, 0};
};

enum class ObsType { Drag, PressureTap, VelocityProfile, SkinFriction, SeparationPoint, ReattachmentLength };

struct Measurement {
    double observed = 0.0, sigma_obs = 0.01, sigma_model = 0.0;
    double totalVariance() const { return sigma_obs * sigma_obs + sigma_model * sigma_model; }
};

struct Observable {
    ObsType type = ObsType::Drag;
    Measurement meas;
    std::string patchName;
    Vec3 location;
    int component = 0;
    double referenceArea = 1.0, refVelocity = 1.0, refPressure = 0.0, refDensity = 1.0;
};

// Observables resolved against one mesh for the device (rsg_problem_set_observations).
struct ResolvedObservations {
    std::vector<rsg_observation> obs;
    std::vector<int32_t> faces;
    std::vector<double> faceX;
};

class ObservationOperator {
public:
    // The mesh-only part of evaluate: nearest cells / faces and patch face lists.
    void resolve(const Mesh& mesh, ResolvedObservations& out) const;
    void addDrag(const std::string& wallPatch, double Cd_obs, double sigma, double refArea, double refVel,
                 double sigma_model = 0.0);
    void addPressureTap(const Vec3& loc, double Cp_obs, double sigma, double refVel, double sigma_model = 0.0,
                        double refPressure = 0.0, double refDensity = 1.0);
    void addVelocityProfile(const Vec3& loc, int comp, double Uobs, double sigma, double sigma_model = 0.0);
    void addSkinFriction(const std::string& wallPatch, const Vec3& loc, double Cf_obs, double sigma,
                         double refVel, double sigma_model = 0.0);
    void addSeparationPoint(const std::string& wallPatch, double xSep_obs, double sigma, double sigma_model = 0.0);
    void addReattachmentLength(const std::string& wallPatch, double xr_obs, double sigma, double sigma_model = 0.0);

    std::vector<double> evaluate(const Mesh& mesh, const FlowFields& f, double nu) const;
    double logLikelihood(const std::vector<double>& pred) const;
    int nObs() const { return (int)obs_.size(); }
    const std::vector<Observable>& observables() const { return obs_; }

private:
    std::vector<Observable> obs_;
};

// Checkerboard energy ratio of a cell field (core/include/Field.hpp:174-230).
double oddEvenEnergyRatio(const Mesh& mesh, const std::vector<double>& phi);

}  // namespace rsh
