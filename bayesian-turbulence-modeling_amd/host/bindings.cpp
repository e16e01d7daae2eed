// bindings.cpp -- the rans_sst_py module (drop-in for the reference's
// python/Bindings.cpp:273-519, incompressible classes): same class, method
// and argument names; the solve runs on the MI355X engine (librsg.so).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <stdexcept>

#include "compressible.hpp"
#include "forward_model.hpp"
#include "mesh.hpp"
#include "sensitivity.hpp"
#include "model.hpp"

namespace py = pybind11;
using namespace rsh;

namespace {

py::array_t<double> vec3Array(const std::vector<Vec3>& v) {
    py::array_t<double> a({(py::ssize_t)v.size(), (py::ssize_t)3});
    auto r = a.mutable_unchecked<2>();
    for (size_t i = 0; i < v.size(); ++i) { r(i, 0) = v[i].x; r(i, 1) = v[i].y; r(i, 2) = v[i].z; }
    return a;
}

py::array_t<double> arr1(const std::vector<double>& v) {
    py::array_t<double> a((py::ssize_t)v.size());
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}

py::array_t<double> arr3(const std::vector<double>& v) {
    py::array_t<double> a({(py::ssize_t)(v.size() / 3), (py::ssize_t)3});
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}

py::dict wallPatchData(const Mesh& m, const std::string& name) {
    int pi = -1;
    for (int p = 0; p < m.nPatches(); ++p) if (m.patch(p).name == name) { pi = p; break; }
    if (pi < 0) throw std::runtime_error("Mesh::wall_patch_data: unknown patch '" + name + "'");
    const Patch& pat = m.patch(pi);
    const int nf = (int)pat.faces.size();
    py::array_t<int> own(nf);
    py::array_t<double> del(nf), area(nf), cen({nf, 3}), nor({nf, 3});
    auto o = own.mutable_unchecked<1>(); auto d = del.mutable_unchecked<1>(); auto a = area.mutable_unchecked<1>();
    auto c = cen.mutable_unchecked<2>(); auto n = nor.mutable_unchecked<2>();
    for (int k = 0; k < nf; ++k) {
        const Face& f = m.face(pat.faces[k]);
        o(k) = f.owner; d(k) = f.delta; a(k) = f.area;
        c(k, 0) = f.center.x; c(k, 1) = f.center.y; c(k, 2) = f.center.z;
        n(k, 0) = f.normal.x; n(k, 1) = f.normal.y; n(k, 2) = f.normal.z;
    }
    py::dict out;
    out["name"] = pat.name; out["type"] = pat.type; out["n_faces"] = nf;
    out["owner"] = own; out["delta"] = del; out["center"] = cen; out["normal"] = nor; out["area"] = area;
    return out;
}

py::dict fieldsDict(const ForwardModel& fm) {
    if (!fm.hasLastFields()) throw std::runtime_error("No fields available — call evaluate() first");
    const FlowFields& f = fm.lastFields();
    py::dict d;
    d["U"] = arr3(f.U); d["p"] = arr1(f.p); d["k"] = arr1(f.k); d["omega"] = arr1(f.omega);
    d["nuT"] = arr1(f.nuT); d["F1"] = arr1(f.F1); d["F2"] = arr1(f.F2); d["Pk"] = arr1(f.Pk);
    return d;
}

// python/Bindings.cpp:247-271: "p" thermodynamic, "p_mechanical" the working variable
py::dict compFieldsDict(const CompressibleForwardModel& fm) {
    if (!fm.hasLastFields()) throw std::runtime_error("No fields available — call evaluate() first");
    const CompressibleFields& F = fm.lastFields();
    const FlowFields& f = F.flow;
    py::dict d;
    d["U"] = arr3(f.U); d["p"] = arr1(f.p); d["p_mechanical"] = arr1(F.pMech); d["T"] = arr1(F.T);
    d["rho"] = arr1(F.rho); d["k"] = arr1(f.k); d["omega"] = arr1(f.omega); d["nuT"] = arr1(f.nuT);
    d["F1"] = arr1(f.F1); d["F2"] = arr1(f.F2); d["Pk"] = arr1(f.Pk);
    return d;
}

std::vector<double> toVec(py::array_t<double, py::array::c_style | py::array::forcecast> a) {
    auto r = a.unchecked<1>();
    std::vector<double> v(r.shape(0));
    for (py::ssize_t i = 0; i < r.shape(0); ++i) v[i] = r(i);
    return v;
}

}  // namespace

PYBIND11_MODULE(rans_sst_py, m) {
    m.doc() = "RANS-SST forward model G(x, theta) on AMD Instinct MI355X (HIP engine librsg.so)";

    py::enum_<EvaluationStatus>(m, "EvaluationStatus")
        .value("Converged", EvaluationStatus::Converged)
        .value("Unconverged", EvaluationStatus::Unconverged)
        .value("DivergenceDetected", EvaluationStatus::Diverged)
        .value("InvalidParameters", EvaluationStatus::InvalidParameters)
        .value("Unknown", EvaluationStatus::Unknown);

    py::class_<EvaluationResult>(m, "EvaluationResult")
        .def_readonly("status", &EvaluationResult::status)
        .def_readonly("log_lik", &EvaluationResult::loglik)
        .def_readonly("predictions", &EvaluationResult::predictions)
        .def_readonly("simple_iters", &EvaluationResult::simpleIters)
        .def_readonly("latency", &EvaluationResult::latency)
        .def_readonly("device_time", &EvaluationResult::deviceTime);

    py::class_<SSTCoefficients>(m, "SSTCoefficients")
        .def(py::init<>())
        .def_readwrite("sigma_k1", &SSTCoefficients::sigma_k1).def_readwrite("sigma_w1", &SSTCoefficients::sigma_w1)
        .def_readwrite("beta1", &SSTCoefficients::beta1).def_readwrite("alpha1", &SSTCoefficients::alpha1)
        .def_readwrite("sigma_k2", &SSTCoefficients::sigma_k2).def_readwrite("sigma_w2", &SSTCoefficients::sigma_w2)
        .def_readwrite("beta2", &SSTCoefficients::beta2).def_readwrite("alpha2", &SSTCoefficients::alpha2)
        .def_readwrite("betaStar", &SSTCoefficients::betaStar).def_readwrite("a1", &SSTCoefficients::a1)
        .def_readwrite("kappa", &SSTCoefficients::kappa);

    py::class_<InferenceParameterSet>(m, "InferenceParameterSet")
        .def(py::init<>())
        .def_readonly("name", &InferenceParameterSet::name)
        .def("n_active", &InferenceParameterSet::nActive)
        .def("active_names", &InferenceParameterSet::activeNames)
        .def("pack", &InferenceParameterSet::pack)
        .def("unpack", &InferenceParameterSet::unpack)
        .def("in_bounds", &InferenceParameterSet::inBounds)
        .def("lower_bounds", &InferenceParameterSet::lowerBounds)
        .def("upper_bounds", &InferenceParameterSet::upperBounds)
        .def_static("a1_betaStar", &InferenceParameterSet::a1_betaStar)
        .def_static("inlet_turb", &InferenceParameterSet::inletTurb)
        .def_static("near_wall4", &InferenceParameterSet::nearWall4)
        .def_static("all11", &InferenceParameterSet::all11)
        .def_static("live10", &InferenceParameterSet::live10)
        .def_static("from_indices", &InferenceParameterSet::fromIndices, py::arg("name"), py::arg("indices"));

    py::class_<Vec3>(m, "Vec3")
        .def(py::init<double, double, double>())
        .def_readwrite("x", &Vec3::x).def_readwrite("y", &Vec3::y).def_readwrite("z", &Vec3::z);

    py::class_<Mesh>(m, "Mesh")
        .def_static("make_channel_2d", py::overload_cast<int, int, double, double>(&Mesh::makeChannel2D),
                    py::arg("nx"), py::arg("ny"), py::arg("Lx"), py::arg("Ly"))
        .def_static("make_plate_2d", &Mesh::makePlate2D, py::arg("nx"), py::arg("ny"), py::arg("Lx"),
                    py::arg("Ly"), py::arg("Re"), py::arg("y_plus_target") = 1.0)
        .def_static("make_channel_2d", py::overload_cast<int, int, double, double, double, double>(&Mesh::makeChannel2D),
                    py::arg("nx"), py::arg("ny"), py::arg("Lx"), py::arg("Ly"), py::arg("Re"),
                    py::arg("yPlusTarget") = 1.0)
        .def_static("make_backward_facing_step_2d",
                    [](int a, int b, int c, int d, double Lu, double Ld, double hs, double H) {
                        return Mesh::makeBackwardFacingStep2D(a, b, c, d, Lu, Ld, hs, H, 0.0, 1.0);
                    },
                    py::arg("nx_up"), py::arg("nx_down"), py::arg("ny_up"), py::arg("ny_down"), py::arg("Lu"),
                    py::arg("Ld"), py::arg("h_s"), py::arg("H"))
        .def_static("make_backward_facing_step_2d", &Mesh::makeBackwardFacingStep2D, py::arg("nx_up"),
                    py::arg("nx_down"), py::arg("ny_up"), py::arg("ny_down"), py::arg("Lu"), py::arg("Ld"),
                    py::arg("h_s"), py::arg("H"), py::arg("Re"), py::arg("yPlusTarget") = 1.0)
        .def_static("make_curved_channel_periodic_2d",
                    [](py::array_t<double, py::array::c_style | py::array::forcecast> x,
                       py::array_t<double, py::array::c_style | py::array::forcecast> yb, double yTop, int ny,
                       double Re, double yPlusTarget) {
                        return Mesh::makeCurvedChannelPeriodic2D(toVec(x), toVec(yb), yTop, ny, Re, yPlusTarget);
                    },
                    py::arg("x_nodes"), py::arg("y_bottom"), py::arg("y_top"), py::arg("ny"), py::arg("Re"),
                    py::arg("yPlusTarget") = 1.0)
        .def_static("make_channel_periodic_2d",
                    [](py::array_t<double, py::array::c_style | py::array::forcecast> x,
                       py::array_t<double, py::array::c_style | py::array::forcecast> y) {
                        return Mesh::makeChannelPeriodic2D(toVec(x), toVec(y));
                    },
                    py::arg("x_nodes"), py::arg("y_nodes"))
        .def_static("geometric_channel_nodes", &Mesh::geometricChannelNodes, py::arg("n_half"), py::arg("ratio"),
                    py::arg("half_height") = 1.0)
        .def_static("make_channel_column_1d",
                    [](py::array_t<double, py::array::c_style | py::array::forcecast> y, double dx) {
                        return Mesh::makeChannelColumn1D(toVec(y), dx);
                    },
                    py::arg("y_nodes"), py::arg("dx") = 1.0)
        .def_static("load_from_file", &Mesh::loadFromFile)
        .def("save_binary", &Mesh::saveBinary)
        .def("compute_wall_distance",
             [](Mesh& m, bool onDevice) {
                 if (onDevice) computeWallDistanceOnDevice(m);
                 else m.computeWallDistance();
             },
             py::arg("on_device") = false,
             "Mesh::computeWallDistance; on_device=True runs the brute-force search as one GPU kernel "
             "(the same distances, bit for bit)")
        .def("compute_wall_distance_timed",
             [](Mesh& m) { return computeWallDistanceOnDevice(m); },
             "compute_wall_distance(on_device=True), returning the kernel's seconds")
        .def("set_patch_type", &Mesh::setPatchType, py::arg("name"), py::arg("type"))
        .def("n_cells", &Mesh::nCells)
        .def("n_faces", &Mesh::nFaces)
        .def("n_nodes", &Mesh::nNodes)
        .def("n_patches", &Mesh::nPatches)
        .def("n_internal_faces", &Mesh::nInternalFaces)
        .def("patch_names", [](const Mesh& m) { std::vector<std::string> o; for (int p = 0; p < m.nPatches(); ++p) o.push_back(m.patch(p).name); return o; })
        .def("patch_types", [](const Mesh& m) { std::vector<std::string> o; for (int p = 0; p < m.nPatches(); ++p) o.push_back(m.patch(p).type); return o; })
        .def("wall_patch_data", &wallPatchData, py::arg("name"))
        .def("cell_centers", [](const Mesh& m) { std::vector<Vec3> v; for (int c = 0; c < m.nCells(); ++c) v.push_back(m.cell(c).center); return vec3Array(v); })
        .def("cell_volumes", [](const Mesh& m) { std::vector<double> v; for (int c = 0; c < m.nCells(); ++c) v.push_back(m.cell(c).volume); return arr1(v); })
        .def("node_coords", [](const Mesh& m) { std::vector<Vec3> v; for (int i = 0; i < m.nNodes(); ++i) v.push_back(m.node(i)); return vec3Array(v); })
        .def("face_centers", [](const Mesh& m) { std::vector<Vec3> v; for (int f = 0; f < m.nFaces(); ++f) v.push_back(m.face(f).center); return vec3Array(v); })
        .def("wall_distance", [](const Mesh& m) { return arr1(m.wallDistance()); });

    py::class_<FlowBoundaryConditions>(m, "FlowBoundaryConditions")
        .def(py::init<>())
        .def_static("channel_defaults", &FlowBoundaryConditions::channelDefaults)
        .def_static("flat_plate_defaults", &FlowBoundaryConditions::flatPlateDefaults)
        .def_static("bfs_defaults", &FlowBoundaryConditions::bfsDefaults)
        .def_static("couette_defaults", &FlowBoundaryConditions::couetteDefaults, py::arg("mesh"), py::arg("Uwall"),
                    py::arg("kIn"), py::arg("omIn"))
        .def("set_velocity_profile",
             [](FlowBoundaryConditions& bc, const Mesh& mesh, const std::string& name,
                py::array_t<double, py::array::c_style | py::array::forcecast> vals) {
                 auto v = vals.unchecked<2>();
                 if (v.shape(1) != 3) throw std::runtime_error("velocity profile must be (n_faces, 3)");
                 std::vector<Vec3> out(v.shape(0));
                 for (py::ssize_t i = 0; i < v.shape(0); ++i) out[i] = Vec3(v(i, 0), v(i, 1), v(i, 2));
                 bc.setVelocityProfile(mesh, name, out);
             },
             py::arg("mesh"), py::arg("name"), py::arg("values"))
        .def("set_k_profile",
             [](FlowBoundaryConditions& bc, const Mesh& mesh, const std::string& name,
                py::array_t<double, py::array::c_style | py::array::forcecast> vals) {
                 bc.setKProfile(mesh, name, toVec(vals));
             },
             py::arg("mesh"), py::arg("name"), py::arg("values"))
        .def("set_omega_profile",
             [](FlowBoundaryConditions& bc, const Mesh& mesh, const std::string& name,
                py::array_t<double, py::array::c_style | py::array::forcecast> vals) {
                 bc.setOmegaProfile(mesh, name, toVec(vals));
             },
             py::arg("mesh"), py::arg("name"), py::arg("values"));

    py::class_<SolverSettings>(m, "SolverSettings")
        .def(py::init<>())
        .def_readwrite("max_iterations", &SolverSettings::maxIterations)
        .def_readwrite("convergence_tol", &SolverSettings::convergenceTol)
        .def_readwrite("divergence_limit", &SolverSettings::divergenceLimit)
        .def_readwrite("alpha_u", &SolverSettings::alphaU)
        .def_readwrite("alpha_p", &SolverSettings::alphaP)
        .def_readwrite("alpha_k", &SolverSettings::alphaK)
        .def_readwrite("alpha_omega", &SolverSettings::alphaOmega)
        .def_readwrite("inner_iterations", &SolverSettings::innerIterations)
        .def_readwrite("inner_tolerance", &SolverSettings::innerTolerance)
        .def_readwrite("turb_start_iter", &SolverSettings::turbStartIter)
        .def_readwrite("turb_update_interval", &SolverSettings::turbUpdateInterval)
        .def_readwrite("nut_floor_iters", &SolverSettings::nuTFloorIters)
        .def_readwrite("rhie_chow_all_meshes", &SolverSettings::rhieChowAllMeshes)
        .def_readwrite("alpha_injection", &SolverSettings::alphaInjection)
        .def_readwrite("body_force", &SolverSettings::bodyForce)
        .def_readwrite("k_min", &SolverSettings::kMin)
        .def_readwrite("omega_min", &SolverSettings::omegaMin)
        .def_readwrite("verbose", &SolverSettings::verbose)
        .def_readwrite("report_interval", &SolverSettings::reportInterval)
        .def_readwrite("alpha_t", &SolverSettings::alphaT)
        .def_readwrite("use_wall_functions", &SolverSettings::useWallFunctions)
        .def_readwrite("von_karman", &SolverSettings::vonKarman)
        .def_readwrite("wall_fn_E", &SolverSettings::wallFnE)
        .def_readwrite("sst_variant", &SolverSettings::sstVariant)
        // bound here although the reference leaves them unbound (FlowFields.hpp:111-113)
        .def_readwrite("pressure_solver", &SolverSettings::pressureSolver)
        .def_readwrite("momentum_solver", &SolverSettings::momentumSolver)
        .def_readwrite("turbulence_solver", &SolverSettings::turbulenceSolver)
        .def_readwrite("k_max", &SolverSettings::kMax)
        // engine knobs
        .def_readwrite("amg_refresh", &SolverSettings::amgRefresh)
        .def_readwrite("coarse_refresh", &SolverSettings::coarseRefresh)
        .def_readwrite("record_history", &SolverSettings::recordHistory);

    py::class_<ObservationOperator>(m, "ObservationOperator")
        .def(py::init<>())
        .def("add_drag", &ObservationOperator::addDrag, py::arg("wall_patch"), py::arg("cd_obs"), py::arg("sigma"),
             py::arg("ref_area"), py::arg("ref_vel"), py::arg("sigma_model") = 0.0)
        .def("add_skin_friction", &ObservationOperator::addSkinFriction, py::arg("wall_patch"), py::arg("location"),
             py::arg("cf_obs"), py::arg("sigma"), py::arg("ref_vel"), py::arg("sigma_model") = 0.0)
        .def("add_velocity_profile", &ObservationOperator::addVelocityProfile, py::arg("location"),
             py::arg("component"), py::arg("u_obs"), py::arg("sigma"), py::arg("sigma_model") = 0.0)
        .def("add_reattachment_length", &ObservationOperator::addReattachmentLength, py::arg("wall_patch"),
             py::arg("xr_obs"), py::arg("sigma"), py::arg("sigma_model") = 0.0)
        .def("n_obs", &ObservationOperator::nObs);

    py::class_<ForwardModel>(m, "ForwardModel")
        .def(py::init<const Mesh&, const InferenceParameterSet&, const ObservationOperator&,
                      const FlowBoundaryConditions&, double, const SolverSettings&, const Vec3&, double, double,
                      double>(),
             py::arg("mesh"), py::arg("param_set"), py::arg("obs_op"), py::arg("bcs"), py::arg("nu"),
             py::arg("settings") = SolverSettings{}, py::arg("u_init") = Vec3{1, 0, 0}, py::arg("p_init") = 0.0,
             py::arg("k_init") = 1e-4, py::arg("omega_init") = 1.0, py::keep_alive<1, 2>())
        .def("evaluate", &ForwardModel::evaluate, py::call_guard<py::gil_scoped_release>())
        .def("evaluate_batch", &ForwardModel::evaluateBatch, py::call_guard<py::gil_scoped_release>())
        .def("submit", [](ForwardModel& fm, const std::vector<double>& theta, bool useCache) {
                 if (!fm.paramSet().inBounds(theta)) throw std::invalid_argument("theta out of bounds");
                 py::gil_scoped_release nogil;
                 return fm.submit(theta, useCache);
             }, py::arg("theta"), py::arg("use_cache") = true,
             "Enqueue a solve on the device engine (warm-started from the nearest cached state unless "
             "use_cache is False); returns a ticket.")
        .def("complete", &ForwardModel::complete, py::arg("ticket"), py::arg("theta"),
             py::arg("store_cache") = true, py::call_guard<py::gil_scoped_release>(),
             "Wait for a ticket and return its EvaluationResult.")
        .def("penalized_log_likelihood", &ForwardModel::penalizedLogLikelihood,
             py::call_guard<py::gil_scoped_release>())
        .def("param_set", &ForwardModel::paramSet, py::return_value_policy::reference)
        .def("has_last_fields", &ForwardModel::hasLastFields)
        .def("last_fields", &fieldsDict)
        .def("last_history", [](const ForwardModel& fm) {
            const auto& h = fm.lastConvergence().history;
            py::array_t<double> a({(py::ssize_t)(h.size() / 6), (py::ssize_t)6});
            std::copy(h.begin(), h.end(), a.mutable_data());
            return a;
        })
        .def("last_inner_iterations", [](const ForwardModel& fm) {
            const auto& c = fm.lastConvergence();
            return std::vector<double>(c.innerIters, c.innerIters + 6);
        })
        .def("set_precision", [](ForwardModel& fm, const std::string& p) {
            if (p == "fp64") fm.setPrecision(RSG_PREC_FP64);
            else if (p == "mixed") fm.setPrecision(RSG_PREC_MIXED);
            else throw std::invalid_argument("precision must be 'fp64' or 'mixed'");
        })
        .def("set_collect_fields", &ForwardModel::setCollectFields, py::arg("on"),
             "False: later solves keep their fields on the device and return predictions, log-likelihood and "
             "status only (no last_fields, no warm-start cache entry).")
        .def("collect_fields", &ForwardModel::collectFields)
        .def("cache_size", [](ForwardModel& fm) { return fm.cache().size(); })
        .def("clear_cache", [](ForwardModel& fm) { fm.cache().clear(); });

    // compressible pressure-based solver (python/Bindings.cpp:585-635)
    py::class_<IdealGasEOS>(m, "IdealGasEOS")
        .def(py::init<>())
        .def_readwrite("gamma", &IdealGasEOS::gamma).def_readwrite("R", &IdealGasEOS::R)
        .def_readwrite("Pr", &IdealGasEOS::Pr).def_readwrite("Pr_T", &IdealGasEOS::Pr_T)
        .def_readwrite("mu_ref", &IdealGasEOS::mu_ref).def_readwrite("T_ref", &IdealGasEOS::T_ref)
        .def_readwrite("S_suth", &IdealGasEOS::S_suth)
        .def("Cp", &IdealGasEOS::Cp)
        .def("Cv", &IdealGasEOS::Cv)
        .def("viscosity", &IdealGasEOS::viscosity, py::arg("T"))
        .def("conductivity", &IdealGasEOS::conductivity, py::arg("mu"), py::arg("muT"))
        .def("density", &IdealGasEOS::density, py::arg("p"), py::arg("T"))
        .def("temperature", &IdealGasEOS::temperature, py::arg("p"), py::arg("rho"))
        .def("pressure", &IdealGasEOS::pressure, py::arg("rho"), py::arg("T"))
        .def("sound_speed", &IdealGasEOS::soundSpeed, py::arg("T"))
        .def("mach_number", &IdealGasEOS::machNumber, py::arg("Umag"), py::arg("T"))
        .def("total_enthalpy", &IdealGasEOS::totalEnthalpy, py::arg("T"), py::arg("Umag2"));

    py::enum_<TBCType>(m, "TBCType")
        .value("Dirichlet", TBCType::Dirichlet)
        .value("Neumann", TBCType::Neumann)
        .value("AdiabaticWall", TBCType::AdiabaticWall);

    py::class_<PatchTBC>(m, "PatchTBC")
        .def(py::init<>())
        .def_readwrite("type", &PatchTBC::type)
        .def_readwrite("value", &PatchTBC::value)
        .def_readwrite("patch_name", &PatchTBC::patchName);

    py::class_<CompressibleBoundaryConditions>(m, "CompressibleBoundaryConditions")
        .def(py::init<>())
        .def_static("channel_defaults", &CompressibleBoundaryConditions::channelDefaults, py::arg("mesh"),
                    py::arg("Uin"), py::arg("T_in"), py::arg("p_out"), py::arg("kIn"), py::arg("omIn"))
        .def_readwrite("temperature_bc", &CompressibleBoundaryConditions::temperatureBC);

    py::class_<CompressibleForwardModel>(m, "CompressibleForwardModel")
        .def(py::init<const Mesh&, const InferenceParameterSet&, const ObservationOperator&,
                      const CompressibleBoundaryConditions&, const IdealGasEOS&, const SolverSettings&, const Vec3&,
                      double, double, double, double>(),
             py::arg("mesh"), py::arg("param_set"), py::arg("obs_op"), py::arg("bcs"), py::arg("eos"),
             py::arg("settings") = SolverSettings{}, py::arg("u_init") = Vec3{1.0, 0.0, 0.0},
             py::arg("p_init") = 101325.0, py::arg("T_init") = 300.0, py::arg("k_init") = 1e-4,
             py::arg("omega_init") = 100.0, py::keep_alive<1, 2>())
        .def("evaluate", &CompressibleForwardModel::evaluate, py::call_guard<py::gil_scoped_release>())
        .def("evaluate_batch", &CompressibleForwardModel::evaluateBatch, py::call_guard<py::gil_scoped_release>(),
             "Many thetas solved concurrently on the device; each result equals evaluate(theta) alone.")
        .def("penalized_log_likelihood", &CompressibleForwardModel::penalizedLogLikelihood,
             py::call_guard<py::gil_scoped_release>())
        .def("param_set", &CompressibleForwardModel::paramSet, py::return_value_policy::reference)
        .def("has_last_fields", &CompressibleForwardModel::hasLastFields)
        .def("last_fields", &compFieldsDict)
        .def("last_history", [](const CompressibleForwardModel& fm) {
            const auto& h = fm.lastConvergence().history;
            py::array_t<double> a({(py::ssize_t)(h.size() / 6), (py::ssize_t)6});
            std::copy(h.begin(), h.end(), a.mutable_data());
            return a;
        }, "Per outer iteration: Ux, Uy, p residuals, T, k, omega change norms (settings.record_history).")
        .def("last_inner_iterations", [](const CompressibleForwardModel& fm) {
            const auto& c = fm.lastConvergence();
            std::vector<double> v(c.innerIters, c.innerIters + 6);
            v.push_back(fm.lastEnergyIterations());
            return v;
        }, "Cumulative inner iterations: Ux, Uy, Uz, p', k, omega, energy.");

    // TangentGradientResult / ParameterSensitivity (python/Bindings.cpp:521-576): the
    // forward-path members -- solve_state, observe, log_lik and the warm-FD gradient.
    py::class_<TangentGradientResult>(m, "TangentGradientResult")
        .def_readonly("d_obs_d_theta", &TangentGradientResult::dObsDTheta)
        .def_readonly("log_lik_gradient", &TangentGradientResult::logLikGradient)
        .def_readonly("krylov_iters", &TangentGradientResult::krylovIters)
        .def_readonly("krylov_rel_res", &TangentGradientResult::krylovRelRes)
        .def_readonly("krylov_converged", &TangentGradientResult::krylovConverged)
        .def_readonly("n_residual_evals", &TangentGradientResult::nResidualEvals)
        .def_readonly("n_colors", &TangentGradientResult::nColors);

    py::class_<ParameterSensitivity>(m, "ParameterSensitivity")
        .def(py::init<const Mesh&, const ObservationOperator&, const FlowBoundaryConditions&, double,
                      const SolverSettings&, const Vec3&, double, double, double>(),
             py::arg("mesh"), py::arg("obs_op"), py::arg("bcs"), py::arg("nu"),
             py::arg("settings") = SolverSettings{}, py::arg("u_init") = Vec3{1, 0, 0}, py::arg("p_init") = 0.0,
             py::arg("k_init") = 1e-4, py::arg("omega_init") = 1.0, py::keep_alive<1, 2>())
        .def("solve_state", &ParameterSensitivity::solveState, py::arg("theta"),
             py::call_guard<py::gil_scoped_release>())
        .def("has_state", &ParameterSensitivity::hasState)
        .def("n_state", &ParameterSensitivity::nState)
        .def("n_cells", &ParameterSensitivity::nCells)
        .def("n_obs", &ParameterSensitivity::nObs)
        .def("observe", &ParameterSensitivity::observe, py::arg("theta"))
        .def("log_lik", &ParameterSensitivity::logLik, py::arg("theta"))
        .def("eta_jacobian_warm_fd", &ParameterSensitivity::etaJacobianWarmFD, py::arg("theta"),
             py::arg("h_rel") = 5e-4, py::arg("h_floor") = 1e-7, py::arg("warm_max_iter") = 0,
             py::arg("warm_tol") = 0.0, py::call_guard<py::gil_scoped_release>())
        .def("state_fields", [](const ParameterSensitivity& ps) {
            py::dict d;
            const FlowFields& f = ps.state();
            d["U"] = py::array_t<double>(f.U.size(), f.U.data());
            d["p"] = py::array_t<double>(f.p.size(), f.p.data());
            d["k"] = py::array_t<double>(f.k.size(), f.k.data());
            d["omega"] = py::array_t<double>(f.omega.size(), f.omega.data());
            d["nuT"] = py::array_t<double>(f.nuT.size(), f.nuT.data());
            d["F1"] = py::array_t<double>(f.F1.size(), f.F1.data());
            d["F2"] = py::array_t<double>(f.F2.size(), f.F2.data());
            d["Pk"] = py::array_t<double>(f.Pk.size(), f.Pk.data());
            d["CDkw"] = py::array_t<double>(f.CDkw.size(), f.CDkw.data());
            d["U_bnd"] = py::array_t<double>(f.Ub.size(), f.Ub.data());
            d["p_bnd"] = py::array_t<double>(f.pb.size(), f.pb.data());
            d["k_bnd"] = py::array_t<double>(f.kb.size(), f.kb.data());
            d["omega_bnd"] = py::array_t<double>(f.omb.size(), f.omb.data());
            d["turb_established"] = f.turbEstablished;
            return d;
        }, "The stored fixed-point state (host copies).");

    m.def("odd_even_energy_ratio",
          [](const Mesh& mesh, const std::vector<double>& values) {
              if ((int)values.size() != mesh.nCells())
                  throw std::invalid_argument("odd_even_energy_ratio requires exactly one value per cell");
              return oddEvenEnergyRatio(mesh, values);
          },
          py::arg("mesh"), py::arg("values"));

    m.def("evaluate_many",
          [](std::vector<ForwardModel*> fms, std::vector<std::vector<double>> thetas) {
              py::gil_scoped_release nogil;
              return evaluateMany(fms, thetas);
          },
          py::arg("models"), py::arg("thetas"),
          "Solve one theta per ForwardModel concurrently on the device; results in input order.");

    m.def("engine_set_launch", [](int iters, int block) { Engine::get().setLaunch(iters, block); },
          py::arg("iters_per_launch"), py::arg("block_size") = 0);
    m.def("engine_set_team", [](int minCells, int blocks, int iters) { Engine::get().setTeam(minCells, blocks, iters); },
          py::arg("min_cells"), py::arg("blocks") = 0, py::arg("iters_per_launch") = 20,
          "Members on meshes of at least min_cells cells are solved by a team of `blocks` workgroups "
          "(0: 16 sqrt(cells / 40000), at least 16) in cooperative launches of iters_per_launch outer iterations; min_cells 0 turns "
          "team solves off.");
    m.def("engine_set_grid", [](int minCells, int iters) { Engine::get().setGrid(minCells, iters); },
          py::arg("min_cells"), py::arg("iters_per_launch") = 20,
          "Members on meshes of at least min_cells cells (checked at admission) run alone on the whole GPU, "
          "every phase of the outer iteration a grid-wide launch, iters_per_launch outer iterations per engine "
          "round; min_cells 0 turns it off.  Takes precedence over team solves.");
    m.def("engine_set_grid_levels", [](int levels) { Engine::get().setGridLevels(levels); }, py::arg("max_levels"),
          "p' AMG levels of kernel-per-phase members (0: the reference's 20-level cap down to 50 rows; k > 1: k "
          "levels, the coarsest inverted on the FP64 matrix cores).");
    m.def("engine_set_graphs", [](bool enable) { Engine::get().setGraphs(enable); }, py::arg("enable"),
          "Replay AMG V-cycles of kernel-per-phase members from hipGraphs (default on); results are identical "
          "either way.");
    m.def("engine_set_time_slice", [](int us) { Engine::get().setTimeSlice(us); }, py::arg("microseconds"));
    m.def("engine_hold_admission", [](bool hold) {
              rsg_context* c = Engine::get().ctx();
              const int rc = rsg_hold_admission(c, hold ? 1 : 0);
              if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
          }, py::arg("hold"),
          "hold=True stops the engine admitting submitted solves until the matching hold=False, so that a batch "
          "submitted in between starts in one launch; never wait for a solve while holding.");
    m.def("engine_set_schedule", [](const std::string& p) {
              if (p == "round_robin") Engine::get().setSchedule(RSG_SCHEDULE_ROUND_ROBIN);
              else if (p == "largest_first") Engine::get().setSchedule(RSG_SCHEDULE_LARGEST_FIRST);
              else throw std::invalid_argument("schedule must be 'round_robin' or 'largest_first'");
          }, py::arg("policy"));
    m.def("ticket_states",
          [](const std::vector<int64_t>& tickets) {
              std::vector<int32_t> st(tickets.size(), 0);
              if (!tickets.empty()) {
                  rsg_context* c = Engine::get().ctx();
                  const int rc = rsg_poll(c, tickets.data(), (int32_t)tickets.size(), st.data());
                  if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
              }
              return st;
          },
          py::arg("tickets"),
          "Engine state of each ticket: 0 pending, 1 running, 2 done, 3 failed (does not block).");
    m.def("engine_stats", []() { return Engine::get().stats(); });
    m.def("engine_device_info",
          []() {
              int32_t out[2] = {0, 0};
              rsg_context* c = Engine::get().ctx();
              const int rc = rsg_device_info(c, out);
              if (rc != RSG_OK) throw std::runtime_error(engineError(c, rc));
              py::dict d;
              d["cu_count"] = out[0];
              d["wavefront_size"] = out[1];
              return d;
          },
          "Compute units and wavefront size of the engine's device.");
    m.def("_shutdown", []() { Engine::get().shutdown(); });
    py::module_::import("atexit").attr("register")(m.attr("_shutdown"));

    py::register_exception_translator([](std::exception_ptr p) {
        try {
            if (p) std::rethrow_exception(p);
        } catch (const std::invalid_argument& e) {
            PyErr_SetString(PyExc_ValueError, e.what());
        }
    });
}
