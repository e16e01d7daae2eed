// Test-infrastructure driver around the reference solver (johnz1104/bayesian-turbulence-modeling).
//
// This file is NOT product code.  It is compiled together with the reference's
// own sources (core/src/Mesh.cpp, core/src/SSTModel.cpp,
// incompressible/src/SIMPLESolver.cpp, read in place under the reference
// checkout) into oracle/_ref/libref.so by oracle/Makefile.  It exposes a small
// extern "C" surface so the tests can (a) pin the C restatement in
// oracle/rsg_oracle.c bit-for-bit and (b) time the reference as the CPU
// baseline.  Only tests/, the smoke check and bench.py's cpu_baseline leg load
// the resulting library.
//
// Every entry returns 0 on success, -1 on a caught C++ exception.

#include "Mesh.hpp"
#include "Field.hpp"
#include "FlowFields.hpp"
#include "BoundaryCondition.hpp"
#include "LinearSolver.hpp"
#include "SSTModel.hpp"
#include "SIMPLESolver.hpp"
#include "StressOperators.hpp"
#include "ObservationOperator.hpp"
#include "CompressibleSIMPLESolver.hpp"

#include <cstdint>
#include <cstring>
#include <cstdio>
#include <string>
#include <vector>
#include <exception>

namespace {

// Settings arrive packed as doubles; layout documented in oracle/rsg_oracle.h
// (ORS_* indices), shared by the oracle restatement and the tests.
enum {
    S_MAX_ITER = 0, S_CONV_TOL, S_DIV_LIMIT, S_ALPHA_U, S_ALPHA_P, S_ALPHA_K,
    S_ALPHA_OM, S_INNER_ITER, S_INNER_TOL, S_TURB_START, S_TURB_INTERVAL,
    S_NUT_FLOOR_ITERS, S_RC_ALL, S_BF_X, S_BF_Y, S_BF_Z, S_K_MIN, S_K_MAX,
    S_OM_MIN, S_WALL_FN, S_VON_KARMAN, S_SST_VARIANT, S_P_SOLVER, S_M_SOLVER,
    S_T_SOLVER, S_COUNT
};

const char* solverName(int code) {
    switch (code) {
        case 0: return "AMG";
        case 1: return "PCG";
        case 2: return "BiCGSTAB";
        case 3: return "GaussSeidel";
        default: return "BiCGSTAB";
    }
}

SolverSettings unpackSettings(const double* s) {
    SolverSettings o;
    o.maxIterations = (int)s[S_MAX_ITER];
    o.convergenceTol = s[S_CONV_TOL];
    o.divergenceLimit = s[S_DIV_LIMIT];
    o.alphaU = s[S_ALPHA_U];
    o.alphaP = s[S_ALPHA_P];
    o.alphaK = s[S_ALPHA_K];
    o.alphaOmega = s[S_ALPHA_OM];
    o.innerIterations = (int)s[S_INNER_ITER];
    o.innerTolerance = s[S_INNER_TOL];
    o.turbStartIter = (int)s[S_TURB_START];
    o.turbUpdateInterval = (int)s[S_TURB_INTERVAL];
    o.nuTFloorIters = (int)s[S_NUT_FLOOR_ITERS];
    o.rhieChowAllMeshes = s[S_RC_ALL] != 0.0;
    o.bodyForce = Vec3(s[S_BF_X], s[S_BF_Y], s[S_BF_Z]);
    o.kMin = s[S_K_MIN];
    o.kMax = s[S_K_MAX];
    o.omegaMin = s[S_OM_MIN];
    o.useWallFunctions = s[S_WALL_FN] != 0.0;
    o.vonKarman = s[S_VON_KARMAN];
    o.sstVariant = (int)s[S_SST_VARIANT];
    o.pressureSolver = solverName((int)s[S_P_SOLVER]);
    o.momentumSolver = solverName((int)s[S_M_SOLVER]);
    o.turbulenceSolver = solverName((int)s[S_T_SOLVER]);
    o.verbose = false;
    return o;
}

SSTCoefficients unpackCoeffs(const double* c) {
    SSTCoefficients o;
    o.sigma_k1 = c[0]; o.sigma_w1 = c[1]; o.beta1 = c[2]; o.alpha1 = c[3];
    o.sigma_k2 = c[4]; o.sigma_w2 = c[5]; o.beta2 = c[6]; o.alpha2 = c[7];
    o.betaStar = c[8]; o.a1 = c[9]; o.kappa = c[10];
    return o;
}

// bc_kind: 0 channelDefaults, 1 couetteDefaults, 2 flatPlateDefaults
FlowBoundaryConditions makeBCs(const Mesh& m, int kind, const double* v) {
    if (kind == 1) return FlowBoundaryConditions::couetteDefaults(m, v[0], v[1], v[2]);
    if (kind == 2) return FlowBoundaryConditions::flatPlateDefaults(m, v[0], v[1], v[2]);
    return FlowBoundaryConditions::channelDefaults(m, v[0], v[1], v[2]);
}

Mesh loadMesh(const char* path) {
    Mesh m = Mesh::loadFromFile(path);
    m.computeWallDistance();
    return m;
}

LinearSystem makeLdu(int n, int nif, const int* own, const int* nbr,
                     const double* diag, const double* upper, const double* lower,
                     const double* src) {
    LinearSystem A(n, nif);
    for (int i = 0; i < n; ++i) { A.diag[i] = diag[i]; A.source[i] = src[i]; }
    for (int f = 0; f < nif; ++f) {
        A.own[f] = own[f]; A.nbr[f] = nbr[f];
        A.upper[f] = upper[f]; A.lower[f] = lower[f];
    }
    return A;
}

void copyScalar(const ScalarField& f, double* out, double* bout) {
    if (out) for (int i = 0; i < f.size(); ++i) out[i] = f[i];
    if (bout) for (int i = 0; i < f.nBnd(); ++i) bout[i] = f.bnd(i);
}

// Runs SIMPLESolver::solve on f and exports the state (shared by the cold
// and the warm entry points).
int solveAndExport(const Mesh& m, SIMPLESolver& solver, FlowFields& f,
                   double* U, double* p, double* k, double* om, double* nuT,
                   double* F1, double* F2, double* Pk, double* CDkw,
                   double* Ub, double* pb, double* kb, double* omb,
                   double* hist, int histCap, int* info) {
    ConvergenceHistory h = solver.solve(f);
    const int n = m.nCells();
    if (U) for (int i = 0; i < n; ++i) { U[3*i] = f.U[i].x; U[3*i+1] = f.U[i].y; U[3*i+2] = f.U[i].z; }
    if (Ub) for (int i = 0; i < f.U.nBnd(); ++i) { Ub[3*i] = f.U.bnd(i).x; Ub[3*i+1] = f.U.bnd(i).y; Ub[3*i+2] = f.U.bnd(i).z; }
    copyScalar(f.p, p, pb); copyScalar(f.k, k, kb); copyScalar(f.omega, om, omb);
    copyScalar(f.nuT, nuT, nullptr); copyScalar(f.F1, F1, nullptr);
    copyScalar(f.F2, F2, nullptr); copyScalar(f.Pk, Pk, nullptr);
    copyScalar(f.CDkw, CDkw, nullptr);
    int ne = (int)h.entries.size();
    if (hist) for (int e = 0; e < ne && e < histCap; ++e) {
        const auto& r = h.entries[e];
        hist[6*e] = r.Ux; hist[6*e+1] = r.Uy; hist[6*e+2] = r.Uz;
        hist[6*e+3] = r.p; hist[6*e+4] = r.k; hist[6*e+5] = r.omega;
    }
    info[0] = h.finalIter; info[1] = h.converged; info[2] = h.diverged; info[3] = ne;
    info[4] = f.turbEstablished ? 1 : 0;
    return 0;
}
}  // namespace

extern "C" {

int ref_settings_count(void) { return S_COUNT; }

// Mesh sizes of an MSH1 file: out[0..3] = nCells, nFaces, nInternalFaces, nPatches
int ref_mesh_sizes(const char* path, int* out) {
    try {
        Mesh m = Mesh::loadFromFile(path);
        out[0] = m.nCells(); out[1] = m.nFaces();
        out[2] = m.nInternalFaces(); out[3] = m.nPatches();
        return 0;
    } catch (const std::exception& e) { std::fprintf(stderr, "ref: %s\n", e.what()); return -1; }
}

int ref_wall_distance(const char* path, double* wd) {
    try {
        Mesh m = loadMesh(path);
        const auto& w = m.wallDistance();
        for (int i = 0; i < m.nCells(); ++i) wd[i] = w[i];
        return 0;
    } catch (const std::exception& e) { std::fprintf(stderr, "ref: %s\n", e.what()); return -1; }
}

// Full SIMPLE solve from a uniform initial state (SIMPLESolver::initUniform).
// init = {Ux, Uy, Uz, p, k, omega}.  Cell outputs may be null.  hist receives
// 6 doubles (Ux, Uy, Uz, p, k, omega) per iteration entry, up to histCap.
// info[0]=finalIter info[1]=converged info[2]=diverged info[3]=nEntries
int ref_simple_solve(const char* path, int bcKind, const double* bcVals, double nu,
                     const double* coeffs, const double* settings, const double* init,
                     double* U, double* p, double* k, double* om, double* nuT,
                     double* F1, double* F2, double* Pk, double* CDkw,
                     double* Ub, double* pb, double* kb, double* omb,
                     double* hist, int histCap, int* info) {
    try {
        Mesh m = loadMesh(path);
        FlowBoundaryConditions bcs = makeBCs(m, bcKind, bcVals);
        SolverSettings s = unpackSettings(settings);
        SSTModel sst(unpackCoeffs(coeffs));
        sst.variant = static_cast<SSTVariant>(s.sstVariant);
        SIMPLESolver solver(m, sst, bcs, nu, s);
        FlowFields f(m);
        solver.initUniform(f, Vec3(init[0], init[1], init[2]), init[3], init[4], init[5]);
        return solveAndExport(m, solver, f, U, p, k, om, nuT, F1, F2, Pk, CDkw, Ub, pb, kb, omb,
                              hist, histCap, info);
    } catch (const std::exception& e) { std::fprintf(stderr, "ref: %s\n", e.what()); return -1; }
}

// Warm start (ForwardModel.cpp:56-60: fields = cached.fields; solve(fields)).
// state: U[3n] Ub[3nB] p pb k kb omega omegab nuT F1 F2 Pk CDkw (13 arrays).
int ref_simple_solve_warm(const char* path, int bcKind, const double* bcVals, double nu,
                          const double* coeffs, const double* settings,
                          const double* const* state, int turbEstablished,
                          double* U, double* p, double* k, double* om, double* nuT,
                          double* F1, double* F2, double* Pk, double* CDkw,
                          double* Ub, double* pb, double* kb, double* omb,
                          double* hist, int histCap, int* info) {
    try {
        Mesh m = loadMesh(path);
        FlowBoundaryConditions bcs = makeBCs(m, bcKind, bcVals);
        SolverSettings s = unpackSettings(settings);
        SSTModel sst(unpackCoeffs(coeffs));
        sst.variant = static_cast<SSTVariant>(s.sstVariant);
        SIMPLESolver solver(m, sst, bcs, nu, s);
        FlowFields f(m);
        const int n = m.nCells(), nB = f.U.nBnd();
        for (int i = 0; i < n; ++i) f.U[i] = Vec3(state[0][3*i], state[0][3*i+1], state[0][3*i+2]);
        for (int i = 0; i < nB; ++i) f.U.bnd(i) = Vec3(state[1][3*i], state[1][3*i+1], state[1][3*i+2]);
        ScalarField* sf[] = {&f.p, &f.k, &f.omega};
        for (int q = 0; q < 3; ++q) {
            for (int i = 0; i < n; ++i) (*sf[q])[i] = state[2 + 2*q][i];
            for (int i = 0; i < nB; ++i) sf[q]->bnd(i) = state[3 + 2*q][i];
        }
        ScalarField* cl[] = {&f.nuT, &f.F1, &f.F2, &f.Pk, &f.CDkw};
        for (int q = 0; q < 5; ++q)
            for (int i = 0; i < n; ++i) (*cl[q])[i] = state[8 + q][i];
        f.turbEstablished = turbEstablished != 0;
        return solveAndExport(m, solver, f, U, p, k, om, nuT, F1, F2, Pk, CDkw, Ub, pb, kb, omb,
                              hist, histCap, info);
    } catch (const std::exception& e) { std::fprintf(stderr, "ref: %s\n", e.what()); return -1; }
}

// One linear solve with the named reference solver on an LDU system.
// solverCode: 0 AMG, 1 PCG, 2 BiCGSTAB, 3 GaussSeidel.
// out[0]=iterations out[1]=finalRes out[2]=initialRes out[3]=converged
int ref_ldu_solve(int solverCode, int n, int nif, const int* own, const int* nbr,
                  const double* diag, const double* upper, const double* lower,
                  const double* src, double* x, int maxIter, double tol, double* out) {
    try {
        LinearSystem A = makeLdu(n, nif, own, nbr, diag, upper, lower, src);
        auto solver = makeSolver(solverName(solverCode));
        std::vector<double> xv(x, x + n);
        SolverResult r = solver->solve(A, xv, maxIter, tol);
        for (int i = 0; i < n; ++i) x[i] = xv[i];
        out[0] = r.iterations; out[1] = r.finalRes; out[2] = r.initialRes; out[3] = r.converged;
        return 0;
    } catch (const std::exception& e) { std::fprintf(stderr, "ref: %s\n", e.what()); return -1; }
}

// The reference AMG hierarchy for an LDU matrix: repeats lduToCSR /
// buildAggregates / galerkinProduct with the constants of AMGSolver::setup.
// levelSizes receives nRows per level (cap maxLev); maps receives the
// concatenated aggregate maps of every non-coarsest level.  Returns nLevels.
int ref_amg_hierarchy(int n, int nif, const int* own, const int* nbr,
                      const double* diag, const double* upper, const double* lower,
                      int* levelSizes, int* levelNnz, int maxLev, int* maps, int mapsCap) {
    try {
        std::vector<double> zero(n, 0.0);
        LinearSystem A = makeLdu(n, nif, own, nbr, diag, upper, lower, zero.data());
        std::vector<CSRMatrix> lev;
        lev.push_back(lduToCSR(A));
        int off = 0;
        for (int l = 0; l < 20; ++l) {
            CSRMatrix& fine = lev[l];
            if (fine.nRows <= 50) break;
            std::vector<int> map; int nc;
            buildAggregates(fine, map, nc);
            if (nc == 0 || nc >= fine.nRows) break;
            for (int i = 0; i < fine.nRows && off < mapsCap; ++i) maps[off++] = map[i];
            CSRMatrix c = galerkinProduct(fine, map, nc);
            lev.push_back(std::move(c));
        }
        int nl = (int)lev.size();
        for (int l = 0; l < nl && l < maxLev; ++l) {
            levelSizes[l] = lev[l].nRows;
            levelNnz[l] = lev[l].rowPtr[lev[l].nRows];
        }
        return nl;
    } catch (const std::exception& e) { std::fprintf(stderr, "ref: %s\n", e.what()); return -1; }
}

// Green-Gauss gradient of a cell field phi with boundary-face values phib.
int ref_green_gauss(const char* path, const double* phi, const double* phib, double* grad) {
    try {
        Mesh m = loadMesh(path);
        ScalarField f(m, "phi");
        for (int i = 0; i < m.nCells(); ++i) f[i] = phi[i];
        for (int i = 0; i < f.nBnd(); ++i) f.bnd(i) = phib[i];
        VectorField g = greenGaussGrad(f);
        for (int i = 0; i < m.nCells(); ++i) { grad[3*i] = g[i].x; grad[3*i+1] = g[i].y; grad[3*i+2] = g[i].z; }
        return 0;
    } catch (const std::exception& e) { std::fprintf(stderr, "ref: %s\n", e.what()); return -1; }
}

// Operator blocks at a fixed state (SIMPLESolver::assemblePreconBlocks):
// momentum-x, k and omega LDU systems plus the relaxed momentum diagonal.
// The state arrays carry cell values and boundary values.  out arrays:
// mom/k/om each = diag[n] upper[nif] lower[nif] source[n] concatenated.
int ref_precon_blocks(const char* path, int bcKind, const double* bcVals, double nu,
                      const double* coeffs, const double* settings,
                      const double* U, const double* Ub, const double* p, const double* pb,
                      const double* k, const double* kb, const double* om, const double* omb,
                      double* mom, double* kout, double* omout, double* aP,
                      double* closure /* nuT F1 F2 Pk CDkw, 5n */) {
    try {
        Mesh m = loadMesh(path);
        FlowBoundaryConditions bcs = makeBCs(m, bcKind, bcVals);
        SolverSettings s = unpackSettings(settings);
        SSTCoefficients c = unpackCoeffs(coeffs);
        SSTModel sst(c);
        SIMPLESolver solver(m, sst, bcs, nu, s);
        FlowFields f(m);
        const int n = m.nCells(), nif = m.nInternalFaces();
        for (int i = 0; i < n; ++i) {
            f.U[i] = Vec3(U[3*i], U[3*i+1], U[3*i+2]);
            f.p[i] = p[i]; f.k[i] = k[i]; f.omega[i] = om[i];
        }
        for (int i = 0; i < f.U.nBnd(); ++i) {
            f.U.bnd(i) = Vec3(Ub[3*i], Ub[3*i+1], Ub[3*i+2]);
            f.p.bnd(i) = pb[i]; f.k.bnd(i) = kb[i]; f.omega.bnd(i) = omb[i];
        }
        LinearSystem Am, Ak, Ao; std::vector<double> ap;
        solver.assemblePreconBlocks(f, c, Am, Ak, Ao, ap);
        auto dump = [&](const LinearSystem& A, double* o) {
            for (int i = 0; i < n; ++i) o[i] = A.diag[i];
            for (int j = 0; j < nif; ++j) o[n + j] = A.upper[j];
            for (int j = 0; j < nif; ++j) o[n + nif + j] = A.lower[j];
            for (int i = 0; i < n; ++i) o[n + 2*nif + i] = A.source[i];
        };
        dump(Am, mom); dump(Ak, kout); dump(Ao, omout);
        for (int i = 0; i < n; ++i) aP[i] = ap[i];
        if (closure) {
            // the closure fields the blocks were built from (computeFields,
            // non-negativity clamp on nuT)
            FlowFields w = f;
            sst.computeFields(m, w.k, w.omega, w.U, nu, w.nuT, w.F1, w.F2, w.Pk, w.CDkw);
            for (int i = 0; i < n; ++i) {
                closure[i] = std::max(w.nuT[i], 0.0);
                closure[n + i] = w.F1[i]; closure[2*n + i] = w.F2[i];
                closure[3*n + i] = w.Pk[i]; closure[4*n + i] = w.CDkw[i];
            }
        }
        return 0;
    } catch (const std::exception& e) { std::fprintf(stderr, "ref: %s\n", e.what()); return -1; }
}

// Momentum residual blocks at a fixed state (SIMPLESolver::assembleResidual):
// R = [Rux | Ruy | Rk | Romega], 4n values.
int ref_residual(const char* path, int bcKind, const double* bcVals, double nu,
                 const double* coeffs, const double* settings,
                 const double* U, const double* Ub, const double* p, const double* pb,
                 const double* k, const double* kb, const double* om, const double* omb,
                 double* R) {
    try {
        Mesh m = loadMesh(path);
        FlowBoundaryConditions bcs = makeBCs(m, bcKind, bcVals);
        SolverSettings s = unpackSettings(settings);
        SSTCoefficients c = unpackCoeffs(coeffs);
        SSTModel sst(c);
        SIMPLESolver solver(m, sst, bcs, nu, s);
        FlowFields f(m);
        const int n = m.nCells();
        for (int i = 0; i < n; ++i) {
            f.U[i] = Vec3(U[3*i], U[3*i+1], U[3*i+2]);
            f.p[i] = p[i]; f.k[i] = k[i]; f.omega[i] = om[i];
        }
        for (int i = 0; i < f.U.nBnd(); ++i) {
            f.U.bnd(i) = Vec3(Ub[3*i], Ub[3*i+1], Ub[3*i+2]);
            f.p.bnd(i) = pb[i]; f.k.bnd(i) = kb[i]; f.omega.bnd(i) = omb[i];
        }
        std::vector<double> r = solver.assembleResidual(f, c);
        for (int i = 0; i < 4 * n; ++i) R[i] = r[i];
        return 0;
    } catch (const std::exception& e) { std::fprintf(stderr, "ref: %s\n", e.what()); return -1; }
}

// Writes a mesh built by one of the reference factories to MSH1.
// kind 0 makeChannel2D(nx,ny,Lx,Ly); 1 makeChannel2D(...,Re,y+);
// 2 makePlate2D(nx,ny,Lx,Ly,Re,y+); 3 makeBackwardFacingStep2D(8 args);
// 4 makeBackwardFacingStep2D(...,Re,y+);
// 5 makeCurvedChannelPeriodic2D(x[nx+1], yb[nx+1], yTop, ny, Re, y+).
// ip / dp carry the integer / real arguments in signature order (kind 5:
// ip = {nx, ny}, dp = {yTop, Re, y+, x..., yb...}).
// Mesh::loadFromFile (MSH1 or an OpenFOAM polyMesh directory) written back
// as MSH1: the reference reader's topology and geometry, byte for byte.
int ref_convert_mesh(const char* in, const char* out) {
    try {
        Mesh m = Mesh::loadFromFile(in);
        m.saveBinary(out);
        return 0;
    } catch (const std::exception& e) { std::fprintf(stderr, "ref: %s\n", e.what()); return -1; }
}

int ref_make_mesh(int kind, const int* ip, const double* dp, const char* path) {
    try {
        Mesh m;
        switch (kind) {
            case 0: m = Mesh::makeChannel2D(ip[0], ip[1], dp[0], dp[1]); break;
            case 1: m = Mesh::makeChannel2D(ip[0], ip[1], dp[0], dp[1], dp[2], dp[3]); break;
            case 2: m = Mesh::makePlate2D(ip[0], ip[1], dp[0], dp[1], dp[2], dp[3]); break;
            case 3: m = Mesh::makeBackwardFacingStep2D(ip[0], ip[1], ip[2], ip[3], dp[0], dp[1], dp[2], dp[3]); break;
            case 4: m = Mesh::makeBackwardFacingStep2D(ip[0], ip[1], ip[2], ip[3], dp[0], dp[1], dp[2], dp[3], dp[4], dp[5]); break;
            case 5: {
                const int nx = ip[0];
                std::vector<double> x(dp + 3, dp + 3 + nx + 1), yb(dp + 4 + nx, dp + 4 + 2 * nx + 1);
                m = Mesh::makeCurvedChannelPeriodic2D(x, yb, dp[0], ip[1], dp[1], dp[2]);
                break;
            }
            default: return -1;
        }
        m.saveBinary(path);
        return 0;
    } catch (const std::exception& e) { std::fprintf(stderr, "ref: %s\n", e.what()); return -1; }
}

// Transpose-stress source (StressOperators.hpp:31-84) for comp, with the
// wall-zero coefficient when wallZero != 0.
int ref_transpose_stress(const char* path, const double* coeff, const double* U, const double* Ub,
                         int comp, int wallZero, double* src) {
    try {
        Mesh m = loadMesh(path);
        ScalarField c(m, "c");
        VectorField u(m, "U");
        for (int i = 0; i < m.nCells(); ++i) { c[i] = coeff[i]; u[i] = Vec3(U[3*i], U[3*i+1], U[3*i+2]); }
        for (int i = 0; i < u.nBnd(); ++i) u.bnd(i) = Vec3(Ub[3*i], Ub[3*i+1], Ub[3*i+2]);
        std::vector<double> zero(m.nCells(), 0.0);
        std::vector<double> s = transposeStressSource(m, c, u, comp, wallZero ? &zero : nullptr);
        for (int i = 0; i < m.nCells(); ++i) src[i] = s[i];
        return 0;
    } catch (const std::exception& e) { std::fprintf(stderr, "ref: %s\n", e.what()); return -1; }
}

// ObservationOperator::evaluate and logLikelihood (ObservationOperator.hpp:
// 159-378) of the reference on the given U (cells, interleaved), U boundary
// values and p.  Per observable: type 0 Drag 1 PressureTap 2 VelocityProfile
// 3 SkinFriction 4 SeparationPoint 5 ReattachmentLength; patch index (or -1),
// location, component, reference area / velocity / pressure / density, and
// the measurement (observed, sigma, sigma_model).
int ref_observe(const char* path, int nObs, const int* type, const int* patch, const double* loc,
                const int* comp, const double* refArea, const double* refVel, const double* refP,
                const double* refRho, const double* observed, const double* sigma, const double* sigmaModel,
                const double* U, const double* Ub, const double* p, double nu, double* pred, double* loglik) {
    try {
        Mesh m = loadMesh(path);
        ObservationOperator op;
        for (int i = 0; i < nObs; ++i) {
            const Vec3 L(loc[3 * i], loc[3 * i + 1], loc[3 * i + 2]);
            const std::string name = patch[i] >= 0 ? m.patch(patch[i]).name : std::string();
            switch (type[i]) {
                case 0: op.addDrag(name, observed[i], sigma[i], refArea[i], refVel[i], sigmaModel[i]); break;
                case 1: op.addPressureTap(L, observed[i], sigma[i], refVel[i], sigmaModel[i], refP[i], refRho[i]); break;
                case 2: op.addVelocityProfile(L, comp[i], observed[i], sigma[i], sigmaModel[i]); break;
                case 3: op.addSkinFriction(name, L, observed[i], sigma[i], refVel[i], sigmaModel[i]); break;
                case 4: op.addSeparationPoint(name, observed[i], sigma[i], sigmaModel[i]); break;
                case 5: op.addReattachmentLength(name, observed[i], sigma[i], sigmaModel[i]); break;
                default: return -1;
            }
        }
        FlowFields f(m);
        for (int i = 0; i < m.nCells(); ++i) {
            f.U[i] = Vec3(U[3 * i], U[3 * i + 1], U[3 * i + 2]);
            f.p[i] = p[i];
        }
        for (int i = 0; i < f.U.nBnd(); ++i) f.U.bnd(i) = Vec3(Ub[3 * i], Ub[3 * i + 1], Ub[3 * i + 2]);
        const std::vector<double> r = op.evaluate(m, f, nu);
        for (int i = 0; i < nObs; ++i) pred[i] = r[i];
        *loglik = op.logLikelihood(m, f, nu);
        return 0;
    } catch (const std::exception& e) { std::fprintf(stderr, "ref: %s\n", e.what()); return -1; }
}

// Pointwise SST functions (SSTModel.cpp:9-118) for closed-form checks.
// which: 0 F1(k,w,y,nu,CDpos) 1 F2(k,w,y,nu) 2 eddyViscosity(k,w,S,F2)
//        3 production(nuT,S,k,w) 4 productionOmega(nuT,S,k,w)
double ref_sst_point(int which, const double* coeffs, int variant, const double* a) {
    SSTModel sst(unpackCoeffs(coeffs));
    sst.variant = static_cast<SSTVariant>(variant);
    switch (which) {
        case 0: return sst.computeF1(a[0], a[1], a[2], a[3], a[4]);
        case 1: return sst.computeF2(a[0], a[1], a[2], a[3]);
        case 2: return sst.eddyViscosity(a[0], a[1], a[2], a[3]);
        case 3: return sst.production(a[0], a[1], a[2], a[3]);
        case 4: return sst.productionOmega(a[0], a[1], a[2], a[3]);
        default: return 0.0;
    }
}

// CompressibleSIMPLESolver::solve (compressible/src/CompressibleSIMPLESolver.cpp:643-1000)
// from initUniform on a mesh file with CompressibleBoundaryConditions::channelDefaults.
// bc = {Uin, T_in, p_out, kIn, omIn}; eos = {gamma, R, Pr, Pr_T, mu_ref, T_ref, S_suth};
// init = {Ux, Uy, Uz, p, T, k, omega}; alphaT the temperature relaxation.
// Outputs (nullable): U[3n], p (mechanical), T, rho, k, omega, nuT, F1, F2, Pk [n];
// hist 6 per iteration (Ux, Uy, p, T, k, omega); info = {finalIter, converged,
// diverged, entries, turbEstablished}.
int ref_comp_solve(const char* path, const double* bc, const double* eos, const double* coeffs,
                   const double* settings, double alphaT, const double* init,
                   double* U, double* p, double* T, double* rho, double* k, double* om, double* nuT,
                   double* F1, double* F2, double* Pk, double* hist, int histCap, int* info) {
    try {
        Mesh m = loadMesh(path);
        CompressibleBoundaryConditions bcs =
            CompressibleBoundaryConditions::channelDefaults(m, bc[0], bc[1], bc[2], bc[3], bc[4]);
        IdealGasEOS e;
        e.gamma = eos[0]; e.R = eos[1]; e.Pr = eos[2]; e.Pr_T = eos[3];
        e.mu_ref = eos[4]; e.T_ref = eos[5]; e.S_suth = eos[6];
        SolverSettings s = unpackSettings(settings);
        s.alphaT = alphaT;
        SSTModel sst(unpackCoeffs(coeffs));
        sst.variant = static_cast<SSTVariant>(s.sstVariant);
        CompressibleSIMPLESolver solver(m, sst, bcs, e, s);
        CompressibleFlowFields f(m);
        solver.initUniform(f, Vec3(init[0], init[1], init[2]), init[3], init[4], init[5], init[6]);
        CompressibleConvergenceHistory h = solver.solve(f);
        const int n = m.nCells();
        if (U) for (int i = 0; i < n; ++i) { U[3*i] = f.U[i].x; U[3*i+1] = f.U[i].y; U[3*i+2] = f.U[i].z; }
        copyScalar(f.p, p, nullptr); copyScalar(f.T, T, nullptr); copyScalar(f.rho, rho, nullptr);
        copyScalar(f.k, k, nullptr); copyScalar(f.omega, om, nullptr); copyScalar(f.nuT, nuT, nullptr);
        copyScalar(f.F1, F1, nullptr); copyScalar(f.F2, F2, nullptr); copyScalar(f.Pk, Pk, nullptr);
        const int ne = (int)h.entries.size();
        if (hist) for (int q = 0; q < ne && q < histCap; ++q) {
            const auto& r = h.entries[q];
            hist[6*q] = r.Ux; hist[6*q+1] = r.Uy; hist[6*q+2] = r.p;
            hist[6*q+3] = r.T; hist[6*q+4] = r.k; hist[6*q+5] = r.omega;
        }
        info[0] = h.finalIter; info[1] = h.converged; info[2] = h.diverged; info[3] = ne;
        info[4] = f.turbEstablished ? 1 : 0;
        return 0;
    } catch (const std::exception& ex) { std::fprintf(stderr, "ref: %s\n", ex.what()); return -1; }
}

}  // extern "C"
