/* rsg.h -- C-ABI of the MI355X RANS forward-model engine (librsg.so).
 *
 * The engine runs the reference's incompressible SIMPLE + SST k-omega solve
 * (reference: incompressible/src/SIMPLESolver.cpp:696-1043) on one GPU,
 * one workgroup per solve, many solves per launch.  The pybind11 module
 * rans_sst_py (host C++) is the drop-in replacement of the reference's
 * module of the same name (python/Bindings.cpp:273-519) and reaches the GPU
 * only through the functions below.  Plain C types, caller-owned host
 * buffers, int status codes; no exceptions cross this boundary.
 *
 * Reference interfaces each group replaces:
 *   rsg_mesh_upload        Mesh as consumed by SIMPLESolver (core/include/Mesh.hpp:84-197)
 *   rsg_problem_create     SIMPLESolver(mesh, sst, bcs, nu, settings) ctor
 *                          (SIMPLESolver.cpp:10-19) minus the coefficients
 *   rsg_submit/poll/wait   SIMPLESolver::initUniform + solve as called from
 *                          ForwardModel::evaluate (incompressible/src/ForwardModel.cpp:47-66)
 *   rsg_fetch              FlowFields + ConvergenceHistory returned by solve
 *                          (core/include/FlowFields.hpp:9-49)
 *   rsg_ldu_solve          ILinearSolver::solve (core/include/LinearSolver.hpp:115-120)
 *   rsg_set_launch / rsg_set_time_slice / rsg_set_schedule / rsg_set_team,
 *   rsg_hold_admission     engine tuning (no reference counterpart: the reference
 *                          solves one member at a time)
 *   rsg_system_solve       AMGSolver::solve / PCGSolver::solve on one large system
 *                          (core/include/LinearSolver.hpp:602-653, :147-186)
 *   rsg_system_hierarchy   AMGSolver::setup (core/include/LinearSolver.hpp:516-550)
 *   rsg_amg_aggregate      buildAggregates (core/include/LinearSolver.hpp:407-443)
 *   rsg_wall_distance      Mesh::computeWallDistance (core/src/Mesh.cpp:345-373)
 *   rsg_problem_set_observations
 *                          ObservationOperator::evaluate / logLikelihood
 *                          (core/include/ObservationOperator.hpp:159-378) as called
 *                          from ForwardModel::evaluate (ForwardModel.cpp:95-97)
 *   rsg_comp_problem_create / rsg_comp_submit / rsg_comp_fetch
 *                          CompressibleSIMPLESolver ctor, initUniform + solve as called
 *                          from CompressibleForwardModel::evaluate
 *                          (compressible/src/CompressibleForwardModel.cpp:36-97)
 */
#ifndef RSG_H
#define RSG_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define RSG_OK 0
#define RSG_ERR_ARG (-1)        /* invalid argument (maps to ValueError) */
#define RSG_ERR_NOMEM (-2)      /* HBM budget or host allocation exhausted */
#define RSG_ERR_DEVICE (-3)     /* HIP runtime error */
#define RSG_ERR_STATE (-4)      /* unknown id / ticket, or wrong state */
#define RSG_ERR_TIMEOUT (-5)
#define RSG_ERR_UNSUPPORTED (-6)

/* patch types (Patch::type strings, Mesh.hpp:75-80) */
#define RSG_PATCH_WALL 0
#define RSG_PATCH_INLET 1
#define RSG_PATCH_OUTLET 2
#define RSG_PATCH_SYMMETRY 3
#define RSG_PATCH_OTHER 4

/* BC types: same numbering as BCType (core/include/BoundaryCondition.hpp:12-18) */
#define RSG_BC_DIRICHLET 0
#define RSG_BC_NEUMANN 1
#define RSG_BC_INLET_VELOCITY 2
#define RSG_BC_OUTLET_PRESSURE 3
#define RSG_BC_WALL_NOSLIP 4
#define RSG_BC_WALL_KOMEGA 5
#define RSG_BC_WALL_MOVING 6
#define RSG_BC_SYMMETRY 7
#define RSG_BC_CYCLIC 8
#define RSG_BC_KEEP 9            /* face in no patch: value never rewritten */

/* linear solver codes (makeSolver names, LinearSolver.hpp:659-665) */
#define RSG_SOLVER_AMG 0
#define RSG_SOLVER_PCG 1
#define RSG_SOLVER_BICGSTAB 2
/* GaussSeidelSolver (core/include/LinearSolver.hpp:258-299): in-place sweeps
 * in cell order, a sequential recurrence run by one GPU thread; for
 * momentum / turbulence / pressure solves and rsg_ldu_solve, always FP64. */
#define RSG_SOLVER_GAUSS_SEIDEL 3

/* precision of a member's inner linear solves.  RSG_PREC_MIXED: every linear
 * solve of the SIMPLE iteration (momentum, p', k, omega) computes the defect
 * b - A x in FP64, solves A d = r in FP32 (FP32 vectors and matrix values,
 * FP32-accumulated reductions, the AMG V-cycle in FP32 around the FP64 dense
 * coarsest inverse), updates x += d in FP64 and repeats until the FP64 defect
 * is below the inner tolerance (iterative refinement); assembly, the fields,
 * the near-wall omega and the residual accounting stay FP64. */
#define RSG_PREC_FP64 0
#define RSG_PREC_MIXED 1

/* ticket states */
#define RSG_TICKET_PENDING 0
#define RSG_TICKET_RUNNING 1
#define RSG_TICKET_DONE 2
#define RSG_TICKET_FAILED 3

typedef struct rsg_context rsg_context;

/* Mesh geometry and topology exactly as the reference stores it: internal
 * faces first [0, n_internal_faces), boundary faces after.  All arrays are
 * host memory, read during the call only. */
typedef struct rsg_mesh_desc {
    int32_t n_cells, n_faces, n_internal_faces, n_patches;
    const int32_t* owner;          /* [n_faces] */
    const int32_t* neighbor;       /* [n_internal_faces] */
    const double* face_normal;     /* [3 n_faces] unit normal, owner -> neighbour */
    const double* face_area;       /* [n_faces] */
    const double* face_delta;      /* [n_faces] |d| */
    const double* face_d;          /* [3 n_faces] owner->neighbour (owner->face centre on boundaries) */
    const double* face_weight;     /* [n_faces] (internal faces used) */
    const double* cell_volume;     /* [n_cells] */
    const double* wall_distance;   /* [n_cells] (Mesh::computeWallDistance) */
    const int32_t* patch_type;     /* [n_patches] RSG_PATCH_* */
    const int32_t* patch_face_ptr; /* [n_patches + 1] */
    const int32_t* patch_faces;    /* face ids of each patch, in patch order */
} rsg_mesh_desc;

/* Boundary conditions resolved per boundary face b = face - n_internal_faces
 * (per-patch BCType plus any per-face profile, BoundaryCondition.hpp:20-32). */
typedef struct rsg_bc_desc {
    const int32_t* vel_type; const double* vel_value;   /* [nB], [3 nB] */
    const int32_t* p_type;   const double* p_value;     /* [nB], [nB] */
    const int32_t* k_type;   const double* k_value;
    const int32_t* om_type;  const double* om_value;
} rsg_bc_desc;

/* SolverSettings (core/include/FlowFields.hpp:53-132) + engine knobs */
typedef struct rsg_settings {
    int32_t max_iterations;
    double convergence_tol, divergence_limit;
    double alpha_u, alpha_p, alpha_k, alpha_omega;
    int32_t inner_iterations;
    double inner_tolerance;
    int32_t turb_start_iter, turb_update_interval, nut_floor_iters;
    int32_t rhie_chow_all_meshes;
    double body_force[3];
    double k_min, k_max, omega_min;
    int32_t use_wall_functions;
    double von_karman;
    int32_t sst_variant;
    int32_t pressure_solver, momentum_solver, turbulence_solver;   /* RSG_SOLVER_* */
    /* engine knobs (no reference counterpart) */
    int32_t amg_refresh;      /* re-aggregate the p' hierarchy every N outer iterations (>=1) */
    int32_t coarse_refresh;   /* refresh the coarsest dense inverse every N outer iterations */
    int32_t record_history;   /* keep the per-iteration residual history */
} rsg_settings;

/* Initial state.  kind 0: uniform (SIMPLESolver::initUniform); kind 1: full
 * fields (warm start from a cached converged state, ForwardModel.cpp:56-60). */
/* OR into rsg_init.kind (rsg_submit): the solve's fields stay on the device
 * and are dropped at retirement; only status, iteration counts, history and
 * the device-evaluated predictions and log-likelihood are kept.  rsg_fetch of
 * such a ticket returns RSG_ERR_ARG unless every field pointer is NULL.  For
 * callers that need theta -> log-likelihood only (MCMC chains): no field copy
 * to the host at retirement, which the engine thread otherwise makes between
 * launches. */
#define RSG_INIT_NO_FIELDS 0x100
typedef struct rsg_init {
    int32_t kind;                 /* 0: uniform state, 1: fields (warm start); | RSG_INIT_NO_FIELDS */
    double U[3], p, k, omega;
    const double *U_cells, *p_cells, *k_cells, *om_cells;   /* [3n], [n] ... */
    const double *nuT, *F1, *F2, *Pk, *CDkw;                  /* [n] */
    const double *U_bnd, *p_bnd, *k_bnd, *om_bnd;             /* [3nB], [nB] ... */
    int32_t turb_established;
} rsg_init;

/* Result of a finished solve; every pointer is caller-owned and may be NULL. */
typedef struct rsg_result {
    int32_t status;           /* RSG_TICKET_DONE or RSG_TICKET_FAILED */
    int32_t final_iter, converged, diverged, turb_established;
    double *U, *p, *k, *omega, *nuT, *F1, *F2, *Pk, *CDkw;
    double *U_bnd, *p_bnd, *k_bnd, *om_bnd;
    double* history;          /* 6 doubles (Ux Uy Uz p k omega) per outer iteration */
    int32_t history_cap;      /* entries available in history */
    int32_t history_len;      /* entries written */
    double inner_iters[6];    /* cumulative inner iterations per equation */
    double t_submit, t_start, t_end;   /* engine clock, seconds */
    double* predictions;      /* observables evaluated on the device (NULL: not copied) */
    int32_t predictions_cap;  /* entries available in predictions */
    int32_t n_predictions;    /* observables of the problem (0: none set) */
    double log_lik;           /* Gaussian log-likelihood of the predictions */
} rsg_result;

int rsg_context_create(int32_t device, uint64_t hbm_budget_bytes, rsg_context** out);
int rsg_context_destroy(rsg_context* ctx);
const char* rsg_last_error(rsg_context* ctx);
/* engine clock (seconds) used by t_submit/t_start/t_end */
double rsg_now(rsg_context* ctx);
/* device properties: out[0] = compute units, out[1] = wavefront size */
int rsg_device_info(rsg_context* ctx, int32_t* out);

int rsg_mesh_upload(rsg_context* ctx, const rsg_mesh_desc* mesh, int32_t* mesh_id);
int rsg_mesh_release(rsg_context* ctx, int32_t mesh_id);
int rsg_problem_create(rsg_context* ctx, int32_t mesh_id, const rsg_bc_desc* bc, double nu,
                       const rsg_settings* settings, int32_t* problem_id);
int rsg_problem_release(rsg_context* ctx, int32_t problem_id);

/* Observables of a problem, resolved against its mesh by the caller
 * (Observable, core/include/ObservationOperator.hpp:20-32).  Every solve of
 * the problem admitted afterwards evaluates them on the device when it
 * finishes (converged, diverged or at the iteration cap), in the reference's
 * operation order, and rsg_fetch returns the predictions and the Gaussian
 * log-likelihood.  index: the nearest cell (pressure tap, velocity; the
 * reference's findNearestCell) or the nearest face of the patch (skin
 * friction, findNearestWallFace; -1 for an empty patch).  Drag, separation
 * and reattachment read faces[face_begin, face_end): the patch's boundary
 * faces in patch order, face_x their centres' x.  variance = sigma^2 +
 * sigma_model^2, floored at 1e-30 as the reference does.  Set once per
 * problem, before its first solve is admitted. */
#define RSG_OBS_DRAG 0
#define RSG_OBS_PRESSURE_TAP 1
#define RSG_OBS_VELOCITY 2
#define RSG_OBS_SKIN_FRICTION 3
#define RSG_OBS_SEPARATION 4
#define RSG_OBS_REATTACHMENT 5
typedef struct rsg_observation {
    int32_t type;                 /* RSG_OBS_* */
    int32_t index;                /* cell or face (see above) */
    int32_t component;            /* velocity component 0, 1, 2 */
    int32_t face_begin, face_end; /* range in faces[] (patch observables) */
    double ref_velocity, ref_area, ref_density, ref_pressure;
    double observed, sigma, sigma_model;
} rsg_observation;
int rsg_problem_set_observations(rsg_context* ctx, int32_t problem_id, int32_t n, const rsg_observation* obs,
                                 int32_t n_faces, const int32_t* faces, const double* face_x);

/* Enqueue one solve; returns immediately.  coeffs: the 11 SST coefficients
 * in InferenceParameterSet order (core/include/InferenceParameters.hpp:12-14). */
int rsg_submit(rsg_context* ctx, int32_t problem_id, const double* coeffs, const rsg_init* init,
               int32_t precision, int64_t* ticket);

/* ---- Compressible pressure-based SIMPLE (ideal gas + energy equation) ----
 * Replaces CompressibleSIMPLESolver (compressible/include/
 * CompressibleSIMPLESolver.hpp:16-115, compressible/src/
 * CompressibleSIMPLESolver.cpp:10-1000) behind CompressibleForwardModel::
 * evaluate (compressible/src/CompressibleForwardModel.cpp:36-97).
 * Temperature BCs per boundary face (TBCType, CompressibleBCs.hpp:10-25:
 * Dirichlet; Neumann and AdiabaticWall both zero-gradient; KEEP for a face in
 * no patch), the gas of IdealGasEOS (IdealGasEOS.hpp:6-15) and alphaT
 * (SolverSettings::alphaT, core/include/FlowFields.hpp:62).  Velocity,
 * pressure, k and omega BCs use rsg_bc_desc as for the incompressible solve;
 * the outlet's prescribed pressure is thermodynamic.  Compressible solves are
 * FP64.  On meshes of up to 4,096 cells each runs as one workgroup of the
 * compressible member kernel (like the incompressible ensemble: persistent
 * time-sliced launches, continuous batching); larger meshes, or meshes of at
 * least rsg_set_grid's min_cells, run kernel-per-phase, at most four at once. */
#define RSG_TBC_DIRICHLET 0
#define RSG_TBC_ZERO_GRADIENT 1
#define RSG_TBC_KEEP 2
typedef struct rsg_comp_desc {
    const int32_t* t_type;    /* [nB] RSG_TBC_* */
    const double* t_value;    /* [nB] Dirichlet temperature */
    double gamma, R, Pr, Pr_T, mu_ref, T_ref, S_suth;
    double alpha_t;           /* temperature under-relaxation, (0, 1] */
} rsg_comp_desc;
/* Uniform initial state of CompressibleSIMPLESolver::initUniform
 * (CompressibleSIMPLESolver.cpp:27-72): p is the THERMODYNAMIC pressure. */
typedef struct rsg_comp_init {
    double U[3], p, T, k, omega;
} rsg_comp_init;
int rsg_comp_problem_create(rsg_context* ctx, int32_t mesh_id, const rsg_bc_desc* bc, const rsg_comp_desc* comp,
                            const rsg_settings* settings, int32_t* problem_id);
/* Enqueue one compressible solve (FP64).  rsg_fetch then returns U, k,
 * omega, nuT, F1, F2, Pk, CDkw, the boundary values, the observables
 * (evaluated with nu = mu(T_0) / rho_0, CompressibleForwardModel.cpp:88-94)
 * and p as the THERMODYNAMIC pressure p_mech - (2/3) rho max(k, 0) (boundary
 * values mechanical); history entries are (Ux, Uy, p, T, k, omega) per outer
 * iteration (CompressibleResidualEntry); final_iter counts outer iterations
 * as CompressibleConvergenceHistory::finalIter. */
int rsg_comp_submit(rsg_context* ctx, int32_t problem_id, const double* coeffs, const rsg_comp_init* init,
                    int64_t* ticket);
/* A finished compressible solve's temperature, density and mechanical
 * pressure (cells), temperature boundary values and energy-equation inner
 * iterations; any pointer may be NULL.  Call before rsg_release. */
int rsg_comp_fetch(rsg_context* ctx, int64_t ticket, double* T, double* rho, double* p_mech, double* T_bnd,
                   double* energy_iters);

int rsg_poll(rsg_context* ctx, const int64_t* tickets, int32_t n, int32_t* state_out);
/* Blocks until every ticket is DONE/FAILED or timeout_s elapses (<0: forever). */
int rsg_wait(rsg_context* ctx, const int64_t* tickets, int32_t n, double timeout_s);
int rsg_fetch(rsg_context* ctx, int64_t ticket, rsg_result* out);
int rsg_release(rsg_context* ctx, int64_t ticket);
/* Engine tuning: outer iterations per persistent launch and workgroup size
 * (256: two members per CU; 512: one member per CU, lower latency and, as
 * measured, higher ensemble throughput; 0, the default: 512). */
int rsg_set_launch(rsg_context* ctx, int32_t iters_per_launch, int32_t block_size);
/* Time-sliced launches: with microseconds > 0 every launch runs each active
 * member for that long of device wall clock (the outer iteration in progress
 * completes), instead of a fixed iters_per_launch, so small members are not
 * held to the pace of the largest one in the launch.  0 (the default) returns
 * to the fixed iteration budget.  Results do not depend on the setting. */
int rsg_set_time_slice(rsg_context* ctx, int32_t microseconds);
/* Which active members a launch runs when more are active than the device
 * holds at once: RSG_SCHEDULE_ROUND_ROBIN (default) rotates a window over the
 * active members in admission order; RSG_SCHEDULE_LARGEST_FIRST runs the
 * members on the largest meshes first (longest-processing-time order, which
 * shortens an ensemble's makespan; smaller members wait).  Results do not
 * depend on the policy. */
/* Team solves: a member whose mesh has at least min_cells cells (0: never)
 * runs alone on `blocks` 512-thread workgroups of one cooperative launch
 * (0: 16 sqrt(cells / 40000) workgroups, at least 16, at most what the device holds
 * at once and 256), iters_per_launch outer iterations per launch -- the same
 * SIMPLE code with its loops spread over the whole team, phase boundaries as
 * grid-wide barriers and reductions combined in one fixed order.  For meshes far beyond one workgroup's reach
 * (10^4-10^6 cells).  Results differ from the one-workgroup solve only by
 * reduction-order round-off; time slices do not apply to team solves. */
int rsg_set_team(rsg_context* ctx, int32_t min_cells, int32_t blocks, int32_t iters_per_launch);

/* Kernel-per-phase solves: a member whose mesh has at least min_cells cells
 * (0: never; checked when the member is admitted) runs alone on the whole GPU,
 * every phase of its outer iteration a grid-wide launch (assembly, gradients
 * and closure one thread per cell; BiCGSTAB / PCG / AMG-PCG on the CSR image
 * of each system with block-partial dot products; the p' AMG hierarchy
 * aggregated on the host every amg_refresh outer iterations, its coarse
 * Galerkin values refreshed on the device every coarse_refresh), at most
 * iters_per_launch outer iterations per engine round.  For meshes of 10^5
 * cells and more; it takes precedence over team solves.  A member submitted
 * with RSG_PREC_MIXED runs its momentum, p' and k solves as FP32 inner solves
 * (FP32 CSR images, FP32 hierarchy, FP64-accumulated dot products) under FP64
 * defects and refinement; its omega solve and the coarsest-level inverse stay
 * FP64. */
int rsg_set_grid(rsg_context* ctx, int32_t min_cells, int32_t iters_per_launch);
/* p' AMG hierarchy of kernel-per-phase members: max_levels > 1 (default 5)
 * stops coarsening at that many levels once the coarsest level has at most
 * 16,384 rows, and that level is inverted on the FP64 matrix cores at every
 * refresh of the coarse values (rsg_dense_inverse's routine) and applied as
 * one dense product per V-cycle; 0 coarsens as the reference does (at most 20
 * levels, down to 50 rows).  Taken by members admitted afterwards.  The
 * reference's AMGSolver (core/include/LinearSolver.hpp:508-512, 516-551)
 * coarsens to 50 rows and sweeps the coarsest level 50 times. */
int rsg_set_grid_levels(rsg_context* ctx, int32_t max_levels);
/* hipGraph replay of AMG V-cycles on the whole-GPU paths (rsg_system_solve and
 * kernel-per-phase members): enable != 0 (the default) captures each V-cycle's
 * launch sequence once per hierarchy and replays it with one graph launch;
 * 0 launches every phase directly.  Results are identical either way. */
int rsg_set_graphs(rsg_context* ctx, int32_t enable);

#define RSG_SCHEDULE_ROUND_ROBIN 0
#define RSG_SCHEDULE_LARGEST_FIRST 1
int rsg_set_schedule(rsg_context* ctx, int32_t policy);
/* Batch submission: hold = 1 stops the engine from admitting pending tickets
 * (members already running go on); hold = 0 releases one hold, and once none
 * is left every pending ticket is admitted together at the next launch
 * boundary.  Holds nest (one release per hold; releasing with none held is
 * RSG_ERR_STATE).  Without it the engine starts a launch with the first
 * ticket submitted, and tickets submitted while that launch runs wait for the
 * next one (up to one time slice).  A caller that submits many solves at
 * once holds, submits them all, then releases; results do not depend on it. */
int rsg_hold_admission(rsg_context* ctx, int32_t hold);
/* Engine counters: out[0] launches, out[1] members admitted, out[2] members retired,
 * out[3] kernel milliseconds (HIP events), out[4] max concurrent members. */
int rsg_stats(rsg_context* ctx, double* out);

/* One LDU linear solve on the device (x in/out, host buffers).
 * stats = {iterations, finalRes, initialRes, converged}. */
int rsg_ldu_solve(rsg_context* ctx, int32_t n, int32_t nif, const int32_t* own, const int32_t* nbr,
                  const double* diag, const double* upper, const double* lower, const double* b,
                  double* x, int32_t solver, int32_t precision, int32_t max_iter, double tol,
                  double* stats);

/* rsg_ldu_solve without any device allocation: the same solve (same
 * arguments, kernels and results) with its device workspace in a fixed
 * 8 MiB static device array instead of the context's arena.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc / hipFree.  Systems whose workspace
 * (mesh image + Krylov vectors + AMG hierarchy, ~1.55 KB per cell on quad
 * meshes) exceeds 8 MiB -- about 5,400 cells -- are rejected with
 * RSG_ERR_ARG and a message naming the bytes needed.  Every call on one GPU
 * uses that GPU's one array (its contents persist from the last call and are
 * rewritten by the next): overlapping calls on one GPU would race, so the
 * library serialises them with a lock per device (the array is per process,
 * so other processes have their own). */
int rsg_ldu_solve_noalloc(rsg_context* ctx, int32_t n, int32_t nif, const int32_t* own, const int32_t* nbr,
                          const double* diag, const double* upper, const double* lower, const double* b,
                          double* x, int32_t solver, int32_t precision, int32_t max_iter, double tol,
                          double* stats);

/* One large LDU linear solve spread over the whole GPU (the stand-alone
 * pressure-type system): PCG from x (in/out, host buffers) preconditioned by
 * the AMG V-cycle (solver RSG_SOLVER_AMG, hierarchy built on the host with the
 * reference's aggregation) or Jacobi (RSG_SOLVER_PCG); tol is relative to
 * ||b - A x0||.  precision RSG_PREC_FP64: FP64 throughout; RSG_PREC_MIXED: FP64
 * defect and update around FP32 inner PCG solves, each asked for the
 * reduction still missing (0.3 margin), at most 1e-5 relative.
 * Replaces AMGSolver::solve / PCGSolver::solve (core/include/LinearSolver.hpp:
 * 602-653, 147-186) on systems too large for one workgroup.
 * A coarsest level above 64 rows (where the greedy aggregation stalls) of at
 * most 16,384 rows is inverted on the FP64 matrix cores once per call and
 * applied as one dense product per V-cycle (the reference: 50 Gauss-Seidel
 * sweeps, LinearSolver.hpp:508-512, 557-562); if a pivot of that inversion is
 * not positive the level is smoothed by 51 Jacobi sweeps instead.
 * stats (10 doubles) = {iterations (inner, summed), finalRes, initialRes,
 *          converged, levels, refinement sweeps, setup seconds, solve seconds,
 *          rows of the coarsest level inverted on the device (0: none),
 *          seconds of that inversion}. */
int rsg_system_solve(rsg_context* ctx, int32_t n, int32_t nif, const int32_t* own, const int32_t* nbr,
                     const double* diag, const double* upper, const double* lower, const double* b, double* x,
                     int32_t solver, int32_t precision, int32_t max_iter, double tol, double* stats);

/* The explicit inverse of one dense SPD matrix (n x n, row-major, host
 * buffers, n <= 16,384) by the device routine that inverts large AMG coarsest
 * levels: blocked Gauss-Jordan (block 64, no pivoting) whose block products run
 * on the FP64 matrix cores.  Stands where the reference solves its coarsest
 * level by Gauss-Seidel sweeps (core/include/LinearSolver.hpp:508-512,
 * 557-562).  RSG_ERR_ARG if a pivot is not positive (the matrix is not SPD).
 * kernel_seconds (nullable) gets the device time of the inversion. */
int rsg_dense_inverse(rsg_context* ctx, int32_t n, const double* a, double* inv, double* kernel_seconds);

/* Host-only: the AMG hierarchy rsg_system_solve builds (AMGSolver::setup,
 * core/include/LinearSolver.hpp:516-550).  level_sizes/level_nnz get up to
 * max_levels entries; maps gets the concatenated fine->aggregate maps of every
 * level but the coarsest (up to maps_cap entries).  Returns the level count,
 * or RSG_ERR_ARG. */
int rsg_system_hierarchy(int32_t n, int32_t nif, const int32_t* own, const int32_t* nbr, const double* diag,
                         const double* upper, const double* lower, int32_t* level_sizes, int32_t* level_nnz,
                         int32_t max_levels, int32_t* maps, int64_t maps_cap);

/* Greedy aggregation of one CSR matrix (host, bit-exact to the reference).
 * Returns the number of aggregates, or a negative status. */
int rsg_amg_aggregate(int32_t n, const int32_t* row_ptr, const int32_t* col_idx,
                      const double* values, const double* diag, int32_t* map);

/* Wall distance of every cell on the device: for each of the n_cells cell
 * centres (cell_centers, 3 doubles x,y,z per cell) the minimum Euclidean
 * distance to the n_wall wall-face centres (wall_points, 3 doubles per face),
 * into wall_dist[n_cells]; with no wall face every distance is 1e10.
 * Bit-identical to Mesh::computeWallDistance (core/src/Mesh.cpp:345-373).
 * Host buffers.  kernel_seconds (nullable) gets the kernel's duration. */
int rsg_wall_distance(rsg_context* ctx, int32_t n_cells, const double* cell_centers, int32_t n_wall,
                      const double* wall_points, double* wall_dist, double* kernel_seconds);

/* Measurement utility (no reference counterpart): one plain device copy of
 * `bytes` bytes (a multiple of 16), read once and written once, on the engine
 * stream; kernel_seconds (nullable) gets its duration.  Run under hardware
 * byte counters it gives the scale between what FETCH_SIZE / WRITE_SIZE
 * report and the bytes really moved. */
int rsg_copy_calibration(rsg_context* ctx, uint64_t bytes, double* kernel_seconds);

#ifdef __cplusplus
}
#endif
#endif
