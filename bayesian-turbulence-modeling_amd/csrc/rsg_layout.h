// rsg_layout.h -- device-resident data layout of the RANS forward solve.
//
// Plain structs of device pointers shared by the HIP kernels (rsg_member.h),
// the C-ABI implementation (rsg_capi.hip) and the sequential emulation build
// used by the CPU test suite.  Everything is structure-of-arrays, FP64 for
// the fields, the assembled systems and the outer defect, int32 for topology;
// the FP32 inner solves of mixed-precision members view scratch vectors
// (Krylov and AMG level vectors) as float arrays of twice the length.
//
// One mesh = one RsgMeshDev (read-only, shared by every member that solves on
// it).  One solve in flight = one RsgMemberDev (its fields, LDU system,
// Krylov vectors and AMG hierarchy), carved out of the context's HBM arena.
#pragma once
#include <stdint.h>

#define RSG_MAX_LEVELS 20     // AMGSolver::maxLevels_ (LinearSolver.hpp:508)
#define RSG_MIN_COARSE 50     // AMGSolver::minCoarseSize_ (LinearSolver.hpp:509)
#define RSG_COARSE_DENSE 64   // coarsest levels up to this size are inverted densely
#define RSG_DENSE_MAX 16384   // larger coarsest levels up to this size: MFMA inverse on the device (2 GiB)
#define RSG_ELL_K 4           // internal faces per cell in the cell-major (ELL) face slots
// Resident solves: at most RSG_RES_MAXN cells per member (LDS exchange
// buffers) and RSG_RES_MAXC coarse AMG rows (LDS coarse-level x).
#ifdef RSG_EMULATE
#define RSG_RES_MAXN 2048
#define RSG_RES_MAXC 2048
#else
#define RSG_RES_MAXN 1536
#define RSG_RES_MAXC 1536
#endif
// LDS copies of the coarse-level ELL images in the one-member-per-CU kernel
#define RSG_RES_CE 7680       // matrix entries (FP64 value + 16-bit column)
#define RSG_RES_CA 3072       // aggregate entries (16-bit)

// boundary-face flag bits (from the face's patch type)
#define RSG_BF_WALL 1
#define RSG_BF_OUTLET 2

// BC type codes, same numbering as BCType (BoundaryCondition.hpp:12-18)
enum {
    RSGBC_DIRICHLET = 0, RSGBC_NEUMANN = 1, RSGBC_INLET_VELOCITY = 2,
    RSGBC_OUTLET_PRESSURE = 3, RSGBC_WALL_NOSLIP = 4, RSGBC_WALL_KOMEGA = 5,
    RSGBC_WALL_MOVING = 6, RSGBC_SYMMETRY = 7, RSGBC_CYCLIC = 8,
    RSGBC_KEEP = 9   // boundary face in no patch: value never rewritten
};

// member status codes
enum {
    RSGM_EMPTY = 0, RSGM_RUNNING = 1, RSGM_CONVERGED = 2, RSGM_DIVERGED = 3,
    RSGM_MAXITER = 4, RSGM_FAILED = 5
};

struct RsgMeshDev {
    int nC, nIF, nF, nB;
    int hasOutlet;           // any non-empty outlet patch (SIMPLESolver.cpp:337-340)
    int nnz0;                // nC + 2 nIF: level-0 CSR entries
    // faces
    const int* own;          // [nF]
    const int* nbr;          // [nIF]
    const double* w;         // [nIF] interpolation weight
    const double* dlt;       // [nF]  max(delta, 1e-20)
    const double* area;      // [nF]
    const double* nx; const double* ny; const double* nz;    // [nF] unit normal
    const double* sfx; const double* sfy; const double* sfz; // [nF] normal * area
    const double* ex; const double* ey; const double* ez;    // [nIF] d / max(delta, 1e-30)
    // cells
    const double* vol;       // [nC]
    const double* wd;        // [nC] wall distance
    const double* pinY;      // [nC] wall-cell pin distance (0: not a wall cell)
    const int* cfPtr;        // [nC+1]
    const int* cfIdx;        // faces of each cell, ascending face index; (f << 1) | isNeighbour
    // boundary faces, indexed b = f - nIF
    const int* bFlags;       // RSG_BF_* bits
    // level-0 CSR pattern of the LDU matrix (rows sorted as lduToCSR does)
    const int* rp0;          // [nC+1]
    const int* ci0;          // [nnz0]
    const int* src0;         // [nnz0]: <0 -> diag[-1-v]; <nIF -> upper[v]; else lower[v-nIF]
    // cell-major ELL image of the off-diagonal pattern: the k-th internal
    // face of cell c (ascending face order) at [k * nC + c]
    // Cell-major face lists, in the order of the cell-face list: the first
    // RSG_ELL_K internal faces of cell c in ELL slots [k * nC + c], further
    // internal faces in an overflow list, then the boundary faces.
    int ellK;                // 1 when no cell overflows the ELL slots (resident solvers apply)
    const int* ellCnt;       // [nC] internal faces in the ELL slots
    const int* ellNb;        // [RSG_ELL_K * nC] neighbouring cell (padding: c)
    const int* ellFace;      // [RSG_ELL_K * nC] (f << 1) | isNeighbour (padding: -1)
    const int* ovPtr;        // [nC+1] overflow internal faces
    const int* ovFace;       // [ovPtr[nC]] (f << 1) | isNeighbour
    const int* ovNb;         // [ovPtr[nC]] neighbouring cell
    const int* bPtr;         // [nC+1] boundary faces of the cell
    const int* bFace;        // [bPtr[nC]] boundary face index f >= nIF, ascending
};

// per-problem boundary conditions, resolved per boundary face
struct RsgBcDev {
    const int* velT; const double* velV;   // [nB], [3 nB]
    const int* pT;   const double* pV;
    const int* kT;   const double* kV;
    const int* omT;  const double* omV;
};

// SolverSettings (FlowFields.hpp:53-132) plus this engine's own knobs
struct RsgSettingsDev {
    int maxIter; double tol, divLimit;
    double aU, aP, aK, aOm;
    int innerIter; double innerTol;
    int turbStart, turbInterval, nutFloorIters, rcAll;
    double bf[3];
    double kMin, kMax, omMin;
    int wallFn; double vonK;
    int variant;
    int pSolver, mSolver, tSolver;   // 0 AMG 1 PCG 2 BiCGSTAB 3 GS(->BiCGSTAB on device)
    int amgRefresh;                  // rebuild the aggregation every N outer iterations
    int coarseRefresh;               // refresh the coarsest dense inverse every N iterations
    int recordHistory;
};

struct RsgScalars {
    int iter;              // next outer iteration index
    int status;            // RSGM_*
    int turbEstablished;
    int initialized;
    int finalIter;
    int amgBuiltIter;      // iteration the hierarchy was last aggregated (-1: never)
    int nLevels;
    int coarseDense;       // coarsest level has a valid dense inverse
    double normUx0, normUy0, normP0;
    double lastK, lastOm;
    double innerIters[6];  // cumulative inner iterations: Ux Uy Uz p k omega
    int histLen;
    int coarseIter;        // iteration of the last coarsest-level inverse
};

struct RsgAmgDev {
    int rowCap, nnzCap;    // capacities of the concatenated arrays
    int* lvN;              // [RSG_MAX_LEVELS+1] rows
    int* lvRow;            // [..] offset of level rows in vectors/maps
    int* lvRp;             // [..] offset of level row pointers (lvRow[l] + l)
    int* lvNz;             // [..] offset of level entries
    int* rp;               // concatenated row pointers
    int* ci;               // concatenated columns
    double* v;             // concatenated values
    double* dg;            // concatenated diagonals
    int* map;              // fine row -> aggregate, at lvRow[l] for l < L-1
    int* aggPtr;           // aggregate -> fine rows (restriction order), at lvRp[l+1]
    int* aggIdx;           // at lvRow[l]
    int* gPtr;             // coarse entry -> fine entries (Galerkin order), at lvNz[l+1] + (l+1)
    int* gIdx;             // fine entry indices, at lvNz[l]
    double* x; double* b; double* r;   // level vectors at lvRow[l]
    double* cinv;          // [RSG_COARSE_DENSE^2] dense inverse of the coarsest level
    int* scr;              // serial-setup scratch [3 * rowCap + 2 * nnzCap]
    // ELL images of the coarse levels (l >= 1) for the resident V-cycle:
    // row i's k-th entry (CSR order) at lvE[l] + k * n_l + i (column -1 pads),
    // and aggregate I's q-th fine row (restriction order) at lvAE[l] + q * n_l + I.
    // lvK[l] / lvAK[l] = 0: that level has no image (capacity) and uses CSR.
    int eCap, aCap;
    int* lvK; int* lvE; int* lvAK; int* lvAE;   // [RSG_MAX_LEVELS+1]
    int* used;             // [2]: entries of eCol/eVal and of aCol in use
    int* eCol;             // [eCap]
    double* eVal;          // [eCap]
    int* aCol;             // [aCap]
};

// Observation operator of one problem (ObservationOperator.hpp:20-32),
// resolved against the mesh on the host: the nearest cell / wall face of each
// point observable and the patch face lists (with face-centre x) of the
// patch observables.  Codes as RSG_OBS_* in rsg.h.
enum { RSGO_DRAG = 0, RSGO_PRESSURE = 1, RSGO_VELOCITY = 2, RSGO_SKIN_FRICTION = 3, RSGO_SEPARATION = 4,
       RSGO_REATTACHMENT = 5 };

struct RsgObsDev {
    int n;
    const int* type;       // [n]
    const int* idx;        // [n] cell (pressure, velocity), face (skin friction; -1: none)
    const int* comp;       // [n] velocity component
    const int* fb;         // [n] first entry of the observable's faces (drag, separation, reattachment)
    const int* fe;         // [n] one past the last
    const double* par;     // [6 n] refVelocity, referenceArea, refDensity, refPressure, observed, variance
    const int* faces;      // patch faces, patch order (boundary faces)
    const double* fx;      // face-centre x of each listed face
};

// Compressible pressure-based SIMPLE (compressible/src/CompressibleSIMPLESolver.cpp):
// the ideal-gas problem data (IdealGasEOS.hpp:6-38, the temperature BCs of
// CompressibleBCs.hpp:10-25 resolved per boundary face) and the fields the
// incompressible member struct does not carry.  p in RsgMemberDev is the
// MECHANICAL pressure p_thermo + (2/3) rho k (the solver's working variable).
enum { RSGT_DIRICHLET = 0, RSGT_ZERO_GRADIENT = 1 };
struct RsgCompDev {
    const int* tT;         // [nB] RSGT_*
    const double* tV;      // [nB] Dirichlet temperature
    double gamma, R, Pr, PrT, muRef, Tref, S;
    double alphaT;
    double Tinit;          // uniform initial temperature
    double rho0, mu0;      // initial density and viscosity (host-evaluated EOS)
    double *T, *Tb, *rho, *mu, *Told;
    double *Ux0, *Uy0, *Uz0;   // velocity at the start of the outer iteration
    double *mdu;           // muEff div U
    double *gt[3];         // grad (muEff div U)
    double *zeroB;         // [nB] zeros (boundary values of muEff div U)
    double *pMech;         // mechanical pressure of a finished solve (p then holds the thermodynamic one)
    double energyIters;    // energy-equation inner iterations so far (member kernel)
};

struct RsgMemberDev {
    const RsgMeshDev* mesh;
    const RsgBcDev* bc;
    const RsgSettingsDev* set;
    double nu;
    double coeffs[11];     // SST coefficients, InferenceParameters.hpp:12-14 order
    double init[6];        // uniform initial state Ux Uy Uz p k omega
    int initFromFields;    // state already uploaded (warm start)
    int planarZ;           // Uz stays exactly zero: its momentum solve is skipped
    int prec;              // inner linear solves: 0 FP64, 1 FP32 with FP64 defect and refinement
    int iterBudget;        // outer iterations allowed in this launch
    long long sliceTicks;  // wall-clock ticks this launch may run the member (0: no time limit)
    // fields (cell) and boundary values
    double *Ux, *Uy, *Uz, *p, *k, *om, *nuT, *F1, *F2, *Pk, *CDkw, *Smag;
    double *Ubx, *Uby, *Ubz, *pb, *kb, *omb;
    // per-cell scratch
    double *aPx, *aPc, *pp, *ppb, *kOld, *omOld;
    double *g[15];         // grad Ux, Uy, Uz, k, omega (x, y, z each)
    double *gp[3];         // grad p (and grad p')
    // LDU system in use
    double *diag, *upper, *lower, *src;
    double *kv[8];         // Krylov vectors
    RsgAmgDev amg;
    double* hist;          // [6 * histCap] (optional)
    int histCap;
    RsgScalars* sc;
    const RsgObsDev* obs;  // observables of the problem (null: none)
    double* pred;          // [obs->n + 1]: predictions, then the log-likelihood
    RsgCompDev* comp;      // compressible member on the member kernel (null: incompressible)
};
