// rsg_member.h -- the SIMPLE + SST k-omega outer iteration of ONE ensemble
// member, executed by ONE workgroup.
//
// Design (MI355X): a member's mesh is small enough (10^2..10^4 cells) that a
// grid-wide launch per phase would be launch- and sync-bound, so a whole
// workgroup owns a member and runs its outer iterations back to back inside a
// persistent launch; an ensemble is a grid of such workgroups.  Every phase is
// a strided loop over cells or faces separated by workgroup barriers, every
// reduction is a fixed-order block reduction (so a member's result never
// depends on which other members share the launch), and every per-cell
// accumulation is a gather over the cell's faces in ascending face order --
// the order in which the reference's face-loop scatter adds contributions --
// so assembly and Green-Gauss gradients round exactly like the reference.
//
// The same source also compiles as plain C++ with RSG_EMULATE defined: one
// "thread" per block, barriers become no-ops and reductions are sequential.
// That build exists only for the CPU test suite (tests/emu); the product
// always runs the HIP build.
//
// Reference being restated (file:line in the reference repository):
//   SIMPLESolver::solve            incompressible/src/SIMPLESolver.cpp:696-1043
//   assembleMomentum               SIMPLESolver.cpp:56-203
//   assemblePressureCorrection     SIMPLESolver.cpp:322-450
//   correctVelocity/Pressure       SIMPLESolver.cpp:456-475
//   assembleKEquation              SIMPLESolver.cpp:483-560
//   assembleOmegaEquation          SIMPLESolver.cpp:567-669
//   initUniform                    SIMPLESolver.cpp:23-51
//   SSTModel::computeFields        core/src/SSTModel.cpp:122-172 (+ :9-96)
//   greenGaussGrad, |S|            core/include/Field.hpp:122-152, :254-267
//   transposeStressSource          core/include/StressOperators.hpp:31-84
//   apply*BC                       core/include/BoundaryCondition.hpp:196-300
//   BiCGSTAB / PCG / AMG-PCG       core/include/LinearSolver.hpp:147-247, :497-656
//   Gauss-Seidel                   core/include/LinearSolver.hpp:258-299
#pragma once
#include "rsg.h"
#include "rsg_layout.h"

#ifdef RSG_EMULATE
#include <math.h>
#define RSG_D static inline
#define RSG_HD static inline
#define RSG_DM inline
#define RSG_SYNC() ((void)0)
#define RSG_BLOCK_SYNC() ((void)0)
#else
#include <hip/hip_runtime.h>
#define RSG_D __device__ static __attribute__((always_inline)) inline
#define RSG_HD __host__ __device__ static inline
#define RSG_DM __device__ __attribute__((always_inline)) inline   // member functions
#ifdef RSG_TEAM
// Team build: one solve spread over every workgroup of a cooperative launch;
// a phase boundary is a grid-wide barrier.
#include <hip/hip_cooperative_groups.h>
#define RSG_SYNC() cooperative_groups::this_grid().sync()
#else
#define RSG_SYNC() __syncthreads()
#endif
#define RSG_BLOCK_SYNC() __syncthreads()
#endif

// Section timers (built only with RSG_PROF): thread 0 of every member adds the
// wall-clock ticks (100 MHz) between consecutive marks to rsg_prof_ticks[slot].
#if defined(RSG_PROF) && !defined(RSG_EMULATE)
__device__ unsigned long long rsg_prof_ticks[32];
__device__ unsigned long long rsg_prof_wg[3 * 4096];   // per workgroup: start, end, CU id
#define RSG_TMARK(slot)                                                          \
    do {                                                                         \
        __syncthreads();                                                         \
        if (C.t.tid == 0) {                                                      \
            const unsigned long long now_ = wall_clock64();                      \
            atomicAdd(&rsg_prof_ticks[slot], now_ - C.tmark);                    \
            C.tmark = now_;                                                      \
        }                                                                        \
    } while (0)
#else
#define RSG_TMARK(slot) ((void)0)
#endif

namespace rsg {

// ------------------------------------------------------------ primitives
RSG_D double dmax(double a, double b) { return (a < b) ? b : a; }   // std::max
RSG_D double dmin(double a, double b) { return (b < a) ? b : a; }   // std::min

struct Thr {
    int tid;        // thread index in the workgroup
    int bs;         // workgroup size
    double* red;    // LDS reduction scratch: 2 phases x 16 waves x 4 values
    double* lds;    // LDS tile for the coarsest-level dense inverse [64 x 64]
    int phase;
    // resident solves (null xb0: not available)
    double *xb0, *xb1;   // LDS exchange buffers [RSG_RES_MAXN] each
    int xsel;
    double* cvec;        // LDS coarse-level x, cvecCap rows
    int cvecCap;
    // LDS copies of the coarse ELL images (null: read them from global memory)
    double* ceVal; short* ceCol; int ceCap;
    short* caCol; int caCap;
    // team build: tid/bs above span the whole team; these are the thread's
    // place in its own workgroup and the workgroup's place in the team
    int ltid, lbs, blk, nblk;
    double* gred;        // global reduction partials [2 phases][RSG_TEAM_MAXB * 8]
};

#define RSG_TEAM_MAXB 256     // workgroups in a team at most

// Block reductions of NV values of type T (double, or float for the FP32
// inner solves of mixed-precision members: FP32 accumulation throughout).
#ifdef RSG_EMULATE
template <int NV, class T> RSG_D void bsum(Thr&, T (&)[NV]) {}
template <int NV, class T> RSG_D void bmax(Thr&, T (&)[NV]) {}
RSG_D double wave_bcast0(double v) { return v; }
#else
template <class T> RSG_D T wsum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
RSG_D double wmax(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = dmax(v, __shfl_xor(v, off, 64));
    return v;
}
#ifndef RSG_TEAM
// Block sum of NV values, result identical in every thread: lane 0 of each
// wave deposits its wave total, then every thread adds the wave totals in
// wave order.  Double-buffered LDS so consecutive calls need one barrier.
template <int NV, class T> RSG_D void bsum(Thr& t, T (&v)[NV]) {
    static_assert(NV <= 4, "block reductions carry at most 4 values");
    const int lane = t.tid & 63, wid = t.tid >> 6, nw = (t.bs + 63) >> 6;
    T* buf = reinterpret_cast<T*>(t.red + t.phase * 64);
    t.phase ^= 1;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = wsum(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) buf[wid * NV + i] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        T s = T(0);
        for (int w = 0; w < nw; ++w) s += buf[w * NV + i];
        v[i] = s;
    }
}
template <int NV> RSG_D void bmax(Thr& t, double (&v)[NV]) {
    const int lane = t.tid & 63, wid = t.tid >> 6, nw = (t.bs + 63) >> 6;
    double* buf = t.red + t.phase * 64;
    t.phase ^= 1;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = wmax(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) buf[wid * NV + i] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double s = buf[i];
        for (int w = 1; w < nw; ++w) s = dmax(s, buf[w * NV + i]);
        v[i] = s;
    }
}

#else
// Team sum / max: each workgroup reduces as above, its thread 0 stores
// the workgroup's value, and after the grid barrier every wave combines the
// workgroup values in one fixed order -- lane l takes workgroups l, l + 64,
// ... in turn, then the xor butterfly across the lanes -- and takes lane 0's
// result, so every thread of the team holds the same value (NaNs included).
// Double-buffered like the workgroup version.
template <int NV, class T> RSG_D void bsum(Thr& t, T (&v)[NV]) {
    static_assert(NV <= 4, "team reductions carry at most 4 values");
    const int lane = t.ltid & 63, wid = t.ltid >> 6, nw = (t.lbs + 63) >> 6;
    T* buf = reinterpret_cast<T*>(t.red + t.phase * 64);
    T* gbuf = reinterpret_cast<T*>(t.gred + t.phase * (RSG_TEAM_MAXB * 8));
    t.phase ^= 1;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = wsum(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) buf[wid * NV + i] = v[i];
    }
    __syncthreads();
    if (t.ltid == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            T s = T(0);
            for (int w = 0; w < nw; ++w) s += buf[w * NV + i];
            gbuf[t.blk * NV + i] = s;
        }
    }
    RSG_SYNC();
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        T s = T(0);
        for (int b = lane; b < t.nblk; b += 64) s += gbuf[b * NV + i];
        v[i] = __shfl(wsum(s), 0, 64);
    }
}
template <int NV> RSG_D void bmax(Thr& t, double (&v)[NV]) {
    static_assert(NV <= 4, "team reductions carry at most 4 values");
    const int lane = t.ltid & 63, wid = t.ltid >> 6, nw = (t.lbs + 63) >> 6;
    double* buf = t.red + t.phase * 64;
    double* gbuf = t.gred + t.phase * (RSG_TEAM_MAXB * 8);
    t.phase ^= 1;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = wmax(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) buf[wid * NV + i] = v[i];
    }
    __syncthreads();
    if (t.ltid == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = buf[i];
            for (int w = 1; w < nw; ++w) s = dmax(s, buf[w * NV + i]);
            gbuf[t.blk * NV + i] = s;
        }
    }
    RSG_SYNC();
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double s = gbuf[i];   // workgroup 0's value: lanes past the last workgroup keep it
        for (int b = lane; b < t.nblk; b += 64) s = dmax(s, gbuf[b * NV + i]);
        v[i] = __shfl(wmax(s), 0, 64);
    }
}
#endif  // RSG_TEAM
#endif

#define FOR_CELLS(c) for (int c = t.tid; c < nC; c += t.bs)
#define FOR_IFACES(f) for (int f = t.tid; f < nIF; f += t.bs)
#define FOR_BFACES(b) for (int b = t.tid; b < nB; b += t.bs)

// Everything a phase needs, hoisted once per launch.
struct Ctx {
    Thr t;
    const RsgMeshDev* m;
    const RsgBcDev* bc;
    const RsgSettingsDev* s;
    RsgMemberDev* M;
    int nC, nIF, nB;
    double nu;
    // SST coefficients, InferenceParameters.hpp:12-14 order
    double sk1, sw1, b1, al1, sk2, sw2, b2, al2, bs, a1, kap;
    unsigned long long tmark;   // section timer (RSG_PROF builds)
};

RSG_D double blend(double p1, double p2, double F1) { return F1 * p1 + (1.0 - F1) * p2; }

// ------------------------------------------------------- cell face lists
// Calls fn(code, other, slot) for every face of cell c in the order of its
// cell-face list -- internal faces by ascending face index, then boundary
// faces ascending, the order in which the reference's face-loop scatter adds
// into the cell -- with code = (f << 1) | isNeighbour, other the cell across
// an internal face (-1 on a boundary face) and slot the face's ELL slot
// (a compile-time constant after unrolling; -1 for overflow and boundary
// faces).  The ELL slots are read in one round of independent loads.  A
// caller that writes global memory per face defers the writes of ELL faces
// to after the call (indexed by slot), so the compiler may issue every ELL
// face's loads before any store.
template <class Fn> RSG_D void cell_faces(const RsgMeshDev* m, int c, Fn&& fn) {
    const int nC = m->nC;
    const int cnt = m->ellCnt[c];
    int code[RSG_ELL_K], nb[RSG_ELL_K];
#pragma unroll
    for (int k = 0; k < RSG_ELL_K; ++k) { code[k] = m->ellFace[k * nC + c]; nb[k] = m->ellNb[k * nC + c]; }
    const int ov0 = m->ovPtr[c], ov1 = m->ovPtr[c + 1], b0 = m->bPtr[c], b1 = m->bPtr[c + 1];
#pragma unroll
    for (int k = 0; k < RSG_ELL_K; ++k)
        if (k < cnt) fn(code[k], nb[k], k);
    for (int e = ov0; e < ov1; ++e) fn(m->ovFace[e], m->ovNb[e], -1);
    for (int e = b0; e < b1; ++e) fn(m->bFace[e] << 1, -1, -1);
}

// Owner-side LDU off-diagonal writes of one cell's ELL faces, deferred.
struct PendingLdu {
    int f[RSG_ELL_K];
    double up[RSG_ELL_K], lo[RSG_ELL_K];
};
RSG_D void pending_clear(PendingLdu& P) {
#pragma unroll
    for (int k = 0; k < RSG_ELL_K; ++k) P.f[k] = -1;
}
RSG_D void pending_put(PendingLdu& P, RsgMemberDev* M, int slot, int f, double up, double lo) {
    if (slot >= 0) { P.f[slot] = f; P.up[slot] = up; P.lo[slot] = lo; }
    else { M->upper[f] = up; M->lower[f] = lo; }
}
RSG_D void pending_flush(const PendingLdu& P, RsgMemberDev* M) {
#pragma unroll
    for (int k = 0; k < RSG_ELL_K; ++k)
        if (P.f[k] >= 0) { M->upper[P.f[k]] = P.up[k]; M->lower[P.f[k]] = P.lo[k]; }
}

// ------------------------------------------------------- Green-Gauss (gather)
// greenGaussGrad Field.hpp:122-152 for NF fields at once, cell c.
template <int NF>
RSG_D void gg_cell(const RsgMeshDev* m, int c, const double* const (&phi)[NF],
                   const double* const (&phib)[NF], double (&g)[NF][3]) {
    const int nIF = m->nIF;
    double pc[NF];
#pragma unroll
    for (int q = 0; q < NF; ++q) { g[q][0] = g[q][1] = g[q][2] = 0.0; pc[q] = phi[q][c]; }
    cell_faces(m, c, [&](int code, int other, int slot) {
        const int f = code >> 1;
        const double sx = m->sfx[f], sy = m->sfy[f], sz = m->sfz[f];
        if (f < nIF) {
            const bool nbSide = code & 1;
            const double w = m->w[f], wm = 1.0 - w;
#pragma unroll
            for (int q = 0; q < NF; ++q) {
                const double po = nbSide ? phi[q][other] : pc[q];
                const double pn = nbSide ? pc[q] : phi[q][other];
                const double pf = po * w + pn * wm;
                const double ax = sx * pf, ay = sy * pf, az = sz * pf;
                // x - a == x + (-a) exactly: select the operand, not the accumulator
                g[q][0] = g[q][0] + (nbSide ? -ax : ax);
                g[q][1] = g[q][1] + (nbSide ? -ay : ay);
                g[q][2] = g[q][2] + (nbSide ? -az : az);
            }
        } else {
#pragma unroll
            for (int q = 0; q < NF; ++q) {
                const double pb = phib[q][f - nIF];
                g[q][0] = g[q][0] + sx * pb; g[q][1] = g[q][1] + sy * pb; g[q][2] = g[q][2] + sz * pb;
            }
        }
    });
    const double V = m->vol[c];
    if (V > 1e-30) {
#pragma unroll
        for (int q = 0; q < NF; ++q) { g[q][0] = g[q][0] / V; g[q][1] = g[q][1] / V; g[q][2] = g[q][2] / V; }
    }
}

// ------------------------------------------------------------------ SST
// SSTModel.cpp:9-26 computeF1
RSG_D double sst_F1(const Ctx& C, double k, double om, double y, double nu, double CDin) {
    double sqrtK = sqrt(dmax(k, 0.0));
    double omSafe = dmax(om, 1e-20);
    double ySafe = dmax(y, 1e-20);
    double y2s = ySafe * ySafe;
    double term1 = sqrtK / (C.bs * omSafe * ySafe);
    double term2 = 500.0 * nu / (y2s * omSafe);
    double CDpos = dmax(CDin, 1e-20);
    double term3 = 4.0 * C.sw2 * dmax(k, 0.0) / (CDpos * y2s);
    double arg1 = dmin(dmax(term1, term2), term3);
    double a4 = arg1 * arg1 * arg1 * arg1;
    return tanh(a4);
}
// SSTModel.cpp:30-44 computeF2
RSG_D double sst_F2(const Ctx& C, double k, double om, double y, double nu) {
    double sqrtK = sqrt(dmax(k, 0.0));
    double omSafe = dmax(om, 1e-20);
    double ySafe = dmax(y, 1e-20);
    double y2 = ySafe * ySafe;
    double term1 = 2.0 * sqrtK / (C.bs * omSafe * ySafe);
    double term2 = 500.0 * nu / (y2 * omSafe);
    double arg2 = dmax(term1, term2);
    return tanh(arg2 * arg2);
}
// SSTModel.cpp:57-65 eddyViscosity
RSG_D double sst_nut(const Ctx& C, double k, double om, double S, double F2) {
    double kSafe = dmax(k, 0.0);
    if (C.s->variant == 1) return kSafe / dmax(om, 1e-20);
    double denom = dmax(C.a1 * om, S * F2);
    denom = dmax(denom, 1e-20);
    return C.a1 * kSafe / denom;
}
// SSTModel.cpp:70-74 production
RSG_D double sst_prod(const Ctx& C, double nuT, double S, double k, double om) {
    double raw = nuT * S * S;
    double lim = 10.0 * C.bs * dmax(k, 0.0) * dmax(om, 1e-20);
    return dmin(raw, lim);
}
// SSTModel.cpp:91-96 productionOmega
RSG_D double sst_prod_om(const Ctx& C, double nuT, double S, double k, double om) {
    double S2 = S * S;
    double lim = 10.0 * C.bs * dmax(k, 0.0) * dmax(om, 1e-20) / dmax(nuT, 1e-30);
    return dmin(S2, lim);
}

// --------------------------------------------------------------- BCs
// applyVelocityBC BoundaryCondition.hpp:196-225 (per-face types; profiles
// resolved by the host into velV)
RSG_D void vel_bc(Ctx& C) {
    const RsgMeshDev* m = C.m; const RsgBcDev* bc = C.bc; RsgMemberDev* M = C.M;
    Thr& t = C.t; const int nB = C.nB, nIF = C.nIF;
    FOR_BFACES(b) {
        const int f = nIF + b, o = m->own[f];
        switch (bc->velT[b]) {
            case RSGBC_WALL_NOSLIP: M->Ubx[b] = 0.0; M->Uby[b] = 0.0; M->Ubz[b] = 0.0; break;
            case RSGBC_WALL_MOVING:
            case RSGBC_INLET_VELOCITY:
            case RSGBC_DIRICHLET:
                M->Ubx[b] = bc->velV[3 * b]; M->Uby[b] = bc->velV[3 * b + 1]; M->Ubz[b] = bc->velV[3 * b + 2];
                break;
            case RSGBC_SYMMETRY: {
                const double ux = M->Ux[o], uy = M->Uy[o], uz = M->Uz[o];
                const double nx = m->nx[f], ny = m->ny[f], nz = m->nz[f];
                const double d = ux * nx + uy * ny + uz * nz;
                M->Ubx[b] = ux - nx * d; M->Uby[b] = uy - ny * d; M->Ubz[b] = uz - nz * d;
                break;
            }
            case RSGBC_KEEP: break;   // face in no patch: keeps its value
            default: M->Ubx[b] = M->Ux[o]; M->Uby[b] = M->Uy[o]; M->Ubz[b] = M->Uz[o]; break;
        }
    }
}
// applyPressureBC :229-245
RSG_D void p_bc(Ctx& C) {
    const RsgMeshDev* m = C.m; const RsgBcDev* bc = C.bc; RsgMemberDev* M = C.M;
    Thr& t = C.t; const int nB = C.nB, nIF = C.nIF;
    FOR_BFACES(b) {
        const int ty = bc->pT[b];
        if (ty == RSGBC_KEEP) continue;
        M->pb[b] = (ty == RSGBC_OUTLET_PRESSURE || ty == RSGBC_DIRICHLET) ? bc->pV[b] : M->p[m->own[nIF + b]];
    }
}
// applyKBC :249-269
RSG_D void k_bc(Ctx& C) {
    const RsgMeshDev* m = C.m; const RsgBcDev* bc = C.bc; RsgMemberDev* M = C.M;
    Thr& t = C.t; const int nB = C.nB, nIF = C.nIF;
    FOR_BFACES(b) {
        const int ty = bc->kT[b];
        if (ty == RSGBC_KEEP) continue;
        M->kb[b] = ty == RSGBC_WALL_KOMEGA ? 0.0 : ty == RSGBC_DIRICHLET ? bc->kV[b] : M->k[m->own[nIF + b]];
    }
}
// applyOmegaBC :275-300
RSG_D void om_bc(Ctx& C, double beta1) {
    const RsgMeshDev* m = C.m; const RsgBcDev* bc = C.bc; RsgMemberDev* M = C.M;
    Thr& t = C.t; const int nB = C.nB, nIF = C.nIF;
    FOR_BFACES(b) {
        const int ty = bc->omT[b], f = nIF + b;
        if (ty == RSGBC_KEEP) continue;
        if (ty == RSGBC_WALL_KOMEGA) {
            const double y1 = m->dlt[f];
            M->omb[b] = 60.0 * C.nu / (beta1 * y1 * y1);
        } else if (ty == RSGBC_DIRICHLET) {
            M->omb[b] = bc->omV[b];
        } else {
            M->omb[b] = M->om[m->own[f]];
        }
    }
}

// ------------------------------------------------------- LDU operations
// y = A x (LinearSystem::matvec LinearSolver.hpp:68-77) for cell c, gathering
// the off-diagonal terms in ascending face order.  T is the arithmetic type:
// double, or float for the FP32 inner solves of mixed-precision members (the
// FP64 coefficients are rounded to FP32 as they are read).
template <class T = double>
RSG_D T ldu_row_f32(const RsgMeshDev* m, const double* diag, const double* upper, const double* lower, const T* x,
                    int c);
template <class T = double>
RSG_D T ldu_row(const RsgMeshDev* m, const double* diag, const double* upper, const double* lower, const T* x,
                int c) {
    if constexpr (sizeof(T) == sizeof(float)) return ldu_row_f32<T>(m, diag, upper, lower, x, c);
    T y = (T)diag[c] * x[c];
    const int nIF = m->nIF;
    if (m->ellK) {
        // every internal face of the cell sits in its ELL slots, in ascending
        // face order: the same terms added in the same order as the face-list
        // loop below, with every slot's loads issued before the first add
        const int nC = m->nC;
        int code[RSG_ELL_K], nb[RSG_ELL_K];
        T cf[RSG_ELL_K], xv[RSG_ELL_K];
#pragma unroll
        for (int k = 0; k < RSG_ELL_K; ++k) { code[k] = m->ellFace[k * nC + c]; nb[k] = m->ellNb[k * nC + c]; }
#pragma unroll
        for (int k = 0; k < RSG_ELL_K; ++k) {
            const int f = code[k] >> 1;
            cf[k] = code[k] < 0 ? T(0) : (T)((code[k] & 1) ? lower[f] : upper[f]);
            xv[k] = x[nb[k]];
        }
#pragma unroll
        for (int k = 0; k < RSG_ELL_K; ++k)
            if (code[k] >= 0) y += cf[k] * xv[k];
        return y;
    }
    for (int e = m->cfPtr[c]; e < m->cfPtr[c + 1]; ++e) {
        const int code = m->cfIdx[e];
        const int f = code >> 1;
        if (f >= nIF) break;   // boundary faces sort after every internal face
        if (code & 1) y += (T)lower[f] * x[m->own[f]];
        else          y += (T)upper[f] * x[m->nbr[f]];
    }
    return y;
}

// The FP32 row product in the form ex x_c + sum a (x_nb - x_c), ex the FP64
// row sum rounded once (the resident solvers' res_row explains why).
template <class T>
RSG_D T ldu_row_f32(const RsgMeshDev* m, const double* diag, const double* upper, const double* lower, const T* x,
                    int c) {
    const int nIF = m->nIF;
    const T xc = x[c];
    double ex = diag[c];
    T y = T(0);
    if (m->ellK) {
        const int nC = m->nC;
        int code[RSG_ELL_K], nb[RSG_ELL_K];
        double cf[RSG_ELL_K];
        T xv[RSG_ELL_K];
#pragma unroll
        for (int k = 0; k < RSG_ELL_K; ++k) { code[k] = m->ellFace[k * nC + c]; nb[k] = m->ellNb[k * nC + c]; }
#pragma unroll
        for (int k = 0; k < RSG_ELL_K; ++k) {
            const int f = code[k] >> 1;
            cf[k] = code[k] < 0 ? 0.0 : ((code[k] & 1) ? lower[f] : upper[f]);
            xv[k] = x[nb[k]];
        }
#pragma unroll
        for (int k = 0; k < RSG_ELL_K; ++k)
            if (code[k] >= 0) { ex += cf[k]; y += (T)cf[k] * (xv[k] - xc); }
    } else {
        for (int e = m->cfPtr[c]; e < m->cfPtr[c + 1]; ++e) {
            const int code = m->cfIdx[e];
            const int f = code >> 1;
            if (f >= nIF) break;
            const double a = (code & 1) ? lower[f] : upper[f];
            ex += a;
            y += (T)a * (x[(code & 1) ? m->own[f] : m->nbr[f]] - xc);
        }
    }
    return (T)ex * xc + y;
}

struct SolveRes { int iterations; double finalRes, initialRes; int converged; };

template <class T> RSG_D T jac(T r, T d) { return (fabs(d) > T(1e-30)) ? r / d : r; }

// Krylov work vectors of the global-memory solvers.  FP64 members use
// M->kv[0..7]; the FP32 inner solves of mixed members carve up to ten float
// vectors from the storage of kv[0..4] (each holds two), which leaves kv[5..7]
// for the FP64 defect and update of the refinement.
template <class T> RSG_D T* kvec(RsgMemberDev* M, int i);
template <> RSG_DM double* kvec<double>(RsgMemberDev* M, int i) { return M->kv[i]; }
template <> RSG_DM float* kvec<float>(RsgMemberDev* M, int i) {
    return reinterpret_cast<float*>(M->kv[i >> 1]) + (i & 1) * M->mesh->nC;
}

// BiCGSTABSolver::solve LinearSolver.hpp:193-245 (Jacobi right
// preconditioning) from the residual r already at x (rr = r.r block-summed,
// every thread holding it): x and r in/out.
template <class T>
RSG_D SolveRes bicgstab_from(Ctx& C, T* x, T* r, T rr, int maxIter, double tol) {
    Thr& t = C.t; const RsgMeshDev* m = C.m; RsgMemberDev* M = C.M; const int nC = C.nC;
    const double *dg = M->diag, *up = M->upper, *lo = M->lower;
    T *rh = kvec<T>(M, 1), *p = kvec<T>(M, 2), *v = kvec<T>(M, 3);
    T *s = kvec<T>(M, 4), *tt_ = kvec<T>(M, 5), *ph = kvec<T>(M, 6), *sh = kvec<T>(M, 7);
    SolveRes res = {0, 0.0, 0.0, 0};
    const T r0 = sqrt(rr);
    res.initialRes = r0;
    if (r0 < T(1e-30)) { res.converged = 1; return res; }
    FOR_CELLS(c) { rh[c] = r[c]; p[c] = T(0); v[c] = T(0); }
    T rho = T(1), alpha = T(1), omega = T(1);
    T rhoN = rr;   // dot(rh, r) with rh == r: the same sum
    for (int it = 0; it < maxIter; ++it) {
        if (fabs(rhoN) < T(1e-30)) break;
        const T beta = (rhoN / (rho + T(1e-30))) * (alpha / (omega + T(1e-30)));
        FOR_CELLS(c) {
            const T pc = r[c] + beta * (p[c] - omega * v[c]);
            p[c] = pc;
            ph[c] = jac<T>(pc, (T)dg[c]);
        }
        RSG_SYNC();
        T a1[1] = {T(0)};
        FOR_CELLS(c) {
            const T vc = ldu_row<T>(m, dg, up, lo, ph, c);
            v[c] = vc;
            a1[0] += rh[c] * vc;
        }
        bsum(t, a1);
        const T rv = a1[0];
        if (fabs(rv) < T(1e-30)) break;
        alpha = rhoN / rv;
        T a2[1] = {T(0)};
        FOR_CELLS(c) {
            const T sc = r[c] - alpha * v[c];
            s[c] = sc;
            sh[c] = jac<T>(sc, (T)dg[c]);
            a2[0] += sc * sc;
        }
        bsum(t, a2);
        const T sn = sqrt(a2[0]);
        if (sn / r0 < tol) {
            FOR_CELLS(c) x[c] += alpha * ph[c];
            res.iterations = it + 1; res.finalRes = sn / r0; res.initialRes = r0; res.converged = 1;
            return res;
        }
        T a3[2] = {T(0), T(0)};
        FOR_CELLS(c) {
            const T tc = ldu_row<T>(m, dg, up, lo, sh, c);
            tt_[c] = tc;
            a3[0] += tc * tc;
            a3[1] += tc * s[c];
        }
        bsum(t, a3);
        omega = (a3[0] > T(1e-30)) ? a3[1] / a3[0] : T(0);
        T a4[2] = {T(0), T(0)};
        FOR_CELLS(c) {
            x[c] += alpha * ph[c] + omega * sh[c];
            const T rc = s[c] - omega * tt_[c];
            r[c] = rc;
            a4[0] += rc * rc;
            a4[1] += rh[c] * rc;
        }
        bsum(t, a4);
        const T rn = sqrt(a4[0]);
        res.iterations = it + 1;
        res.finalRes = rn / r0;
        if (res.finalRes < tol) { res.converged = 1; return res; }
        if (fabs(omega) < T(1e-30)) break;
        rho = rhoN;
        rhoN = a4[1];
    }
    return res;
}

// FP64 defect r = b - A x of the member's LDU system, returns r.r (block sum).
RSG_D double ldu_defect(Ctx& C, const double* x, double* r) {
    Thr& t = C.t; const RsgMeshDev* m = C.m; RsgMemberDev* M = C.M; const int nC = C.nC;
    double acc[1] = {0.0};
    FOR_CELLS(c) {
        const double rc = M->src[c] - ldu_row<double>(m, M->diag, M->upper, M->lower, x, c);
        r[c] = rc;
        acc[0] += rc * rc;
    }
    bsum(t, acc);
    return acc[0];
}

// One in-place Gauss-Seidel sweep of the member's LDU system in cell order,
// run by one thread (each unknown takes the latest neighbour values:
// GaussSeidelSolver::solve, LinearSolver.hpp:281-288).  A cell's internal
// faces are visited in ascending face order -- the order in which the
// reference builds the cell's adjacency -- so the subtractions round alike.
RSG_D void gs_sweep(const RsgMeshDev* m, const RsgMemberDev* M, double* x) {
    const int nC = m->nC, nIF = m->nIF;
    for (int c = 0; c < nC; ++c) {
        const double d = M->diag[c];
        if (fabs(d) <= 1e-30) continue;
        double s = M->src[c];
        for (int e = m->cfPtr[c]; e < m->cfPtr[c + 1]; ++e) {
            const int code = m->cfIdx[e];
            const int f = code >> 1;
            if (f >= nIF) break;   // boundary faces sort after every internal face
            if (code & 1) s -= M->lower[f] * x[m->own[f]];
            else          s -= M->upper[f] * x[m->nbr[f]];
        }
        x[c] = s / d;
    }
}

// Gauss-Seidel solver (GaussSeidelSolver, LinearSolver.hpp:258-299): initialRes
// = ||b||, sweeps until ||b - A x|| / ||b|| < tol (||b|| < 1e-30 counts as 1).
// The sweep is a sequential recurrence (thread 0); the residual and its norm
// use the whole workgroup.  FP64 whatever the member's precision.
RSG_D SolveRes gauss_seidel(Ctx& C, double* x, int maxIter, double tol) {
    Thr& t = C.t; RsgMemberDev* M = C.M; const int nC = C.nC;
    double* r = M->kv[0];
    SolveRes res = {0, 0.0, 0.0, 0};
    double acc[1] = {0.0};
    FOR_CELLS(c) acc[0] += M->src[c] * M->src[c];
    bsum(t, acc);
    double b0 = sqrt(acc[0]);
    res.initialRes = b0;
    if (b0 < 1e-30) b0 = 1.0;
    for (int it = 0; it < maxIter; ++it) {
        if (t.tid == 0) gs_sweep(C.m, M, x);
        RSG_SYNC();
        const double rr = ldu_defect(C, x, r);
        res.iterations = it + 1;
        res.finalRes = sqrt(rr) / b0;
        if (res.finalRes < tol) { res.converged = 1; break; }
    }
    return res;
}

// BiCGSTAB on the member's system from x (FP64).
RSG_D SolveRes bicgstab(Ctx& C, double* x, int maxIter, double tol) {
    double* r = C.M->kv[0];
    const double rr = ldu_defect(C, x, r);
    return bicgstab_from<double>(C, x, r, rr, maxIter, tol);
}

// ------------------------------------------------------------------ AMG
// Serial hierarchy construction, run by thread 0 while the workgroup waits.
// Aggregation is the reference's greedy pass bit for bit (buildAggregates
// LinearSolver.hpp:407-443); the Galerkin pattern and the gather lists
// reproduce galerkinProduct's summation order (:446-485: coarse entry (I,J)
// sums fine entries in ascending fine row, then ascending entry).
RSG_HD int amg_aggregate(int n, const int* rp, const int* ci, const double* v, const double* dg, int* map) {
    int nc = 0;
    for (int i = 0; i < n; ++i) map[i] = -1;
    for (int i = 0; i < n; ++i) {
        if (map[i] >= 0) continue;
        const int agg = nc++;
        map[i] = agg;
        const double aii = fabs(dg[i]);
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            const int c = ci[j];
            if (c == i || map[c] >= 0) continue;
            const double ajj = fabs(dg[c]);
            const double strength = fabs(v[j]) / sqrt(aii * ajj + 1e-30);
            if (strength > 0.25) map[c] = agg;
        }
    }
    for (int i = 0; i < n; ++i) {
        if (map[i] >= 0) continue;
        double best = -1.0;
        int bestAgg = 0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            const int c = ci[j];
            if (c == i && map[c] < 0) continue;
            if (map[c] >= 0 && fabs(v[j]) > best) { best = fabs(v[j]); bestAgg = map[c]; }
        }
        map[i] = bestAgg;
    }
    return nc;
}

// Builds levels 1.. from level 0 (whose values must be current).
// ELL image of coarse level lc (rows crp/cci/cv) and of its aggregates
// (aP/aI over the finer level), appended at the running offsets.
#ifdef RSG_EMULATE
static int rsg_emu_no_images = 0;   // emulation build only: force the CSR V-cycle (tests)
#endif
RSG_HD void amg_level_images(RsgAmgDev& A, int lc, int nc, const int* crp, const int* cci, const double* cv,
                             const int* aP, const int* aI, int& eNext, int& aNext) {
    int K = 0, AK = 0;
#ifdef RSG_EMULATE
    if (rsg_emu_no_images) { A.lvK[lc] = 0; A.lvAK[lc] = 0; return; }
#endif
    for (int I = 0; I < nc; ++I) {
        K = (crp[I + 1] - crp[I] > K) ? crp[I + 1] - crp[I] : K;
        AK = (aP[I + 1] - aP[I] > AK) ? aP[I + 1] - aP[I] : AK;
    }
    A.lvK[lc] = 0; A.lvAK[lc] = 0;
    if (eNext + (long)K * nc <= A.eCap) {
        A.lvK[lc] = K; A.lvE[lc] = eNext;
        for (int k = 0; k < K; ++k)
            for (int I = 0; I < nc; ++I) {
                const int e = crp[I] + k;
                const bool in = e < crp[I + 1];
                A.eCol[eNext + k * nc + I] = in ? cci[e] : -1;
                A.eVal[eNext + k * nc + I] = in ? cv[e] : 0.0;
            }
        eNext += K * nc;
    }
    if (aNext + (long)AK * nc <= A.aCap) {
        A.lvAK[lc] = AK; A.lvAE[lc] = aNext;
        for (int q = 0; q < AK; ++q)
            for (int I = 0; I < nc; ++I) {
                const int e = aP[I] + q;
                A.aCol[aNext + q * nc + I] = e < aP[I + 1] ? aI[e] : -1;
            }
        aNext += AK * nc;
    }
}

RSG_D void amg_setup_serial(const RsgMeshDev* m, RsgAmgDev& A) {
    int L = 1;
    int eNext = 0, aNext = 0;
    for (int l = 0; l <= RSG_MAX_LEVELS; ++l) { A.lvK[l] = 0; A.lvAK[l] = 0; }
    for (int l = 0; l < RSG_MAX_LEVELS; ++l) {
        const int n = A.lvN[l];
        if (n <= RSG_MIN_COARSE) break;
        const int* rp = A.rp + A.lvRp[l];
        const int* ci = A.ci + A.lvNz[l];
        const double* v = A.v + A.lvNz[l];
        const double* dg = A.dg + A.lvRow[l];
        int* map = A.map + A.lvRow[l];
        const int nnz = rp[n];
        const int nc = amg_aggregate(n, rp, ci, v, dg, map);
        if (nc == 0 || nc >= n) break;
        // capacity of the next level (coarse entries never exceed fine ones)
        const int rowNext = A.lvRow[l] + n;
        const int nzNext = A.lvNz[l] + nnz;
        if (rowNext + nc > A.rowCap || nzNext + nnz > A.nnzCap) break;
        A.lvRow[l + 1] = rowNext;
        A.lvRp[l + 1] = rowNext + (l + 1);
        A.lvNz[l + 1] = nzNext;
        A.lvN[l + 1] = nc;
        // aggregate -> fine rows (ascending): counting sort of map
        int* aP = A.aggPtr + A.lvRp[l + 1];
        int* aI = A.aggIdx + A.lvRow[l];
        for (int I = 0; I <= nc; ++I) aP[I] = 0;
        for (int i = 0; i < n; ++i) aP[map[i] + 1]++;
        for (int I = 0; I < nc; ++I) aP[I + 1] += aP[I];
        int* fill = A.scr;                  // [nc]
        for (int I = 0; I < nc; ++I) fill[I] = aP[I];
        for (int i = 0; i < n; ++i) aI[fill[map[i]]++] = i;
        // coarse pattern, row by row
        int* mark = A.scr;                  // [nc] (reuses fill)
        int* pos = A.scr + A.rowCap;        // [nc]
        int* cols = A.scr + 2 * A.rowCap;   // [row length <= nc]
        int* f2c = A.scr + 3 * A.rowCap;    // [nnz] fine entry -> coarse entry
        for (int J = 0; J < nc; ++J) mark[J] = -1;
        int* crp = A.rp + A.lvRp[l + 1];
        int* cci = A.ci + A.lvNz[l + 1];
        int cnz = 0;
        for (int I = 0; I < nc; ++I) {
            int ncol = 0;
            for (int q = aP[I]; q < aP[I + 1]; ++q) {
                const int i = aI[q];
                for (int j = rp[i]; j < rp[i + 1]; ++j) {
                    const int J = map[ci[j]];
                    if (mark[J] != I) { mark[J] = I; cols[ncol++] = J; }
                }
            }
            for (int a = 1; a < ncol; ++a) {
                const int key = cols[a];
                int b = a - 1;
                while (b >= 0 && cols[b] > key) { cols[b + 1] = cols[b]; --b; }
                cols[b + 1] = key;
            }
            crp[I] = cnz;
            for (int a = 0; a < ncol; ++a) { pos[cols[a]] = cnz + a; cci[cnz + a] = cols[a]; }
            for (int q = aP[I]; q < aP[I + 1]; ++q) {
                const int i = aI[q];
                for (int j = rp[i]; j < rp[i + 1]; ++j) f2c[j] = pos[map[ci[j]]];
            }
            cnz += ncol;
        }
        crp[nc] = cnz;
        // coarse entry -> fine entries, stable in fine entry order
        int* gP = A.gPtr + A.lvNz[l + 1] + (l + 1);
        int* gI = A.gIdx + A.lvNz[l];
        for (int e = 0; e <= cnz; ++e) gP[e] = 0;
        for (int j = 0; j < nnz; ++j) gP[f2c[j] + 1]++;
        for (int e = 0; e < cnz; ++e) gP[e + 1] += gP[e];
        int* gfill = A.scr + 3 * A.rowCap + A.nnzCap;   // [cnz]
        for (int e = 0; e < cnz; ++e) gfill[e] = gP[e];
        for (int j = 0; j < nnz; ++j) gI[gfill[f2c[j]]++] = j;
        // coarse values (serially here; refreshed in parallel every iteration)
        double* cv = A.v + A.lvNz[l + 1];
        double* cdg = A.dg + A.lvRow[l + 1];
        for (int I = 0; I < nc; ++I) cdg[I] = 0.0;
        for (int I = 0; I < nc; ++I)
            for (int e = crp[I]; e < crp[I + 1]; ++e) {
                double s = 0.0;
                for (int q = gP[e]; q < gP[e + 1]; ++q) s += v[gI[q]];
                cv[e] = s;
                if (cci[e] == I) cdg[I] = s;
            }
        amg_level_images(A, l + 1, nc, crp, cci, cv, aP, aI, eNext, aNext);
        L = l + 2;
    }
    A.used[0] = eNext; A.used[1] = aNext;
    A.lvN[RSG_MAX_LEVELS] = L;   // level count
}

// Parallel refresh of level 0's CSR values from the current LDU matrix and,
// with coarse set, of every coarser level's Galerkin values.
RSG_D void amg_refresh_values(Ctx& C, bool coarse) {
    Thr& t = C.t; const RsgMeshDev* m = C.m; RsgMemberDev* M = C.M;
    RsgAmgDev& A = M->amg;
    const int nC = C.nC, nIF = C.nIF, nnz0 = m->nnz0;
    for (int e = t.tid; e < nnz0; e += t.bs) {
        const int s = m->src0[e];
        A.v[e] = s < 0 ? M->diag[-1 - s] : (s < nIF ? M->upper[s] : M->lower[s - nIF]);
    }
    FOR_CELLS(c) A.dg[c] = M->diag[c];
    RSG_SYNC();
    if (!coarse) return;
    const int L = A.lvN[RSG_MAX_LEVELS];
    for (int l = 0; l + 1 < L; ++l) {
        const int nc = A.lvN[l + 1];
        const int* crp = A.rp + A.lvRp[l + 1];
        const int* cci = A.ci + A.lvNz[l + 1];
        const int* gP = A.gPtr + A.lvNz[l + 1] + (l + 1);
        const int* gI = A.gIdx + A.lvNz[l];
        const double* fv = A.v + A.lvNz[l];
        double* cv = A.v + A.lvNz[l + 1];
        double* cdg = A.dg + A.lvRow[l + 1];
        const int K = A.lvK[l + 1];
        double* ev = A.eVal + A.lvE[l + 1];
        for (int I = t.tid; I < nc; I += t.bs) {
            double d = 0.0;
            for (int e = crp[I]; e < crp[I + 1]; ++e) {
                double s = 0.0;
                for (int q = gP[e]; q < gP[e + 1]; ++q) s += fv[gI[q]];
                cv[e] = s;
                if (K > 0) ev[(e - crp[I]) * nc + I] = s;
                if (cci[e] == I) d = s;
            }
            cdg[I] = d;
        }
        RSG_SYNC();
    }
}

// Dense inverse of the coarsest level (Gauss-Jordan, no pivoting: the level
// is a Galerkin image of an SPD M-matrix).  Built in LDS by wave 0, lane j
// owning column j; LDS accesses of one wave execute in program order, so the
// in-place elimination needs no workgroup barrier.  The inverse is written
// to global memory for the V-cycle's coarse solve.
// The dense Gauss-Jordan inverse of the coarsest level in LDS by the threads
// of one workgroup (ltid, lbs: their place in it); returns 0 on a zero pivot.
RSG_D int coarse_dense_block(double* G, const int* rp, const int* ci, const double* v, int n, int ltid, int lbs) {
    const int DN = RSG_COARSE_DENSE;
    for (int idx = ltid; idx < n * DN; idx += lbs) G[idx] = 0.0;
    RSG_BLOCK_SYNC();
    for (int i = ltid; i < n; i += lbs)
        for (int e = rp[i]; e < rp[i + 1]; ++e) G[i * DN + ci[e]] += v[e];
    RSG_BLOCK_SYNC();
    int ok = 1;
#ifdef RSG_EMULATE
    for (int k = 0; k < n; ++k) {
        const double piv = G[k * DN + k];
        if (!(fabs(piv) > 1e-300)) { ok = 0; break; }
        for (int j = 0; j < n; ++j) G[k * DN + j] = (j == k ? 1.0 : G[k * DN + j]) / piv;
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            const double fct = G[i * DN + k];
            for (int j = 0; j < n; ++j) G[i * DN + j] = (j == k ? 0.0 : G[i * DN + j]) - fct * G[k * DN + j];
        }
    }
#else
    // Lane j owns column j; wave w owns rows w, w + nw, ...  Within a row the
    // factor G[i][k] is read before any lane of the (same) wave writes it; the
    // pivot row is read by every wave before the barrier and rewritten after.
    {
        const int j = ltid & 63, w = ltid >> 6, nw = lbs >> 6;
        for (int k = 0; k < n; ++k) {
            const double piv = G[k * DN + k];
            if (!(fabs(piv) > 1e-300)) { ok = 0; break; }   // same value in every thread
            const double rk = (j < n) ? (j == k ? 1.0 : G[k * DN + j]) / piv : 0.0;
            __syncthreads();
            for (int i = w; i < n; i += nw) {
                if (i == k) {
                    if (j < n) G[k * DN + j] = rk;
                    continue;
                }
                const double fct = G[i * DN + k];
                __builtin_amdgcn_wave_barrier();
                if (j < n) G[i * DN + j] = (j == k ? 0.0 : G[i * DN + j]) - fct * rk;
                __builtin_amdgcn_wave_barrier();
            }
            __syncthreads();
        }
    }
#endif
    RSG_BLOCK_SYNC();
    return ok;
}

RSG_D void amg_coarse_inverse(Ctx& C) {
    Thr& t = C.t; RsgMemberDev* M = C.M; RsgAmgDev& A = M->amg;
    const int L = A.lvN[RSG_MAX_LEVELS];
    const int n = A.lvN[L - 1];
    const int DN = RSG_COARSE_DENSE;
    if (n > DN) { if (t.tid == 0) M->sc->coarseDense = 0; RSG_SYNC(); return; }
    const int* rp = A.rp + A.lvRp[L - 1];
    const int* ci = A.ci + A.lvNz[L - 1];
    const double* v = A.v + A.lvNz[L - 1];
#ifdef RSG_TEAM
    // the first workgroup of the team factors; the others wait at the barrier
    if (t.blk == 0) {
        const int ok = coarse_dense_block(t.lds, rp, ci, v, n, t.ltid, t.lbs);
        for (int idx = t.ltid; idx < n * DN; idx += t.lbs) A.cinv[idx] = t.lds[idx];
        if (t.ltid == 0) M->sc->coarseDense = ok;
    }
#else
    const int ok = coarse_dense_block(t.lds, rp, ci, v, n, t.tid, t.bs);
    for (int idx = t.tid; idx < n * DN; idx += t.bs) A.cinv[idx] = t.lds[idx];
    if (t.tid == 0) M->sc->coarseDense = ok;
#endif
    RSG_SYNC();
}

// V-cycle preconditioner z = M^{-1} r (AMGSolver::vcycle :553-589 with
// damped-Jacobi smoothing in place of the sequential Gauss-Seidel sweep, and
// an exact dense coarsest solve in place of 50 Gauss-Seidel sweeps).
// s -/+ sum over CSR entries e0..e1-1 of v[e] x[ci[e]], term by term in CSR
// order (the same roundings as the plain loop); the loads of four entries
// are issued together.  T: arithmetic type (values rounded to it as read).
template <bool SUB, class T>
RSG_D T csr_row_acc(const int* ci, const double* v, const T* x, int e0, int e1, T s) {
    for (int e = e0; e < e1; e += 4) {
        int c[4]; T a[4], xv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool ok = e + k < e1;
            c[k] = ok ? ci[e + k] : 0;
            a[k] = ok ? (T)v[e + k] : T(0);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) xv[k] = (e + k < e1) ? x[c[k]] : T(0);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (e + k < e1) { if (SUB) s -= a[k] * xv[k]; else s += a[k] * xv[k]; }
    }
    return s;
}

template <class T>
RSG_D void csr_jacobi(Thr& t, int n, const int* rp, const int* ci, const double* v, const double* dg,
                      const T* b, const T* xin, T* xout, bool xzero) {
    const T wJ = T(2.0 / 3.0);
    for (int i = t.tid; i < n; i += t.bs) {
        T s = b[i];
        if (!xzero) s = csr_row_acc<true, T>(ci, v, xin, rp[i], rp[i + 1], s);
        const T d = (T)dg[i];
        const T x0 = xzero ? T(0) : xin[i];
        xout[i] = (fabs(d) > T(1e-30)) ? x0 + wJ * s / d : x0;
    }
}

// Residual row i of a CSR level: b_i - sum of the row's entries times x, in
// CSR order.
template <class T>
RSG_D T csr_row_residual(const int* rp, const int* ci, const double* v, const T* b, const T* x, int i) {
    const T s = csr_row_acc<false, T>(ci, v, x, rp[i], rp[i + 1], T(0));
    return b[i] - s;
}

// Level vectors of the hierarchy in type T: FP64 members use A.x / A.b / A.r
// as doubles; FP32 inner solves view the same storage as floats (the level
// offsets A.lvRow index elements either way).
template <class T> RSG_D T* lvec(double* base) { return reinterpret_cast<T*>(base); }

// Level 0's right-hand side is rin itself.  Going down, each coarse row sums
// the residuals of its aggregate's rows straight from x (a residual pass, a
// barrier and the restriction, rounded the same, without the barrier); going
// up, the first post-smoothing sweep reads every x value as the prolonged
// x[j] + xc[map[j]], so the prolongation pass and its barrier go away.  The
// coarsest level's dense inverse is applied in FP64 whatever T is.
template <class T>
RSG_D void amg_vcycle(Ctx& C, const T* rin, T* z) {
    Thr& t = C.t; RsgMemberDev* M = C.M; RsgAmgDev& A = M->amg;
    const int L = A.lvN[RSG_MAX_LEVELS];
    const T wJ = T(2.0 / 3.0);
    T* Ax = lvec<T>(A.x); T* Ab = lvec<T>(A.b); T* Ar = lvec<T>(A.r);
    for (int l = 0; l + 1 < L; ++l) {
        const int n = A.lvN[l];
        const int* rp = A.rp + A.lvRp[l]; const int* ci = A.ci + A.lvNz[l];
        const double* v = A.v + A.lvNz[l]; const double* dg = A.dg + A.lvRow[l];
        T* x = Ax + A.lvRow[l]; T* r = Ar + A.lvRow[l];
        const T* b = (l == 0) ? rin : Ab + A.lvRow[l];
        csr_jacobi<T>(t, n, rp, ci, v, dg, b, x, r, true);
        RSG_SYNC();
        csr_jacobi<T>(t, n, rp, ci, v, dg, b, r, x, false);
        RSG_SYNC();
        const int nc = A.lvN[l + 1];
        const int* aP = A.aggPtr + A.lvRp[l + 1];
        const int* aI = A.aggIdx + A.lvRow[l];
        T* cb = Ab + A.lvRow[l + 1];
        for (int I = t.tid; I < nc; I += t.bs) {
            T s = T(0);
            for (int q = aP[I]; q < aP[I + 1]; ++q) s += csr_row_residual<T>(rp, ci, v, b, x, aI[q]);
            cb[I] = s;
        }
    }
    RSG_SYNC();
    {   // coarsest level
        const int l = L - 1, n = A.lvN[l];
        T* x = (L == 1) ? z : Ax + A.lvRow[l];
        const T* b = (l == 0) ? rin : Ab + A.lvRow[l];
        if (M->sc->coarseDense) {
            const int DN = RSG_COARSE_DENSE;
            for (int i = t.tid; i < n; i += t.bs) {
                double s = 0.0;
                for (int j = 0; j < n; ++j) s += A.cinv[i * DN + j] * (double)b[j];
                x[i] = (T)s;
            }
            RSG_SYNC();
        } else {
            const int* rp = A.rp + A.lvRp[l]; const int* ci = A.ci + A.lvNz[l];
            const double* v = A.v + A.lvNz[l]; const double* dg = A.dg + A.lvRow[l];
            T* r = Ar + A.lvRow[l];
            csr_jacobi<T>(t, n, rp, ci, v, dg, b, x, x, true);
            RSG_SYNC();
            for (int s = 0; s < 25; ++s) {
                csr_jacobi<T>(t, n, rp, ci, v, dg, b, x, r, false);
                RSG_SYNC();
                csr_jacobi<T>(t, n, rp, ci, v, dg, b, r, x, false);
                RSG_SYNC();
            }
        }
    }
    for (int l = L - 2; l >= 0; --l) {
        const int n = A.lvN[l];
        const int* rp = A.rp + A.lvRp[l]; const int* ci = A.ci + A.lvNz[l];
        const double* v = A.v + A.lvNz[l]; const double* dg = A.dg + A.lvRow[l];
        T* x = Ax + A.lvRow[l]; T* r = Ar + A.lvRow[l];
        const T* b = (l == 0) ? rin : Ab + A.lvRow[l];
        const int* map = A.map + A.lvRow[l];
        const T* cx = Ax + A.lvRow[l + 1];
        for (int i = t.tid; i < n; i += t.bs) {
            T s = b[i];
            for (int e = rp[i]; e < rp[i + 1]; ++e) s -= (T)v[e] * (x[ci[e]] + cx[map[ci[e]]]);
            const T d = (T)dg[i];
            const T x0 = x[i] + cx[map[i]];
            r[i] = (fabs(d) > T(1e-30)) ? x0 + wJ * s / d : x0;
        }
        RSG_SYNC();
        T* out = (l == 0) ? z : x;
        csr_jacobi<T>(t, n, rp, ci, v, dg, b, r, out, false);
        RSG_SYNC();
    }
}

// PCG (:147-186) with Jacobi or AMG V-cycle preconditioning (:602-653), from
// the residual r already at x (rr = r.r block-summed): x and r in/out.
template <class T>
RSG_D SolveRes pcg_from(Ctx& C, T* x, T* r, T rr, int maxIter, double tol, bool useAmg) {
    Thr& t = C.t; const RsgMeshDev* m = C.m; RsgMemberDev* M = C.M; const int nC = C.nC;
    const double *dg = M->diag, *up = M->upper, *lo = M->lower;
    T *z = kvec<T>(M, 1), *p = kvec<T>(M, 2), *Ap = kvec<T>(M, 3);
    SolveRes res = {0, 0.0, 0.0, 0};
    const T r0 = sqrt(rr);
    res.initialRes = r0;
    if (r0 < T(1e-30)) { res.converged = 1; return res; }
    T rzv[1] = {T(0)};
    if (useAmg) {
        amg_vcycle<T>(C, r, z);
        FOR_CELLS(c) { p[c] = z[c]; rzv[0] += r[c] * z[c]; }
    } else {
        FOR_CELLS(c) { const T zc = jac<T>(r[c], (T)dg[c]); z[c] = zc; p[c] = zc; rzv[0] += r[c] * zc; }
    }
    bsum(t, rzv);
    T rz = rzv[0];
    for (int it = 0; it < maxIter; ++it) {
        T a1[1] = {T(0)};
        FOR_CELLS(c) { const T q = ldu_row<T>(m, dg, up, lo, p, c); Ap[c] = q; a1[0] += p[c] * q; }
        bsum(t, a1);
        const T pAp = a1[0];
        if (fabs(pAp) < T(1e-30)) break;
        const T alpha = rz / pAp;
        T a2[1] = {T(0)};
        FOR_CELLS(c) {
            x[c] += alpha * p[c];
            const T rc = r[c] - alpha * Ap[c];
            r[c] = rc;
            a2[0] += rc * rc;
        }
        bsum(t, a2);
        const T rn = sqrt(a2[0]);
        res.iterations = it + 1;
        res.finalRes = rn / r0;
        if (res.finalRes < tol) { res.converged = 1; return res; }
        T a3[1] = {T(0)};
        if (useAmg) {
            amg_vcycle<T>(C, r, z);
            FOR_CELLS(c) a3[0] += r[c] * z[c];
        } else {
            FOR_CELLS(c) { const T zc = jac<T>(r[c], (T)dg[c]); z[c] = zc; a3[0] += r[c] * zc; }
        }
        bsum(t, a3);
        const T rzNew = a3[0];
        const T beta = rzNew / (rz + T(1e-30));
        FOR_CELLS(c) p[c] = z[c] + beta * p[c];
        rz = rzNew;
        RSG_SYNC();
    }
    return res;
}

RSG_D SolveRes pcg(Ctx& C, double* x, int maxIter, double tol, bool useAmg) {
    double* r = C.M->kv[0];
    const double rr = ldu_defect(C, x, r);
    return pcg_from<double>(C, x, r, rr, maxIter, tol, useAmg);
}

// Mixed precision on the global-memory path: FP64 defect r = b - A x, an FP32
// inner solve of A d = r from d = 0 (BiCGSTAB, or PCG with Jacobi / the
// V-cycle) asked for the reduction still missing, x += d in FP64, and the
// FP64 defect again -- iterative refinement until the true FP64 residual is
// below tol relative to the first defect (the reference's inner tolerance,
// LinearSolver.hpp:147-247), at most RSG_MIXED_SWEEPS sweeps and maxIter
// inner iterations in all.  initialRes is the FP64 first defect, as in the
// FP64 solvers, so the outer iteration's residual accounting is unchanged.
#define RSG_MIXED_SWEEPS 4
// Whether a member's inner solves run mixed: fixed at compile time in the
// kernels built for one precision (PREC 0 or 1), read from the member
// otherwise (PREC -1), so an FP64 kernel carries no FP32 solver code.
template <int PREC> RSG_D bool is_mixed(const Ctx& C) { return PREC < 0 ? C.M->prec != 0 : PREC == 1; }
RSG_D SolveRes lsolve_mixed_global(Ctx& C, int code, double* x, bool amg, int maxIter, double tol) {
    Thr& t = C.t; RsgMemberDev* M = C.M; const int nC = C.nC;
    double* r = M->kv[5];
    float* rf = kvec<float>(M, 0);
    float* d = kvec<float>(M, 8);
    SolveRes res = {0, 0.0, 0.0, 0};
    double rn = sqrt(ldu_defect(C, x, r));
    const double r0 = rn;
    res.initialRes = r0;
    if (r0 < 1e-30) { res.converged = 1; return res; }
    for (int sweep = 0; sweep < RSG_MIXED_SWEEPS && res.iterations < maxIter; ++sweep) {
        // the FP32 solve sees the defect scaled to unit norm (see lsolve_mixed_res)
        const double inv = 1.0 / rn;
        FOR_CELLS(c) { rf[c] = (float)(r[c] * inv); d[c] = 0.0f; }
        RSG_SYNC();
        const float rr = 1.0f;
        const double tolIn = tol * r0 / rn;
        const SolveRes in = code == 2 ? bicgstab_from<float>(C, d, rf, rr, maxIter - res.iterations, tolIn)
                                      : pcg_from<float>(C, d, rf, rr, maxIter - res.iterations, tolIn, code == 0 && amg);
        res.iterations += in.iterations;
        RSG_SYNC();
        FOR_CELLS(c) x[c] += rn * (double)d[c];
        RSG_SYNC();
        const double rnNew = sqrt(ldu_defect(C, x, r));
        res.finalRes = rnNew / r0;
        if (res.finalRes < tol) { res.converged = 1; break; }
        if (!(rnNew < rn) || in.iterations == 0) break;   // the FP32 solve stalls: stop refining
        rn = rnNew;
    }
    return res;
}

}  // namespace rsg
#include "rsg_resident.h"
namespace rsg {

// One linear solve: the resident solvers when the member's cells fit the
// workgroup's registers (and, for the V-cycle, its coarse vectors fit LDS),
// the global-memory solvers otherwise.
// allowMixed false keeps this solve in FP64 for a mixed member.
template <int CMAX, int PREC = -1>
RSG_D SolveRes lsolve(Ctx& C, int code, double* x, bool amgReady, bool allowMixed = true) {
    const RsgSettingsDev* s = C.s;
    if (code == RSG_SOLVER_GAUSS_SEIDEL) return gauss_seidel(C, x, s->innerIter, s->innerTol);
    const bool amg = code == 0 && amgReady;
    if (res_fits(C, CMAX) && (!amg || res_amg_fits(C)))
        return lsolve_res<CMAX, PREC>(C, code, x, amg, s->innerIter, s->innerTol, allowMixed);
    if (allowMixed && is_mixed<PREC>(C)) return lsolve_mixed_global(C, code, x, amgReady, s->innerIter, s->innerTol);
    if (code == 0) return pcg(C, x, s->innerIter, s->innerTol, amgReady);
    if (code == 1) return pcg(C, x, s->innerIter, s->innerTol, false);
    return bicgstab(C, x, s->innerIter, s->innerTol);
}

// ------------------------------------------------------------ assembly
RSG_D double* ucomp(RsgMemberDev* M, int comp) { return comp == 0 ? M->Ux : comp == 1 ? M->Uy : M->Uz; }
RSG_D double* ubcomp(RsgMemberDev* M, int comp) { return comp == 0 ? M->Ubx : comp == 1 ? M->Uby : M->Ubz; }

// Convection-diffusion face coefficients shared by the momentum, k and omega
// assemblies: upper = -(Df + cNeg), lower = -(Df + cPos).
RSG_D void conv_diff_face(const RsgMeshDev* m, const RsgMemberDev* M, int f, int o, int n, double nuEff_o,
                          double nuEff_n, double& up, double& lo) {
    const double Sf = m->area[f], delta = m->dlt[f], w = m->w[f], wm = 1.0 - w;
    const double nuEff_f = w * nuEff_o + (1.0 - w) * nuEff_n;
    const double Df = nuEff_f * Sf / delta;
    const double ufx = M->Ux[o] * w + M->Ux[n] * wm;
    const double ufy = M->Uy[o] * w + M->Uy[n] * wm;
    const double ufz = M->Uz[o] * w + M->Uz[n] * wm;
    const double mFlux = (ufx * m->nx[f] + ufy * m->ny[f] + ufz * m->nz[f]) * Sf;
    const double cPos = dmax(mFlux, 0.0), cNeg = dmax(-mFlux, 0.0);
    up = -(Df + cNeg);
    lo = -(Df + cPos);
}

// boundary-face convection-diffusion of a scalar with face value phib
// (outflow / inflow branches, SIMPLESolver.cpp:132-140)
RSG_D void conv_diff_bface(const RsgMeshDev* m, const RsgMemberDev* M, int f, int b, double nuEff,
                           double phib, double& ddg, double& dsrc) {
    const double Sf = m->area[f], delta = m->dlt[f];
    const double Db = nuEff * Sf / delta;
    const double mFlux = (M->Ubx[b] * m->nx[f] + M->Uby[b] * m->ny[f] + M->Ubz[b] * m->nz[f]) * Sf;
    const bool out = mFlux >= 0;
    ddg = out ? mFlux + Db : Db;
    dsrc = out ? Db * phib : (Db - mFlux) * phib;
}

// assembleMomentum SIMPLESolver.cpp:56-203 for one component.  Requires
// current grad U in M->g[0..8] (transpose stress) and grad p in M->gp.
RSG_D void assemble_momentum(Ctx& C, int comp) {
    Thr& t = C.t; const RsgMeshDev* m = C.m; RsgMemberDev* M = C.M;
    const int nC = C.nC, nIF = C.nIF;
    const double nu = C.nu;
    // column `comp` of grad U: (dUx/dx_comp, dUy/dx_comp, dUz/dx_comp)
    const double* gc0 = M->g[0 * 3 + comp];
    const double* gc1 = M->g[1 * 3 + comp];
    const double* gc2 = M->g[2 * 3 + comp];
    const double* phiOld = ucomp(M, comp);
    const double* Ub = ubcomp(M, comp);
    const double fb = C.s->bf[comp];
    const double a = C.s->aU;
    const int wallFn = C.s->wallFn;
    const double* gpc = M->gp[comp];
    // One pass per cell: each internal face's coefficients and transpose-stress
    // flux are evaluated from the cell on either side with the same operands
    // (so both sides see identical values); the owner side stores the LDU
    // off-diagonals.
    FOR_CELLS(c) {
        // Accumulators start at +0 and never become -0, so a +0 increment is
        // exact; every branch produces increments and one add per accumulator
        // follows (no branch-dependent accumulator address).
        double dg = 0.0, src = 0.0, ts = 0.0;
        PendingLdu pend;
        pending_clear(pend);
        cell_faces(m, c, [&](int code, int other, int slot) {
            const int f = code >> 1;
            double ddg, dsrc = 0.0, dts;
            if (f < nIF) {
                const bool nbSide = code & 1;
                const int o = nbSide ? other : c, n = nbSide ? c : other;
                double up, lo;
                conv_diff_face(m, M, f, o, n, nu + M->nuT[o], nu + M->nuT[n], up, lo);
                // transpose-stress face flux, StressOperators.hpp:50-59
                const double w = m->w[f], wm = 1.0 - w;
                const double cf = w * M->nuT[o] + (1.0 - w) * M->nuT[n];
                const double gx = gc0[o] * w + gc0[n] * wm;
                const double gy = gc1[o] * w + gc1[n] * wm;
                const double gz = gc2[o] * w + gc2[n] * wm;
                const double fs = cf * (gx * m->nx[f] + gy * m->ny[f] + gz * m->nz[f]) * m->area[f];
                if (!nbSide) pending_put(pend, M, slot, f, up, lo);
                ddg = nbSide ? -up : -lo;
                dts = nbSide ? -fs : fs;
            } else {
                const int b = f - nIF;
                const int isWall = m->bFlags[b] & RSG_BF_WALL;
                const double nuEff = (isWall && !wallFn) ? nu : nu + M->nuT[c];
                conv_diff_bface(m, M, f, b, nuEff, Ub[b], ddg, dsrc);
                // transpose stress boundary term with the wall coefficient zeroed
                const double cfb = isWall ? 0.0 : M->nuT[c];
                dts = cfb * (gc0[c] * m->nx[f] + gc1[c] * m->ny[f] + gc2[c] * m->nz[f]) * m->area[f];
            }
            dg += ddg; src += dsrc; ts += dts;
        });
        pending_flush(pend, M);
        const double vol = m->vol[c];
        src -= gpc[c] * vol;
        if (fb != 0.0) src += fb * vol;
        src += ts;
        src += (1.0 - a) / a * dg * phiOld[c];
        dg /= a;
        M->diag[c] = dg;
        M->src[c] = src;
        if (comp == 0) M->aPx[c] = dg;
    }
    RSG_SYNC();
}

// assemblePressureCorrection SIMPLESolver.cpp:322-450 (aP = relaxed x-diag).
// grad p must be current in M->gp when the Rhie-Chow term is active.
RSG_D void assemble_pcorr(Ctx& C) {
    Thr& t = C.t; const RsgMeshDev* m = C.m; RsgMemberDev* M = C.M;
    const int nC = C.nC, nIF = C.nIF;
    const bool rc = !m->hasOutlet || C.s->rcAll;
    const double* aP = M->aPx;
    FOR_CELLS(c) {
        double dg = 0.0, src = 0.0;
        PendingLdu pend;
        pending_clear(pend);
        cell_faces(m, c, [&](int code, int other, int slot) {
            const int f = code >> 1;
            double ddg, dsrc;
            if (f < nIF) {
                const bool nbSide = code & 1;
                const int o = nbSide ? other : c, n = nbSide ? c : other;
                const double Sf = m->area[f], delta = m->dlt[f], w = m->w[f], wm = 1.0 - w;
                const double Vo = m->vol[o], Vn = m->vol[n];
                const double dP_o = (fabs(aP[o]) > 1e-30) ? Vo / aP[o] : 0.0;
                const double dP_n = (fabs(aP[n]) > 1e-30) ? Vn / aP[n] : 0.0;
                const double dP_f = w * dP_o + (1.0 - w) * dP_n;
                const double coeff = dP_f * Sf / delta;
                const double ufx = M->Ux[o] * w + M->Ux[n] * wm;
                const double ufy = M->Uy[o] * w + M->Uy[n] * wm;
                const double ufz = M->Uz[o] * w + M->Uz[n] * wm;
                double massFlux = (ufx * m->nx[f] + ufy * m->ny[f] + ufz * m->nz[f]) * Sf;
                if (rc) {
                    const double gbx = M->gp[0][o] * w + M->gp[0][n] * wm;
                    const double gby = M->gp[1][o] * w + M->gp[1][n] * wm;
                    const double gbz = M->gp[2][o] * w + M->gp[2][n] * wm;
                    const double ge = gbx * m->ex[f] + gby * m->ey[f] + gbz * m->ez[f];
                    massFlux += -dP_f * Sf * ((M->p[n] - M->p[o]) / delta - ge);
                }
                if (!nbSide) pending_put(pend, M, slot, f, -coeff, -coeff);
                ddg = -(-coeff);
                dsrc = nbSide ? massFlux : -massFlux;
            } else {
                const int b = f - nIF;
                const double Sf = m->area[f];
                const double mf = (M->Ubx[b] * m->nx[f] + M->Uby[b] * m->ny[f] + M->Ubz[b] * m->nz[f]) * Sf;
                dsrc = -mf;
                ddg = 0.0;
                if (m->bFlags[b] & RSG_BF_OUTLET) {
                    const double dP_o = (fabs(aP[c]) > 1e-30) ? m->vol[c] / aP[c] : 0.0;
                    ddg = dP_o * Sf / m->dlt[f];
                }
            }
            dg += ddg; src += dsrc;   // +0 increments are exact (accumulators are never -0)
        });
        pending_flush(pend, M);
        if (c == 0 && !m->hasOutlet) dg += fabs(dg) > 1e-30 ? dg : 1.0;
        M->diag[c] = dg;
        M->src[c] = src;
    }
    RSG_SYNC();
}

// assembleKEquation (:483-560) when isOmega == false, assembleOmegaEquation
// (:567-669) when true.
RSG_D void assemble_turb(Ctx& C, bool isOmega) {
    Thr& t = C.t; const RsgMeshDev* m = C.m; RsgMemberDev* M = C.M;
    const int nC = C.nC, nIF = C.nIF;
    const double nu = C.nu;
    const double s1 = isOmega ? C.sw1 : C.sk1, s2 = isOmega ? C.sw2 : C.sk2;
    const double* phi = isOmega ? M->om : M->k;
    const double* phib = isOmega ? M->omb : M->kb;
    const double a = isOmega ? C.s->aOm : C.s->aK;
    FOR_CELLS(c) {
        double dg = 0.0, src = 0.0;
        PendingLdu pend;
        pending_clear(pend);
        cell_faces(m, c, [&](int code, int other, int slot) {
            const int f = code >> 1;
            double ddg, dsrc = 0.0;
            if (f < nIF) {
                const bool nbSide = code & 1;
                const int o = nbSide ? other : c, n = nbSide ? c : other;
                const double w = m->w[f];
                const double F1f = w * M->F1[o] + (1.0 - w) * M->F1[n];
                const double sg = blend(s1, s2, F1f);
                double up, lo;
                conv_diff_face(m, M, f, o, n, nu + sg * M->nuT[o], nu + sg * M->nuT[n], up, lo);
                if (!nbSide) pending_put(pend, M, slot, f, up, lo);
                ddg = nbSide ? -up : -lo;
            } else {
                const int b = f - nIF;
                const double sg = blend(s1, s2, M->F1[c]);
                conv_diff_bface(m, M, f, b, nu + sg * M->nuT[c], phib[b], ddg, dsrc);
            }
            dg += ddg; src += dsrc;   // +0 increments are exact (accumulators are never -0)
        });
        pending_flush(pend, M);
        const double vol = m->vol[c];
        if (!isOmega) {
            src += M->Pk[c] * vol;
            const double destruction = C.bs * dmax(M->om[c], 1e-20);
            dg += destruction * vol;
        } else {
            const double F1 = M->F1[c];
            const double omC = dmax(M->om[c], 1e-20);
            const double alphaB = blend(C.al1, C.al2, F1);
            src += alphaB * sst_prod_om(C, M->nuT[c], M->Smag[c], M->k[c], M->om[c]) * vol;
            const double betaB = blend(C.b1, C.b2, F1);
            dg += betaB * omC * vol;
            src += (1.0 - F1) * M->CDkw[c] * vol;
        }
        src += (1.0 - a) / a * dg * phi[c];
        dg /= a;
        M->diag[c] = dg;
        M->src[c] = src;
        if (isOmega) M->omOld[c] = phi[c]; else M->kOld[c] = phi[c];
    }
    RSG_SYNC();
}

// velocity gradients of the listed components into M->g
RSG_D void grad_u(Ctx& C, int c0, int c1) {
    Thr& t = C.t; const RsgMeshDev* m = C.m; RsgMemberDev* M = C.M; const int nC = C.nC;
    for (int comp = c0; comp < c1; ++comp) {
        const double* const phi[1] = {ucomp(M, comp)};
        const double* const phib[1] = {ubcomp(M, comp)};
        FOR_CELLS(c) {
            double g[1][3];
            gg_cell<1>(m, c, phi, phib, g);
            M->g[comp * 3 + 0][c] = g[0][0]; M->g[comp * 3 + 1][c] = g[0][1]; M->g[comp * 3 + 2][c] = g[0][2];
        }
    }
}

RSG_D void grad_scalar(Ctx& C, const double* phi_, const double* phib_, double** out) {
    Thr& t = C.t; const RsgMeshDev* m = C.m; const int nC = C.nC;
    const double* const phi[1] = {phi_};
    const double* const phib[1] = {phib_};
    FOR_CELLS(c) {
        double g[1][3];
        gg_cell<1>(m, c, phi, phib, g);
        out[0][c] = g[0][0]; out[1][c] = g[0][1]; out[2][c] = g[0][2];
    }
}

// --------------------------------------------------------------- init
// initUniform SIMPLESolver.cpp:23-51 (+ applyAllBCs with beta1 = 0.075)
RSG_D void init_uniform_fields(Ctx& C) {
    Thr& t = C.t; const RsgMeshDev* m = C.m; RsgMemberDev* M = C.M;
    const int nC = C.nC, nB = C.nB;
    const double* I = M->init;
    FOR_CELLS(c) {
        M->Ux[c] = I[0]; M->Uy[c] = I[1]; M->Uz[c] = I[2];
        M->p[c] = I[3]; M->k[c] = I[4];
        M->F1[c] = 1.0; M->F2[c] = 1.0; M->Pk[c] = 0.0; M->CDkw[c] = 0.0;
        const double y = dmax(m->wd[c], 1e-20);
        const double omWall = 60.0 * C.nu / (C.b1 * y * y);
        const double om = dmax(omWall, I[5]);
        M->om[c] = om;
        M->nuT[c] = C.a1 * I[4] / dmax(C.a1 * om, 1e-20);
    }
    FOR_BFACES(b) {
        M->Ubx[b] = I[0]; M->Uby[b] = I[1]; M->Ubz[b] = I[2];
        M->pb[b] = I[3]; M->kb[b] = I[4]; M->omb[b] = I[5];
    }
}
RSG_D void init_uniform(Ctx& C) {
    init_uniform_fields(C);
    RSG_SYNC();
    vel_bc(C); p_bc(C); k_bc(C); om_bc(C, 0.075);
    RSG_SYNC();
}

// Step 1 of the outer iteration for cell c: Green-Gauss gradients of U, k,
// omega, |S| and the SST closure (computeFields, SSTModel.cpp:122-172), with
// the startup eddy-viscosity floor nuTMin.  kbGrad: boundary values of k for
// its gradient (null: the member's kb; the compressible solver hands the
// closure a copy of k whose boundary values are zero).
RSG_D void sst_closure_cell(Ctx& C, int c, double nuTMin, const double* kbGrad = nullptr) {
    const RsgMeshDev* m = C.m; RsgMemberDev* M = C.M; const RsgSettingsDev* st = C.s;
    const double nu = C.nu;
    const double* const phi[5] = {M->Ux, M->Uy, M->Uz, M->k, M->om};
    const double* const phib[5] = {M->Ubx, M->Uby, M->Ubz, kbGrad ? kbGrad : M->kb, M->omb};
    double g[5][3];
    gg_cell<5>(m, c, phi, phib, g);
#pragma unroll
    for (int q = 0; q < 3; ++q) { M->g[q * 3][c] = g[q][0]; M->g[q * 3 + 1][c] = g[q][1]; M->g[q * 3 + 2][c] = g[q][2]; }
    const double S11 = g[0][0], S22 = g[1][1], S33 = g[2][2];
    const double S12 = 0.5 * (g[0][1] + g[1][0]);
    const double S13 = 0.5 * (g[0][2] + g[2][0]);
    const double S23 = 0.5 * (g[1][2] + g[2][1]);
    const double Sm = sqrt(2.0 * (S11 * S11 + S22 * S22 + S33 * S33
                                  + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23)));
    M->Smag[c] = Sm;
    const double kc = dmax(M->k[c], 0.0);
    const double wc = dmax(M->om[c], 1e-20);
    const double y = dmax(m->wd[c], 1e-20);
    const double dotkw = g[3][0] * g[4][0] + g[3][1] * g[4][1] + g[3][2] * g[4][2];
    const double CDkw = 2.0 * C.sw2 / wc * dotkw;
    M->CDkw[c] = CDkw;
    const double CDpos = dmax(CDkw, 1e-20);
    const double F1 = (st->variant == 2) ? 1.0 : sst_F1(C, kc, wc, y, nu, CDpos);
    const double F2 = sst_F2(C, kc, wc, y, nu);
    M->F1[c] = F1; M->F2[c] = F2;
    const double nuT = sst_nut(C, kc, wc, Sm, F2);
    M->Pk[c] = sst_prod(C, nuT, Sm, kc, wc);
    M->nuT[c] = dmax(nuT, nuTMin);
}

// Step 4 (correctVelocity / correctPressure, SIMPLESolver.cpp:456-475) for cell c.
RSG_D void correct_cell(Ctx& C, int c) {
    const RsgMeshDev* m = C.m; RsgMemberDev* M = C.M;
    const double rAP = (fabs(M->aPx[c]) > 1e-30) ? 1.0 / M->aPx[c] : 0.0;
    const double vol = m->vol[c];
    M->Ux[c] -= rAP * M->gp[0][c] * vol;
    M->Uy[c] -= rAP * M->gp[1][c] * vol;
    M->Uz[c] -= rAP * M->gp[2][c] * vol;
    M->p[c] += C.s->aP * M->pp[c];
}

// Step 5's wall-cell omega pin (SIMPLESolver.cpp:884-908) for cell c and its
// terms of the field-change norms (:910-922): mx = max |k|, max |omega|,
// max |k - kOld|, max |omega - omegaOld| (running).
RSG_D void pin_cell(Ctx& C, int c, double (&mx)[4]) {
    const RsgMeshDev* m = C.m; RsgMemberDev* M = C.M; const RsgSettingsDev* st = C.s;
    const double nu = C.nu;
    const double y1 = m->pinY[c];
    if (y1 > 0.0) {
        const double omRes = 60.0 * nu / (C.b1 * y1 * y1);
        if (st->wallFn) {
            const double sbs = sqrt(C.bs);
            const double k_p = dmax(M->k[c], 1e-30);
            const double uTau = sqrt(sqrt(C.bs) * k_p);
            const double omLog = uTau / (sbs * st->vonK * y1);
            M->om[c] = sqrt(omRes * omRes + omLog * omLog);
        } else {
            M->om[c] = omRes;
        }
    }
    mx[0] = dmax(mx[0], fabs(M->k[c]));
    mx[1] = dmax(mx[1], fabs(M->om[c]));
    mx[2] = dmax(mx[2], fabs(M->k[c] - M->kOld[c]));
    mx[3] = dmax(mx[3], fabs(M->om[c] - M->omOld[c]));
}

// The per-launch context of member M for thread t: mesh, BCs, settings, nu
// and the 11 SST coefficients (InferenceParameters.hpp:12-14 order).
RSG_D void ctx_init(Ctx& C, Thr t, RsgMemberDev* M) {
    C.t = t; C.m = M->mesh; C.bc = M->bc; C.s = M->set; C.M = M;
    C.nC = C.m->nC; C.nIF = C.m->nIF; C.nB = C.m->nB; C.nu = M->nu;
    C.sk1 = M->coeffs[0]; C.sw1 = M->coeffs[1]; C.b1 = M->coeffs[2]; C.al1 = M->coeffs[3];
    C.sk2 = M->coeffs[4]; C.sw2 = M->coeffs[5]; C.b2 = M->coeffs[6]; C.al2 = M->coeffs[7];
    C.bs = M->coeffs[8]; C.a1 = M->coeffs[9]; C.kap = M->coeffs[10];
}

// ------------------------------------------------------- outer iteration
struct Scal {
    int iter, status, turbEstablished, finalIter, amgBuiltIter, coarseIter;
    double normUx0, normUy0, normP0, lastK, lastOm;
};

template <int CMAX, int PREC = -1>
RSG_D void simple_iteration(Ctx& C, Scal& S) {
    Thr& t = C.t; const RsgMeshDev* m = C.m; RsgMemberDev* M = C.M; const RsgSettingsDev* st = C.s;
    const int nC = C.nC, nIF = C.nIF, nB = C.nB;
    const double nu = C.nu;
    const int iter = S.iter;
    const bool turbActive = iter >= st->turbStart;
    const bool turbUpdate = turbActive && ((iter - st->turbStart) % st->turbInterval == 0);

    // 1. SST closure (computeFields) + frozen |S| + startup nuT floor
    if (turbUpdate) {
        if (!S.turbEstablished && iter >= st->turbStart + st->nutFloorIters) S.turbEstablished = 1;
        const double nuTMin = S.turbEstablished ? 0.0 : 0.1 * nu;
        FOR_CELLS(c) sst_closure_cell(C, c, nuTMin);
    } else {
        grad_u(C, 0, 3);
    }
    grad_scalar(C, M->p, M->pb, M->gp);
    RSG_SYNC();
    RSG_TMARK(0);

    // 2. momentum components, each assembled from the latest U
    SolveRes rU0 = {0, 0.0, 0.0, 0}, rU1 = rU0, rU2 = rU0;   // named, not indexed: no private-memory array
    for (int comp = 0; comp < 3; ++comp) {
        if (comp == 2 && M->planarZ) break;   // Uz stays 0: the z system's residual is exactly 0
        if (comp > 0) { grad_u(C, comp - 1, comp); RSG_SYNC(); }
        assemble_momentum(C, comp);
        RSG_TMARK(1);
        const SolveRes r = lsolve<CMAX, PREC>(C, st->mSolver, ucomp(M, comp), false);
        if (comp == 0) rU0 = r; else if (comp == 1) rU1 = r; else rU2 = r;
        RSG_SYNC();
        RSG_TMARK(2);
    }
    vel_bc(C);
    RSG_SYNC();

    // 3. pressure correction
    assemble_pcorr(C);
    RSG_TMARK(3);
    bool amgReady = false;
    if (st->pSolver == 0) {
        // Aggregation every amgRefresh iterations; the coarse Galerkin values
        // and the coarsest inverse every coarseRefresh iterations (between
        // refreshes the V-cycle's coarse levels lag the current matrix, which
        // changes the preconditioner, not the system being solved).
        const bool rebuild = S.amgBuiltIter < 0 || iter - S.amgBuiltIter >= st->amgRefresh;
        const bool coarse = rebuild || iter - S.coarseIter >= st->coarseRefresh;
        if (rebuild) {
            amg_refresh_values(C, false);
            RSG_TMARK(4);
            if (t.tid == 0) amg_setup_serial(m, M->amg);
            RSG_SYNC();
            S.amgBuiltIter = iter;
            RSG_TMARK(5);
        } else {
            // level 0's CSR copy feeds only the global-memory V-cycle and the Galerkin products
            const bool resident = res_fits(C, CMAX) && res_amg_fits(C);
            if (coarse || !resident) amg_refresh_values(C, coarse);
            RSG_TMARK(4);
        }
        if (coarse) {
            amg_coarse_inverse(C);
            S.coarseIter = iter;
            RSG_TMARK(6);
        }
        amgReady = true;
    }
    FOR_CELLS(c) M->pp[c] = 0.0;
    RSG_SYNC();
    SolveRes rP = lsolve<CMAX, PREC>(C, st->pSolver, M->pp, amgReady);
    RSG_SYNC();
    RSG_TMARK(7);
    FOR_BFACES(b) M->ppb[b] = (m->bFlags[b] & RSG_BF_OUTLET) ? 0.0 : M->pp[m->own[nIF + b]];
    RSG_SYNC();

    // 4. correctVelocity / correctPressure
    grad_scalar(C, M->pp, M->ppb, M->gp);
    FOR_CELLS(c) correct_cell(C, c);
    RSG_SYNC();
    vel_bc(C); p_bc(C);
    RSG_SYNC();
    RSG_TMARK(8);

    // 5. turbulence transport
    SolveRes rK = {0, 0.0, 0.0, 0}, rOm = {0, 0.0, 0.0, 0};
    if (turbUpdate) {
        assemble_turb(C, false);
        RSG_TMARK(9);
        rK = lsolve<CMAX, PREC>(C, st->tSolver, M->k, false);
        RSG_SYNC();
        RSG_TMARK(10);
        FOR_CELLS(c) M->k[c] = dmax(st->kMin, dmin(st->kMax, M->k[c]));
        RSG_SYNC();
        k_bc(C);
        RSG_SYNC();
        assemble_turb(C, true);
        RSG_TMARK(11);
        // The omega system stays FP64 in mixed members: it carries the
        // near-wall treatment (wall cells pinned at 60 nu / (beta1 y^2), 10^5-10^8
        // times the interior omega), and an FP32 solve cannot resolve the
        // interior corrections next to those rows -- its noise floor holds the
        // outer residuals above 1e-8 (measured: a Re_tau 181 member stalls at
        // p' ~5e-8 with FP32 omega solves and converges in the FP64 solve's
        // iteration count, within 2e-11 of it, with FP64 ones).
        rOm = lsolve<CMAX, PREC>(C, st->tSolver, M->om, false, false);
        RSG_SYNC();
        RSG_TMARK(12);
        FOR_CELLS(c) M->om[c] = dmax(st->omMin, dmin(1e15, M->om[c]));
        RSG_SYNC();
        om_bc(C, C.b1);
        RSG_SYNC();
        // wall-cell omega pin (:884-908) and field-change norms (:910-922)
        double mx[4] = {1e-30, 1e-30, 0.0, 0.0};
        FOR_CELLS(c) pin_cell(C, c, mx);
        bmax(t, mx);
        S.lastK = mx[2] / mx[0];
        S.lastOm = mx[3] / mx[1];
    }

    // 6. residual bookkeeping and convergence / divergence tests (:934-1035)
    const double eUx = rU0.initialRes, eUy = rU1.initialRes, eUz = rU2.initialRes;
    const double eP = rP.initialRes;
    const double eK = turbActive ? S.lastK : 0.0, eOm = turbActive ? S.lastOm : 0.0;
    if (t.tid == 0) {
        RsgScalars* sc = M->sc;
        sc->innerIters[0] += rU0.iterations; sc->innerIters[1] += rU1.iterations;
        sc->innerIters[2] += rU2.iterations; sc->innerIters[3] += rP.iterations;
        sc->innerIters[4] += rK.iterations; sc->innerIters[5] += rOm.iterations;
        if (st->recordHistory && M->hist && iter < M->histCap) {
            double* h = M->hist + 6 * (long)iter;
            h[0] = eUx; h[1] = eUy; h[2] = eUz; h[3] = eP; h[4] = eK; h[5] = eOm;
            sc->histLen = iter + 1;
        }
    }
    if (iter == 0) {
        S.normUx0 = dmax(eUx, 1e-30); S.normUy0 = dmax(eUy, 1e-30); S.normP0 = dmax(eP, 1e-30);
    } else if (iter < 50) {
        S.normP0 = dmax(S.normP0, eP);
    }
    const double normMom0 = dmax(S.normUx0, S.normUy0);
    const double nUx = normMom0 < 1e-20 ? 0.0 : eUx / normMom0;
    const double nUy = normMom0 < 1e-20 ? 0.0 : eUy / normMom0;
    const double nP = S.normP0 < 1e-20 ? 0.0 : eP / S.normP0;
    double maxRes = dmax(dmax(nUx, nUy), nP);
    if (turbActive) maxRes = dmax(dmax(maxRes, S.lastK), S.lastOm);
    if (!isfinite(eUx) || !isfinite(eUy) || !isfinite(eUz) || !isfinite(eP) || !isfinite(eK) ||
        !isfinite(eOm) || !isfinite(S.lastK) || !isfinite(S.lastOm))
        maxRes = INFINITY;
    const bool turbScheduled = st->turbStart < st->maxIter;
    S.iter = iter + 1;
    if (maxRes < st->tol && iter > 0 && (!turbScheduled || (turbActive && S.turbEstablished))) {
        S.status = RSGM_CONVERGED; S.finalIter = iter;
    } else if (isnan(maxRes) || isinf(maxRes) || maxRes > st->divLimit) {
        S.status = RSGM_DIVERGED; S.finalIter = iter;
    } else if (S.iter >= st->maxIter) {
        S.status = RSGM_MAXITER; S.finalIter = st->maxIter;
    }
    RSG_SYNC();
    RSG_TMARK(13);
}

// Runs up to M->iterBudget outer iterations of one member, and no longer than
// M->sliceTicks of wall clock when that is set (the iteration in progress
// finishes; a member's results do not depend on where launches split it).
// ObservationOperator::evaluate and logLikelihood (ObservationOperator.hpp:
// 159-378) on the member's final fields, by one thread in the reference's
// operation order (so the predictions are bit-identical to the host's).  The
// mesh-only searches (nearest cell, nearest face of the patch) and the patch
// face lists come resolved from the host; pred[n] gets the log-likelihood.
RSG_D void observe_member(const RsgMemberDev* M) {
    const RsgObsDev* O = M->obs;
    const RsgMeshDev* m = M->mesh;
    const int nIF = m->nIF;
    const double nu = M->nu;
    double ll = 0.0;
    for (int i = 0; i < O->n; ++i) {
        const double* par = O->par + 6 * i;
        const double refU = par[0];
        double out = 0.0;
        switch (O->type[i]) {
            case RSGO_DRAG: {                                   // computeDrag :186-223
                double force = 0.0;
                for (int j = O->fb[i]; j < O->fe[i]; ++j) {
                    const int f = O->faces[j], ow = m->own[f];
                    const double Sf = m->area[f];
                    force += M->p[ow] * m->nx[f] * Sf;
                    const double tau_x = nu * (M->Ux[ow] - M->Ubx[f - nIF]) / m->dlt[f];
                    force += tau_x * Sf;
                }
                const double dyn = 0.5 * refU * refU;
                out = force / (dyn * par[1]);
                break;
            }
            case RSGO_PRESSURE: {                               // computePressureAt :261-271
                const double dyn = 0.5 * par[2] * refU * refU;
                out = (M->p[O->idx[i]] - par[3]) / dmax(dyn, 1e-20);
                break;
            }
            case RSGO_VELOCITY: {                               // computeVelocityAt :251-258
                const int c = O->idx[i], q = O->comp[i];
                out = q == 0 ? M->Ux[c] : q == 1 ? M->Uy[c] : M->Uz[c];
                break;
            }
            case RSGO_SKIN_FRICTION: {                          // computeSkinFrictionAt :226-248
                const int f = O->idx[i];
                if (f < 0) break;
                const int ow = m->own[f];
                const double ux = M->Ux[ow], uy = M->Uy[ow], uz = M->Uz[ow];
                const double nx = m->nx[f], ny = m->ny[f], nz = m->nz[f];
                const double dn = ux * nx + uy * ny + uz * nz;
                const double tx = ux - nx * dn, ty = uy - ny * dn, tz = uz - nz * dn;
                const double tau = nu * sqrt(tx * tx + ty * ty + tz * tz) / m->dlt[f];
                out = tau / dmax(0.5 * refU * refU, 1e-20);
                break;
            }
            case RSGO_SEPARATION:                               // computeSeparationPoint :277-307
            case RSGO_REATTACHMENT: {                           // computeReattachmentLength :311-340
                const bool sep = O->type[i] == RSGO_SEPARATION;
                out = sep ? 0.0 : (O->fe[i] > O->fb[i] ? O->fx[O->fe[i] - 1] : 0.0);
                bool havePrev = false;
                double prevTau = sep ? 1.0 : -1.0, prevX = 0.0;
                for (int j = O->fb[i]; j < O->fe[i]; ++j) {
                    const int f = O->faces[j], ow = m->own[f];
                    const double ux = M->Ux[ow], uy = M->Uy[ow], uz = M->Uz[ow];
                    const double nx = m->nx[f];
                    const double dn = ux * nx + uy * m->ny[f] + uz * m->nz[f];
                    const double tau = nu * (ux - nx * dn) / m->dlt[f];
                    const double x = O->fx[j];
                    if (sep && havePrev && prevTau > 0 && tau <= 0) {
                        const double denom = prevTau - tau;
                        const double frac = (fabs(denom) > 1e-30) ? prevTau / denom : 0.0;
                        out = prevX + frac * (x - prevX);
                        break;
                    }
                    if (!sep && havePrev && prevTau < 0 && tau >= 0) {
                        const double denom = tau - prevTau;
                        const double frac = (fabs(denom) > 1e-30) ? (-prevTau) / denom : 0.0;
                        out = prevX + frac * (x - prevX);
                        break;
                    }
                    prevTau = tau; prevX = x; havePrev = true;
                }
                break;
            }
            default: break;
        }
        M->pred[i] = out;
        const double diff = out - par[4];                       // logLikelihood :364-378
        ll += -0.5 * diff * diff / par[5];
    }
    M->pred[O->n] = ll;
}

template <int CMAX, int PREC = -1>
RSG_D void run_member(Thr t, RsgMemberDev* M) {
    Ctx C;
    ctx_init(C, t, M);
    RsgScalars* sc = M->sc;
    if (sc->status != RSGM_RUNNING) return;
    if (!sc->initialized) {
        if (!M->initFromFields) init_uniform(C);
        // level-0 pattern of the AMG hierarchy
        RsgAmgDev& A = M->amg;
        const RsgMeshDev* m = C.m;
        for (int i = C.t.tid; i <= m->nC; i += C.t.bs) A.rp[i] = m->rp0[i];
        for (int e = C.t.tid; e < m->nnz0; e += C.t.bs) A.ci[e] = m->ci0[e];
        if (C.t.tid == 0) {
            A.lvN[0] = m->nC; A.lvRow[0] = 0; A.lvRp[0] = 0; A.lvNz[0] = 0;
            A.lvN[RSG_MAX_LEVELS] = 1;
        }
        RSG_SYNC();
        if (C.t.tid == 0) { sc->initialized = 1; sc->amgBuiltIter = -1; }
        RSG_SYNC();
    }
    Scal S;
    S.iter = sc->iter; S.status = sc->status; S.turbEstablished = sc->turbEstablished;
    S.finalIter = sc->finalIter; S.amgBuiltIter = sc->amgBuiltIter; S.coarseIter = sc->coarseIter;
    S.normUx0 = sc->normUx0; S.normUy0 = sc->normUy0; S.normP0 = sc->normP0;
    S.lastK = sc->lastK; S.lastOm = sc->lastOm;
    RSG_SYNC();
#if defined(RSG_PROF) && !defined(RSG_EMULATE)
    C.tmark = wall_clock64();
#endif
    // the coarsest dense inverse lives in LDS: restore it for this launch
    // (a team reads it from global memory)
#ifndef RSG_TEAM
    if (S.amgBuiltIter >= 0 && sc->coarseDense) {
        const RsgAmgDev& A = M->amg;
        const int nL = A.lvN[A.lvN[RSG_MAX_LEVELS] - 1];
        for (int idx = C.t.tid; idx < nL * RSG_COARSE_DENSE; idx += C.t.bs) C.t.lds[idx] = A.cinv[idx];
        RSG_SYNC();
    }
#endif
#ifndef RSG_EMULATE
    const long long tStart = wall_clock64();
#endif
    for (int b = 0; b < M->iterBudget && S.status == RSGM_RUNNING; ++b) {
        simple_iteration<CMAX, PREC>(C, S);
#if !defined(RSG_EMULATE) && !defined(RSG_TEAM)
        // time slice: thread 0 reads the clock and every thread follows its
        // decision, so the workgroup leaves the loop together (a team runs
        // its fixed iteration budget)
        if (M->sliceTicks > 0) {
            if (C.t.tid == 0) C.t.red[0] = (wall_clock64() - tStart >= M->sliceTicks) ? 1.0 : 0.0;
            RSG_SYNC();
            const bool stop = C.t.red[0] != 0.0;
            RSG_SYNC();
            if (stop) break;
        }
#endif
    }
    // a finished member's observables, from the fields every thread wrote
    if (M->obs && S.status != RSGM_RUNNING) {
        RSG_SYNC();
        if (C.t.tid == 0) observe_member(M);
    }
    if (C.t.tid == 0) {
        sc->iter = S.iter; sc->status = S.status; sc->turbEstablished = S.turbEstablished;
        sc->finalIter = S.finalIter; sc->amgBuiltIter = S.amgBuiltIter; sc->coarseIter = S.coarseIter;
        sc->normUx0 = S.normUx0; sc->normUy0 = S.normUy0; sc->normP0 = S.normP0;
        sc->lastK = S.lastK; sc->lastOm = S.lastOm;
        sc->nLevels = M->amg.lvN[RSG_MAX_LEVELS];
    }
}

// One linear solve (ILinearSolver::solve, LinearSolver.hpp:115-120) on the
// LDU system already stored in M (diag/upper/lower/src), x = M->pp in/out.
// solver: RSG_SOLVER_* codes (0 AMG-PCG, 1 Jacobi-PCG, 2 BiCGSTAB, 3 Gauss-Seidel).
template <int CMAX>
RSG_D SolveRes ldu_solve_member(Thr t, RsgMemberDev* M, int solver, int maxIter, double tol, int* nLevels) {
    Ctx C;
    C.t = t; C.m = M->mesh; C.bc = M->bc; C.s = M->set; C.M = M;
    C.nC = C.m->nC; C.nIF = C.m->nIF; C.nB = C.m->nB; C.nu = 0.0;
    bool amg = false;
    if (solver == 0) {
        RsgAmgDev& A = M->amg;
        const RsgMeshDev* m = C.m;
        for (int i = t.tid; i <= m->nC; i += t.bs) A.rp[i] = m->rp0[i];
        for (int e = t.tid; e < m->nnz0; e += t.bs) A.ci[e] = m->ci0[e];
        if (t.tid == 0) { A.lvN[0] = m->nC; A.lvRow[0] = 0; A.lvRp[0] = 0; A.lvNz[0] = 0; A.lvN[RSG_MAX_LEVELS] = 1; }
        RSG_SYNC();
        amg_refresh_values(C, false);
        if (t.tid == 0) amg_setup_serial(m, A);
        RSG_SYNC();
        amg_coarse_inverse(C);
        amg = true;
    }
    SolveRes r;
    if (solver == RSG_SOLVER_GAUSS_SEIDEL) r = gauss_seidel(C, M->pp, maxIter, tol);
    else if (res_fits(C, CMAX) && (!amg || res_amg_fits(C))) r = lsolve_res<CMAX>(C, solver, M->pp, amg, maxIter, tol);
    else if (is_mixed<-1>(C)) r = lsolve_mixed_global(C, solver, M->pp, amg, maxIter, tol);
    else if (solver == 2) r = bicgstab(C, M->pp, maxIter, tol);
    else r = pcg(C, M->pp, maxIter, tol, amg);
    RSG_SYNC();
    *nLevels = M->amg.lvN[RSG_MAX_LEVELS];
    return r;
}

// One linear solve of the member's current LDU system into x (in/out) on one
// workgroup: the member path's solvers, for a kernel-per-phase member whose
// mesh is small enough that a grid-wide Krylov iteration (a host round trip
// per dot product) costs more than one workgroup's.  AMG (solver 0):
// amgMode bit 4 copies the level-0 pattern (first use), bit 1 aggregates the
// hierarchy anew, bit 2 refreshes its coarse values and the coarsest
// inverse; otherwise the LDS copy of the coarsest inverse is restored.
template <int CMAX>
RSG_D SolveRes member_solve_x(Thr t, RsgMemberDev* M, int solver, double* x, int amgMode, int maxIter,
                              double tol) {
    Ctx C;
    ctx_init(C, t, M);
    bool amg = false;
    if (solver == 0) {
        RsgAmgDev& A = M->amg;
        const RsgMeshDev* m = C.m;
        if (amgMode & 4) {
            for (int i = t.tid; i <= m->nC; i += t.bs) A.rp[i] = m->rp0[i];
            for (int e = t.tid; e < m->nnz0; e += t.bs) A.ci[e] = m->ci0[e];
            if (t.tid == 0) { A.lvN[0] = m->nC; A.lvRow[0] = 0; A.lvRp[0] = 0; A.lvNz[0] = 0; A.lvN[RSG_MAX_LEVELS] = 1; }
            RSG_SYNC();
        }
        const bool rebuild = (amgMode & 1) != 0, coarse = rebuild || (amgMode & 2) != 0;
        if (rebuild) {
            amg_refresh_values(C, false);
            if (t.tid == 0) amg_setup_serial(m, A);
            RSG_SYNC();
        } else {
            const bool resident = res_fits(C, CMAX) && res_amg_fits(C);
            if (coarse || !resident) amg_refresh_values(C, coarse);
        }
        if (coarse) {
            amg_coarse_inverse(C);
        } else if (M->sc->coarseDense) {
            const int nL = A.lvN[A.lvN[RSG_MAX_LEVELS] - 1];
            for (int idx = t.tid; idx < nL * RSG_COARSE_DENSE; idx += t.bs) t.lds[idx] = A.cinv[idx];
            RSG_SYNC();
        }
        amg = true;
    }
    SolveRes r;
    if (solver == RSG_SOLVER_GAUSS_SEIDEL) r = gauss_seidel(C, x, maxIter, tol);
    else if (res_fits(C, CMAX) && (!amg || res_amg_fits(C))) r = lsolve_res<CMAX>(C, solver, x, amg, maxIter, tol, false);
    else if (solver == 2) r = bicgstab(C, x, maxIter, tol);
    else r = pcg(C, x, maxIter, tol, amg);
    RSG_SYNC();
    return r;
}

#include "rsg_comp_dev.h"

#undef FOR_CELLS
#undef FOR_IFACES
#undef FOR_BFACES
}  // namespace rsg
