// rsg_big.hip -- one large member (10^5..10^7 cells) solved by kernel-per-phase
// launches over the whole GPU.
//
// The ensemble kernels give a member one workgroup; a member far beyond one
// workgroup's reach instead runs every phase of the outer iteration as its
// own grid-wide launch: the per-cell and per-face code of rsg_member.h
// (assembly, Green-Gauss, SST closure, boundary values, corrections, the wall
// omega pin) with one thread per cell, and the linear solves on the CSR image
// of each assembled LDU system with the grid kernels of rsg_grid.h
// (deterministic block-partial dot products).  Between phases the GPU drains
// one launch and starts the next (a few microseconds), instead of the
// cooperative team's grid barrier (tens of microseconds), and each phase
// streams its arrays at HBM rate.
//
// The p' AMG hierarchy is aggregated on the host (the reference's greedy
// pass, rsg::grid::coarsen) from the p' matrix of the outer iteration that
// needs it, every amg_refresh outer iterations; level 0's values follow the
// current matrix every iteration and the coarse Galerkin values are refreshed
// on the device every coarse_refresh iterations (gather lists in the
// reference's summation order), like the member path's device AMG.
//
// Reference restated (as rsg_member.h): SIMPLESolver::solve
// (incompressible/src/SIMPLESolver.cpp:731-1036) phase by phase, the
// BiCGSTAB / PCG / AMG-PCG solvers of core/include/LinearSolver.hpp:147-247,
// 497-656.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "rsg.h"
#include "rsg_layout.h"
#include "rsg_prep.h"
#include "rsg_member.h"
#include "rsg_grid.h"
#include "rsg_big.h"

// a member's linear solve on one workgroup (rsg_chunk.hip)
hipError_t rsg_member_solve_launch(RsgMemberDev* M, int solver, double* x, int amgMode, int maxIter, double tol,
                                   double* out, hipStream_t s);

namespace rsg {
namespace big {

using grid::kBlock;

__device__ inline Ctx gctx(RsgMemberDev* M) {
    Thr t;
    t.tid = blockIdx.x * blockDim.x + threadIdx.x;
    t.bs = gridDim.x * blockDim.x;
    t.red = nullptr; t.lds = nullptr; t.phase = 0;
    t.xb0 = nullptr; t.xb1 = nullptr; t.xsel = 0; t.cvec = nullptr; t.cvecCap = 0;
    t.ceVal = nullptr; t.ceCol = nullptr; t.ceCap = 0; t.caCol = nullptr; t.caCap = 0;
    t.ltid = threadIdx.x; t.lbs = blockDim.x; t.blk = blockIdx.x; t.nblk = gridDim.x; t.gred = nullptr;
    Ctx C;
    ctx_init(C, t, M);
    return C;
}

#define BIG_CELLS(c) for (int c = C.t.tid; c < C.nC; c += C.t.bs)
#define BIG_BFACES(b) for (int b = C.t.tid; b < C.nB; b += C.t.bs)

// initUniform's field values (SIMPLESolver.cpp:23-51); the boundary values follow in k_bcs
__global__ __launch_bounds__(kBlock) void k_init(RsgMemberDev* M) {
    Ctx C = gctx(M);
    init_uniform_fields(C);
}

// boundary values: bit 1 velocity, 2 pressure, 4 k, 8 omega (with beta1)
__global__ __launch_bounds__(kBlock) void k_bcs(RsgMemberDev* M, int mask, double beta1) {
    Ctx C = gctx(M);
    if (mask & 1) vel_bc(C);
    if (mask & 2) p_bc(C);
    if (mask & 4) k_bc(C);
    if (mask & 8) om_bc(C, beta1);
}

// step 1: SST closure (or the velocity gradients alone) and grad p
__global__ __launch_bounds__(kBlock) void k_step1(RsgMemberDev* M, int turbUpdate, double nuTMin) {
    Ctx C = gctx(M);
    if (turbUpdate) {
        BIG_CELLS(c) sst_closure_cell(C, c, nuTMin);
    } else {
        grad_u(C, 0, 3);
    }
    grad_scalar(C, M->p, M->pb, M->gp);
}

__global__ __launch_bounds__(kBlock) void k_grad_u(RsgMemberDev* M, int c0, int c1) {
    Ctx C = gctx(M);
    grad_u(C, c0, c1);
}

__global__ __launch_bounds__(kBlock) void k_asm_mom(RsgMemberDev* M, int comp) {
    Ctx C = gctx(M);
    assemble_momentum(C, comp);
}

__global__ __launch_bounds__(kBlock) void k_asm_pc(RsgMemberDev* M) {
    Ctx C = gctx(M);
    assemble_pcorr(C);
    BIG_CELLS(c) M->pp[c] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_asm_turb(RsgMemberDev* M, int isOmega) {
    Ctx C = gctx(M);
    assemble_turb(C, isOmega != 0);
}

// p' boundary values (zero at outlets)
__global__ __launch_bounds__(kBlock) void k_ppb(RsgMemberDev* M) {
    Ctx C = gctx(M);
    const RsgMeshDev* m = C.m;
    BIG_BFACES(b) M->ppb[b] = (m->bFlags[b] & RSG_BF_OUTLET) ? 0.0 : M->pp[m->own[C.nIF + b]];
}

// step 4: grad p' is current; velocity and pressure corrections
__global__ __launch_bounds__(kBlock) void k_grad_pp(RsgMemberDev* M) {
    Ctx C = gctx(M);
    grad_scalar(C, M->pp, M->ppb, M->gp);
}

__global__ __launch_bounds__(kBlock) void k_correct(RsgMemberDev* M) {
    Ctx C = gctx(M);
    BIG_CELLS(c) correct_cell(C, c);
}

// k / omega clamps after their solves (SIMPLESolver.cpp:861, :871)
__global__ __launch_bounds__(kBlock) void k_clamp(RsgMemberDev* M, int isOmega) {
    Ctx C = gctx(M);
    const RsgSettingsDev* st = C.s;
    if (isOmega) { BIG_CELLS(c) M->om[c] = dmax(st->omMin, dmin(1e15, M->om[c])); }
    else { BIG_CELLS(c) M->k[c] = dmax(st->kMin, dmin(st->kMax, M->k[c])); }
}

// wall omega pin and the block maxima of the change norms' terms
__global__ __launch_bounds__(kBlock) void k_pin(RsgMemberDev* M, double* part) {
    __shared__ double sh[4][kBlock / 64];
    Ctx C = gctx(M);
    double mx[4] = {1e-30, 1e-30, 0.0, 0.0};
    BIG_CELLS(c) pin_cell(C, c, mx);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double v = mx[q];
        for (int off = 32; off >= 1; off >>= 1) v = dmax(v, __shfl_xor(v, off, 64));
        if (lane == 0) sh[q][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = sh[threadIdx.x][0];
        for (int q = 1; q < kBlock / 64; ++q) v = dmax(v, sh[threadIdx.x][q]);
        part[4 * blockIdx.x + threadIdx.x] = v;
    }
}

// out[q] = max over the nb block partials of value q (one workgroup)
__global__ __launch_bounds__(kBlock) void k_max4(int nb, const double* part, double* out) {
    __shared__ double sh[4][kBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double v = part[q];
        for (int b = threadIdx.x; b < nb; b += kBlock) v = dmax(v, part[4 * b + q]);
        for (int off = 32; off >= 1; off >>= 1) v = dmax(v, __shfl_xor(v, off, 64));
        if (lane == 0) sh[q][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = sh[threadIdx.x][0];
        for (int q = 1; q < kBlock / 64; ++q) v = dmax(v, sh[threadIdx.x][q]);
        out[threadIdx.x] = v;
    }
}

// CSR values (level-0 pattern of the mesh, lduToCSR order) of the member's
// current LDU system, and its diagonal
__global__ __launch_bounds__(kBlock) void k_csr_fill(RsgMemberDev* M, double* v, double* dg) {
    Ctx C = gctx(M);
    const RsgMeshDev* m = C.m;
    const int nnz = m->nnz0, nIF = C.nIF;
    for (int e = C.t.tid; e < nnz; e += C.t.bs) {
        const int s = m->src0[e];
        v[e] = s < 0 ? M->diag[-1 - s] : (s < nIF ? M->upper[s] : M->lower[s - nIF]);
    }
    BIG_CELLS(c) dg[c] = M->diag[c];
}

// coarse Galerkin values of one level: entry E sums its fine entries (gather
// list, ascending fine entry = the reference's summation order)
__global__ __launch_bounds__(kBlock) void k_galerkin(int nnzc, const int* gP, const int* gI, const double* fv,
                                                     double* cv) {
    const int E = blockIdx.x * kBlock + threadIdx.x;
    if (E >= nnzc) return;
    double s = 0.0;
    for (int q = gP[E]; q < gP[E + 1]; ++q) s += fv[gI[q]];
    cv[E] = s;
}

__global__ __launch_bounds__(kBlock) void k_diag_of(int n, const int* rp, const int* ci, const double* v,
                                                    double* dg) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; ++e)
        if (ci[e] == i) d = v[e];
    dg[i] = d;
}

// dense Gauss-Jordan inverse of the coarsest level (<= 64 rows), one workgroup
__global__ __launch_bounds__(kBlock) void k_coarse_inverse(int n, const int* rp, const int* ci, const double* v,
                                                           double* cinv, int* ok) {
    __shared__ double G[RSG_COARSE_DENSE * RSG_COARSE_DENSE];
    const int good = coarse_dense_block(G, rp, ci, v, n, threadIdx.x, blockDim.x);
    for (int idx = threadIdx.x; idx < n * RSG_COARSE_DENSE; idx += blockDim.x) cinv[idx] = G[idx];
    if (threadIdx.x == 0) *ok = good;
}

// BiCGSTAB phases (the recurrence of rsg_member.h's bicgstab); scalars come
// in FP64 and are rounded to T once
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bicg_p(int n, double beta, double omega, const T* r, const T* v,
                                                   const T* dg, T* p, T* ph) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const T pc = r[i] + (T)beta * (p[i] - (T)omega * v[i]);
    p[i] = pc;
    ph[i] = jac<T>(pc, dg[i]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bicg_s(int n, double alpha, const T* r, const T* v, const T* dg, T* s,
                                                   T* sh) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const T sc = r[i] - (T)alpha * v[i];
    s[i] = sc;
    sh[i] = jac<T>(sc, dg[i]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bicg_x(int n, double alpha, double omega, const T* ph, const T* sh,
                                                   const T* s, const T* t, T* x, T* r) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    x[i] += (T)alpha * ph[i] + (T)omega * sh[i];
    r[i] = s[i] - (T)omega * t[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpy_t(int n, double a, const T* p, T* x) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] += (T)a * p[i];
}

// FP32 image of an FP64 CSR level in row_ax's form (off-diagonal values, the
// FP64 row sum rounded once, the diagonal)
__global__ __launch_bounds__(kBlock) void k_to_f32(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                   const double* __restrict__ v, const double* __restrict__ dg,
                                                   float* __restrict__ vf, float* __restrict__ exf,
                                                   float* __restrict__ dgf) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double rs = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
        rs += v[e];
        vf[e] = (ci[e] == i) ? 0.0f : (float)v[e];
    }
    exf[i] = (float)rs;
    dgf[i] = (float)dg[i];
}

// bf = r / rn (the defect scaled to unit norm), ef = 0
__global__ __launch_bounds__(kBlock) void k_scale_f32(int n, const double* r, double inv, float* bf, float* ef) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    bf[i] = (float)(r[i] * inv);
    ef[i] = 0.0f;
}

// x += rn ef (FP64)
__global__ __launch_bounds__(kBlock) void k_refine(int n, double rn, const float* ef, double* x) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] += rn * (double)ef[i];
}

__global__ __launch_bounds__(kBlock) void k_copy(int n, const double* a, double* b) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// one Gauss-Seidel sweep of the member's LDU system, one thread (sequential)
__global__ void k_gs_sweep(RsgMemberDev* M, double* x) {
    if (blockIdx.x == 0 && threadIdx.x == 0) gs_sweep(M->mesh, M, x);
}

__global__ void k_observe(RsgMemberDev* M) {
    if (blockIdx.x == 0 && threadIdx.x == 0) observe_member(M);
}

#undef BIG_CELLS
#undef BIG_BFACES

inline int cells_grid(int n) { return std::max(1, std::min((n + kBlock - 1) / kBlock, 1 << 16)); }

}  // namespace big
}  // namespace rsg

using namespace rsg;
using namespace rsg::big;

constexpr int kSolveSlots = 8;   // one-workgroup solves of an outer iteration (memberSolve members)

// Host state of one large member.
struct RsgBig {
    RsgMemberDev hM{};            // host copy of the member struct (device pointers)
    RsgMemberDev* dM = nullptr;   // the struct on the device
    RsgSettingsDev set{};
    const PrepMesh* pm = nullptr;
    int nC = 0, nnz0 = 0;
    const int* rp0 = nullptr;     // device level-0 pattern (the mesh's)
    const int* ci0 = nullptr;
    hipStream_t s = nullptr;
    bool ownStream = false;
    grid::DevAlloc A;             // work vectors, scratch
    double *v0 = nullptr, *dg0 = nullptr, *vz = nullptr, *vq = nullptr;
    grid::Scratch S;
    // p' AMG: hierarchy and the device gather lists of its coarse levels
    std::unique_ptr<grid::DevAlloc> HA;
    grid::Hierarchy<double> H;
    std::vector<int*> gP, gI;
    int* dOk = nullptr;
    bool haveH = false;
    int maxLevels = 0;            // 0: the reference's cap; else levels before an MFMA-inverted coarsest
    // mixed precision (a member submitted with RSG_PREC_MIXED): FP32 images of
    // the level-0 system and of the p' hierarchy, FP32 Krylov vectors
    bool mixed = false;
    double* r64 = nullptr;        // FP64 defect
    float *v0f = nullptr, *ex0f = nullptr, *dg0f = nullptr;
    float* fw = nullptr;          // 12 FP32 vectors of nC
    std::unique_ptr<grid::DevAlloc> HFA;
    grid::Hierarchy<float> Hf;
    grid::VcycleGraphs VG, VGf;   // p' V-cycles as hipGraphs (FP64, FP32 mirror)
    grid::DenseWork DW;           // the coarsest level's inverse when it is above 64 rows
    bool bigDense = false;
    double denseMs = 0.0;         // device time of the inversions so far
    // compressible member (rsg_comp.h): its extra fields and per-iteration maxima
    bool comp = false;
    RsgCompDev hQ{};
    RsgCompDev* dQ = nullptr;
    double* compPart = nullptr;
    double compTIters = 0.0;      // energy-equation inner iterations so far
    // small meshes: every linear solve one workgroup-wide launch of the member
    // path's solvers (results collected once per outer iteration)
    bool memberSolve = false;
    double* solveOut = nullptr;   // [4 * kSolveSlots] device
    double* hsolve = nullptr;     // pinned host copy
    // outer-iteration state (the device member's Scal, kept on the host)
    RsgScalars sc{};
    std::vector<double> hist;
    std::string err;
    int initialized = 0;
};

namespace {

template <typename T>
hipError_t dot_read(RsgBig& B, int n, const T* a, const T* b, const T* c, const T* d, double* out2) {
    grid::dot2<T>(n, a, b, c, d, B.S, B.S.dots, B.s);
    hipError_t e = hipMemcpyAsync(B.S.hdots, B.S.dots, 2 * sizeof(double), hipMemcpyDeviceToHost, B.s);
    if (e == hipSuccess) e = hipStreamSynchronize(B.s);
    out2[0] = B.S.hdots[0];
    out2[1] = c ? B.S.hdots[1] : 0.0;
    return e;
}

// BiCGSTAB (LinearSolver.hpp:193-245, Jacobi right preconditioning) on the
// CSR system (level-0 pattern, values v / ex / dg in precision T), x in/out;
// w: eight work vectors of n.
template <typename T>
hipError_t bicgstab_grid(RsgBig& B, const T* v, const T* ex, const T* dg, const T* b, T* x, T* const* w, int maxIter,
                         double tol, SolveRes& res) {
    const int n = B.nC;
    T *r = w[0], *rh = w[1], *p = w[2], *vv = w[3], *sv = w[4], *tv = w[5], *ph = w[6], *sh = w[7];
    const dim3 g(grid::blocks(n)), bl(kBlock);
    res = SolveRes{0, 0.0, 0.0, 0};
    hipLaunchKernelGGL(grid::k_residual<T>, g, bl, 0, B.s, n, B.rp0, B.ci0, v, ex, b, (const T*)x, r);
    double d2[2];
    hipError_t e = dot_read<T>(B, n, r, r, nullptr, nullptr, d2);
    if (e != hipSuccess) return e;
    const double r0 = std::sqrt(d2[0]);
    res.initialRes = r0;
    if (r0 < 1e-30) { res.converged = 1; return hipSuccess; }
    hipLaunchKernelGGL((grid::k_convert<T, T>), g, bl, 0, B.s, n, (const T*)r, rh);
    hipLaunchKernelGGL(grid::k_fill<T>, g, bl, 0, B.s, n, p, T(0));
    hipLaunchKernelGGL(grid::k_fill<T>, g, bl, 0, B.s, n, vv, T(0));
    double rho = 1.0, alpha = 1.0, omega = 1.0, rhoN = d2[0];
    for (int it = 0; it < maxIter; ++it) {
        if (std::fabs(rhoN) < 1e-30) break;
        const double beta = (rhoN / (rho + 1e-30)) * (alpha / (omega + 1e-30));
        hipLaunchKernelGGL(k_bicg_p<T>, g, bl, 0, B.s, n, beta, omega, (const T*)r, (const T*)vv, dg, p, ph);
        hipLaunchKernelGGL(grid::k_spmv<T>, g, bl, 0, B.s, n, B.rp0, B.ci0, v, ex, (const T*)ph, vv);
        if ((e = dot_read<T>(B, n, rh, vv, nullptr, nullptr, d2)) != hipSuccess) return e;
        const double rv = d2[0];
        if (std::fabs(rv) < 1e-30) break;
        alpha = rhoN / rv;
        hipLaunchKernelGGL(k_bicg_s<T>, g, bl, 0, B.s, n, alpha, (const T*)r, (const T*)vv, dg, sv, sh);
        if ((e = dot_read<T>(B, n, sv, sv, nullptr, nullptr, d2)) != hipSuccess) return e;
        const double sn = std::sqrt(d2[0]);
        if (sn / r0 < tol) {
            hipLaunchKernelGGL(k_axpy_t<T>, g, bl, 0, B.s, n, alpha, (const T*)ph, x);
            res.iterations = it + 1; res.finalRes = sn / r0; res.converged = 1;
            return hipGetLastError();
        }
        hipLaunchKernelGGL(grid::k_spmv<T>, g, bl, 0, B.s, n, B.rp0, B.ci0, v, ex, (const T*)sh, tv);
        if ((e = dot_read<T>(B, n, tv, tv, tv, sv, d2)) != hipSuccess) return e;
        omega = (d2[0] > 1e-30) ? d2[1] / d2[0] : 0.0;
        hipLaunchKernelGGL(k_bicg_x<T>, g, bl, 0, B.s, n, alpha, omega, (const T*)ph, (const T*)sh, (const T*)sv,
                           (const T*)tv, x, r);
        if ((e = dot_read<T>(B, n, r, r, rh, r, d2)) != hipSuccess) return e;
        const double rn = std::sqrt(d2[0]);
        res.iterations = it + 1;
        res.finalRes = rn / r0;
        if (res.finalRes < tol) { res.converged = 1; break; }
        if (std::fabs(omega) < 1e-30) break;
        rho = rhoN;
        rhoN = d2[1];
    }
    return hipGetLastError();
}

// Coarse levels and gather lists of the p' hierarchy, aggregated on the host
// from the current level-0 matrix (values in B.v0 / B.dg0).
hipError_t build_amg(RsgBig& B) {
    const PrepMesh& pm = *B.pm;
    std::vector<grid::Level> H(1);
    grid::Level& L0 = H[0];
    L0.n = pm.nC; L0.nnz = pm.nnz0;
    L0.rp = pm.rp0; L0.ci = pm.ci0;
    L0.v.resize(pm.nnz0); L0.dg.resize(pm.nC);
    hipError_t e = hipMemcpyAsync(L0.v.data(), B.v0, 8 * (size_t)pm.nnz0, hipMemcpyDeviceToHost, B.s);
    if (e == hipSuccess) e = hipMemcpyAsync(L0.dg.data(), B.dg0, 8 * (size_t)pm.nC, hipMemcpyDeviceToHost, B.s);
    if (e == hipSuccess) e = hipStreamSynchronize(B.s);
    if (e != hipSuccess) return e;
    // the reference's greedy aggregation while it coarsens; where it stalls on
    // a level larger than the dense coarsest size (isotropic Galerkin
    // stencils: every coupling is at most a quarter of the diagonal, so none
    // is "strong"), two rounds of pairwise matching instead -- a stalled level
    // would otherwise end the hierarchy unfactored and the V-cycle would lose
    // its coarse correction
    for (int l = 0; l + 1 < RSG_MAX_LEVELS; ++l) {
        grid::Level& f = H.back();
        if (f.n <= RSG_MIN_COARSE) break;
        // a capped hierarchy ends as soon as its coarsest level can be inverted
        if (B.maxLevels > 0 && (int)H.size() >= B.maxLevels && f.n <= RSG_DENSE_MAX) break;
        grid::Level c;
        const bool ok = grid::coarsen(f, c);
        if ((!ok || c.n > 0.75 * f.n) && f.n > RSG_COARSE_DENSE) {
            const std::vector<int> refMap = f.map;
            grid::Level c2;
            if (grid::coarsen_pairwise(f, c2)) c = std::move(c2);
            else if (ok) f.map = refMap;
            else break;
        } else if (!ok) {
            break;
        }
        if (c.n <= 0 || c.n >= f.n || (int)f.map.size() != f.n) { B.err = "AMG coarsening failed"; return hipErrorUnknown; }
        H.push_back(std::move(c));
    }
    // gather lists: coarse entry (I, J) <- fine entries of the aggregate rows
    // of I whose column's aggregate is J, ascending fine entry
    std::vector<std::vector<int>> gPh(H.size()), gIh(H.size());
    for (size_t l = 1; l < H.size(); ++l) {
        const grid::Level& f = H[l - 1];
        const grid::Level& c = H[l];
        std::vector<int> f2c(f.nnz, -1);
        for (int I = 0; I < c.n; ++I)
            for (int q = c.aP[I]; q < c.aP[I + 1]; ++q) {
                const int i = c.aI[q];
                for (int e2 = f.rp[i]; e2 < f.rp[i + 1]; ++e2) {
                    const int J = f.map[f.ci[e2]];
                    const int* lo = c.ci.data() + c.rp[I];
                    const int* hi = c.ci.data() + c.rp[I + 1];
                    const int* at = std::lower_bound(lo, hi, J);
                    if (at == hi || *at != J) { B.err = "AMG gather list: coarse entry not found"; return hipErrorUnknown; }
                    f2c[e2] = (int)(at - c.ci.data());
                }
            }
        for (int e2 = 0; e2 < f.nnz; ++e2)
            if (f2c[e2] < 0) { B.err = "AMG gather list: fine entry without a coarse entry"; return hipErrorUnknown; }
        std::vector<int>& P = gPh[l];
        std::vector<int>& I2 = gIh[l];
        P.assign(c.nnz + 1, 0);
        for (int e2 = 0; e2 < f.nnz; ++e2) P[f2c[e2] + 1]++;
        for (int E = 0; E < c.nnz; ++E) P[E + 1] += P[E];
        I2.assign(f.nnz, 0);
        std::vector<int> fill(P.begin(), P.end() - 1);
        for (int e2 = 0; e2 < f.nnz; ++e2) I2[fill[f2c[e2]]++] = e2;
    }
    std::vector<double> G;
    const bool dense = H.back().n <= RSG_COARSE_DENSE && grid::dense_inverse(H.back(), G);
    B.DW.release();
    B.VG.clear(); B.VGf.clear();      // the graphs hold the old hierarchy's addresses
    B.HA.reset(new grid::DevAlloc());
    B.H = grid::Hierarchy<double>();
    e = grid::upload<double>(H, G, dense, *B.HA, B.H, B.s);
    if (e != hipSuccess) return e;
    // a coarsest level above 64 rows: explicit inverse on the matrix cores,
    // recomputed with every refresh of the coarse values
    B.bigDense = false;
    B.DW = grid::DenseWork();
    if (!dense && H.size() > 1 && H.back().n <= RSG_DENSE_MAX) {
        if (!B.DW.take(H.back().n, *B.HA)) return B.HA->err;
        B.bigDense = true;
    }
    B.gP.assign(H.size(), nullptr); B.gI.assign(H.size(), nullptr);
    for (size_t l = 1; l < H.size(); ++l) {
        B.gP[l] = B.HA->take<int>(gPh[l].size());
        B.gI[l] = B.HA->take<int>(gIh[l].size());
        if (B.HA->err == hipSuccess)
            B.HA->err = hipMemcpyAsync(B.gP[l], gPh[l].data(), 4 * gPh[l].size(), hipMemcpyHostToDevice, B.s);
        if (B.HA->err == hipSuccess)
            B.HA->err = hipMemcpyAsync(B.gI[l], gIh[l].data(), 4 * gIh[l].size(), hipMemcpyHostToDevice, B.s);
    }
    if (B.HA->err == hipSuccess) B.HA->err = hipStreamSynchronize(B.s);
    if (B.HA->err == hipSuccess && B.mixed) {
        // FP32 mirror of the hierarchy: the same patterns, FP32 values in
        // row_ax's form and FP32 level vectors (filled by refresh_amg)
        B.HFA.reset(new grid::DevAlloc());
        grid::DevAlloc& F = *B.HFA;
        B.Hf = grid::Hierarchy<float>();
        B.Hf.lv.resize(B.H.lv.size());
        for (size_t l = 0; l < B.H.lv.size(); ++l) {
            const grid::DevLevel<double>& d = B.H.lv[l];
            grid::DevLevel<float>& f = B.Hf.lv[l];
            f.n = d.n; f.nnz = d.nnz; f.rp = d.rp; f.ci = d.ci; f.map = d.map; f.aP = d.aP; f.aI = d.aI;
            f.v = F.take<float>(d.nnz); f.ex = F.take<float>(d.n); f.dg = F.take<float>(d.n);
            f.x = F.take<float>(d.n); f.b = F.take<float>(d.n); f.r = F.take<float>(d.n);
        }
        if (B.H.dense) { B.Hf.cinv = F.take<float>(RSG_COARSE_DENSE * RSG_COARSE_DENSE); B.Hf.dense = true; }
        if (F.err != hipSuccess) return F.err;
    }
    B.haveH = B.HA->err == hipSuccess;
    return B.HA->err;
}

// FP32 mirror of hierarchy levels [l0, l1) and of the coarsest inverse
void mirror_f32(RsgBig& B, int l0, int l1) {
    for (int l = l0; l < l1; ++l) {
        const grid::DevLevel<double>& d = B.H.lv[l];
        grid::DevLevel<float>& f = B.Hf.lv[l];
        hipLaunchKernelGGL(k_to_f32, dim3(grid::blocks(d.n)), dim3(kBlock), 0, B.s, d.n, (const int*)d.rp,
                           (const int*)d.ci, (const double*)d.v, (const double*)d.dg, f.v, f.ex, f.dg);
    }
    if (l1 == (int)B.H.lv.size()) {
        B.Hf.dense = B.H.dense;
        if (B.H.dense)
            hipLaunchKernelGGL((grid::k_convert<float, double>), dim3(grid::blocks(RSG_COARSE_DENSE * RSG_COARSE_DENSE)),
                               dim3(kBlock), 0, B.s, RSG_COARSE_DENSE * RSG_COARSE_DENSE, (const double*)B.H.cinv,
                               B.Hf.cinv);
        B.Hf.dinv = B.H.dinv;          // the large coarsest inverse stays FP64
        B.Hf.dinvN = B.H.dinvN;
    }
}

// level-0 values of the hierarchy from the current p' matrix; with `coarse`
// the Galerkin values of every coarser level and the coarsest inverse too
hipError_t refresh_amg(RsgBig& B, bool coarse) {
    grid::Hierarchy<double>& H = B.H;
    hipLaunchKernelGGL(k_csr_fill, dim3(cells_grid(std::max(B.nC, B.nnz0))), dim3(kBlock), 0, B.s, B.dM,
                       H.lv[0].v, H.lv[0].dg);
    if (!coarse) {
        if (B.mixed) mirror_f32(B, 0, 1);
        return hipGetLastError();
    }
    const int L = (int)H.lv.size();
    for (int l = 1; l < L; ++l) {
        grid::DevLevel<double>& c = H.lv[l];
        hipLaunchKernelGGL(k_galerkin, dim3(grid::blocks(c.nnz)), dim3(kBlock), 0, B.s, c.nnz, (const int*)B.gP[l],
                           (const int*)B.gI[l], (const double*)H.lv[l - 1].v, c.v);
        hipLaunchKernelGGL(k_diag_of, dim3(grid::blocks(c.n)), dim3(kBlock), 0, B.s, c.n, (const int*)c.rp,
                           (const int*)c.ci, (const double*)c.v, c.dg);
    }
    if (H.dense) {
        const grid::DevLevel<double>& c = H.lv[L - 1];
        hipLaunchKernelGGL(k_coarse_inverse, dim3(1), dim3(kBlock), 0, B.s, c.n, (const int*)c.rp,
                           (const int*)c.ci, (const double*)c.v, H.cinv, B.dOk);
        int ok = 0;
        hipError_t e = hipMemcpyAsync(&ok, B.dOk, sizeof ok, hipMemcpyDeviceToHost, B.s);
        if (e == hipSuccess) e = hipStreamSynchronize(B.s);
        if (e != hipSuccess) return e;
        if (!ok) H.dense = false;     // zero pivot: the coarsest level is smoothed instead
    }
    if (B.bigDense) {
        const grid::DevLevel<double>& c = H.lv[L - 1];
        bool ok = false;
        double ms = 0.0;
        hipError_t e = grid::dense_invert(B.DW, c.rp, c.ci, c.v, B.s, &ok, &ms);
        if (e != hipSuccess) return e;
        B.denseMs += ms;
        H.dinv = ok ? B.DW.M : nullptr;     // a pivot not positive: smoothed instead
        H.dinvN = B.DW.N;
    }
    if (B.mixed) mirror_f32(B, 0, L);
    return hipGetLastError();
}

// Gauss-Seidel (rsg_member.h's gauss_seidel): sequential sweeps on one
// thread, residual and norms grid-wide.
hipError_t gs_grid(RsgBig& B, double* x, int maxIter, double tol, SolveRes& res) {
    const int n = B.nC;
    const dim3 g(grid::blocks(n)), bl(kBlock);
    double* r = B.hM.kv[0];
    double d2[2];
    res = SolveRes{0, 0.0, 0.0, 0};
    hipError_t e = dot_read<double>(B, n, B.hM.src, B.hM.src, nullptr, nullptr, d2);
    if (e != hipSuccess) return e;
    double b0 = std::sqrt(d2[0]);
    res.initialRes = b0;
    if (b0 < 1e-30) b0 = 1.0;
    for (int it = 0; it < maxIter; ++it) {
        hipLaunchKernelGGL(k_gs_sweep, dim3(1), dim3(1), 0, B.s, B.dM, x);
        hipLaunchKernelGGL(grid::k_residual<double>, g, bl, 0, B.s, n, B.rp0, B.ci0, (const double*)B.v0,
                           (const double*)nullptr, (const double*)B.hM.src, (const double*)x, r);
        if ((e = dot_read<double>(B, n, r, r, nullptr, nullptr, d2)) != hipSuccess) return e;
        res.iterations = it + 1;
        res.finalRes = std::sqrt(d2[0]) / b0;
        if (res.finalRes < tol) { res.converged = 1; break; }
    }
    return hipGetLastError();
}

// FP32 inner solve from zero of the level-0 system (FP32 image) or, for
// p' with the AMG ready, of the FP32 hierarchy: A ef = bf.
hipError_t inner_f32(RsgBig& B, int code, bool amg, const float* bf, float* ef, int maxIter, double tol,
                     SolveRes& res) {
    const int n = B.nC;
    float* w[8];
    for (int j = 0; j < 8; ++j) w[j] = B.fw + (size_t)(j + 2) * n;
    if (code == 2) return bicgstab_grid<float>(B, B.v0f, B.ex0f, B.dg0f, bf, ef, w, maxIter, tol, res);
    grid::Hierarchy<float> J;
    if (!amg) {
        J.lv.resize(1);
        grid::DevLevel<float>& d = J.lv[0];
        d.n = n; d.nnz = B.nnz0;
        d.rp = const_cast<int*>(B.rp0); d.ci = const_cast<int*>(B.ci0); d.v = B.v0f; d.ex = B.ex0f; d.dg = B.dg0f;
    }
    grid::PcgResult pr;
    hipError_t e = grid::pcg<float>(amg ? B.Hf : J, bf, ef, w[0], w[1], w[2], w[3], maxIter, tol, amg, B.S, B.s, pr,
                                    &B.VGf);
    res = SolveRes{pr.iterations, pr.finalRes, pr.initialRes, pr.converged};
    return e;
}

// Mixed-precision solve (rsg_member.h's lsolve_mixed_global): FP64 defect of
// the FP64 system, FP32 inner solve of the defect scaled to unit norm, FP64
// update; at most RSG_MIXED_SWEEPS sweeps and maxIter inner iterations in all.
hipError_t solve_mixed(RsgBig& B, int code, double* x, bool amg, int maxIter, double tol, SolveRes& res) {
    const int n = B.nC;
    const dim3 g(grid::blocks(n)), bl(kBlock);
    const double* v = amg ? B.H.lv[0].v : B.v0;
    float* bf = B.fw;
    float* ef = B.fw + n;
    if (!amg)
        hipLaunchKernelGGL(k_to_f32, g, bl, 0, B.s, n, B.rp0, B.ci0, (const double*)B.v0, (const double*)B.dg0,
                           B.v0f, B.ex0f, B.dg0f);
    res = SolveRes{0, 0.0, 0.0, 0};
    double d2[2];
    auto defect = [&](double& rn) {
        hipLaunchKernelGGL(grid::k_residual<double>, g, bl, 0, B.s, n, B.rp0, B.ci0, v, (const double*)nullptr,
                           (const double*)B.hM.src, (const double*)x, B.r64);
        hipError_t e = dot_read<double>(B, n, B.r64, B.r64, nullptr, nullptr, d2);
        rn = std::sqrt(d2[0]);
        return e;
    };
    double rn = 0.0;
    hipError_t e = defect(rn);
    if (e != hipSuccess) return e;
    const double r0 = rn;
    res.initialRes = r0;
    if (r0 < 1e-30) { res.converged = 1; return hipSuccess; }
    for (int sweep = 0; sweep < RSG_MIXED_SWEEPS && res.iterations < maxIter; ++sweep) {
        hipLaunchKernelGGL(k_scale_f32, g, bl, 0, B.s, n, (const double*)B.r64, 1.0 / rn, bf, ef);
        SolveRes in{0, 0.0, 0.0, 0};
        e = inner_f32(B, code, amg, bf, ef, maxIter - res.iterations, tol * r0 / rn, in);
        if (e != hipSuccess) return e;
        res.iterations += in.iterations;
        hipLaunchKernelGGL(k_refine, g, bl, 0, B.s, n, rn, (const float*)ef, x);
        double rnNew = 0.0;
        if ((e = defect(rnNew)) != hipSuccess) return e;
        res.finalRes = rnNew / r0;
        if (res.finalRes < tol) { res.converged = 1; break; }
        if (!(rnNew < rn) || in.iterations == 0) break;   // the FP32 solve stalls: stop refining
        rn = rnNew;
    }
    return hipGetLastError();
}

// One linear solve of the member's current system into x (device); a mixed
// member's solves run solve_mixed unless allowMixed is false (omega).
hipError_t solve(RsgBig& B, int code, double* x, bool amgReady, SolveRes& res, bool allowMixed = true) {
    const RsgSettingsDev& st = B.set;
    const int n = B.nC;
    RsgMemberDev& M = B.hM;
    hipLaunchKernelGGL(k_csr_fill, dim3(cells_grid(std::max(n, B.nnz0))), dim3(kBlock), 0, B.s, B.dM, B.v0, B.dg0);
    const bool amg = code == 0 && amgReady;
    if (code == RSG_SOLVER_GAUSS_SEIDEL) return gs_grid(B, x, st.innerIter, st.innerTol, res);
    if (B.mixed && allowMixed) return solve_mixed(B, code, x, amg, st.innerIter, st.innerTol, res);
    if (code == 2) {
        double* w[8];
        for (int j = 0; j < 8; ++j) w[j] = M.kv[j];
        return bicgstab_grid<double>(B, B.v0, nullptr, B.dg0, M.src, x, w, st.innerIter, st.innerTol, res);
    }
    grid::Hierarchy<double> J;                    // Jacobi: level 0 alone
    if (!amg) {
        J.lv.resize(1);
        grid::DevLevel<double>& d = J.lv[0];
        d.n = n; d.nnz = B.nnz0;
        d.rp = const_cast<int*>(B.rp0); d.ci = const_cast<int*>(B.ci0); d.v = B.v0; d.dg = B.dg0;
    }
    grid::PcgResult pr;
    hipError_t e = grid::pcg<double>(amg ? B.H : J, M.src, x, M.kv[0], B.vz, M.kv[2], B.vq, st.innerIter,
                                     st.innerTol, amg, B.S, B.s, pr, &B.VG);
    res = SolveRes{pr.iterations, pr.finalRes, pr.initialRes, pr.converged};
    return e;
}

// std::max's comparison, as the device dmax (NaN operands behave alike)
inline double hmax(double a, double b) { return (a < b) ? b : a; }

// One SIMPLE outer iteration (SIMPLESolver.cpp:731-1036), the phases of
// rsg_member.h's simple_iteration as grid launches.
hipError_t outer_iteration(RsgBig& B) {
    RsgScalars& sc = B.sc;
    const RsgSettingsDev& st = B.set;
    RsgMemberDev& M = B.hM;
    const int iter = sc.iter;
    const int n = B.nC;
    const dim3 gc(cells_grid(n)), bl(kBlock);
    const bool turbActive = iter >= st.turbStart;
    const bool turbUpdate = turbActive && ((iter - st.turbStart) % st.turbInterval == 0);
    hipError_t e = hipSuccess;
    // 1. SST closure + grad p
    double nuTMin = 0.0;
    if (turbUpdate) {
        if (!sc.turbEstablished && iter >= st.turbStart + st.nutFloorIters) sc.turbEstablished = 1;
        nuTMin = sc.turbEstablished ? 0.0 : 0.1 * M.nu;
    }
    hipLaunchKernelGGL(k_step1, gc, bl, 0, B.s, B.dM, turbUpdate ? 1 : 0, nuTMin);
    // 2. momentum
    SolveRes rU[3] = {{0, 0.0, 0.0, 0}, {0, 0.0, 0.0, 0}, {0, 0.0, 0.0, 0}};
    double* ucomp[3] = {M.Ux, M.Uy, M.Uz};
    for (int comp = 0; comp < 3 && e == hipSuccess; ++comp) {
        if (comp == 2 && M.planarZ) break;
        if (comp > 0) hipLaunchKernelGGL(k_grad_u, gc, bl, 0, B.s, B.dM, comp - 1, comp);
        hipLaunchKernelGGL(k_asm_mom, gc, bl, 0, B.s, B.dM, comp);
        e = solve(B, st.mSolver, ucomp[comp], false, rU[comp]);
    }
    hipLaunchKernelGGL(k_bcs, gc, bl, 0, B.s, B.dM, 1, 0.0);
    // 3. pressure correction
    hipLaunchKernelGGL(k_asm_pc, gc, bl, 0, B.s, B.dM);
    bool amgReady = false;
    if (e == hipSuccess && st.pSolver == 0) {
        const bool rebuild = sc.amgBuiltIter < 0 || iter - sc.amgBuiltIter >= st.amgRefresh || !B.haveH;
        const bool coarse = rebuild || iter - sc.coarseIter >= st.coarseRefresh;
        if (rebuild) {
            hipLaunchKernelGGL(k_csr_fill, dim3(cells_grid(std::max(n, B.nnz0))), bl, 0, B.s, B.dM, B.v0, B.dg0);
            e = build_amg(B);
            sc.amgBuiltIter = iter;
        }
        if (e == hipSuccess) e = refresh_amg(B, coarse);
        if (coarse) sc.coarseIter = iter;
        sc.nLevels = (int)B.H.lv.size();
        sc.coarseDense = (B.H.dense || B.H.dinv) ? 1 : 0;
        amgReady = true;
    }
    SolveRes rP{0, 0.0, 0.0, 0};
    if (e == hipSuccess) e = solve(B, st.pSolver, M.pp, amgReady, rP);
    hipLaunchKernelGGL(k_ppb, gc, bl, 0, B.s, B.dM);
    // 4. corrections
    hipLaunchKernelGGL(k_grad_pp, gc, bl, 0, B.s, B.dM);
    hipLaunchKernelGGL(k_correct, gc, bl, 0, B.s, B.dM);
    hipLaunchKernelGGL(k_bcs, gc, bl, 0, B.s, B.dM, 1 | 2, 0.0);
    // 5. turbulence
    SolveRes rK{0, 0.0, 0.0, 0}, rOm{0, 0.0, 0.0, 0};
    if (turbUpdate && e == hipSuccess) {
        hipLaunchKernelGGL(k_asm_turb, gc, bl, 0, B.s, B.dM, 0);
        e = solve(B, st.tSolver, M.k, false, rK);
        hipLaunchKernelGGL(k_clamp, gc, bl, 0, B.s, B.dM, 0);
        hipLaunchKernelGGL(k_bcs, gc, bl, 0, B.s, B.dM, 4, 0.0);
        hipLaunchKernelGGL(k_asm_turb, gc, bl, 0, B.s, B.dM, 1);
        if (e == hipSuccess) e = solve(B, st.tSolver, M.om, false, rOm, false);   // omega stays FP64
        hipLaunchKernelGGL(k_clamp, gc, bl, 0, B.s, B.dM, 1);
        hipLaunchKernelGGL(k_bcs, gc, bl, 0, B.s, B.dM, 8, M.coeffs[2]);
        hipLaunchKernelGGL(k_pin, gc, bl, 0, B.s, B.dM, B.S.part);
        hipLaunchKernelGGL(k_max4, dim3(1), dim3(kBlock), 0, B.s, (int)gc.x, (const double*)B.S.part, B.S.dots);
        if (e == hipSuccess) e = hipMemcpyAsync(B.S.hdots, B.S.dots, 4 * sizeof(double), hipMemcpyDeviceToHost, B.s);
        if (e == hipSuccess) e = hipStreamSynchronize(B.s);
        sc.lastK = B.S.hdots[2] / B.S.hdots[0];
        sc.lastOm = B.S.hdots[3] / B.S.hdots[1];
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return e;
    // 6. residual bookkeeping and the convergence / divergence tests (:934-1035)
    const double eUx = rU[0].initialRes, eUy = rU[1].initialRes, eUz = rU[2].initialRes, eP = rP.initialRes;
    const double eK = turbActive ? sc.lastK : 0.0, eOm = turbActive ? sc.lastOm : 0.0;
    sc.innerIters[0] += rU[0].iterations; sc.innerIters[1] += rU[1].iterations; sc.innerIters[2] += rU[2].iterations;
    sc.innerIters[3] += rP.iterations; sc.innerIters[4] += rK.iterations; sc.innerIters[5] += rOm.iterations;
    if (st.recordHistory && iter < M.histCap) {
        const double h[6] = {eUx, eUy, eUz, eP, eK, eOm};
        B.hist.insert(B.hist.end(), h, h + 6);
        sc.histLen = iter + 1;
    }
    if (iter == 0) {
        sc.normUx0 = hmax(eUx, 1e-30); sc.normUy0 = hmax(eUy, 1e-30); sc.normP0 = hmax(eP, 1e-30);
    } else if (iter < 50) {
        sc.normP0 = hmax(sc.normP0, eP);
    }
    const double normMom0 = hmax(sc.normUx0, sc.normUy0);
    const double nUx = normMom0 < 1e-20 ? 0.0 : eUx / normMom0;
    const double nUy = normMom0 < 1e-20 ? 0.0 : eUy / normMom0;
    const double nP = sc.normP0 < 1e-20 ? 0.0 : eP / sc.normP0;
    double maxRes = hmax(hmax(nUx, nUy), nP);
    if (turbActive) maxRes = hmax(hmax(maxRes, sc.lastK), sc.lastOm);
    if (!std::isfinite(eUx) || !std::isfinite(eUy) || !std::isfinite(eUz) || !std::isfinite(eP) ||
        !std::isfinite(eK) || !std::isfinite(eOm) || !std::isfinite(sc.lastK) || !std::isfinite(sc.lastOm))
        maxRes = INFINITY;
    const bool turbScheduled = st.turbStart < st.maxIter;
    sc.iter = iter + 1;
    if (maxRes < st.tol && iter > 0 && (!turbScheduled || (turbActive && sc.turbEstablished))) {
        sc.status = RSGM_CONVERGED; sc.finalIter = iter;
    } else if (std::isnan(maxRes) || std::isinf(maxRes) || maxRes > st.divLimit) {
        sc.status = RSGM_DIVERGED; sc.finalIter = iter;
    } else if (sc.iter >= st.maxIter) {
        sc.status = RSGM_MAXITER; sc.finalIter = st.maxIter;
    }
    return hipSuccess;
}

#include "rsg_comp.h"

}  // namespace

RsgBig* rsg_big_create(const RsgMemberDev& hM, RsgMemberDev* dM, const RsgSettingsDev& set, const PrepMesh& pm,
                       const RsgScalars& sc0, hipStream_t s, int maxLevels, bool graphs, const RsgCompDev* comp,
                       std::string& err) {
    std::unique_ptr<RsgBig> B(new RsgBig());
    B->maxLevels = maxLevels;
    B->VG.enabled = B->VGf.enabled = graphs;
    B->hM = hM; B->dM = dM; B->set = set; B->pm = &pm; B->nC = pm.nC; B->nnz0 = pm.nnz0;
    B->sc = sc0;
    // the member's own stream (default priority): members run concurrently,
    // one host thread each; the engine stream's staging must be done first
    hipError_t e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&B->s, hipStreamNonBlocking);
    if (e != hipSuccess) { err = std::string("big member: ") + hipGetErrorString(e); return nullptr; }
    B->ownStream = true;
    RsgMeshDev mesh;
    e = hipMemcpyAsync(&mesh, hM.mesh, sizeof mesh, hipMemcpyDeviceToHost, B->s);
    if (e == hipSuccess) e = hipStreamSynchronize(B->s);
    if (e != hipSuccess) { err = std::string("big member: ") + hipGetErrorString(e); rsg_big_destroy(B.release()); return nullptr; }
    B->rp0 = mesh.rp0; B->ci0 = mesh.ci0;
    const int nb = std::max(grid::blocks(std::max(pm.nC, pm.nnz0)), cells_grid(pm.nC));
    B->v0 = B->A.take<double>(pm.nnz0); B->dg0 = B->A.take<double>(pm.nC);
    B->vz = B->A.take<double>(pm.nC); B->vq = B->A.take<double>(pm.nC);
    B->S.part = B->A.take<double>(4 * (size_t)nb + 4);
    B->S.dots = B->A.take<double>(16);
    B->dOk = B->A.take<int>(1);
    B->mixed = hM.prec != 0;
    if (B->mixed) {
        B->r64 = B->A.take<double>(pm.nC);
        B->v0f = B->A.take<float>(pm.nnz0); B->ex0f = B->A.take<float>(pm.nC); B->dg0f = B->A.take<float>(pm.nC);
        B->fw = B->A.take<float>(10 * (size_t)pm.nC);
    }
    if (comp) {
        // the compressible fields; the problem's data (temperature BCs, gas) as given
        B->comp = true;
        RsgCompDev& Q = B->hQ;
        Q = *comp;
        const size_t nC = pm.nC, nB = std::max(pm.nB, 1);
        Q.T = B->A.take<double>(nC); Q.Tb = B->A.take<double>(nB); Q.rho = B->A.take<double>(nC);
        Q.mu = B->A.take<double>(nC); Q.Told = B->A.take<double>(nC);
        Q.Ux0 = B->A.take<double>(nC); Q.Uy0 = B->A.take<double>(nC); Q.Uz0 = B->A.take<double>(nC);
        Q.mdu = B->A.take<double>(nC);
        for (int q = 0; q < 3; ++q) Q.gt[q] = B->A.take<double>(nC);
        Q.zeroB = B->A.take<double>(nB);
        Q.pMech = B->A.take<double>(nC);
        B->compPart = B->A.take<double>((size_t)CQ_COUNT * cells_grid(pm.nC));
        B->memberSolve = hM.amg.rp != nullptr;   // the workspace carries the member path's AMG arrays
        B->solveOut = B->A.take<double>(4 * kSolveSlots);
        B->dQ = B->A.take<RsgCompDev>(1);
        if (B->A.err == hipSuccess) B->A.err = hipMemsetAsync(Q.zeroB, 0, 8 * nB, B->s);
        if (B->A.err == hipSuccess)
            B->A.err = hipMemsetAsync(B->compPart, 0, 8 * (size_t)CQ_COUNT * cells_grid(pm.nC), B->s);
        if (B->A.err == hipSuccess) B->A.err = hipMemcpyAsync(B->dQ, &Q, sizeof Q, hipMemcpyHostToDevice, B->s);
    }
    if (B->A.err == hipSuccess) B->A.err = hipHostMalloc((void**)&B->S.hdots, 16 * sizeof(double), hipHostMallocDefault);
    if (B->A.err == hipSuccess && B->memberSolve)
        B->A.err = hipHostMalloc((void**)&B->hsolve, 4 * kSolveSlots * sizeof(double), hipHostMallocDefault);
    if (B->A.err == hipSuccess) B->A.err = hipMemcpyAsync(dM, &hM, sizeof hM, hipMemcpyHostToDevice, B->s);
    if (B->A.err == hipSuccess) B->A.err = hipStreamSynchronize(B->s);
    if (B->A.err != hipSuccess) {
        err = std::string("big member: ") + hipGetErrorString(B->A.err);
        rsg_big_destroy(B.release());
        return nullptr;
    }
    return B.release();
}

void rsg_big_destroy(RsgBig* B) {
    if (!B) return;
    if (B->ownStream) (void)hipStreamSynchronize(B->s);
    B->VG.clear(); B->VGf.clear();
    B->DW.release();
    if (B->S.hdots) (void)hipHostFree(B->S.hdots);
    if (B->hsolve) (void)hipHostFree(B->hsolve);
    hipStream_t s = B->ownStream ? B->s : nullptr;
    delete B;                     // frees the member's work arrays
    if (s) (void)hipStreamDestroy(s);
}

int rsg_big_run(RsgBig* B, int iters, RsgScalars& out, std::string& err) {
    hipError_t e = hipSuccess;
    if (!B->sc.initialized) {
        const dim3 gc(cells_grid(B->nC)), bl(kBlock);
        if (B->comp) {
            // initUniform (CompressibleSIMPLESolver.cpp:27-72): fields, then
            // applyAllCompressibleBCs with the initial viscosity and beta1 = 0.075
            hipLaunchKernelGGL(kc_init, gc, bl, 0, B->s, B->dM, B->dQ);
            hipLaunchKernelGGL(k_bcs, gc, bl, 0, B->s, B->dM, 15, 0.075);
            hipLaunchKernelGGL(kc_bnd, gc, bl, 0, B->s, B->dM, B->dQ, 1, 1);
        } else if (!B->hM.initFromFields) {
            hipLaunchKernelGGL(k_init, gc, bl, 0, B->s, B->dM);
            hipLaunchKernelGGL(k_bcs, gc, bl, 0, B->s, B->dM, 15, 0.075);
        }
        B->sc.initialized = 1;
        B->sc.amgBuiltIter = -1;
        e = hipGetLastError();
    }
    for (int k = 0; k < iters && B->sc.status == RSGM_RUNNING && e == hipSuccess; ++k)
        e = B->comp ? comp_outer_iteration(*B) : outer_iteration(*B);
    if (e == hipSuccess && B->sc.status != RSGM_RUNNING) {
        // the finished member's history and scalars on the device (collected
        // like any member's), then its observables
        if (!B->hist.empty())
            e = hipMemcpyAsync(B->hM.hist, B->hist.data(), 8 * B->hist.size(), hipMemcpyHostToDevice, B->s);
        if (e == hipSuccess) e = hipMemcpyAsync(B->hM.sc, &B->sc, sizeof B->sc, hipMemcpyHostToDevice, B->s);
        if (e == hipSuccess && B->comp) {
            hipLaunchKernelGGL(kc_export, dim3(cells_grid(B->nC)), dim3(kBlock), 0, B->s, B->dM, B->dQ);
            hipLaunchKernelGGL(kc_export_nu, dim3(1), dim3(64), 0, B->s, B->dM, B->dQ);
            e = hipGetLastError();
        }
        if (e == hipSuccess && B->hM.obs) {
            hipLaunchKernelGGL(k_observe, dim3(1), dim3(64), 0, B->s, B->dM);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(B->s);
    out = B->sc;
    if (e != hipSuccess) {
        err = std::string("big member: ") + (B->err.empty() ? std::string(hipGetErrorString(e)) : B->err);
        return RSG_ERR_DEVICE;
    }
    return RSG_OK;
}

int rsg_big_comp_fields(RsgBig* B, double* T, double* rho, double* pMech, double* Tb, double* energyIters,
                        std::string& err) {
    if (!B->comp) { err = "not a compressible member"; return RSG_ERR_STATE; }
    const RsgCompDev& Q = B->hQ;
    const size_t n = B->nC, nB = B->pm->nB;
    hipError_t e = hipSuccess;
    if (T) e = hipMemcpyAsync(T, Q.T, 8 * n, hipMemcpyDeviceToHost, B->s);
    if (e == hipSuccess && rho) e = hipMemcpyAsync(rho, Q.rho, 8 * n, hipMemcpyDeviceToHost, B->s);
    if (e == hipSuccess && pMech) e = hipMemcpyAsync(pMech, Q.pMech, 8 * n, hipMemcpyDeviceToHost, B->s);
    if (e == hipSuccess && Tb && nB > 0) e = hipMemcpyAsync(Tb, Q.Tb, 8 * nB, hipMemcpyDeviceToHost, B->s);
    if (e == hipSuccess) e = hipStreamSynchronize(B->s);
    if (energyIters) *energyIters = B->compTIters;
    if (e != hipSuccess) { err = std::string("compressible member: ") + hipGetErrorString(e); return RSG_ERR_DEVICE; }
    return RSG_OK;
}
