// rsg_chunk.h -- the member kernels: the persistent SIMPLE chunk
// (k_simple_chunk, one workgroup per member, many members per launch) and the
// stand-alone LDU solve (k_ldu_solve), with the LDS layout and the descriptor
// staging they share.  rsg_chunk.hip is compiled once per (workgroup size,
// precision) pair; each object instantiates its kernel and exports plain
// launchers to rsg_capi.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "rsg_layout.h"
#include "rsg_member.h"

namespace {

// Widest cells-per-thread instantiation of the resident solvers for a block
// size (it bounds the kernel's register allocation; RSG_RES_MAXN cells at most).
constexpr int cmax_for(int bs) { return bs >= 1024 ? 1 : (bs >= 512 ? 3 : 6); }

// Workgroup LDS: reduction scratch, coarsest dense inverse, the resident
// solvers' two exchange buffers and the coarse-level x (~69 KB, so two
// member workgroups share a CU).
struct LdsArea {
    double red[128];
    double lds[RSG_COARSE_DENSE * RSG_COARSE_DENSE];
    double xb[2 * RSG_RES_MAXN];
    double cv[RSG_RES_MAXC];
};
// One member per CU: room for LDS copies of the coarse AMG levels (~150 KB).
struct LdsAreaWide {
    LdsArea a;
    double ceVal[RSG_RES_CE];
    short ceCol[RSG_RES_CE];
    short caCol[RSG_RES_CA];
};

__device__ rsg::Thr make_thr(LdsArea& L, int bs) {
    rsg::Thr t;
    t.tid = threadIdx.x; t.bs = bs; t.red = L.red; t.lds = L.lds; t.phase = 0;
    t.ltid = threadIdx.x; t.lbs = bs; t.blk = 0; t.nblk = 1; t.gred = nullptr;
    t.xb0 = L.xb; t.xb1 = L.xb + RSG_RES_MAXN; t.xsel = 0;
    t.cvec = L.cv; t.cvecCap = RSG_RES_MAXC;
    t.ceVal = nullptr; t.ceCol = nullptr; t.ceCap = 0; t.caCol = nullptr; t.caCap = 0;
    return t;
}

__device__ rsg::Thr make_thr(LdsAreaWide& L, int bs) {
    rsg::Thr t = make_thr(L.a, bs);
    t.ceVal = L.ceVal; t.ceCol = L.ceCol; t.ceCap = RSG_RES_CE;
    t.caCol = L.caCol; t.caCap = RSG_RES_CA;
    return t;
}

// 256 threads: two members per CU share the LDS; 512: one member per CU.
template <int BS> struct LdsFor { typedef LdsArea type; };
template <> struct LdsFor<512> { typedef LdsAreaWide type; };

// The member's descriptors (every phase dereferences them) and its AMG
// level tables, copied into LDS for the launch so that these lookups cost an
// LDS access instead of an L2 round trip.  The tables are written by the
// hierarchy setup and copied back at the end of the launch.
constexpr int kLvTabs = 8;   // lvN lvRow lvRp lvNz lvK lvE lvAK lvAE
struct LdsDesc {
    RsgMemberDev M;
    RsgMeshDev mesh;
    RsgBcDev bc;
    RsgSettingsDev set;
    int lv[kLvTabs][RSG_MAX_LEVELS + 1];
    int used[2];
    int* lvG[kLvTabs];       // the tables' global homes
    int* usedG;
};

__device__ void copy_words(void* dst, const void* src, int bytes) {
    int* d = static_cast<int*>(dst);
    const int* s = static_cast<const int*>(src);
    for (int i = threadIdx.x; i < bytes / 4; i += blockDim.x) d[i] = s[i];
}

__device__ RsgMemberDev* stage_descriptors(LdsDesc& D, const RsgMemberDev* Mg) {
    static_assert(sizeof(RsgMemberDev) % 4 == 0 && sizeof(RsgMeshDev) % 4 == 0 &&
                  sizeof(RsgBcDev) % 4 == 0 && sizeof(RsgSettingsDev) % 4 == 0, "word-sized descriptors");
    copy_words(&D.M, Mg, sizeof(RsgMemberDev));
    __syncthreads();
    // a stand-alone linear solve (k_ldu_solve) has no BCs or settings: null pointers stay null
    copy_words(&D.mesh, D.M.mesh, sizeof(RsgMeshDev));
    if (D.M.bc) copy_words(&D.bc, D.M.bc, sizeof(RsgBcDev));
    if (D.M.set) copy_words(&D.set, D.M.set, sizeof(RsgSettingsDev));
    RsgAmgDev& A = D.M.amg;
    int* g[kLvTabs] = {A.lvN, A.lvRow, A.lvRp, A.lvNz, A.lvK, A.lvE, A.lvAK, A.lvAE};
    for (int q = 0; q < kLvTabs; ++q) copy_words(D.lv[q], g[q], 4 * (RSG_MAX_LEVELS + 1));
    copy_words(D.used, A.used, 8);
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < kLvTabs; ++q) D.lvG[q] = g[q];
        D.usedG = A.used;
        D.M.mesh = &D.mesh;
        if (D.M.bc) D.M.bc = &D.bc;
        if (D.M.set) D.M.set = &D.set;
        A.lvN = D.lv[0]; A.lvRow = D.lv[1]; A.lvRp = D.lv[2]; A.lvNz = D.lv[3];
        A.lvK = D.lv[4]; A.lvE = D.lv[5]; A.lvAK = D.lv[6]; A.lvAE = D.lv[7];
        A.used = D.used;
    }
    __syncthreads();
    return &D.M;
}

__device__ void unstage_tables(LdsDesc& D) {
    __syncthreads();
    for (int q = 0; q < kLvTabs; ++q) copy_words(D.lvG[q], D.lv[q], 4 * (RSG_MAX_LEVELS + 1));
    copy_words(D.usedG, D.used, 8);
}

// PREC: 0 every member of the launch is FP64, 1 every member is mixed (the
// other precision's solver code is not built into the kernel).
template <int BS, int PREC>
__global__ __launch_bounds__(BS) void k_simple_chunk(RsgMemberDev* members, RsgScalars* statusOut) {
    __shared__ typename LdsFor<BS>::type L;
    __shared__ LdsDesc D;
    rsg::Thr t = make_thr(L, BS);
#ifdef RSG_PROF
    const unsigned long long t0 = wall_clock64();
#endif
    RsgMemberDev* M = stage_descriptors(D, members + blockIdx.x);
    rsg::run_member<cmax_for(BS), PREC>(t, M);
    unstage_tables(D);
#ifdef RSG_PROF
    if (threadIdx.x == 0 && blockIdx.x < 4096) {
        rsg_prof_wg[3 * blockIdx.x] = t0;
        rsg_prof_wg[3 * blockIdx.x + 1] = wall_clock64();
        rsg_prof_wg[3 * blockIdx.x + 2] = (unsigned long long)__smid();
    }
#endif
    __syncthreads();
    if (threadIdx.x == 0) statusOut[blockIdx.x] = *M->sc;
}

// One linear solve on the LDU system already stored in the workspace.
template <int BS>
__global__ __launch_bounds__(BS) void k_ldu_solve(RsgMemberDev* Mg, int solver, int maxIter, double tol,
                                                  double* stats) {
    __shared__ LdsArea L;
    __shared__ LdsDesc D;
    rsg::Thr t = make_thr(L, BS);
    RsgMemberDev* M = stage_descriptors(D, Mg);
    int nl = 0;
    rsg::SolveRes r = rsg::ldu_solve_member<cmax_for(BS)>(t, M, solver, maxIter, tol, &nl);
    unstage_tables(D);
    if (t.tid == 0) {
        stats[0] = r.iterations; stats[1] = r.finalRes; stats[2] = r.initialRes; stats[3] = r.converged;
        stats[4] = nl;
    }
}

// Compressible members (M->comp set), one workgroup each, like
// k_simple_chunk: persistent over the launch's iteration budget or slice.
template <int BS>
__global__ __launch_bounds__(BS) void k_comp_chunk(RsgMemberDev* members, RsgScalars* statusOut) {
    __shared__ typename LdsFor<BS>::type L;
    __shared__ LdsDesc D;
    rsg::Thr t = make_thr(L, BS);
    RsgMemberDev* M = stage_descriptors(D, members + blockIdx.x);
    rsg::run_comp_member<cmax_for(BS)>(t, M);
    unstage_tables(D);
    __syncthreads();
    if (threadIdx.x == 0) statusOut[blockIdx.x] = *M->sc;
}

// One linear solve of a kernel-per-phase member's current LDU system into x
// on one workgroup (member_solve_x); out = iterations, final / initial
// residual, converged.
template <int BS>
__global__ __launch_bounds__(BS) void k_member_solve(RsgMemberDev* Mg, int solver, double* x, int amgMode,
                                                     int maxIter, double tol, double* out) {
    __shared__ LdsArea L;
    __shared__ LdsDesc D;
    rsg::Thr t = make_thr(L, BS);
    RsgMemberDev* M = stage_descriptors(D, Mg);
    rsg::SolveRes r = rsg::member_solve_x<cmax_for(BS)>(t, M, solver, x, amgMode, maxIter, tol);
    unstage_tables(D);
    if (t.tid == 0) { out[0] = r.iterations; out[1] = r.finalRes; out[2] = r.initialRes; out[3] = r.converged; }
}

}  // namespace
