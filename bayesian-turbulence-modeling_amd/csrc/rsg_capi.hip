// rsg_capi.hip -- librsg.so: the C-ABI of include/rsg.h on one MI355X.
//
// Execution model
//   * one HBM arena per context (hipMalloc once); meshes, boundary-condition
//     blocks and member workspaces are carved from it by a first-fit
//     allocator, so admission is by HBM footprint and no hipMalloc/hipFree
//     happens while solves are in flight;
//   * an engine thread runs continuous batching: it admits pending solves
//     into free slots, launches one persistent chunk kernel over every active
//     member (one workgroup per member, `itersPerLaunch` outer iterations
//     each), reads back the compact per-slot status, retires converged /
//     diverged / exhausted members individually and admits new ones between
//     launches.  Members never interact, so a result does not depend on what
//     else is in the launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "rsg.h"
#include "rsg_layout.h"
#include "rsg_prep.h"
#include "rsg_member.h"
#include "rsg_grid.h"
#include "rsg_geom.h"
#include "rsg_big.h"

// rsg_team.hip
int rsg_team_capacity(int device);
hipError_t rsg_team_launch(RsgMemberDev* dM, RsgScalars* dOut, double* gred, int blocks, hipStream_t s);

// Workspace of rsg_ldu_solve_noalloc: one stand-alone LDU solve's device block
// (mesh image, member struct, member workspace) lives here instead of in the
// context's arena.  Each GPU holds its own copy until the process ends; host
// code reaches it only through the per-device address that
// rsg_context_create resolves with hipGetSymbolAddress (ctx->scratchDev).
#define RSG_NOALLOC_SCRATCH_BYTES (8u << 20)
static __device__ __align__(256) char scratch[RSG_NOALLOC_SCRATCH_BYTES];

namespace {
// one lock per device: calls through any context of the process on the same
// GPU take turns on that GPU's scratch array
std::mutex g_scratch_mu[64];
}  // namespace

// ------------------------------------------------------------------ kernels
namespace {

// Team sizing when rsg_set_team's blocks is 0: a phase's work falls as
// cells / workgroups while its grid barrier costs more with every workgroup,
// so the best team grows as the square root of the cell count (measured on
// bump channels: 16 workgroups best at 40k cells, 32-40 at 160k, 128 ahead of
// 256 at 2M).
constexpr double kTeamBlocksAt40k = 16.0;
constexpr int kTeamMinBlocks = 16;

}  // namespace

// rsg_chunk.hip, one object per (workgroup size, precision): the member kernels
hipError_t rsg_chunk_launch_256_0(int, RsgMemberDev*, RsgScalars*, hipStream_t);
hipError_t rsg_chunk_launch_256_1(int, RsgMemberDev*, RsgScalars*, hipStream_t);
hipError_t rsg_chunk_launch_512_0(int, RsgMemberDev*, RsgScalars*, hipStream_t);
hipError_t rsg_chunk_launch_512_1(int, RsgMemberDev*, RsgScalars*, hipStream_t);
int rsg_chunk_occupancy_256_0();
int rsg_chunk_occupancy_256_1();
int rsg_chunk_occupancy_512_0();
int rsg_chunk_occupancy_512_1();
hipError_t rsg_ldu_launch(RsgMemberDev* M, int solver, int maxIter, double tol, double* stats, hipStream_t s);

hipError_t rsg_comp_chunk_launch(int grid, RsgMemberDev* members, RsgScalars* statusOut, hipStream_t s);
int rsg_comp_chunk_occupancy();
static hipError_t rsg_chunk_launch(int bs, int prec, int grid, RsgMemberDev* members, RsgScalars* statusOut,
                                   hipStream_t s) {
    if (bs == 512) return prec ? rsg_chunk_launch_512_1(grid, members, statusOut, s)
                               : rsg_chunk_launch_512_0(grid, members, statusOut, s);
    return prec ? rsg_chunk_launch_256_1(grid, members, statusOut, s)
                : rsg_chunk_launch_256_0(grid, members, statusOut, s);
}

// Member workgroups of one shape a CU holds at once (the smaller of the two
// precisions' kernels).
static int rsg_chunk_occupancy(int bs) {
    const int a = bs == 512 ? rsg_chunk_occupancy_512_0() : rsg_chunk_occupancy_256_0();
    const int b = bs == 512 ? rsg_chunk_occupancy_512_1() : rsg_chunk_occupancy_256_1();
    return std::min(a, b);
}

namespace {

// --------------------------------------------------------------- arena
struct Arena {
    char* base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> freeBlocks;   // offset -> size
    size_t used = 0;
    bool alloc(size_t bytes, size_t& off) {
        bytes = (bytes + 255) & ~size_t(255);
        for (auto it = freeBlocks.begin(); it != freeBlocks.end(); ++it) {
            if (it->second >= bytes) {
                off = it->first;
                size_t rest = it->second - bytes;
                freeBlocks.erase(it);
                if (rest) freeBlocks[off + bytes] = rest;
                used += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        used -= bytes;
        auto it = freeBlocks.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != freeBlocks.end() && it->first + it->second == nx->first) {
            it->second += nx->second; freeBlocks.erase(nx);
        }
        if (it != freeBlocks.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) { pv->second += it->second; freeBlocks.erase(it); }
        }
    }
};

struct MeshEntry {
    rsg::PrepMesh prep;
    size_t off = 0, bytes = 0;      // arena block: arrays + the RsgMeshDev struct
    RsgMeshDev* devStruct = nullptr;
    int users = 0;
};

struct ProblemEntry {
    int meshId = 0;
    double nu = 0.0;
    bool bcZeroZ = false;           // every boundary face prescribes zero z velocity
    rsg_settings settings{};
    size_t off = 0, bytes = 0;      // arena block: BC arrays + RsgBcDev + RsgSettingsDev
    RsgBcDev* bcDev = nullptr;
    RsgSettingsDev* setDev = nullptr;
    int users = 0;
    int nObs = 0;                   // observables (rsg_problem_set_observations)
    size_t obsOff = 0, obsBytes = 0;
    RsgObsDev* obsDev = nullptr;
    // compressible problem (rsg_comp_problem_create): gas, alphaT, the
    // temperature BC arrays (arena block compOff)
    bool comp = false;
    RsgCompDev compDev{};
    size_t compOff = 0, compBytes = 0;
};

struct Ticket {
    int64_t id = 0;
    int problemId = 0;
    double coeffs[11] = {0};
    double init[6] = {0};
    int warm = 0;
    std::vector<double> warmCells;   // U(3n) p k om nuT F1 F2 Pk CDkw
    std::vector<double> warmBnd;     // U(3nB) p k om
    int warmTurb = 0;
    int precision = 0;
    bool noFields = false;           // RSG_INIT_NO_FIELDS: fields stay on the device
    int state = RSG_TICKET_PENDING;
    // results (host copies)
    RsgScalars sc{};
    std::vector<double> cells, bnd, hist, pred;
    double tSubmit = 0, tStart = 0, tEnd = 0;
    // device workspace
    size_t off = 0, bytes = 0;
    RsgMemberDev M{};
    // kernel-per-phase members (mesh of at least gridMinCells cells)
    bool grid = false;
    RsgMemberDev* dM = nullptr;      // the member struct on the device
    RsgBig* big = nullptr;
    // compressible solve: initial T / rho / mu, results T, rho, p_mech (cells), T (boundary)
    bool comp = false;
    RsgCompDev compDev{};            // member-kernel compressible solve: its arrays (arena) and struct
    RsgCompDev* dComp = nullptr;
    double compT0 = 0, compRho0 = 0, compMu0 = 0;
    std::vector<double> compCells, compTb;
    double compEnergyIters = 0;
};

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

// kernel-per-phase members run concurrently, at most this many at once
constexpr size_t kGridParallel = 4;
// compressible members on meshes up to this many cells run on the member
// kernel (one workgroup each, k_comp_chunk), larger ones kernel-per-phase;
// on the kernel-per-phase path (rsg_set_grid) members up to this size solve
// their linear systems on one workgroup between the grid-wide assembly launches
constexpr int kCompMemberSolveCells = 4096;

struct rsg_context {
    int device = 0;
    hipStream_t stream = nullptr;
    Arena arena;
    std::mutex mu;
    std::condition_variable cvEngine, cvDone;
    std::string lastError;
    std::chrono::steady_clock::time_point t0;
    std::map<int, std::unique_ptr<MeshEntry>> meshes;
    std::map<int, std::unique_ptr<ProblemEntry>> problems;
    std::map<int64_t, std::unique_ptr<Ticket>> tickets;
    std::deque<int64_t> pending;
    std::vector<int64_t> active;
    int holds = 0;           // rsg_hold_admission depth: > 0 admits no pending ticket
    int nextMesh = 1, nextProblem = 1;
    int64_t nextTicket = 1;
    int itersPerLaunch = 200;
    int blockSize = 0;       // 0: automatic (512)
    int sliceUs = 0;         // > 0: each launch runs every member for this long (no iteration cap)
    int clockKHz = 0;        // device wall clock (wall_clock64) frequency
    // team solves (rsg_set_team): members on meshes of at least teamMinCells
    // cells run alone on teamBlocks workgroups (0: sized by the cell count),
    // teamIters outer iterations a launch
    int teamMinCells = 0, teamBlocks = 0, teamIters = 20, teamCap = 0;
    // kernel-per-phase solves (rsg_set_grid): members on meshes of at least
    // gridMinCells cells (0: never), gridIters outer iterations per engine round
    int gridMinCells = 0, gridIters = 20;
    int gridLevels = 5;      // p' AMG levels of kernel-per-phase members (0: the reference's cap)
    int useGraphs = 1;       // V-cycles replayed from hipGraphs (rsg_set_graphs)
    double* dGred = nullptr;
    RsgMemberDev* dTeamM = nullptr;
    RsgScalars* dTeamOut = nullptr;
    int resident256 = 0, resident512 = 0;   // member workgroups the device holds at once, per shape
    int residentComp = 0;                    // compressible member workgroups (k_comp_chunk)
    size_t rot = 0;          // first active member of the next launch (rotating window)
    int schedule = RSG_SCHEDULE_ROUND_ROBIN;
    int maxSlots = 4096;
    std::thread engine;
    bool stop = false;
    bool engineError = false;
    // device staging for launches
    RsgMemberDev* dMembers = nullptr;
    RsgScalars* dStatus = nullptr;
    int dCap = 0;
    std::vector<RsgMemberDev> hMembers;
    std::vector<RsgScalars> hStatus;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // stats
    double nLaunch = 0, nAdmitted = 0, nRetired = 0, kernelMs = 0, maxActive = 0;
    int cuCount = 0, warpSize = 64;
    char* scratchDev = nullptr;   // this device's copy of the no-allocation scratch array
};

namespace {

int set_err(rsg_context* c, int code, const std::string& msg) {
    c->lastError = msg;
    return code;
}

#define HIPCHK(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_err(ctx, RSG_ERR_DEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

// Staging slots for the device-side descriptor structs that follow their
// arrays in one allocation.
constexpr size_t kMeshStructSlot = (sizeof(RsgMeshDev) + 255) & ~size_t(255);
static_assert(sizeof(RsgBcDev) <= 256, "problem staging reserves 256 bytes for RsgBcDev");

// host pointers of the result region of a member workspace
void member_init_scalars(RsgScalars& s, int turbEstablished) {
    memset(&s, 0, sizeof s);
    s.status = RSGM_RUNNING;
    s.amgBuiltIter = -1;
    s.coarseIter = -1000000000;
    s.normUx0 = s.normUy0 = s.normP0 = 1.0;
    s.lastK = s.lastOm = 1.0;
    s.turbEstablished = turbEstablished;
}

// Host arrays src[q] (n doubles each) to device arrays dst[q]: one copy of
// the device span when the destinations lie at a constant stride (the member
// workspace carves them back to back, 256-byte aligned; the alignment gaps
// inside the span belong to the same workspace), else one copy per array.
hipError_t scatter_from_host(double* const* dst, int k, int n, const double* const* src) {
    const ptrdiff_t st = k > 1 ? dst[1] - dst[0] : n;
    bool strided = st >= n;
    for (int q = 1; q < k && strided; ++q) strided = dst[q] == dst[0] + q * st;
    if (strided) {
        std::vector<double> span((size_t)(k - 1) * st + n, 0.0);
        for (int q = 0; q < k; ++q) memcpy(span.data() + (size_t)q * st, src[q], 8 * (size_t)n);
        return hipMemcpy(dst[0], span.data(), 8 * span.size(), hipMemcpyHostToDevice);
    }
    hipError_t e = hipSuccess;
    for (int q = 0; q < k && e == hipSuccess; ++q)
        e = hipMemcpy(dst[q], src[q], 8 * (size_t)n, hipMemcpyHostToDevice);
    return e;
}

// The compressible fields of a member-kernel solve (RsgCompDev's arrays);
// T, rho and p_mech first and back to back, so that retirement copies them at once.
void carve_comp(rsg::Carver& cv, int nC, int nB, RsgCompDev& Q) {
    const size_t n = nC, nb = std::max(nB, 1);
    Q.T = cv.take<double>(n); Q.rho = cv.take<double>(n); Q.pMech = cv.take<double>(n);
    Q.Tb = cv.take<double>(nb);
    Q.mu = cv.take<double>(n); Q.Told = cv.take<double>(n);
    Q.Ux0 = cv.take<double>(n); Q.Uy0 = cv.take<double>(n); Q.Uz0 = cv.take<double>(n);
    Q.mdu = cv.take<double>(n);
    for (int q = 0; q < 3; ++q) Q.gt[q] = cv.take<double>(n);
    Q.zeroB = cv.take<double>(nb);
}

// Admit one ticket (ctx->mu held).  Returns false when the arena is full.
bool admit(rsg_context* ctx, Ticket& T, std::string& err) {
    ProblemEntry& P = *ctx->problems[T.problemId];
    MeshEntry& ME = *ctx->meshes[P.meshId];
    const rsg::PrepMesh& pm = ME.prep;
    const int histCap = P.settings.record_history ? P.settings.max_iterations : 0;
    T.grid = (T.comp && pm.nC > kCompMemberSolveCells) || (ctx->gridMinCells > 0 && pm.nC >= ctx->gridMinCells);
    const bool compMember = T.comp && !T.grid;
    // lean workspace (no member-path AMG arrays) unless a compressible member
    // on a small mesh solves on one workgroup
    const bool lean = T.grid && !(T.comp && pm.nC <= kCompMemberSolveCells);
    rsg::Carver sz;
    {
        RsgMemberDev tmp;
        rsg::carve_member(sz, pm.nC, pm.nIF, pm.nB, pm.nnz0, histCap, tmp, lean);
        if (P.nObs > 0) sz.take<double>((size_t)P.nObs + 1);
        if (T.grid) sz.take<RsgMemberDev>(1);
        if (compMember) { RsgCompDev q; carve_comp(sz, pm.nC, pm.nB, q); sz.take<RsgCompDev>(1); }
    }
    size_t bytes = sz.off;
    size_t off;
    if (!ctx->arena.alloc(bytes, off)) return false;
    T.off = off; T.bytes = bytes;
    RsgMemberDev& M = T.M;
    memset(&M, 0, sizeof M);
    rsg::Carver cv; cv.base = ctx->arena.base + off;
    rsg::carve_member(cv, pm.nC, pm.nIF, pm.nB, pm.nnz0, histCap, M, lean);
    if (P.nObs > 0) { M.obs = P.obsDev; M.pred = cv.take<double>((size_t)P.nObs + 1); }
    T.dM = T.grid ? cv.take<RsgMemberDev>(1) : nullptr;
    if (compMember) {
        RsgCompDev& Q = T.compDev;
        Q = P.compDev;
        Q.Tinit = T.compT0; Q.rho0 = T.compRho0; Q.mu0 = T.compMu0; Q.energyIters = 0.0;
        carve_comp(cv, pm.nC, pm.nB, Q);
        T.dComp = cv.take<RsgCompDev>(1);
        M.comp = T.dComp;
    }
    M.mesh = ME.devStruct; M.bc = P.bcDev; M.set = P.setDev;
    M.nu = T.comp ? T.compMu0 / T.compRho0 : P.nu;   // compressible: the initial mu / rho
    for (int i = 0; i < 11; ++i) M.coeffs[i] = T.coeffs[i];
    for (int i = 0; i < 6; ++i) M.init[i] = T.init[i];
    M.initFromFields = T.warm;
    M.prec = T.precision;
    M.planarZ = rsg::planar_z(pm, P.bcZeroZ, &P.settings, T.warm ? 0.0 : T.init[2],
                              T.warm ? T.warmCells.data() : nullptr,
                              T.warm && pm.nB > 0 ? T.warmBnd.data() : nullptr) ? 1 : 0;
    RsgScalars s;
    member_init_scalars(s, T.warm ? T.warmTurb : 0);
    hipError_t e = hipMemcpyAsync(M.sc, &s, sizeof s, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && compMember) {
        e = hipMemsetAsync(T.compDev.zeroB, 0, 8 * (size_t)std::max(pm.nB, 1), ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(T.dComp, &T.compDev, sizeof(RsgCompDev), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);   // T.compDev is the source
    }
    if (e == hipSuccess && T.warm) {
        const int n = pm.nC, nB = pm.nB;
        double* cdst[] = {M.Ux, M.Uy, M.Uz, M.p, M.k, M.om, M.nuT, M.F1, M.F2, M.Pk, M.CDkw};
        // warm cells arrive as U interleaved (3n) then scalars; de-interleave on host
        std::vector<double> tmp(3 * (size_t)n);
        for (int c = 0; c < n; ++c) { tmp[c] = T.warmCells[3 * c]; tmp[n + c] = T.warmCells[3 * c + 1]; tmp[2 * n + c] = T.warmCells[3 * c + 2]; }
        const double* csrc[11];
        for (int q = 0; q < 3; ++q) csrc[q] = tmp.data() + (size_t)q * n;
        for (int q = 3; q < 11; ++q) csrc[q] = T.warmCells.data() + 3 * (size_t)n + (size_t)(q - 3) * n;
        e = scatter_from_host(cdst, 11, n, csrc);
        if (e == hipSuccess && nB > 0) {
            std::vector<double> tb(3 * (size_t)nB);
            for (int b = 0; b < nB; ++b) { tb[b] = T.warmBnd[3 * b]; tb[nB + b] = T.warmBnd[3 * b + 1]; tb[2 * nB + b] = T.warmBnd[3 * b + 2]; }
            double* bdst[] = {M.Ubx, M.Uby, M.Ubz, M.pb, M.kb, M.omb};
            const double* bsrc[6];
            for (int q = 0; q < 3; ++q) bsrc[q] = tb.data() + (size_t)q * nB;
            for (int q = 3; q < 6; ++q) bsrc[q] = T.warmBnd.data() + 3 * (size_t)nB + (size_t)(q - 3) * nB;
            e = scatter_from_host(bdst, 6, nB, bsrc);
        }
        T.warmCells.clear(); T.warmCells.shrink_to_fit();
        T.warmBnd.clear(); T.warmBnd.shrink_to_fit();
    }
    if (e != hipSuccess) {
        ctx->arena.release(off, bytes);
        err = std::string("admit: ") + hipGetErrorString(e);
        return false;
    }
    P.users++;
    T.state = RSG_TICKET_RUNNING;
    T.tStart = now_s(ctx->t0);
    return true;
}

// Copies a finished member's state to host (ctx->mu not held).
// k device arrays of n doubles each to dst[q * n ..]: one copy of their span
// when they lie at a constant stride (the member workspace carves them back
// to back, 256-byte aligned), else one copy per array.  A retirement costs a
// few synchronous copies instead of seventeen.
hipError_t gather_to_host(double* const* src, int k, int n, double* dst) {
    const ptrdiff_t st = k > 1 ? src[1] - src[0] : n;
    bool strided = st >= n;
    for (int q = 1; q < k && strided; ++q) strided = src[q] == src[0] + q * st;
    if (strided) {
        std::vector<double> span((size_t)(k - 1) * st + n);
        const hipError_t e = hipMemcpy(span.data(), src[0], 8 * span.size(), hipMemcpyDeviceToHost);
        for (int q = 0; q < k && e == hipSuccess; ++q)
            memcpy(dst + (size_t)q * n, span.data() + (size_t)q * st, 8 * (size_t)n);
        return e;
    }
    hipError_t e = hipSuccess;
    for (int q = 0; q < k && e == hipSuccess; ++q)
        e = hipMemcpy(dst + (size_t)q * n, src[q], 8 * (size_t)n, hipMemcpyDeviceToHost);
    return e;
}

hipError_t collect(Ticket& T, const rsg::PrepMesh& pm, const rsg_settings& st, int nObs) {
    const int n = pm.nC, nB = pm.nB;
    RsgMemberDev& M = T.M;
    hipError_t e = hipSuccess;
    if (!T.noFields) {
        T.cells.assign(11 * (size_t)n, 0.0);
        T.bnd.assign(6 * (size_t)(nB > 0 ? nB : 1), 0.0);
        double* csrc[] = {M.Ux, M.Uy, M.Uz, M.p, M.k, M.om, M.nuT, M.F1, M.F2, M.Pk, M.CDkw};
        e = gather_to_host(csrc, 11, n, T.cells.data());
        double* bsrc[] = {M.Ubx, M.Uby, M.Ubz, M.pb, M.kb, M.omb};
        if (e == hipSuccess && nB > 0) e = gather_to_host(bsrc, 6, nB, T.bnd.data());
    }
    if (e == hipSuccess && st.record_history && M.hist) {
        const int len = std::min(T.sc.histLen, M.histCap);
        T.hist.resize(6 * (size_t)std::max(len, 1));
        if (len > 0) e = hipMemcpy(T.hist.data(), M.hist, 48 * (size_t)len, hipMemcpyDeviceToHost);
        T.hist.resize(6 * (size_t)len);
    }
    if (e == hipSuccess && M.obs) {
        T.pred.resize((size_t)nObs + 1);
        e = hipMemcpy(T.pred.data(), M.pred, 8 * T.pred.size(), hipMemcpyDeviceToHost);
    }
    return e;
}

// A finished member-kernel compressible solve's T, rho, p_mech, T boundary
// values and energy-equation iterations to host (ctx->mu not held).
hipError_t collect_comp(Ticket& T, const rsg::PrepMesh& pm) {
    const size_t n = pm.nC, nB = pm.nB;
    const RsgCompDev& Q = T.compDev;
    T.compCells.assign(3 * n, 0.0);
    T.compTb.assign(std::max<size_t>(nB, 1), 0.0);
    double* src[] = {Q.T, Q.rho, Q.pMech};
    hipError_t e = gather_to_host(src, 3, (int)n, T.compCells.data());
    if (e == hipSuccess && nB > 0) e = hipMemcpy(T.compTb.data(), Q.Tb, 8 * nB, hipMemcpyDeviceToHost);
    RsgCompDev d;
    if (e == hipSuccess) e = hipMemcpy(&d, T.dComp, sizeof d, hipMemcpyDeviceToHost);
    if (e == hipSuccess) T.compEnergyIters = d.energyIters;
    return e;
}

void engine_loop(rsg_context* ctx) {
    hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(ctx->mu);
    while (true) {
        ctx->cvEngine.wait(lk, [&] {
            return ctx->stop || (!ctx->pending.empty() && ctx->holds == 0) || !ctx->active.empty();
        });
        if (ctx->stop) break;
        // admission (HBM footprint first-fit, FIFO), none while a submitter holds it
        std::string err;
        while (ctx->holds == 0 && !ctx->pending.empty() && (int)ctx->active.size() < ctx->maxSlots) {
            int64_t id = ctx->pending.front();
            Ticket& T = *ctx->tickets[id];
            if (!admit(ctx, T, err)) {
                if (!err.empty()) { T.state = RSG_TICKET_FAILED; ctx->pending.pop_front(); ctx->lastError = err; ctx->cvDone.notify_all(); err.clear(); continue; }
                if (ctx->active.empty()) {   // cannot ever fit
                    T.state = RSG_TICKET_FAILED; ctx->pending.pop_front();
                    ctx->lastError = "member workspace exceeds the HBM budget";
                    ctx->cvDone.notify_all();
                    continue;
                }
                break;
            }
            ctx->pending.pop_front();
            ctx->active.push_back(id);
            ctx->nAdmitted += 1;
        }
        if (ctx->active.empty()) continue;
        ctx->maxActive = std::max(ctx->maxActive, (double)ctx->active.size());
        // kernel-per-phase members: the first kGridParallel in admission
        // order run gridIters outer iterations each this round, concurrently
        // with one another and with the member launch below (own streams and
        // host threads); the others share the member launches
        std::vector<int64_t> gridIds, rest;
        for (int64_t id : ctx->active) {
            if (!ctx->tickets[id]->grid) rest.push_back(id);
            else if (gridIds.size() < kGridParallel) gridIds.push_back(id);
        }
        const int na = (int)rest.size();
        if (na > ctx->dCap) {
            if (ctx->dMembers) hipFree(ctx->dMembers);
            if (ctx->dStatus) hipFree(ctx->dStatus);
            int cap = std::max(na, 2 * ctx->dCap);
            hipMalloc(&ctx->dMembers, sizeof(RsgMemberDev) * cap);
            hipMalloc(&ctx->dStatus, sizeof(RsgScalars) * cap);
            ctx->dCap = cap;
        }
        // Automatic: 512-thread member workgroups, one per CU.  Two 256-thread
        // workgroups per CU were measured to deliver fewer member iterations
        // per second at every ensemble size (512 members of 1536 cells: 171k
        // against 256k per second).
        const int bs = ctx->blockSize ? ctx->blockSize : 512;
        // One launch holds at most as many members as the device runs at
        // once: a launch of more would run them in rounds, and its last
        // round would leave most CUs idle.  The window rotates over the
        // active members so that every one of them keeps advancing.
        const int resident = bs == 256 ? ctx->resident256 : ctx->resident512;
        const int nl = resident > 0 ? std::min(na, resident) : na;
        std::vector<int64_t> batch(nl);
        if (nl == 0) {
        } else if (ctx->schedule == RSG_SCHEDULE_LARGEST_FIRST && nl < na) {
            // longest-processing-time first: the members on the largest
            // meshes (the longest solves of an ensemble) hold the device,
            // the others fill the places they free (ties: admission order)
            std::vector<int64_t> order = rest;
            auto cells = [&](int64_t id) {
                return ctx->meshes[ctx->problems[ctx->tickets[id]->problemId]->meshId]->prep.nC;
            };
            std::stable_sort(order.begin(), order.end(),
                             [&](int64_t a, int64_t b) { return cells(a) > cells(b); });
            std::copy(order.begin(), order.begin() + nl, batch.begin());
        } else {
            const size_t r0 = ctx->rot % (size_t)na;
            for (int i = 0; i < nl; ++i) batch[i] = rest[(r0 + i) % na];
            ctx->rot = r0 + nl;
        }
        // members on meshes of at least teamMinCells cells run as teams, one
        // cooperative launch each, before the workgroup-per-member launch
        std::vector<int64_t> team;
        if (ctx->teamMinCells > 0) {
            std::vector<int64_t> reg;
            for (int64_t id : batch) {
                const int nC = ctx->meshes[ctx->problems[ctx->tickets[id]->problemId]->meshId]->prep.nC;
                (nC >= ctx->teamMinCells ? team : reg).push_back(id);
            }
            batch.swap(reg);
        }
        const int nr = (int)batch.size();
        std::vector<RsgMemberDev> teamM(team.size());
        std::vector<RsgScalars> teamSt(team.size());
        std::vector<int> teamNb(team.size());
        for (size_t q = 0; q < team.size(); ++q) {
            Ticket& T = *ctx->tickets[team[q]];
            T.M.iterBudget = ctx->teamIters;
            T.M.sliceTicks = 0;
            teamM[q] = T.M;
            // team size: the set one, else 16 sqrt(cells / 40,000) workgroups
            // (at least kTeamMinBlocks)
            const int nC = ctx->meshes[ctx->problems[T.problemId]->meshId]->prep.nC;
            const int want = ctx->teamBlocks > 0
                                 ? ctx->teamBlocks
                                 : std::max(kTeamMinBlocks, (int)std::lround(kTeamBlocksAt40k * std::sqrt(nC / 40000.0)));
            teamNb[q] = std::min(want, ctx->teamCap);
        }
        // FP64 members first, then mixed ones, then compressible ones: each
        // group runs its own kernel (one launch each, back to back)
        auto group = [&](int64_t id) {
            const Ticket& T = *ctx->tickets[id];
            return T.comp ? 2 : (T.precision == 0 ? 0 : 1);
        };
        std::stable_sort(batch.begin(), batch.end(), [&](int64_t a, int64_t b) { return group(a) < group(b); });
        int n64 = 0, nmix = 0;
        for (int64_t id : batch) { n64 += group(id) == 0; nmix += group(id) == 1; }
        ctx->hMembers.resize(nr);
        ctx->hStatus.resize(nr);
        for (int i = 0; i < nr; ++i) {
            Ticket& T = *ctx->tickets[batch[i]];
            if (ctx->sliceUs > 0 && ctx->clockKHz > 0) {
                T.M.iterBudget = 1 << 30;
                T.M.sliceTicks = (long long)ctx->sliceUs * ctx->clockKHz / 1000;
            } else {
                T.M.iterBudget = ctx->itersPerLaunch;
                T.M.sliceTicks = 0;
            }
            ctx->hMembers[i] = T.M;
        }
        std::vector<RsgMemberDev> gridM(gridIds.size());
        std::vector<RsgSettingsDev> gridSet(gridIds.size());
        std::vector<const rsg::PrepMesh*> gridMesh(gridIds.size());
        std::vector<RsgCompDev> gridComp(gridIds.size());
        const int gridLevels = ctx->gridLevels;
        const bool useGraphs = ctx->useGraphs != 0;
        for (size_t q = 0; q < gridIds.size(); ++q) {
            Ticket& T = *ctx->tickets[gridIds[q]];
            ProblemEntry& P = *ctx->problems[T.problemId];
            gridM[q] = T.M;
            gridSet[q] = rsg::settings_dev(&P.settings);
            gridMesh[q] = &ctx->meshes[P.meshId]->prep;
            if (T.comp) {
                gridComp[q] = P.compDev;
                gridComp[q].Tinit = T.compT0; gridComp[q].rho0 = T.compRho0; gridComp[q].mu0 = T.compMu0;
            }
        }
        lk.unlock();
        hipError_t e = hipSuccess;
        float ms = 0.f;
        std::vector<int64_t> gridDone;
        std::string gridErr;
        for (size_t q = 0; q < gridIds.size() && e == hipSuccess; ++q) {
            Ticket& T = *ctx->tickets[gridIds[q]];
            std::string berr;
            if (!T.big) {
                RsgScalars s0;
                member_init_scalars(s0, T.warm ? T.warmTurb : 0);
                T.big = rsg_big_create(gridM[q], T.dM, gridSet[q], *gridMesh[q], s0, ctx->stream, gridLevels,
                                       useGraphs, T.comp ? &gridComp[q] : nullptr, berr);
                if (!T.big) { e = hipErrorUnknown; gridErr = berr; break; }
            }
        }
        // kernel-per-phase members: each on its own stream from its own host
        // thread (a small member's phases leave most CUs idle, and every
        // Krylov iteration waits on the host), joined after the member launch
        const auto round0 = std::chrono::steady_clock::now();
        std::vector<int> gridRc(gridIds.size(), RSG_OK);
        std::vector<std::string> gridErrs(gridIds.size());
        std::vector<Ticket*> gridTk;
        for (int64_t id : gridIds) gridTk.push_back(ctx->tickets[id].get());
        std::vector<std::thread> gridTh;
        if (e == hipSuccess) {
            const int dev = ctx->device, iters = ctx->gridIters;
            for (size_t q = 0; q < gridIds.size(); ++q)
                gridTh.emplace_back([&gridTk, &gridRc, &gridErrs, q, dev, iters]() {
                    Ticket& T = *gridTk[q];
                    if (hipSetDevice(dev) != hipSuccess) { gridRc[q] = RSG_ERR_DEVICE; gridErrs[q] = "hipSetDevice"; return; }
                    gridRc[q] = rsg_big_run(T.big, iters, T.sc, gridErrs[q]);
                });
        }
        if (!team.empty() && ctx->teamCap <= 0) e = hipErrorInvalidConfiguration;
        if (!team.empty() && e == hipSuccess && !ctx->dGred) {
            e = hipMalloc(&ctx->dGred, sizeof(double) * 2 * RSG_TEAM_MAXB * 8);
            if (e == hipSuccess) e = hipMalloc(&ctx->dTeamM, sizeof(RsgMemberDev));
            if (e == hipSuccess) e = hipMalloc(&ctx->dTeamOut, sizeof(RsgScalars));
        }
        for (size_t q = 0; q < team.size() && e == hipSuccess; ++q) {
            e = hipMemcpyAsync(ctx->dTeamM, &teamM[q], sizeof(RsgMemberDev), hipMemcpyHostToDevice, ctx->stream);
            if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
            if (e == hipSuccess) e = rsg_team_launch(ctx->dTeamM, ctx->dTeamOut, ctx->dGred, teamNb[q], ctx->stream);
            if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(&teamSt[q], ctx->dTeamOut, sizeof(RsgScalars), hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            float tm = 0.f;
            if (e == hipSuccess && hipEventElapsedTime(&tm, ctx->ev0, ctx->ev1) == hipSuccess) ms += tm;
        }
        if (nr > 0 && e == hipSuccess) {
            e = hipMemcpyAsync(ctx->dMembers, ctx->hMembers.data(), sizeof(RsgMemberDev) * nr,
                               hipMemcpyHostToDevice, ctx->stream);
            hipEventRecord(ctx->ev0, ctx->stream);
            hipError_t el = hipSuccess;
            if (n64 > 0) el = rsg_chunk_launch(bs, 0, n64, ctx->dMembers, ctx->dStatus, ctx->stream);
            if (el == hipSuccess && nmix > 0)
                el = rsg_chunk_launch(bs, 1, nmix, ctx->dMembers + n64, ctx->dStatus + n64, ctx->stream);
            if (el == hipSuccess && nr > n64 + nmix)
                el = rsg_comp_chunk_launch(nr - n64 - nmix, ctx->dMembers + n64 + nmix, ctx->dStatus + n64 + nmix,
                                           ctx->stream);
            hipEventRecord(ctx->ev1, ctx->stream);
            if (e == hipSuccess) e = el;
            if (e == hipSuccess)
                e = hipMemcpyAsync(ctx->hStatus.data(), ctx->dStatus, sizeof(RsgScalars) * nr, hipMemcpyDeviceToHost,
                                   ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            float cm = 0.f;
            if (e == hipSuccess && hipEventElapsedTime(&cm, ctx->ev0, ctx->ev1) == hipSuccess) ms += cm;
        }
        for (std::thread& t : gridTh) t.join();
        if (!gridTh.empty())   // the round's wall time: the launches above overlap the grid members
            ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - round0).count();
        for (size_t q = 0; q < gridIds.size() && e == hipSuccess; ++q) {
            if (gridRc[q] != RSG_OK) { e = hipErrorUnknown; gridErr = gridErrs[q]; break; }
            if (gridTk[q]->sc.status != RSGM_RUNNING) gridDone.push_back(gridIds[q]);
        }
        // retire finished members
        std::vector<int64_t> done;
        for (int64_t id : gridDone) {
            if (e != hipSuccess) break;
            Ticket& T = *ctx->tickets[id];
            ProblemEntry& P = *ctx->problems[T.problemId];
            e = collect(T, ctx->meshes[P.meshId]->prep, P.settings, P.nObs);
            if (e == hipSuccess && T.comp) {
                const rsg::PrepMesh& pm = ctx->meshes[P.meshId]->prep;
                T.compCells.assign(3 * (size_t)pm.nC, 0.0);
                T.compTb.assign((size_t)std::max(pm.nB, 1), 0.0);
                const size_t nC = pm.nC;
                if (rsg_big_comp_fields(T.big, T.compCells.data(), T.compCells.data() + nC,
                                        T.compCells.data() + 2 * nC, T.compTb.data(), &T.compEnergyIters,
                                        gridErr) != RSG_OK)
                    e = hipErrorUnknown;
            }
            rsg_big_destroy(T.big);
            T.big = nullptr;
            done.push_back(id);
        }
        for (size_t q = 0; q < team.size() && e == hipSuccess; ++q) {
            Ticket& T = *ctx->tickets[team[q]];
            T.sc = teamSt[q];
            if (T.sc.status != RSGM_RUNNING) {
                ProblemEntry& P = *ctx->problems[T.problemId];
                e = collect(T, ctx->meshes[P.meshId]->prep, P.settings, P.nObs);
                done.push_back(team[q]);
            }
        }
        for (int i = 0; i < nr && e == hipSuccess; ++i) {
            Ticket& T = *ctx->tickets[batch[i]];
            T.sc = ctx->hStatus[i];
            if (T.sc.status != RSGM_RUNNING) {
                ProblemEntry& P = *ctx->problems[T.problemId];
                e = collect(T, ctx->meshes[P.meshId]->prep, P.settings, P.nObs);
                if (e == hipSuccess && T.comp) e = collect_comp(T, ctx->meshes[P.meshId]->prep);
                done.push_back(batch[i]);
            }
        }
        lk.lock();
        ctx->nLaunch += 1;
        ctx->kernelMs += ms;
        if (e != hipSuccess) {
            ctx->lastError = gridErr.empty() ? std::string("engine: ") + hipGetErrorString(e) : gridErr;
            ctx->engineError = true;
            for (int64_t id : ctx->active) ctx->tickets[id]->state = RSG_TICKET_FAILED;
            ctx->active.clear();
            ctx->cvDone.notify_all();
            break;
        }
        for (int64_t id : done) {
            Ticket& T = *ctx->tickets[id];
            ProblemEntry& P = *ctx->problems[T.problemId];
            ctx->arena.release(T.off, T.bytes);
            P.users--;
            T.state = RSG_TICKET_DONE;
            T.tEnd = now_s(ctx->t0);
            ctx->active.erase(std::find(ctx->active.begin(), ctx->active.end(), id));
            ctx->nRetired += 1;
        }
        if (!done.empty()) ctx->cvDone.notify_all();
    }
}

}  // namespace

// ------------------------------------------------------------------ C-ABI
extern "C" {

int rsg_context_create(int32_t device, uint64_t hbm_budget_bytes, rsg_context** out) {
    if (!out) return RSG_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RSG_ERR_DEVICE;
    if (device < 0 || device >= ndev) return RSG_ERR_ARG;
    std::unique_ptr<rsg_context> c(new rsg_context());
    c->device = device;
    c->t0 = std::chrono::steady_clock::now();
    if (hipSetDevice(device) != hipSuccess) return RSG_ERR_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return RSG_ERR_DEVICE;
    c->cuCount = prop.multiProcessorCount;
    c->warpSize = prop.warpSize;
    if (hipDeviceGetAttribute(&c->clockKHz, hipDeviceAttributeWallClockRate, device) != hipSuccess) c->clockKHz = 0;
    c->teamCap = rsg_team_capacity(device);
    {
        // the no-allocation solve's workspace: this device's copy of the
        // scratch array (host code uses this address, never the symbol)
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return RSG_ERR_DEVICE;
        c->scratchDev = static_cast<char*>(p);
    }
    c->resident256 = rsg_chunk_occupancy(256) * c->cuCount;
    c->resident512 = rsg_chunk_occupancy(512) * c->cuCount;
    c->residentComp = rsg_comp_chunk_occupancy() * c->cuCount;
    if (hbm_budget_bytes == 0) hbm_budget_bytes = 8ull << 30;
    if (hipMalloc(&c->arena.base, hbm_budget_bytes) != hipSuccess) return RSG_ERR_NOMEM;
    c->arena.size = hbm_budget_bytes;
    c->arena.freeBlocks[0] = hbm_budget_bytes;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return RSG_ERR_DEVICE;
    hipEventCreate(&c->ev0);
    hipEventCreate(&c->ev1);
    rsg_context* ctx = c.release();
    ctx->engine = std::thread(engine_loop, ctx);
    *out = ctx;
    return RSG_OK;
}

int rsg_context_destroy(rsg_context* ctx) {
    if (!ctx) return RSG_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->stop = true;
    }
    ctx->cvEngine.notify_all();
    if (ctx->engine.joinable()) ctx->engine.join();
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->tickets)
        if (kv.second->big) { rsg_big_destroy(kv.second->big); kv.second->big = nullptr; }
    if (ctx->dMembers) hipFree(ctx->dMembers);
    if (ctx->dStatus) hipFree(ctx->dStatus);
    if (ctx->dGred) hipFree(ctx->dGred);
    if (ctx->dTeamM) hipFree(ctx->dTeamM);
    if (ctx->dTeamOut) hipFree(ctx->dTeamOut);
    if (ctx->arena.base) hipFree(ctx->arena.base);
    hipEventDestroy(ctx->ev0);
    hipEventDestroy(ctx->ev1);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return RSG_OK;
}

const char* rsg_last_error(rsg_context* ctx) { return ctx ? ctx->lastError.c_str() : "null context"; }

double rsg_now(rsg_context* ctx) { return ctx ? now_s(ctx->t0) : 0.0; }

int rsg_device_info(rsg_context* ctx, int32_t* out) {
    if (!ctx || !out) return RSG_ERR_ARG;
    out[0] = ctx->cuCount; out[1] = ctx->warpSize;
    return RSG_OK;
}

int rsg_mesh_upload(rsg_context* ctx, const rsg_mesh_desc* mesh, int32_t* mesh_id) {
    if (!ctx || !mesh || !mesh_id) return RSG_ERR_ARG;
    std::unique_ptr<MeshEntry> ME(new MeshEntry());
    std::string err;
    if (!rsg::prep_mesh(mesh, ME->prep, err)) return set_err(ctx, RSG_ERR_ARG, err);
    rsg::MeshOffsets mo = rsg::mesh_offsets(ME->prep);
    const size_t bytes = mo.total + kMeshStructSlot;   // arrays, then the RsgMeshDev struct
    std::vector<char> stage(bytes, 0);
    rsg::mesh_stage(ME->prep, mo, stage.data());
    std::lock_guard<std::mutex> lk(ctx->mu);
    size_t off;
    if (!ctx->arena.alloc(bytes, off)) return set_err(ctx, RSG_ERR_NOMEM, "mesh exceeds the HBM budget");
    char* dbase = ctx->arena.base + off;
    RsgMeshDev view = rsg::mesh_view(ME->prep, mo, dbase);
    memcpy(stage.data() + mo.total, &view, sizeof view);
    hipSetDevice(ctx->device);
    hipError_t e = hipMemcpy(dbase, stage.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) { ctx->arena.release(off, bytes); return set_err(ctx, RSG_ERR_DEVICE, hipGetErrorString(e)); }
    ME->off = off; ME->bytes = bytes;
    ME->devStruct = reinterpret_cast<RsgMeshDev*>(dbase + mo.total);
    int id = ctx->nextMesh++;
    ctx->meshes[id] = std::move(ME);
    *mesh_id = id;
    return RSG_OK;
}

int rsg_mesh_release(rsg_context* ctx, int32_t mesh_id) {
    if (!ctx) return RSG_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->meshes.find(mesh_id);
    if (it == ctx->meshes.end()) return set_err(ctx, RSG_ERR_STATE, "unknown mesh id");
    if (it->second->users > 0) return set_err(ctx, RSG_ERR_STATE, "mesh still used by a problem");
    ctx->arena.release(it->second->off, it->second->bytes);
    ctx->meshes.erase(it);
    return RSG_OK;
}

int rsg_problem_create(rsg_context* ctx, int32_t mesh_id, const rsg_bc_desc* bc, double nu,
                       const rsg_settings* settings, int32_t* problem_id) {
    if (!ctx || !bc || !settings || !problem_id) return RSG_ERR_ARG;
    std::string err;
    if (!rsg::check_settings(settings, err)) return set_err(ctx, RSG_ERR_ARG, err);
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->meshes.find(mesh_id);
    if (it == ctx->meshes.end()) return set_err(ctx, RSG_ERR_STATE, "unknown mesh id");
    const int nB = it->second->prep.nB;
    if (nB > 0 && (!bc->vel_type || !bc->vel_value || !bc->p_type || !bc->p_value || !bc->k_type ||
                   !bc->k_value || !bc->om_type || !bc->om_value))
        return set_err(ctx, RSG_ERR_ARG, "bc descriptor: missing array");
    std::unique_ptr<ProblemEntry> P(new ProblemEntry());
    P->meshId = mesh_id; P->nu = nu; P->settings = *settings;
    P->bcZeroZ = nB == 0 || rsg::bc_zero_z(bc, nB);
    rsg::BcOffsets bo = rsg::bc_offsets(std::max(nB, 1));
    const size_t bytes = bo.total + 512 + sizeof(RsgBcDev) + sizeof(RsgSettingsDev);
    std::vector<char> stage(bytes, 0);
    if (nB > 0) rsg::bc_stage(bc, nB, bo, stage.data());
    size_t off;
    if (!ctx->arena.alloc(bytes, off)) return set_err(ctx, RSG_ERR_NOMEM, "problem exceeds the HBM budget");
    char* dbase = ctx->arena.base + off;
    RsgBcDev bview = rsg::bc_view(bo, dbase);
    RsgSettingsDev sdev = rsg::settings_dev(settings);
    const size_t oB = bo.total, oS = bo.total + 256;
    memcpy(stage.data() + oB, &bview, sizeof bview);
    memcpy(stage.data() + oS, &sdev, sizeof sdev);
    hipSetDevice(ctx->device);
    hipError_t e = hipMemcpy(dbase, stage.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) { ctx->arena.release(off, bytes); return set_err(ctx, RSG_ERR_DEVICE, hipGetErrorString(e)); }
    P->off = off; P->bytes = bytes;
    P->bcDev = reinterpret_cast<RsgBcDev*>(dbase + oB);
    P->setDev = reinterpret_cast<RsgSettingsDev*>(dbase + oS);
    it->second->users++;
    int id = ctx->nextProblem++;
    ctx->problems[id] = std::move(P);
    *problem_id = id;
    return RSG_OK;
}

int rsg_problem_release(rsg_context* ctx, int32_t problem_id) {
    if (!ctx) return RSG_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->problems.find(problem_id);
    if (it == ctx->problems.end()) return set_err(ctx, RSG_ERR_STATE, "unknown problem id");
    if (it->second->users > 0) return set_err(ctx, RSG_ERR_STATE, "problem has solves in flight");
    for (auto& kv : ctx->tickets)
        if (kv.second->problemId == problem_id && kv.second->state == RSG_TICKET_PENDING)
            return set_err(ctx, RSG_ERR_STATE, "problem has pending solves");
    ctx->meshes[it->second->meshId]->users--;
    ctx->arena.release(it->second->off, it->second->bytes);
    if (it->second->obsBytes) ctx->arena.release(it->second->obsOff, it->second->obsBytes);
    if (it->second->compBytes) ctx->arena.release(it->second->compOff, it->second->compBytes);
    ctx->problems.erase(it);
    return RSG_OK;
}

int rsg_problem_set_observations(rsg_context* ctx, int32_t problem_id, int32_t n, const rsg_observation* obs,
                                 int32_t n_faces, const int32_t* faces, const double* face_x) {
    if (!ctx || n < 0 || (n > 0 && !obs) || n_faces < 0 || (n_faces > 0 && (!faces || !face_x))) return RSG_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->problems.find(problem_id);
    if (it == ctx->problems.end()) return set_err(ctx, RSG_ERR_STATE, "unknown problem id");
    ProblemEntry& P = *it->second;
    if (P.obsBytes) return set_err(ctx, RSG_ERR_STATE, "observations already set for this problem");
    if (P.users > 0) return set_err(ctx, RSG_ERR_STATE, "problem has solves in flight");
    const rsg::PrepMesh& pm = ctx->meshes[P.meshId]->prep;
    for (int j = 0; j < n_faces; ++j)
        if (faces[j] < pm.nIF || faces[j] >= pm.nF) return set_err(ctx, RSG_ERR_ARG, "observation face is not a boundary face");
    for (int i = 0; i < n; ++i) {
        const rsg_observation& o = obs[i];
        const bool point = o.type == RSG_OBS_PRESSURE_TAP || o.type == RSG_OBS_VELOCITY;
        const bool patch = o.type == RSG_OBS_DRAG || o.type == RSG_OBS_SEPARATION || o.type == RSG_OBS_REATTACHMENT;
        if (o.type < 0 || o.type > RSG_OBS_REATTACHMENT) return set_err(ctx, RSG_ERR_ARG, "unknown observation type");
        if (point && (o.index < 0 || o.index >= pm.nC)) return set_err(ctx, RSG_ERR_ARG, "observation cell out of range");
        if (o.type == RSG_OBS_SKIN_FRICTION && (o.index < -1 || (o.index >= 0 && (o.index < pm.nIF || o.index >= pm.nF))))
            return set_err(ctx, RSG_ERR_ARG, "skin-friction face is not a boundary face");
        if (o.type == RSG_OBS_VELOCITY && (o.component < 0 || o.component > 2))
            return set_err(ctx, RSG_ERR_ARG, "velocity component must be 0, 1 or 2");
        if (patch && (o.face_begin < 0 || o.face_end < o.face_begin || o.face_end > n_faces))
            return set_err(ctx, RSG_ERR_ARG, "observation face range out of bounds");
    }
    if (n == 0) return RSG_OK;
    // image: RsgObsDev, then type idx comp fb fe [n] ints, par [6n], faces [n_faces], fx [n_faces]
    rsg::Carver cv;
    cv.take<RsgObsDev>(1);
    for (int q = 0; q < 5; ++q) cv.take<int>(n);
    cv.take<double>(6 * (size_t)n);
    cv.take<int>(std::max(n_faces, 1));
    cv.take<double>(std::max(n_faces, 1));
    const size_t bytes = cv.off;
    size_t off;
    if (!ctx->arena.alloc(bytes, off)) return set_err(ctx, RSG_ERR_NOMEM, "observations exceed the HBM budget");
    char* dbase = ctx->arena.base + off;
    std::vector<char> stage(bytes, 0);
    rsg::Carver hv; hv.base = stage.data();
    rsg::Carver dv; dv.base = dbase;
    RsgObsDev* hs = hv.take<RsgObsDev>(1); RsgObsDev* ds = dv.take<RsgObsDev>(1);
    int* hi[5]; const int* di[5];
    for (int q = 0; q < 5; ++q) { hi[q] = hv.take<int>(n); di[q] = dv.take<int>(n); }
    double* hp = hv.take<double>(6 * (size_t)n); const double* dp = dv.take<double>(6 * (size_t)n);
    int* hf = hv.take<int>(std::max(n_faces, 1)); const int* df = dv.take<int>(std::max(n_faces, 1));
    double* hx = hv.take<double>(std::max(n_faces, 1)); const double* dx = dv.take<double>(std::max(n_faces, 1));
    for (int i = 0; i < n; ++i) {
        const rsg_observation& o = obs[i];
        hi[0][i] = o.type; hi[1][i] = o.index; hi[2][i] = o.component; hi[3][i] = o.face_begin; hi[4][i] = o.face_end;
        double* pr = hp + 6 * (size_t)i;
        pr[0] = o.ref_velocity; pr[1] = o.ref_area; pr[2] = o.ref_density; pr[3] = o.ref_pressure;
        pr[4] = o.observed;
        double var = o.sigma * o.sigma + o.sigma_model * o.sigma_model;   // Measurement::totalVariance
        pr[5] = var < 1e-30 ? 1e-30 : var;
    }
    for (int j = 0; j < n_faces; ++j) { hf[j] = faces[j]; hx[j] = face_x[j]; }
    hs->n = n; hs->type = di[0]; hs->idx = di[1]; hs->comp = di[2]; hs->fb = di[3]; hs->fe = di[4];
    hs->par = dp; hs->faces = df; hs->fx = dx;
    hipSetDevice(ctx->device);
    hipError_t e = hipMemcpy(dbase, stage.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) { ctx->arena.release(off, bytes); return set_err(ctx, RSG_ERR_DEVICE, hipGetErrorString(e)); }
    P.nObs = n; P.obsOff = off; P.obsBytes = bytes; P.obsDev = ds;
    return RSG_OK;
}

int rsg_submit(rsg_context* ctx, int32_t problem_id, const double* coeffs, const rsg_init* init,
               int32_t precision, int64_t* ticket) {
    if (!ctx || !coeffs || !init || !ticket) return RSG_ERR_ARG;
    if (precision != RSG_PREC_FP64 && precision != RSG_PREC_MIXED)
        return set_err(ctx, RSG_ERR_ARG, "unknown precision code");
    std::unique_ptr<Ticket> T(new Ticket());
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (ctx->engineError) return set_err(ctx, RSG_ERR_DEVICE, ctx->lastError);
        auto it = ctx->problems.find(problem_id);
        if (it == ctx->problems.end()) return set_err(ctx, RSG_ERR_STATE, "unknown problem id");
        if (it->second->comp) return set_err(ctx, RSG_ERR_ARG, "compressible problem: submit with rsg_comp_submit");
        const rsg::PrepMesh& pm = ctx->meshes[it->second->meshId]->prep;
        T->problemId = problem_id;
        for (int i = 0; i < 11; ++i) T->coeffs[i] = coeffs[i];
        T->init[0] = init->U[0]; T->init[1] = init->U[1]; T->init[2] = init->U[2];
        T->init[3] = init->p; T->init[4] = init->k; T->init[5] = init->omega;
        T->precision = precision;
        T->noFields = (init->kind & RSG_INIT_NO_FIELDS) != 0;
        const int kind = init->kind & ~RSG_INIT_NO_FIELDS;
        if (kind == 1) {
            const int n = pm.nC, nB = pm.nB;
            if (!init->U_cells || !init->p_cells || !init->k_cells || !init->om_cells || !init->nuT ||
                !init->F1 || !init->F2 || !init->Pk || !init->CDkw ||
                (nB > 0 && (!init->U_bnd || !init->p_bnd || !init->k_bnd || !init->om_bnd)))
                return set_err(ctx, RSG_ERR_ARG, "warm start: missing field array");
            T->warm = 1;
            T->warmTurb = init->turb_established;
            T->warmCells.resize(11 * (size_t)n);
            memcpy(T->warmCells.data(), init->U_cells, 24 * (size_t)n);
            const double* sc[] = {init->p_cells, init->k_cells, init->om_cells, init->nuT, init->F1,
                                  init->F2, init->Pk, init->CDkw};
            for (int q = 0; q < 8; ++q) memcpy(T->warmCells.data() + 3 * (size_t)n + (size_t)q * n, sc[q], 8 * (size_t)n);
            T->warmBnd.resize(6 * (size_t)std::max(nB, 1));
            if (nB > 0) {
                memcpy(T->warmBnd.data(), init->U_bnd, 24 * (size_t)nB);
                memcpy(T->warmBnd.data() + 3 * (size_t)nB, init->p_bnd, 8 * (size_t)nB);
                memcpy(T->warmBnd.data() + 4 * (size_t)nB, init->k_bnd, 8 * (size_t)nB);
                memcpy(T->warmBnd.data() + 5 * (size_t)nB, init->om_bnd, 8 * (size_t)nB);
            }
        } else if (kind != 0) {
            return set_err(ctx, RSG_ERR_ARG, "init kind must be 0 (uniform) or 1 (fields), optionally with "
                                             "RSG_INIT_NO_FIELDS");
        }
        T->id = ctx->nextTicket++;
        T->tSubmit = now_s(ctx->t0);
        *ticket = T->id;
        ctx->pending.push_back(T->id);
        ctx->tickets[T->id] = std::move(T);
    }
    ctx->cvEngine.notify_all();
    return RSG_OK;
}

int rsg_comp_problem_create(rsg_context* ctx, int32_t mesh_id, const rsg_bc_desc* bc, const rsg_comp_desc* comp,
                            const rsg_settings* settings, int32_t* problem_id) {
    if (!ctx || !bc || !comp || !settings || !problem_id) return RSG_ERR_ARG;
    if (!(comp->gamma > 1.0) || !(comp->R > 0.0) || !(comp->Pr > 0.0) || !(comp->Pr_T > 0.0) ||
        !(comp->mu_ref > 0.0) || !(comp->T_ref > 0.0) || !(comp->S_suth >= 0.0) || !(comp->alpha_t > 0.0) ||
        !(comp->alpha_t <= 1.0))
        return set_err(ctx, RSG_ERR_ARG, "compressible problem: gas constants or alpha_t out of range");
    int nB = 0;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        auto it = ctx->meshes.find(mesh_id);
        if (it == ctx->meshes.end()) return set_err(ctx, RSG_ERR_STATE, "unknown mesh id");
        nB = it->second->prep.nB;
    }
    if (nB > 0 && (!comp->t_type || !comp->t_value))
        return set_err(ctx, RSG_ERR_ARG, "compressible problem: missing temperature BC array");
    for (int b = 0; b < nB; ++b)
        if (comp->t_type[b] < RSG_TBC_DIRICHLET || comp->t_type[b] > RSG_TBC_KEEP)
            return set_err(ctx, RSG_ERR_ARG, "compressible problem: unknown temperature BC type");
    int32_t pid = -1;
    const int rc = rsg_problem_create(ctx, mesh_id, bc, 0.0, settings, &pid);
    if (rc != RSG_OK) return rc;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ProblemEntry& P = *ctx->problems[pid];
    const size_t nb1 = (size_t)std::max(nB, 1);
    const size_t oV = (4 * nb1 + 255) & ~size_t(255);
    const size_t bytes = oV + 8 * nb1;
    std::vector<char> stage(bytes, 0);
    for (int b = 0; b < nB; ++b) {
        const int32_t t = comp->t_type[b] == RSG_TBC_DIRICHLET ? RSGT_DIRICHLET
                          : comp->t_type[b] == RSG_TBC_ZERO_GRADIENT ? RSGT_ZERO_GRADIENT : 2;
        memcpy(stage.data() + 4 * b, &t, 4);
    }
    if (nB > 0) memcpy(stage.data() + oV, comp->t_value, 8 * (size_t)nB);
    size_t off;
    hipError_t e = hipErrorOutOfMemory;
    if (ctx->arena.alloc(bytes, off)) {
        e = hipMemcpy(ctx->arena.base + off, stage.data(), bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) ctx->arena.release(off, bytes);
    }
    if (e != hipSuccess) {
        ctx->meshes[P.meshId]->users--;
        ctx->arena.release(P.off, P.bytes);
        ctx->problems.erase(pid);
        return set_err(ctx, e == hipErrorOutOfMemory ? RSG_ERR_NOMEM : RSG_ERR_DEVICE,
                       "compressible problem: temperature BCs do not fit");
    }
    P.comp = true;
    P.compOff = off; P.compBytes = bytes;
    RsgCompDev& Q = P.compDev;
    Q.tT = reinterpret_cast<const int*>(ctx->arena.base + off);
    Q.tV = reinterpret_cast<const double*>(ctx->arena.base + off + oV);
    Q.gamma = comp->gamma; Q.R = comp->R; Q.Pr = comp->Pr; Q.PrT = comp->Pr_T;
    Q.muRef = comp->mu_ref; Q.Tref = comp->T_ref; Q.S = comp->S_suth;
    Q.alphaT = comp->alpha_t;
    *problem_id = pid;
    return RSG_OK;
}

int rsg_comp_submit(rsg_context* ctx, int32_t problem_id, const double* coeffs, const rsg_comp_init* init,
                    int64_t* ticket) {
    if (!ctx || !coeffs || !init || !ticket) return RSG_ERR_ARG;
    if (!(init->T > 0.0) || !(init->p > 0.0) || !std::isfinite(init->k) || !std::isfinite(init->omega))
        return set_err(ctx, RSG_ERR_ARG, "compressible initial state: T and p must be positive");
    std::unique_ptr<Ticket> T(new Ticket());
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (ctx->engineError) return set_err(ctx, RSG_ERR_DEVICE, ctx->lastError);
        auto it = ctx->problems.find(problem_id);
        if (it == ctx->problems.end()) return set_err(ctx, RSG_ERR_STATE, "unknown problem id");
        if (!it->second->comp) return set_err(ctx, RSG_ERR_ARG, "not a compressible problem");
        const RsgCompDev& Q = it->second->compDev;
        T->problemId = problem_id;
        for (int i = 0; i < 11; ++i) T->coeffs[i] = coeffs[i];
        // initUniform (CompressibleSIMPLESolver.cpp:27-72): density from the
        // thermodynamic state, the stored (mechanical) pressure, Sutherland mu
        const double rho0 = init->p / (Q.R * init->T);
        const double ratio = init->T / Q.Tref;
        const double mu0 = Q.muRef * ratio * std::sqrt(ratio) * (Q.Tref + Q.S) / (init->T + Q.S);
        T->init[0] = init->U[0]; T->init[1] = init->U[1]; T->init[2] = init->U[2];
        T->init[3] = init->p + (2.0 / 3.0) * rho0 * std::max(init->k, 0.0);
        T->init[4] = init->k; T->init[5] = init->omega;
        T->comp = true;
        T->compT0 = init->T; T->compRho0 = rho0; T->compMu0 = mu0;
        T->precision = RSG_PREC_FP64;
        T->id = ctx->nextTicket++;
        T->tSubmit = now_s(ctx->t0);
        *ticket = T->id;
        ctx->pending.push_back(T->id);
        ctx->tickets[T->id] = std::move(T);
    }
    ctx->cvEngine.notify_all();
    return RSG_OK;
}

int rsg_comp_fetch(rsg_context* ctx, int64_t ticket, double* T_cells, double* rho, double* p_mech, double* T_bnd,
                   double* energy_iters) {
    if (!ctx) return RSG_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->tickets.find(ticket);
    if (it == ctx->tickets.end()) return set_err(ctx, RSG_ERR_STATE, "unknown ticket");
    Ticket& T = *it->second;
    if (!T.comp) return set_err(ctx, RSG_ERR_ARG, "not a compressible solve");
    if (T.state != RSG_TICKET_DONE) return set_err(ctx, RSG_ERR_STATE, "ticket not finished");
    const rsg::PrepMesh& pm = ctx->meshes[ctx->problems[T.problemId]->meshId]->prep;
    const size_t n = pm.nC;
    if (T_cells) memcpy(T_cells, T.compCells.data(), 8 * n);
    if (rho) memcpy(rho, T.compCells.data() + n, 8 * n);
    if (p_mech) memcpy(p_mech, T.compCells.data() + 2 * n, 8 * n);
    if (T_bnd && pm.nB > 0) memcpy(T_bnd, T.compTb.data(), 8 * (size_t)pm.nB);
    if (energy_iters) *energy_iters = T.compEnergyIters;
    return RSG_OK;
}

int rsg_poll(rsg_context* ctx, const int64_t* tickets, int32_t n, int32_t* state_out) {
    if (!ctx || (n > 0 && (!tickets || !state_out))) return RSG_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (int i = 0; i < n; ++i) {
        auto it = ctx->tickets.find(tickets[i]);
        if (it == ctx->tickets.end()) return set_err(ctx, RSG_ERR_STATE, "unknown ticket");
        state_out[i] = it->second->state;
    }
    return RSG_OK;
}

int rsg_wait(rsg_context* ctx, const int64_t* tickets, int32_t n, double timeout_s) {
    if (!ctx || (n > 0 && !tickets)) return RSG_ERR_ARG;
    std::unique_lock<std::mutex> lk(ctx->mu);
    for (int i = 0; i < n; ++i)
        if (ctx->tickets.find(tickets[i]) == ctx->tickets.end()) return set_err(ctx, RSG_ERR_STATE, "unknown ticket");
    auto finished = [&] {
        for (int i = 0; i < n; ++i) {
            int s = ctx->tickets[tickets[i]]->state;
            if (s != RSG_TICKET_DONE && s != RSG_TICKET_FAILED) return false;
        }
        return true;
    };
    if (timeout_s < 0) { ctx->cvDone.wait(lk, finished); return RSG_OK; }
    if (!ctx->cvDone.wait_for(lk, std::chrono::duration<double>(timeout_s), finished))
        return set_err(ctx, RSG_ERR_TIMEOUT, "wait timed out");
    return RSG_OK;
}

int rsg_fetch(rsg_context* ctx, int64_t ticket, rsg_result* out) {
    if (!ctx || !out) return RSG_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->tickets.find(ticket);
    if (it == ctx->tickets.end()) return set_err(ctx, RSG_ERR_STATE, "unknown ticket");
    Ticket& T = *it->second;
    out->status = T.state;
    out->t_submit = T.tSubmit; out->t_start = T.tStart; out->t_end = T.tEnd;
    if (T.state == RSG_TICKET_FAILED) return RSG_OK;
    if (T.state != RSG_TICKET_DONE) return set_err(ctx, RSG_ERR_STATE, "ticket not finished");
    const rsg::PrepMesh& pm = ctx->meshes[ctx->problems[T.problemId]->meshId]->prep;
    const int n = pm.nC, nB = pm.nB;
    out->final_iter = T.sc.finalIter;
    out->converged = T.sc.status == RSGM_CONVERGED;
    out->diverged = T.sc.status == RSGM_DIVERGED;
    out->turb_established = T.sc.turbEstablished;
    for (int i = 0; i < 6; ++i) out->inner_iters[i] = T.sc.innerIters[i];
    if (T.noFields && (out->U || out->p || out->k || out->omega || out->nuT || out->F1 || out->F2 || out->Pk ||
                       out->CDkw || out->U_bnd || out->p_bnd || out->k_bnd || out->om_bnd))
        return set_err(ctx, RSG_ERR_ARG, "fields of a solve submitted with RSG_INIT_NO_FIELDS were not collected: "
                                         "pass NULL field pointers");
    const double* c = T.cells.data();
    if (out->U) for (int i = 0; i < n; ++i) { out->U[3 * i] = c[i]; out->U[3 * i + 1] = c[n + i]; out->U[3 * i + 2] = c[2 * n + i]; }
    double* dst[] = {out->p, out->k, out->omega, out->nuT, out->F1, out->F2, out->Pk, out->CDkw};
    for (int q = 0; q < 8; ++q) if (dst[q]) memcpy(dst[q], c + (size_t)(3 + q) * n, 8 * (size_t)n);
    const double* b = T.bnd.data();
    if (out->U_bnd) for (int i = 0; i < nB; ++i) { out->U_bnd[3 * i] = b[i]; out->U_bnd[3 * i + 1] = b[nB + i]; out->U_bnd[3 * i + 2] = b[2 * nB + i]; }
    double* bd[] = {out->p_bnd, out->k_bnd, out->om_bnd};
    for (int q = 0; q < 3; ++q) if (bd[q] && nB > 0) memcpy(bd[q], b + (size_t)(3 + q) * nB, 8 * (size_t)nB);
    out->history_len = 0;
    if (out->history && out->history_cap > 0) {
        int len = std::min<int>((int)(T.hist.size() / 6), out->history_cap);
        memcpy(out->history, T.hist.data(), 48 * (size_t)len);
        out->history_len = len;
    }
    out->n_predictions = 0;
    out->log_lik = 0.0;
    if (!T.pred.empty()) {
        const int nObs = (int)T.pred.size() - 1;
        out->n_predictions = nObs;
        out->log_lik = T.pred[nObs];
        if (out->predictions && out->predictions_cap >= nObs) memcpy(out->predictions, T.pred.data(), 8 * (size_t)nObs);
    }
    return RSG_OK;
}

int rsg_release(rsg_context* ctx, int64_t ticket) {
    if (!ctx) return RSG_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->tickets.find(ticket);
    if (it == ctx->tickets.end()) return set_err(ctx, RSG_ERR_STATE, "unknown ticket");
    if (it->second->state == RSG_TICKET_RUNNING) return set_err(ctx, RSG_ERR_STATE, "ticket still running");
    if (it->second->state == RSG_TICKET_PENDING) {
        ctx->pending.erase(std::find(ctx->pending.begin(), ctx->pending.end(), ticket));
    }
    ctx->tickets.erase(it);
    return RSG_OK;
}

int rsg_set_launch(rsg_context* ctx, int32_t iters_per_launch, int32_t block_size) {
    if (!ctx || iters_per_launch <= 0) return RSG_ERR_ARG;
    if (block_size != 0 && block_size != 256 && block_size != 512)
        return set_err(ctx, RSG_ERR_ARG, "block size must be 0 (automatic), 256 or 512");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->itersPerLaunch = iters_per_launch;
    ctx->blockSize = block_size;
    return RSG_OK;
}

int rsg_set_team(rsg_context* ctx, int32_t min_cells, int32_t blocks, int32_t iters_per_launch) {
    if (!ctx) return RSG_ERR_ARG;
    if (min_cells < 0 || blocks < 0 || blocks > RSG_TEAM_MAXB || iters_per_launch < 1)
        return set_err(ctx, RSG_ERR_ARG, "team: min_cells >= 0, 0 <= blocks <= 256, iters_per_launch >= 1");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (min_cells > 0 && ctx->teamCap <= 0) return set_err(ctx, RSG_ERR_UNSUPPORTED, "team launches are not available");
    ctx->teamMinCells = min_cells; ctx->teamBlocks = blocks; ctx->teamIters = iters_per_launch;
    return RSG_OK;
}

int rsg_set_grid(rsg_context* ctx, int32_t min_cells, int32_t iters_per_launch) {
    if (!ctx) return RSG_ERR_ARG;
    if (min_cells < 0) return set_err(ctx, RSG_ERR_ARG, "min_cells must be >= 0");
    if (iters_per_launch <= 0) return set_err(ctx, RSG_ERR_ARG, "iters_per_launch must be > 0");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->gridMinCells = min_cells; ctx->gridIters = iters_per_launch;
    return RSG_OK;
}

int rsg_set_grid_levels(rsg_context* ctx, int32_t max_levels) {
    if (!ctx) return RSG_ERR_ARG;
    if (max_levels < 0 || max_levels == 1 || max_levels > RSG_MAX_LEVELS)
        return set_err(ctx, RSG_ERR_ARG, "max_levels must be 0 or 2..20");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->gridLevels = max_levels;
    return RSG_OK;
}

int rsg_set_graphs(rsg_context* ctx, int32_t enable) {
    if (!ctx) return RSG_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->useGraphs = enable ? 1 : 0;
    return RSG_OK;
}

int rsg_set_schedule(rsg_context* ctx, int32_t policy) {
    if (!ctx) return RSG_ERR_ARG;
    if (policy != RSG_SCHEDULE_ROUND_ROBIN && policy != RSG_SCHEDULE_LARGEST_FIRST)
        return set_err(ctx, RSG_ERR_ARG, "unknown schedule policy");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->schedule = policy;
    return RSG_OK;
}

int rsg_set_time_slice(rsg_context* ctx, int32_t microseconds) {
    if (!ctx || microseconds < 0) return RSG_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (microseconds > 0 && ctx->clockKHz <= 0)
        return set_err(ctx, RSG_ERR_UNSUPPORTED, "device wall clock rate unknown: time slices unavailable");
    ctx->sliceUs = microseconds;
    return RSG_OK;
}

int rsg_hold_admission(rsg_context* ctx, int32_t hold) {
    if (!ctx || (hold != 0 && hold != 1)) return RSG_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (hold) ctx->holds += 1;
        else if (ctx->holds > 0) ctx->holds -= 1;
        else return set_err(ctx, RSG_ERR_STATE, "rsg_hold_admission: no hold to release");
    }
    ctx->cvEngine.notify_all();
    return RSG_OK;
}

int rsg_stats(rsg_context* ctx, double* out) {
    if (!ctx || !out) return RSG_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    out[0] = ctx->nLaunch; out[1] = ctx->nAdmitted; out[2] = ctx->nRetired;
    out[3] = ctx->kernelMs; out[4] = ctx->maxActive;
    return RSG_OK;
}

}  // extern "C"

namespace {

// Argument checks and host-side topology of one stand-alone LDU solve.
int ldu_prepare(rsg_context* ctx, int32_t n, int32_t nif, const int32_t* own, const int32_t* nbr,
                const double* diag, const double* upper, const double* lower, const double* b, double* x,
                int32_t solver, int32_t precision, double* stats, rsg::PrepMesh& pm) {
    if (!ctx || n <= 0 || nif < 0 || !own || (nif > 0 && !nbr) || !diag || (nif > 0 && (!upper || !lower)) ||
        !b || !x || !stats)
        return RSG_ERR_ARG;
    if (precision != RSG_PREC_FP64 && precision != RSG_PREC_MIXED)
        return set_err(ctx, RSG_ERR_ARG, "unknown precision code");
    if (solver != RSG_SOLVER_AMG && solver != RSG_SOLVER_PCG && solver != RSG_SOLVER_BICGSTAB &&
        solver != RSG_SOLVER_GAUSS_SEIDEL)
        return set_err(ctx, RSG_ERR_ARG, "unknown solver code");
    for (int f = 0; f < nif; ++f)
        if (own[f] < 0 || own[f] >= n || nbr[f] < 0 || nbr[f] >= n) return set_err(ctx, RSG_ERR_ARG, "face index out of range");
    std::string err;
    if (!rsg::prep_topology(n, nif, own, nbr, pm, err)) return set_err(ctx, RSG_ERR_ARG, err);
    return RSG_OK;
}

// Device bytes of one stand-alone LDU solve laid out from `base`: the mesh
// image and its struct, the member struct, the member workspace.
size_t ldu_block_bytes(const rsg::PrepMesh& pm, const rsg::MeshOffsets& mo) {
    const size_t mbytes = mo.total + kMeshStructSlot;
    const size_t wbytes = rsg::member_bytes(pm.nC, pm.nIF, pm.nB, pm.nnz0, 0);
    return mbytes + 256 + sizeof(RsgMemberDev) + 256 + wbytes + 256;
}

// Runs one stand-alone LDU solve in the device block at `base` (ctx->mu
// held): stages the topology and the system, launches the one-workgroup
// solve and copies x and the statistics back.  Allocates no device memory.
int ldu_solve_at(rsg_context* ctx, char* base, const rsg::PrepMesh& pm, const rsg::MeshOffsets& mo,
                 const double* diag, const double* upper, const double* lower, const double* b, double* x,
                 int32_t solver, int32_t precision, int32_t max_iter, double tol, double* stats) {
    const int n = pm.nC, nif = pm.nIF;
    const size_t mbytes = mo.total + kMeshStructSlot;
    hipSetDevice(ctx->device);
    std::vector<char> stage(mbytes, 0);
    rsg::mesh_stage(pm, mo, stage.data());
    RsgMeshDev mview = rsg::mesh_view(pm, mo, base);
    memcpy(stage.data() + mo.total, &mview, sizeof mview);
    char* mstruct = base + mo.total;
    RsgMemberDev* dM = reinterpret_cast<RsgMemberDev*>(base + ((mbytes + 255) & ~size_t(255)));
    char* wbase = reinterpret_cast<char*>(dM) + ((sizeof(RsgMemberDev) + 255) & ~size_t(255));
    RsgMemberDev M;
    memset(&M, 0, sizeof M);
    rsg::Carver cv; cv.base = wbase;
    rsg::carve_member(cv, pm.nC, pm.nIF, pm.nB, pm.nnz0, 0, M);
    M.mesh = reinterpret_cast<RsgMeshDev*>(mstruct);
    M.prec = precision;
    hipError_t e = hipMemcpy(base, stage.data(), mbytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dM, &M, sizeof M, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(M.diag, diag, 8 * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(M.src, b, 8 * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(M.pp, x, 8 * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess && nif > 0) e = hipMemcpy(M.upper, upper, 8 * (size_t)nif, hipMemcpyHostToDevice);
    if (e == hipSuccess && nif > 0) e = hipMemcpy(M.lower, lower, 8 * (size_t)nif, hipMemcpyHostToDevice);
    RsgScalars s; member_init_scalars(s, 0);
    if (e == hipSuccess) e = hipMemcpy(M.sc, &s, sizeof s, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        // the stats land in grad-p scratch, which no linear solver touches
        e = rsg_ldu_launch(dM, solver, max_iter, tol, M.gp[0], ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    double st[5] = {0, 0, 0, 0, 0};
    if (e == hipSuccess) e = hipMemcpy(st, M.gp[0], sizeof st, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(x, M.pp, 8 * (size_t)n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_err(ctx, RSG_ERR_DEVICE, hipGetErrorString(e));
    for (int i = 0; i < 4; ++i) stats[i] = st[i];
    return RSG_OK;
}

}  // namespace


extern "C" {

int rsg_ldu_solve(rsg_context* ctx, int32_t n, int32_t nif, const int32_t* own, const int32_t* nbr,
                  const double* diag, const double* upper, const double* lower, const double* b,
                  double* x, int32_t solver, int32_t precision, int32_t max_iter, double tol,
                  double* stats) {
    rsg::PrepMesh pm;
    int rc = ldu_prepare(ctx, n, nif, own, nbr, diag, upper, lower, b, x, solver, precision, stats, pm);
    if (rc != RSG_OK) return rc;
    rsg::MeshOffsets mo = rsg::mesh_offsets(pm);
    const size_t total = ldu_block_bytes(pm, mo);
    std::lock_guard<std::mutex> lk(ctx->mu);
    size_t off;
    if (!ctx->arena.alloc(total, off)) return set_err(ctx, RSG_ERR_NOMEM, "system exceeds the HBM budget");
    rc = ldu_solve_at(ctx, ctx->arena.base + off, pm, mo, diag, upper, lower, b, x, solver, precision, max_iter, tol,
                      stats);
    ctx->arena.release(off, total);
    return rc;
}

int rsg_ldu_solve_noalloc(rsg_context* ctx, int32_t n, int32_t nif, const int32_t* own, const int32_t* nbr,
                          const double* diag, const double* upper, const double* lower, const double* b,
                          double* x, int32_t solver, int32_t precision, int32_t max_iter, double tol,
                          double* stats) {
    rsg::PrepMesh pm;
    int rc = ldu_prepare(ctx, n, nif, own, nbr, diag, upper, lower, b, x, solver, precision, stats, pm);
    if (rc != RSG_OK) return rc;
    if (!ctx->scratchDev) return set_err(ctx, RSG_ERR_DEVICE, "the scratch array's device address is unknown");
    rsg::MeshOffsets mo = rsg::mesh_offsets(pm);
    const size_t total = ldu_block_bytes(pm, mo);
    if (total > RSG_NOALLOC_SCRATCH_BYTES)
        return set_err(ctx, RSG_ERR_ARG, "rsg_ldu_solve_noalloc: the system needs " + std::to_string(total) +
                                             " bytes of device workspace, more than the " +
                                             std::to_string(RSG_NOALLOC_SCRATCH_BYTES) +
                                             "-byte scratch array (rsg_ldu_solve has no such limit)");
    std::lock_guard<std::mutex> dk(g_scratch_mu[ctx->device & 63]);
    std::lock_guard<std::mutex> lk(ctx->mu);
    return ldu_solve_at(ctx, ctx->scratchDev, pm, mo, diag, upper, lower, b, x, solver, precision, max_iter, tol,
                        stats);
}

int rsg_system_solve(rsg_context* ctx, int32_t n, int32_t nif, const int32_t* own, const int32_t* nbr,
                     const double* diag, const double* upper, const double* lower, const double* b, double* x,
                     int32_t solver, int32_t precision, int32_t max_iter, double tol, double* stats) {
    using namespace rsg::grid;
    if (!ctx || n <= 0 || nif < 0 || !own || (nif > 0 && !nbr) || !diag || (nif > 0 && (!upper || !lower)) ||
        !b || !x || !stats || max_iter < 0)
        return RSG_ERR_ARG;
    if (solver != RSG_SOLVER_AMG && solver != RSG_SOLVER_PCG)
        return set_err(ctx, RSG_ERR_ARG, "solver must be RSG_SOLVER_AMG or RSG_SOLVER_PCG");
    if (precision != RSG_PREC_FP64 && precision != RSG_PREC_MIXED)
        return set_err(ctx, RSG_ERR_ARG, "unknown precision code");
    for (int f = 0; f < nif; ++f)
        if (own[f] < 0 || own[f] >= n || nbr[f] < 0 || nbr[f] >= n) return set_err(ctx, RSG_ERR_ARG, "face index out of range");
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<Level> H;
    build_hierarchy(n, nif, own, nbr, diag, upper, lower, solver == RSG_SOLVER_AMG ? RSG_MAX_LEVELS : 0, H);
    std::vector<double> G;
    const bool dense = solver == RSG_SOLVER_AMG && H.back().n <= 64 && dense_inverse(H.back(), G);
    // a stalled coarsest level (above 64 rows): MFMA inverse on the device
    const bool bigDense = solver == RSG_SOLVER_AMG && !dense && H.size() > 1 && H.back().n <= RSG_DENSE_MAX;
    int denseRows = 0;
    double denseMs = 0.0;
    const auto t1 = std::chrono::steady_clock::now();
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, RSG_ERR_DEVICE, "hipSetDevice");
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return set_err(ctx, RSG_ERR_DEVICE, "stream");
    hipError_t e = hipSuccess;
    PcgResult res;
    int refinements = 0;
    {
        DevAlloc A;
        Scratch S;
        S.part = A.take<double>(2 * (size_t)blocks(n) + 2);
        S.dots = A.take<double>(8);
        if (A.err == hipSuccess) A.err = hipHostMalloc((void**)&S.hdots, 8 * sizeof(double), hipHostMallocDefault);
        Hierarchy<double> H64;
        VcycleGraphs VG64, VG32;
        VG64.enabled = VG32.enabled = ctx->useGraphs != 0;
        if (precision == RSG_PREC_FP64) {
            upload<double>(H, G, dense, A, H64, s);
        } else {
            std::vector<Level> top(1, H[0]);   // FP64 level 0 for the defect
            std::vector<double> none;
            upload<double>(top, none, false, A, H64, s);
        }
        const double* dinv = nullptr;
        int dinvN = 0;
        if (bigDense && A.err == hipSuccess) {
            double* inv = nullptr;
            bool ok = false;
            A.err = dense_inverse_device(H.back(), A, s, &inv, &dinvN, &ok, &denseMs);
            if (ok) { dinv = inv; denseRows = H.back().n; }
        }
        if (precision == RSG_PREC_FP64) { H64.dinv = dinv; H64.dinvN = dinvN; }
        double* b64 = A.take<double>(n);
        double* x64 = A.take<double>(n);
        if (A.err == hipSuccess) A.err = hipMemcpyAsync(b64, b, 8 * (size_t)n, hipMemcpyHostToDevice, s);
        if (A.err == hipSuccess) A.err = hipMemcpyAsync(x64, x, 8 * (size_t)n, hipMemcpyHostToDevice, s);
        if (precision == RSG_PREC_FP64) {
            double* r = A.take<double>(n); double* z = A.take<double>(n);
            double* p = A.take<double>(n); double* q = A.take<double>(n);
            if (A.err == hipSuccess)
                A.err = pcg<double>(H64, b64, x64, r, z, p, q, max_iter, tol, solver == RSG_SOLVER_AMG, S, s, res,
                                    &VG64);
        } else {
            // FP64 defect / refinement around FP32 inner PCG solves; each inner
            // solve reduces its defect by what is still missing (with a 0.3
            // margin), but never asks FP32 for more than 1e-5
            Hierarchy<float> H32;
            upload<float>(H, G, dense, A, H32, s);
            H32.dinv = dinv; H32.dinvN = dinvN;
            double* r64 = A.take<double>(n);
            float* b32 = A.take<float>(n); float* e32 = A.take<float>(n);
            float* r = A.take<float>(n); float* z = A.take<float>(n);
            float* p = A.take<float>(n); float* q = A.take<float>(n);
            const DevLevel<double>& A0 = H64.lv[0];
            const dim3 gr(blocks(n)), bl(kBlock);
            double r0 = -1.0;
            int inner = 0;
            for (int k = 0; k < 30 && A.err == hipSuccess; ++k) {
                hipLaunchKernelGGL(k_residual<double>, gr, bl, 0, s, n, A0.rp, A0.ci, A0.v, (const double*)nullptr, (const double*)b64,
                                   (const double*)x64, r64);
                dot2<double>(n, r64, r64, (const double*)nullptr, (const double*)nullptr, S, S.dots + 2, s);
                A.err = hipMemcpyAsync(S.hdots, S.dots, 4 * sizeof(double), hipMemcpyDeviceToHost, s);
                if (A.err == hipSuccess) A.err = hipStreamSynchronize(s);
                if (A.err != hipSuccess) break;
                const double rn = std::sqrt(S.hdots[2]);
                if (r0 < 0) { r0 = rn; res.initialRes = rn; }
                res.finalRes = r0 > 0 ? rn / r0 : 0.0;
                if (r0 < 1e-30 || res.finalRes < tol) { res.converged = 1; break; }
                if (inner >= max_iter) break;
                hipLaunchKernelGGL((k_convert<float, double>), gr, bl, 0, s, n, (const double*)r64, b32);
                hipLaunchKernelGGL(k_fill<float>, gr, bl, 0, s, n, e32, 0.0f);
                PcgResult ir;
                const double innerTol = std::max(1e-5, 0.3 * tol / res.finalRes);
                A.err = pcg<float>(H32, b32, e32, r, z, p, q, max_iter - inner, innerTol, solver == RSG_SOLVER_AMG,
                                   S, s, ir, &VG32);
                inner += ir.iterations;
                hipLaunchKernelGGL(k_add<float>, gr, bl, 0, s, n, (const float*)e32, x64);
                ++refinements;
            }
            res.iterations = inner;
        }
        if (A.err == hipSuccess) A.err = hipMemcpyAsync(x, x64, 8 * (size_t)n, hipMemcpyDeviceToHost, s);
        if (A.err == hipSuccess) A.err = hipStreamSynchronize(s);
        e = A.err;
        if (S.hdots) (void)hipHostFree(S.hdots);
    }
    (void)hipStreamDestroy(s);
    const auto t2 = std::chrono::steady_clock::now();
    if (e != hipSuccess) return set_err(ctx, RSG_ERR_DEVICE, hipGetErrorString(e));
    stats[0] = res.iterations; stats[1] = res.finalRes; stats[2] = res.initialRes; stats[3] = res.converged;
    stats[4] = (double)H.size(); stats[5] = refinements;
    stats[6] = std::chrono::duration<double>(t1 - t0).count();
    stats[7] = std::chrono::duration<double>(t2 - t1).count();
    stats[8] = denseRows; stats[9] = 1e-3 * denseMs;
    return RSG_OK;
}

int rsg_dense_inverse(rsg_context* ctx, int32_t n, const double* a, double* inv, double* kernel_seconds) {
    using namespace rsg::grid;
    if (!ctx || n <= 0 || n > RSG_DENSE_MAX || !a || !inv) return RSG_ERR_ARG;
    // the matrix as a CSR level (every entry, zeros included, keeps the pattern dense)
    Level L;
    L.n = n; L.nnz = n * n;
    L.rp.resize(n + 1); L.ci.resize((size_t)n * n); L.v.assign(a, a + (size_t)n * n);
    for (int i = 0; i <= n; ++i) L.rp[i] = i * n;
    for (size_t q = 0; q < (size_t)n * n; ++q) L.ci[q] = (int)(q % n);
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, RSG_ERR_DEVICE, "hipSetDevice");
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return set_err(ctx, RSG_ERR_DEVICE, "stream");
    hipError_t e = hipSuccess;
    bool ok = false;
    double ms = 0.0;
    {
        DevAlloc A;
        double* M = nullptr;
        int N = 0;
        e = dense_inverse_device(L, A, s, &M, &N, &ok, &ms);
        if (e == hipSuccess && ok)
            e = hipMemcpy2DAsync(inv, 8 * (size_t)n, M, 8 * (size_t)N, 8 * (size_t)n, n, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) return set_err(ctx, RSG_ERR_DEVICE, hipGetErrorString(e));
    if (!ok) return set_err(ctx, RSG_ERR_ARG, "matrix is not positive definite (a pivot was not positive)");
    if (kernel_seconds) *kernel_seconds = 1e-3 * ms;
    return RSG_OK;
}

int rsg_system_hierarchy(int32_t n, int32_t nif, const int32_t* own, const int32_t* nbr, const double* diag,
                         const double* upper, const double* lower, int32_t* level_sizes, int32_t* level_nnz,
                         int32_t max_levels, int32_t* maps, int64_t maps_cap) {
    using namespace rsg::grid;
    if (n <= 0 || nif < 0 || !own || (nif > 0 && !nbr) || !diag || (nif > 0 && (!upper || !lower)) || max_levels < 0)
        return RSG_ERR_ARG;
    for (int f = 0; f < nif; ++f)
        if (own[f] < 0 || own[f] >= n || nbr[f] < 0 || nbr[f] >= n) return RSG_ERR_ARG;
    std::vector<Level> H;
    build_hierarchy(n, nif, own, nbr, diag, upper, lower, RSG_MAX_LEVELS, H);
    int64_t off = 0;
    for (size_t l = 0; l < H.size(); ++l) {
        if ((int)l < max_levels) {
            if (level_sizes) level_sizes[l] = H[l].n;
            if (level_nnz) level_nnz[l] = H[l].nnz;
        }
        if (maps)
            for (size_t i = 0; i < H[l].map.size() && off < maps_cap; ++i) maps[off++] = H[l].map[i];
    }
    return (int)H.size();
}

int rsg_amg_aggregate(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* values,
                      const double* diag, int32_t* map) {
    if (n < 0 || (n > 0 && (!row_ptr || !col_idx || !values || !diag || !map))) return RSG_ERR_ARG;
    for (int i = 0; i < n; ++i)
        for (int j = row_ptr[i]; j < row_ptr[i + 1]; ++j)
            if (col_idx[j] < 0 || col_idx[j] >= n) return RSG_ERR_ARG;
    return rsg::amg_aggregate(n, row_ptr, col_idx, values, diag, map);
}

int rsg_wall_distance(rsg_context* ctx, int32_t n_cells, const double* cell_centers, int32_t n_wall,
                      const double* wall_points, double* wall_dist, double* kernel_seconds) {
    using namespace rsg::geom;
    if (!ctx || n_cells < 0 || n_wall < 0 || (n_cells > 0 && (!cell_centers || !wall_dist)) ||
        (n_wall > 0 && !wall_points))
        return RSG_ERR_ARG;
    if (kernel_seconds) *kernel_seconds = 0.0;
    if (n_cells == 0) return RSG_OK;
    if (n_wall == 0) {                                   // Mesh.cpp:358-362
        std::fill(wall_dist, wall_dist + n_cells, 1e10);
        return RSG_OK;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, RSG_ERR_DEVICE, "hipSetDevice");
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return set_err(ctx, RSG_ERR_DEVICE, "stream");
    hipError_t e = hipSuccess;
    float ms = 0.0f;
    {
        rsg::grid::DevAlloc A;
        double* cc = A.take<double>(3 * (size_t)n_cells);
        double* wp = A.take<double>(3 * (size_t)n_wall);
        double* out = A.take<double>((size_t)n_cells);
        hipEvent_t ev[2] = {nullptr, nullptr};
        if (A.err == hipSuccess) A.err = hipEventCreate(&ev[0]);
        if (A.err == hipSuccess) A.err = hipEventCreate(&ev[1]);
        if (A.err == hipSuccess) A.err = hipMemcpyAsync(cc, cell_centers, 24 * (size_t)n_cells, hipMemcpyHostToDevice, s);
        if (A.err == hipSuccess) A.err = hipMemcpyAsync(wp, wall_points, 24 * (size_t)n_wall, hipMemcpyHostToDevice, s);
        if (A.err == hipSuccess) A.err = hipEventRecord(ev[0], s);
        if (A.err == hipSuccess) {
            hipLaunchKernelGGL(k_wall_distance, dim3(wall_blocks(n_cells)), dim3(kWallBlock), 0, s, n_cells, cc,
                               n_wall, wp, out);
            A.err = hipGetLastError();
        }
        if (A.err == hipSuccess) A.err = hipEventRecord(ev[1], s);
        if (A.err == hipSuccess) A.err = hipMemcpyAsync(wall_dist, out, 8 * (size_t)n_cells, hipMemcpyDeviceToHost, s);
        if (A.err == hipSuccess) A.err = hipStreamSynchronize(s);
        if (A.err == hipSuccess) A.err = hipEventElapsedTime(&ms, ev[0], ev[1]);
        for (hipEvent_t v : ev)
            if (v) (void)hipEventDestroy(v);
        e = A.err;
    }
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) return set_err(ctx, RSG_ERR_DEVICE, hipGetErrorString(e));
    if (kernel_seconds) *kernel_seconds = 1e-3 * ms;
    return RSG_OK;
}

// Plain device copy of `bytes` bytes (16-byte vector loads and stores,
// grid-stride), used to scale hardware byte counters: a run of it under
// FETCH_SIZE / WRITE_SIZE reads what the counters report for a known traffic.
__global__ void k_copy_calibration(const double2* __restrict__ a, double2* __restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

int rsg_copy_calibration(rsg_context* ctx, uint64_t bytes, double* kernel_seconds) {
    if (!ctx || bytes < 16 || (bytes & 15)) return RSG_ERR_ARG;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, RSG_ERR_DEVICE, "hipSetDevice");
    const size_t n = bytes / 16;
    float ms = 0.0f;
    hipError_t e;
    {
        rsg::grid::DevAlloc A;
        double2* a = A.take<double2>(n);
        double2* b = A.take<double2>(n);
        hipEvent_t ev[2] = {nullptr, nullptr};
        if (A.err == hipSuccess) A.err = hipMemsetAsync(a, 0, bytes, ctx->stream);
        if (A.err == hipSuccess) A.err = hipEventCreate(&ev[0]);
        if (A.err == hipSuccess) A.err = hipEventCreate(&ev[1]);
        if (A.err == hipSuccess) A.err = hipEventRecord(ev[0], ctx->stream);
        if (A.err == hipSuccess) {
            const int blocks = (int)std::min<size_t>((n + 255) / 256, 8192);
            hipLaunchKernelGGL(k_copy_calibration, dim3(blocks), dim3(256), 0, ctx->stream, a, b, n);
            A.err = hipGetLastError();
        }
        if (A.err == hipSuccess) A.err = hipEventRecord(ev[1], ctx->stream);
        if (A.err == hipSuccess) A.err = hipStreamSynchronize(ctx->stream);
        if (A.err == hipSuccess) A.err = hipEventElapsedTime(&ms, ev[0], ev[1]);
        for (hipEvent_t v : ev)
            if (v) (void)hipEventDestroy(v);
        e = A.err;
    }
    if (e != hipSuccess) return set_err(ctx, RSG_ERR_DEVICE, hipGetErrorString(e));
    if (kernel_seconds) *kernel_seconds = 1e-3 * ms;
    return RSG_OK;
}

#ifdef RSG_PROF
// Section timers of the instrumented build: seconds per slot, then reset.
int rsg_prof_read(double* out) {
    unsigned long long t[32];
    if (hipMemcpyFromSymbol(t, HIP_SYMBOL(rsg_prof_ticks), sizeof t) != hipSuccess) return RSG_ERR_DEVICE;
    for (int i = 0; i < 32; ++i) out[i] = 1e-8 * (double)t[i];
    const unsigned long long z[32] = {};
    if (hipMemcpyToSymbol(HIP_SYMBOL(rsg_prof_ticks), z, sizeof z) != hipSuccess) return RSG_ERR_DEVICE;
    return RSG_OK;
}

// Per-workgroup start / end ticks and CU id of the last SIMPLE launch.
int rsg_prof_wg_read(uint64_t* out, int n) {
    if (n > 4096) n = 4096;
    if (hipMemcpyFromSymbol(out, HIP_SYMBOL(rsg_prof_wg), 24 * (size_t)n) != hipSuccess) return RSG_ERR_DEVICE;
    return RSG_OK;
}
#endif

}  // extern "C"
