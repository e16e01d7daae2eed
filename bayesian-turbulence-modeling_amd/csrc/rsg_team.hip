// rsg_team.hip -- one solve on many workgroups: the member SIMPLE code
// (rsg_member.h) built with RSG_TEAM and run by every workgroup of one
// cooperative launch, for meshes too large for one workgroup.
//
// In this build the member code's cell / face loops stride over the whole
// team (thread index and size of the launch), a phase boundary is a grid-wide
// barrier, block reductions become team reductions (workgroup values combined
// in workgroup order, identical in every thread), the coarsest dense inverse
// is factored by the first workgroup, and the linear solves take the
// global-memory paths (no LDS images, no register-resident vectors).  The
// member's descriptors and AMG tables are read from global memory.
#define RSG_TEAM 1
#include <hip/hip_runtime.h>

#include "rsg.h"
#include "rsg_layout.h"
#include "rsg_prep.h"
#include "rsg_member.h"

namespace {

constexpr int kTeamBlock = 512;

__global__ __launch_bounds__(kTeamBlock) void k_simple_team(RsgMemberDev* M, RsgScalars* out, double* gred) {
    __shared__ double red[128];
    __shared__ double lds[RSG_COARSE_DENSE * RSG_COARSE_DENSE];
    rsg::Thr t;
    t.tid = blockIdx.x * kTeamBlock + threadIdx.x;
    t.bs = gridDim.x * kTeamBlock;
    t.red = red; t.lds = lds; t.phase = 0;
    t.xb0 = nullptr; t.xb1 = nullptr; t.xsel = 0;
    t.cvec = nullptr; t.cvecCap = 0;
    t.ceVal = nullptr; t.ceCol = nullptr; t.ceCap = 0; t.caCol = nullptr; t.caCap = 0;
    t.ltid = threadIdx.x; t.lbs = kTeamBlock; t.blk = blockIdx.x; t.nblk = gridDim.x; t.gred = gred;
    rsg::run_member<3>(t, M);
    RSG_SYNC();
    if (t.tid == 0) *out = *M->sc;
}

}  // namespace

// Workgroups one team may hold: every one must be resident at once.
int rsg_team_capacity(int device) {
    int per = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_simple_team, kTeamBlock, 0) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    const int n = per * cus;
    return n < RSG_TEAM_MAXB ? n : RSG_TEAM_MAXB;
}

hipError_t rsg_team_launch(RsgMemberDev* dM, RsgScalars* dOut, double* gred, int blocks, hipStream_t s) {
    void* args[] = {&dM, &dOut, &gred};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_simple_team), dim3(blocks), dim3(kTeamBlock),
                                      args, 0, s);
}
