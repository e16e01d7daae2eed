// rsg_chunk.hip -- one instantiation of the member kernels, for the workgroup
// size RSG_CHUNK_BS and precision RSG_CHUNK_PREC given on the command line
// (the build compiles this file once per pair, in parallel).
#include "rsg_chunk.h"

#ifndef RSG_CHUNK_BS
#error "RSG_CHUNK_BS (256 or 512) must be defined"
#endif
#ifndef RSG_CHUNK_PREC
#error "RSG_CHUNK_PREC (0 FP64, 1 mixed) must be defined"
#endif
#define RSG_CAT3_(a, b, c) a##b##_##c
#define RSG_CAT3(a, b, c) RSG_CAT3_(a, b, c)

hipError_t RSG_CAT3(rsg_chunk_launch_, RSG_CHUNK_BS, RSG_CHUNK_PREC)(int grid, RsgMemberDev* members,
                                                                     RsgScalars* statusOut, hipStream_t s) {
    hipLaunchKernelGGL((k_simple_chunk<RSG_CHUNK_BS, RSG_CHUNK_PREC>), dim3(grid), dim3(RSG_CHUNK_BS), 0, s, members,
                       statusOut);
    return hipGetLastError();
}

// Member workgroups of this kernel one CU holds at once.
int RSG_CAT3(rsg_chunk_occupancy_, RSG_CHUNK_BS, RSG_CHUNK_PREC)() {
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_simple_chunk<RSG_CHUNK_BS, RSG_CHUNK_PREC>,
                                                     RSG_CHUNK_BS, 0) != hipSuccess)
        return 0;
    return per;
}

#if RSG_CHUNK_BS == 256 && RSG_CHUNK_PREC == 0
// The stand-alone LDU solve (rsg_ldu_solve) runs one 256-thread workgroup;
// its precision is read from the member at run time.
hipError_t rsg_ldu_launch(RsgMemberDev* M, int solver, int maxIter, double tol, double* stats, hipStream_t s) {
    hipLaunchKernelGGL(k_ldu_solve<256>, dim3(1), dim3(256), 0, s, M, solver, maxIter, tol, stats);
    return hipGetLastError();
}
// A kernel-per-phase member's linear solve on one 256-thread workgroup.
hipError_t rsg_member_solve_launch(RsgMemberDev* M, int solver, double* x, int amgMode, int maxIter, double tol,
                                   double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_member_solve<256>, dim3(1), dim3(256), 0, s, M, solver, x, amgMode, maxIter, tol, out);
    return hipGetLastError();
}
#endif

#if RSG_CHUNK_BS == 512 && RSG_CHUNK_PREC == 0
// Compressible members: 512-thread workgroups, FP64.
hipError_t rsg_comp_chunk_launch(int grid, RsgMemberDev* members, RsgScalars* statusOut, hipStream_t s) {
    hipLaunchKernelGGL(k_comp_chunk<512>, dim3(grid), dim3(512), 0, s, members, statusOut);
    return hipGetLastError();
}
int rsg_comp_chunk_occupancy() {
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_comp_chunk<512>, 512, 0) != hipSuccess) return 0;
    return per;
}
#endif
