// rsg_comp.h -- compressible pressure-based SIMPLE (ideal gas, energy
// equation, density-weighted fluxes) on the kernel-per-phase path.
//
// Included by rsg_big.hip inside its anonymous namespace: a compressible
// member on the kernel-per-phase path is an RsgBig whose outer iteration is
// comp_outer_iteration below.  Every phase of rsg_comp_dev.h (the reference
// citations are there) is a grid-wide launch over cells, and every linear
// solve is rsg_big.hip's solve() on the CSR image of the assembled LDU system
// (BiCGSTAB / PCG / AMG-PCG / Gauss-Seidel as the settings name them; the
// energy equation always BiCGSTAB).  On a small mesh (a workspace carved with
// the member path's AMG arrays: memberSolve) each solve is instead one
// workgroup-wide launch of the member path's solvers on the LDU system, and
// the host waits once per outer iteration instead of once per dot product.

// slots of the one-workgroup solves' results (memberSolve members)
enum { SLOT_UX = 0, SLOT_UY, SLOT_UZ, SLOT_P, SLOT_T, SLOT_K, SLOT_OM };

// reduction slots of the per-iteration maxima (block partials in B.compPart)
enum {
    CQ_UVAL = 0, CQ_UDIFF, CQ_TVAL, CQ_TDIFF, CQ_KVAL, CQ_OMVAL, CQ_KDIFF, CQ_OMDIFF, CQ_PMAX, CQ_BAD, CQ_COUNT
};

// Block maxima of NQ per-thread values into part[CQ_COUNT * block + q0 + q].
// Every thread of the block calls it.
template <int NQ>
__device__ inline void comp_block_max(const double (&v)[NQ], double* part, int q0) {
    __shared__ double sh[NQ][kBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double x = v[q];
        for (int off = 32; off >= 1; off >>= 1) x = dmax(x, __shfl_xor(x, off, 64));
        if (lane == 0) sh[q][w] = x;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        double x = sh[threadIdx.x][0];
        for (int q = 1; q < kBlock / 64; ++q) x = dmax(x, sh[threadIdx.x][q]);
        part[CQ_COUNT * blockIdx.x + q0 + threadIdx.x] = x;
    }
}

// One grid-wide launch per phase of rsg_comp_dev.h (one thread per cell).
__global__ __launch_bounds__(kBlock) void kc_init(RsgMemberDev* M, RsgCompDev* Q) {
    Ctx C = gctx(M);
    comp_init_fields(C, Q);
}
__global__ __launch_bounds__(kBlock) void kc_state(RsgMemberDev* M, RsgCompDev* Q, int start, int tbc, int mech) {
    Ctx C = gctx(M);
    comp_state(C, Q, start, tbc, mech);
}
__global__ __launch_bounds__(kBlock) void kc_bnd(RsgMemberDev* M, RsgCompDev* Q, int tbc, int mech) {
    Ctx C = gctx(M);
    comp_bnd(C, Q, tbc, mech);
}
// velocity and pressure BCs after the corrections, outlet mechanicalized
__global__ __launch_bounds__(kBlock) void kc_bnd_vp(RsgMemberDev* M, RsgCompDev* Q) {
    Ctx C = gctx(M);
    vel_bc(C);
    p_bc(C);
    comp_outlet_mech(C, Q);   // the same face-to-thread mapping as p_bc
}
__global__ __launch_bounds__(kBlock) void kc_step1(RsgMemberDev* M, RsgCompDev* Q, int turbUpdate, int established) {
    Ctx C = gctx(M);
    comp_closure(C, Q, turbUpdate, established);
}
__global__ __launch_bounds__(kBlock) void kc_divu(RsgMemberDev* M, RsgCompDev* Q) {
    Ctx C = gctx(M);
    comp_divu(C, Q);
}
__global__ __launch_bounds__(kBlock) void kc_gtr(RsgMemberDev* M, RsgCompDev* Q) {
    Ctx C = gctx(M);
    comp_gtr(C, Q);
}
__global__ __launch_bounds__(kBlock) void kc_asm_mom(RsgMemberDev* M, RsgCompDev* Q, int comp) {
    Ctx C = gctx(M);
    comp_asm_mom(C, Q, comp);
}
__global__ __launch_bounds__(kBlock) void kc_asm_pc(RsgMemberDev* M, RsgCompDev* Q) {
    Ctx C = gctx(M);
    comp_asm_pc(C, Q);
}
__global__ __launch_bounds__(kBlock) void kc_asm_e(RsgMemberDev* M, RsgCompDev* Q, double* part) {
    Ctx C = gctx(M);
    double mx[2] = {1e-30, 0.0};
    comp_asm_e(C, Q, mx);
    comp_block_max<2>(mx, part, CQ_UVAL);
}
__global__ __launch_bounds__(kBlock) void kc_tpost(RsgMemberDev* M, RsgCompDev* Q, double* part) {
    Ctx C = gctx(M);
    double mx[2] = {1e-30, 0.0};
    comp_tpost(C, Q, mx);
    comp_block_max<2>(mx, part, CQ_TVAL);
}
__global__ __launch_bounds__(kBlock) void kc_asm_turb(RsgMemberDev* M, RsgCompDev* Q, int isOmega) {
    Ctx C = gctx(M);
    comp_asm_turb(C, Q, isOmega);
}
__global__ __launch_bounds__(kBlock) void kc_ombc(RsgMemberDev* M, RsgCompDev* Q, double beta1) {
    Ctx C = gctx(M);
    comp_ombc(C, Q, beta1);
}
__global__ __launch_bounds__(kBlock) void kc_pin(RsgMemberDev* M, RsgCompDev* Q, double* part) {
    Ctx C = gctx(M);
    double mx[4] = {1e-30, 1e-30, 0.0, 0.0};
    comp_pin(C, Q, mx);
    comp_block_max<4>(mx, part, CQ_KVAL);
}
__global__ __launch_bounds__(kBlock) void kc_valid(RsgMemberDev* M, RsgCompDev* Q, double* part) {
    Ctx C = gctx(M);
    double mx[2] = {0.0, 0.0};
    comp_valid(C, Q, mx);
    comp_block_max<2>(mx, part, CQ_PMAX);
}

// out[q] = max over the nb block partials of slot q (one workgroup)
__global__ __launch_bounds__(kBlock) void kc_reduce(int nb, const double* part, double* out) {
    __shared__ double sh[CQ_COUNT][kBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int q = 0; q < CQ_COUNT; ++q) {
        double v = part[q];
        for (int b = threadIdx.x; b < nb; b += kBlock) v = dmax(v, part[CQ_COUNT * b + q]);
        for (int off = 32; off >= 1; off >>= 1) v = dmax(v, __shfl_xor(v, off, 64));
        if (lane == 0) sh[q][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < CQ_COUNT) {
        double v = sh[threadIdx.x][0];
        for (int q = 1; q < kBlock / 64; ++q) v = dmax(v, sh[threadIdx.x][q]);
        out[threadIdx.x] = v;
    }
}

// A finished solve's export (comp_export) and the member's effective nu for
// the observables, written by thread 0 after every thread's export.
__global__ __launch_bounds__(kBlock) void kc_export(RsgMemberDev* M, RsgCompDev* Q) {
    Ctx C = gctx(M);
    comp_export(C, Q);
}
__global__ void kc_export_nu(RsgMemberDev* M, RsgCompDev* Q) {
    if (blockIdx.x == 0 && threadIdx.x == 0) M->nu = comp_visc(Q, Q->T[0]) / dmax(Q->rho[0], 1e-30);
}

// ---------------------------------------------------------------- host
// Reads the CQ_COUNT maxima of the partials the iteration's kernels wrote
// (and, for a memberSolve member, the results of its solves).
hipError_t comp_maxima(RsgBig& B, int nb, double* out) {
    hipLaunchKernelGGL(kc_reduce, dim3(1), dim3(kBlock), 0, B.s, nb, (const double*)B.compPart, B.S.dots);
    hipError_t e = hipMemcpyAsync(B.S.hdots, B.S.dots, CQ_COUNT * sizeof(double), hipMemcpyDeviceToHost, B.s);
    if (e == hipSuccess && B.memberSolve)
        e = hipMemcpyAsync(B.hsolve, B.solveOut, 4 * kSolveSlots * sizeof(double), hipMemcpyDeviceToHost, B.s);
    if (e == hipSuccess) e = hipStreamSynchronize(B.s);
    for (int q = 0; q < CQ_COUNT; ++q) out[q] = B.S.hdots[q];
    return e;
}

// The iteration's linear solve into x: on the grid (synchronous, res filled)
// or, for a memberSolve member, one workgroup-wide launch whose result lands
// in `slot` (res filled after comp_maxima).
hipError_t comp_solve(RsgBig& B, int code, double* x, bool amgReady, int amgMode, int slot, SolveRes& res) {
    if (!B.memberSolve) return solve(B, code, x, amgReady, res, false);
    const RsgSettingsDev& st = B.set;
    return rsg_member_solve_launch(B.dM, code, x, amgMode, st.innerIter, st.innerTol, B.solveOut + 4 * slot, B.s);
}
inline void slot_result(const RsgBig& B, int slot, SolveRes& r) {
    const double* o = B.hsolve + 4 * slot;
    r.iterations = (int)o[0]; r.finalRes = o[1]; r.initialRes = o[2]; r.converged = (int)o[3];
}

inline bool finite_res(const SolveRes& r) { return std::isfinite(r.initialRes) && std::isfinite(r.finalRes); }

// One outer iteration of CompressibleSIMPLESolver::solve (:686-998).
hipError_t comp_outer_iteration(RsgBig& B) {
    RsgScalars& sc = B.sc;
    const RsgSettingsDev& st = B.set;
    RsgMemberDev& M = B.hM;
    RsgCompDev& Q = B.hQ;
    RsgMemberDev* dM = B.dM;
    RsgCompDev* dQ = B.dQ;
    const int iter = sc.iter;
    const int n = B.nC;
    const dim3 gc(cells_grid(n)), bl(kBlock);
    const int nb = (int)gc.x;
    const bool turbActive = iter >= st.turbStart;
    const bool turbUpdate = turbActive && ((iter - st.turbStart) % st.turbInterval == 0);
    hipError_t e = hipSuccess;
    // 1. thermodynamic state: rho, mu, the starting velocity, outlet pressure
    hipLaunchKernelGGL(kc_state, gc, bl, 0, B.s, dM, dQ, 1, 0, 1);
    // 2. SST closure with the local viscosity, grad p
    if (turbUpdate && !sc.turbEstablished && iter >= st.turbStart + st.nutFloorIters) sc.turbEstablished = 1;
    hipLaunchKernelGGL(kc_step1, gc, bl, 0, B.s, dM, dQ, turbUpdate ? 1 : 0, sc.turbEstablished);
    // 3. momentum
    SolveRes rU[3] = {{0, 0.0, 0.0, 0}, {0, 0.0, 0.0, 0}, {0, 0.0, 0.0, 0}};
    double* ucompv[3] = {M.Ux, M.Uy, M.Uz};
    for (int comp = 0; comp < 3 && e == hipSuccess; ++comp) {
        if (comp == 2 && M.planarZ) break;
        if (comp > 0) hipLaunchKernelGGL(k_grad_u, gc, bl, 0, B.s, dM, comp - 1, comp);
        hipLaunchKernelGGL(kc_divu, gc, bl, 0, B.s, dM, dQ);
        hipLaunchKernelGGL(kc_gtr, gc, bl, 0, B.s, dM, dQ);
        hipLaunchKernelGGL(kc_asm_mom, gc, bl, 0, B.s, dM, dQ, comp);
        e = comp_solve(B, st.mSolver, ucompv[comp], false, 0, SLOT_UX + comp, rU[comp]);
    }
    hipLaunchKernelGGL(k_bcs, gc, bl, 0, B.s, dM, 1, 0.0);
    // 4. pressure correction (density-weighted)
    hipLaunchKernelGGL(kc_asm_pc, gc, bl, 0, B.s, dM, dQ);
    bool amgReady = false;
    int amgMode = 0;
    if (e == hipSuccess && st.pSolver == 0 && B.memberSolve) {
        // the member path's device AMG (aggregated on one thread of the solve)
        const bool rebuild = sc.amgBuiltIter < 0 || iter - sc.amgBuiltIter >= st.amgRefresh;
        const bool coarse = rebuild || iter - sc.coarseIter >= st.coarseRefresh;
        amgMode = (sc.amgBuiltIter < 0 ? 4 : 0) | (rebuild ? 1 : 0) | (coarse ? 2 : 0);
        if (rebuild) sc.amgBuiltIter = iter;
        if (coarse) sc.coarseIter = iter;
        amgReady = true;
    } else if (e == hipSuccess && st.pSolver == 0) {
        const bool rebuild = sc.amgBuiltIter < 0 || iter - sc.amgBuiltIter >= st.amgRefresh || !B.haveH;
        const bool coarse = rebuild || iter - sc.coarseIter >= st.coarseRefresh;
        if (rebuild) {
            hipLaunchKernelGGL(k_csr_fill, dim3(cells_grid(std::max(n, B.nnz0))), bl, 0, B.s, dM, B.v0, B.dg0);
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
    if (e == hipSuccess) e = comp_solve(B, st.pSolver, M.pp, amgReady, amgMode, SLOT_P, rP);
    hipLaunchKernelGGL(k_ppb, gc, bl, 0, B.s, dM);
    hipLaunchKernelGGL(k_grad_pp, gc, bl, 0, B.s, dM);
    hipLaunchKernelGGL(k_correct, gc, bl, 0, B.s, dM);
    hipLaunchKernelGGL(kc_bnd_vp, gc, bl, 0, B.s, dM, dQ);
    // 5. energy (BiCGSTAB), T >= 1, temperature BC, density
    hipLaunchKernelGGL(kc_asm_e, gc, bl, 0, B.s, dM, dQ, B.compPart);
    SolveRes rT{0, 0.0, 0.0, 0};
    if (e == hipSuccess) e = comp_solve(B, 2, Q.T, false, 0, SLOT_T, rT);
    hipLaunchKernelGGL(kc_tpost, gc, bl, 0, B.s, dM, dQ, B.compPart);
    hipLaunchKernelGGL(kc_state, gc, bl, 0, B.s, dM, dQ, 0, 1, 1);
    // 6. turbulence
    SolveRes rK{0, 0.0, 0.0, 0}, rOm{0, 0.0, 0.0, 0};
    if (turbUpdate && e == hipSuccess) {
        hipLaunchKernelGGL(kc_asm_turb, gc, bl, 0, B.s, dM, dQ, 0);
        e = comp_solve(B, st.tSolver, M.k, false, 0, SLOT_K, rK);
        hipLaunchKernelGGL(k_clamp, gc, bl, 0, B.s, dM, 0);
        hipLaunchKernelGGL(k_bcs, gc, bl, 0, B.s, dM, 4, 0.0);
        hipLaunchKernelGGL(kc_asm_turb, gc, bl, 0, B.s, dM, dQ, 1);
        if (e == hipSuccess) e = comp_solve(B, st.tSolver, M.om, false, 0, SLOT_OM, rOm);
        hipLaunchKernelGGL(k_clamp, gc, bl, 0, B.s, dM, 1);
        hipLaunchKernelGGL(kc_ombc, gc, bl, 0, B.s, dM, dQ, M.coeffs[2]);
        hipLaunchKernelGGL(kc_pin, gc, bl, 0, B.s, dM, dQ, B.compPart);
    }
    // 7. state validity, max |p|, and every maximum of the iteration at once
    hipLaunchKernelGGL(kc_valid, gc, bl, 0, B.s, dM, dQ, B.compPart);
    double mx[CQ_COUNT];
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = comp_maxima(B, nb, mx);
    if (e != hipSuccess) return e;
    if (B.memberSolve) {
        for (int c = 0; c < 3; ++c)
            if (!(c == 2 && M.planarZ)) slot_result(B, SLOT_UX + c, rU[c]);
        slot_result(B, SLOT_P, rP);
        slot_result(B, SLOT_T, rT);
        if (turbUpdate) { slot_result(B, SLOT_K, rK); slot_result(B, SLOT_OM, rOm); }
    }
    const double uChange = mx[CQ_UDIFF] / mx[CQ_UVAL];
    const double tChange = mx[CQ_TDIFF] / mx[CQ_TVAL];
    if (turbUpdate) {
        sc.lastK = mx[CQ_KDIFF] / mx[CQ_KVAL];
        sc.lastOm = mx[CQ_OMDIFF] / mx[CQ_OMVAL];
    }
    // residual entry (Ux, Uy, p, T, k, omega) and the tests of :934-998
    const double eUx = rU[0].initialRes, eUy = rU[1].initialRes, eP = rP.initialRes;
    const double eK = turbActive ? sc.lastK : 0.0, eOm = turbActive ? sc.lastOm : 0.0;
    sc.innerIters[0] += rU[0].iterations; sc.innerIters[1] += rU[1].iterations; sc.innerIters[2] += rU[2].iterations;
    sc.innerIters[3] += rP.iterations; sc.innerIters[4] += rK.iterations; sc.innerIters[5] += rOm.iterations;
    B.compTIters += rT.iterations;
    if (st.recordHistory && iter < M.histCap) {
        const double h[6] = {eUx, eUy, eP, tChange, eK, eOm};
        B.hist.insert(B.hist.end(), h, h + 6);
        sc.histLen = iter + 1;
    }
    sc.iter = iter + 1;
    bool diverged = !std::isfinite(eUx) || !std::isfinite(eUy) || !std::isfinite(eP) || !std::isfinite(tChange) ||
                    !std::isfinite(eK) || !std::isfinite(eOm) || !std::isfinite(uChange) || !finite_res(rU[0]) ||
                    !finite_res(rU[1]) || !finite_res(rU[2]) || !finite_res(rP) || !finite_res(rT) ||
                    (turbUpdate && (!finite_res(rK) || !finite_res(rOm))) || mx[CQ_BAD] != 0.0;
    if (!diverged) diverged = mx[CQ_PMAX] > st.divLimit;
    if (diverged) {
        sc.status = RSGM_DIVERGED; sc.finalIter = iter + 1;
        return hipSuccess;
    }
    double maxRes = hmax(hmax(hmax(hmax(eUx, eUy), eP), uChange), tChange);
    if (turbActive) maxRes = hmax(hmax(maxRes, sc.lastK), sc.lastOm);
    const bool turbScheduled = st.turbStart < st.maxIter;
    if (maxRes < st.tol && iter > 0 && (!turbScheduled || (turbActive && sc.turbEstablished))) {
        // the accepted state's density from the final k, re-validated
        hipLaunchKernelGGL(kc_state, gc, bl, 0, B.s, dM, dQ, 0, 0, 1);
        hipLaunchKernelGGL(kc_valid, gc, bl, 0, B.s, dM, dQ, B.compPart);
        e = hipGetLastError();
        if (e == hipSuccess) e = comp_maxima(B, nb, mx);
        if (e != hipSuccess) return e;
        sc.status = mx[CQ_BAD] != 0.0 ? RSGM_DIVERGED : RSGM_CONVERGED;
        sc.finalIter = iter + 1;
    } else if (sc.iter >= st.maxIter) {
        sc.status = RSGM_MAXITER; sc.finalIter = st.maxIter;
    }
    return hipSuccess;
}
