// rsg_comp.h -- compressible pressure-based SIMPLE (ideal gas, energy
// equation, density-weighted fluxes) on the kernel-per-phase path.
//
// Included by rsg_big.hip inside its anonymous namespace: a compressible
// member is an RsgBig whose outer iteration is comp_outer_iteration below.
// Every phase is a grid-wide launch over cells (the cell-gather assembly of
// rsg_member.h: internal faces by ascending face index, then boundary faces,
// the order of the reference's face-loop scatter), and every linear solve is
// rsg_big.hip's solve() on the CSR image of the assembled LDU system
// (BiCGSTAB / PCG / AMG-PCG / Gauss-Seidel as the settings name them; the
// energy equation always BiCGSTAB).  On a small mesh (a workspace carved with
// the member path's AMG arrays: memberSolve) each solve is instead one
// workgroup-wide launch of the member path's solvers on the LDU system, and
// the host waits once per outer iteration instead of once per dot product.
//
// Reference restated: CompressibleSIMPLESolver (compressible/src/
// CompressibleSIMPLESolver.cpp): initUniform :27-72, updateDensity :125-138,
// mechanicalizeOutletPressure :82-98, assembleMomentum :144-313,
// assemblePressureCorrection :321-372, assembleEnergy :378-450,
// assembleKEquation :456-527, assembleOmegaEquation :531-608, correctVelocity /
// correctPressure :612-645, stateIsValid :100-121, solve :666-1000; the
// ideal-gas relations of IdealGasEOS.hpp:6-38; the per-cell-viscosity SST
// closure (core/src/SSTModel.cpp:176-203) and omega wall value
// (core/include/BoundaryCondition.hpp:305-331); the wall-aware transpose
// stress (core/include/StressOperators.hpp:31-84).

// slots of the one-workgroup solves' results (memberSolve members)
enum { SLOT_UX = 0, SLOT_UY, SLOT_UZ, SLOT_P, SLOT_T, SLOT_K, SLOT_OM };

// reduction slots of the per-iteration maxima (block partials in B.compPart)
enum {
    CQ_UVAL = 0, CQ_UDIFF, CQ_TVAL, CQ_TDIFF, CQ_KVAL, CQ_OMVAL, CQ_KDIFF, CQ_OMDIFF, CQ_PMAX, CQ_BAD, CQ_COUNT
};

#define COMP_CELLS(c) for (int c = C.t.tid; c < C.nC; c += C.t.bs)
#define COMP_BFACES(b) for (int b = C.t.tid; b < C.nB; b += C.t.bs)

// Sutherland viscosity (IdealGasEOS.hpp:18-22)
__device__ inline double comp_visc(const RsgCompDev* Q, double T) {
    const double ratio = T / Q->Tref;
    return Q->muRef * ratio * sqrt(ratio) * (Q->Tref + Q->S) / (T + Q->S);
}
__device__ inline double comp_cp(const RsgCompDev* Q) { return Q->gamma * Q->R / (Q->gamma - 1.0); }
// conductivity mu Cp / Pr + muT Cp / Pr_T (IdealGasEOS.hpp:25-27)
__device__ inline double comp_lambda(const RsgCompDev* Q, double mu, double muT) {
    const double Cp = comp_cp(Q);
    return mu * Cp / Q->Pr + muT * Cp / Q->PrT;
}
// rho = p_mech / (R max(T, 1) + (2/3) max(k, 0)) (updateDensity)
__device__ inline double comp_rho(const RsgCompDev* Q, double p, double T, double k) {
    return p / (Q->R * dmax(T, 1.0) + (2.0 / 3.0) * dmax(k, 0.0));
}
// outlet faces whose prescribed (thermodynamic) pressure is mechanicalized
__device__ inline bool comp_outlet_p(const Ctx& C, int b) {
    const int ty = C.bc->pT[b];
    return (C.m->bFlags[b] & RSG_BF_OUTLET) && (ty == RSGBC_OUTLET_PRESSURE || ty == RSGBC_DIRICHLET);
}

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

// initUniform's fields beyond the incompressible ones (T, rho, mu, T boundary)
__global__ __launch_bounds__(kBlock) void kc_init(RsgMemberDev* M, RsgCompDev* Q) {
    Ctx C = gctx(M);
    init_uniform_fields(C);
    COMP_CELLS(c) { Q->T[c] = Q->Tinit; Q->rho[c] = Q->rho0; Q->mu[c] = Q->mu0; }
    COMP_BFACES(b) Q->Tb[b] = Q->Tinit;
}

// updateDensity (+ updateViscosity and the iteration's starting velocity when
// `start`), the temperature BC when `tbc`, and the mechanicalized outlet
// pressure when `mech` (its owner density evaluated like the cell's).
__global__ __launch_bounds__(kBlock) void kc_state(RsgMemberDev* M, RsgCompDev* Q, int start, int tbc, int mech) {
    Ctx C = gctx(M);
    COMP_CELLS(c) {
        Q->rho[c] = comp_rho(Q, M->p[c], Q->T[c], M->k[c]);
        if (start) {
            Q->mu[c] = comp_visc(Q, Q->T[c]);
            Q->Ux0[c] = M->Ux[c]; Q->Uy0[c] = M->Uy[c]; Q->Uz0[c] = M->Uz[c];
        }
    }
    const int nIF = C.nIF;
    COMP_BFACES(b) {
        const int o = C.m->own[nIF + b];
        if (tbc) {
            const int ty = Q->tT[b];
            if (ty == RSGT_DIRICHLET) Q->Tb[b] = Q->tV[b];
            else if (ty == RSGT_ZERO_GRADIENT) Q->Tb[b] = Q->T[o];
        }
        if (mech && comp_outlet_p(C, b))
            M->pb[b] = C.bc->pV[b] + (2.0 / 3.0) * comp_rho(Q, M->p[o], Q->T[o], M->k[o]) * dmax(M->k[o], 0.0);
    }
}

// temperature BC and mechanicalized outlet from the stored density
__global__ __launch_bounds__(kBlock) void kc_bnd(RsgMemberDev* M, RsgCompDev* Q, int tbc, int mech) {
    Ctx C = gctx(M);
    const int nIF = C.nIF;
    COMP_BFACES(b) {
        const int o = C.m->own[nIF + b];
        if (tbc) {
            const int ty = Q->tT[b];
            if (ty == RSGT_DIRICHLET) Q->Tb[b] = Q->tV[b];
            else if (ty == RSGT_ZERO_GRADIENT) Q->Tb[b] = Q->T[o];
        }
        if (mech && comp_outlet_p(C, b))
            M->pb[b] = C.bc->pV[b] + (2.0 / 3.0) * Q->rho[o] * dmax(M->k[o], 0.0);
    }
}

// velocity and pressure BCs after the corrections, outlet mechanicalized
// (correctVelocity / correctPressure)
__global__ __launch_bounds__(kBlock) void kc_bnd_vp(RsgMemberDev* M, RsgCompDev* Q) {
    Ctx C = gctx(M);
    vel_bc(C);
    p_bc(C);   // same face-to-thread mapping as the loop below
    const int nIF = C.nIF;
    COMP_BFACES(b) {
        if (comp_outlet_p(C, b)) {
            const int o = C.m->own[nIF + b];
            M->pb[b] = C.bc->pV[b] + (2.0 / 3.0) * Q->rho[o] * dmax(M->k[o], 0.0);
        }
    }
}

// step 2: SST closure with the local kinematic viscosity mu / rho (or the
// velocity gradients alone) and grad p (mechanical)
__global__ __launch_bounds__(kBlock) void kc_step1(RsgMemberDev* M, RsgCompDev* Q, int turbUpdate, int established) {
    Ctx C = gctx(M);
    if (turbUpdate) {
        const double nuTMin = established ? 0.0 : 0.1 * (Q->mu[0] / dmax(Q->rho[0], 1e-30));
        COMP_CELLS(c) {
            C.nu = Q->mu[c] / dmax(Q->rho[c], 1e-30);
            sst_closure_cell(C, c, nuTMin);
        }
    } else {
        grad_u(C, 0, 3);
    }
    grad_scalar(C, M->p, M->pb, M->gp);
}

// muEff div U per cell (current velocity gradients)
__global__ __launch_bounds__(kBlock) void kc_divu(RsgMemberDev* M, RsgCompDev* Q) {
    Ctx C = gctx(M);
    COMP_CELLS(c) {
        const double me = Q->mu[c] + Q->rho[c] * M->nuT[c];
        Q->mdu[c] = me * (M->g[0][c] + M->g[4][c] + M->g[8][c]);
    }
}

__global__ __launch_bounds__(kBlock) void kc_gtr(RsgMemberDev* M, RsgCompDev* Q) {
    Ctx C = gctx(M);
    grad_scalar(C, Q->mdu, Q->zeroB, Q->gt);
}

// assembleMomentum for one component: density-weighted upwind convection,
// central diffusion with muEff = mu + rho nuT (molecular mu at resolved
// walls), -grad p V, the transpose stress (wall coefficient: molecular mu)
// and -(2/3) grad(muEff div U) V, under-relaxation; aPx keeps the x diagonal.
__global__ __launch_bounds__(kBlock) void kc_asm_mom(RsgMemberDev* M, RsgCompDev* Q, int comp) {
    Ctx C = gctx(M);
    const RsgMeshDev* m = C.m;
    const int nIF = C.nIF;
    const double* gc0 = M->g[0 * 3 + comp];
    const double* gc1 = M->g[1 * 3 + comp];
    const double* gc2 = M->g[2 * 3 + comp];
    const double* phiOld = ucomp(M, comp);
    const double* Ub = ubcomp(M, comp);
    const double a = C.s->aU;
    const int wallFn = C.s->wallFn;
    const double* gpc = M->gp[comp];
    const double* gtc = Q->gt[comp];
    const double* rho = Q->rho;
    const double* mu = Q->mu;
    const double* nuT = M->nuT;
    COMP_CELLS(c) {
        double dg = 0.0, src = 0.0, ts = 0.0;
        PendingLdu pend;
        pending_clear(pend);
        cell_faces(m, c, [&](int code, int other, int slot) {
            const int f = code >> 1;
            double ddg, dsrc = 0.0, dts;
            const double Sf = m->area[f], delta = m->dlt[f];
            if (f < nIF) {
                const bool nbSide = code & 1;
                const int o = nbSide ? other : c, n = nbSide ? c : other;
                const double w = m->w[f], wm = 1.0 - w;
                const double rho_f = w * rho[o] + (1.0 - w) * rho[n];
                const double mu_f = w * mu[o] + (1.0 - w) * mu[n];
                const double nuT_f = w * nuT[o] + (1.0 - w) * nuT[n];
                const double Df = (mu_f + rho_f * nuT_f) * Sf / delta;
                const double ufx = M->Ux[o] * w + M->Ux[n] * wm;
                const double ufy = M->Uy[o] * w + M->Uy[n] * wm;
                const double ufz = M->Uz[o] * w + M->Uz[n] * wm;
                const double mFlux = rho_f * (ufx * m->nx[f] + ufy * m->ny[f] + ufz * m->nz[f]) * Sf;
                const double cPos = dmax(mFlux, 0.0), cNeg = dmax(-mFlux, 0.0);
                const double up = -(Df + cNeg), lo = -(Df + cPos);
                // transpose stress, coefficient muEff interpolated to the face
                const double meo = mu[o] + rho[o] * nuT[o], men = mu[n] + rho[n] * nuT[n];
                const double cf = w * meo + (1.0 - w) * men;
                const double gx = gc0[o] * w + gc0[n] * wm;
                const double gy = gc1[o] * w + gc1[n] * wm;
                const double gz = gc2[o] * w + gc2[n] * wm;
                const double fs = cf * (gx * m->nx[f] + gy * m->ny[f] + gz * m->nz[f]) * Sf;
                if (!nbSide) pending_put(pend, M, slot, f, up, lo);
                ddg = nbSide ? -up : -lo;
                dts = nbSide ? -fs : fs;
            } else {
                const int b = f - nIF;
                const int isWall = m->bFlags[b] & RSG_BF_WALL;
                const double muFull = mu[c] + rho[c] * nuT[c];
                const double muEff = (isWall && !wallFn) ? mu[c] : muFull;
                const double Db = muEff * Sf / delta;
                const double mFlux = rho[c] * (M->Ubx[b] * m->nx[f] + M->Uby[b] * m->ny[f] + M->Ubz[b] * m->nz[f]) * Sf;
                const bool out = mFlux >= 0;
                ddg = out ? mFlux + Db : Db;
                dsrc = out ? Db * Ub[b] : (Db - mFlux) * Ub[b];
                const double cfb = isWall ? mu[c] : muFull;
                dts = cfb * (gc0[c] * m->nx[f] + gc1[c] * m->ny[f] + gc2[c] * m->nz[f]) * Sf;
            }
            dg += ddg; src += dsrc; ts += dts;
        });
        pending_flush(pend, M);
        const double vol = m->vol[c];
        src -= gpc[c] * vol;
        src += ts - (2.0 / 3.0) * gtc[c] * vol;
        src += (1.0 - a) / a * dg * phiOld[c];
        dg /= a;
        M->diag[c] = dg;
        M->src[c] = src;
        if (comp == 0) M->aPx[c] = dg;
    }
}

// assemblePressureCorrection: coefficients rho V / aP (x diagonal), the
// density-weighted mass-flux imbalance, outlet faces in the diagonal; no
// Rhie-Chow term and no reference-pressure pin.  p' starts at zero.
__global__ __launch_bounds__(kBlock) void kc_asm_pc(RsgMemberDev* M, RsgCompDev* Q) {
    Ctx C = gctx(M);
    const RsgMeshDev* m = C.m;
    const int nIF = C.nIF;
    const double* aP = M->aPx;
    const double* rho = Q->rho;
    COMP_CELLS(c) {
        double dg = 0.0, src = 0.0;
        PendingLdu pend;
        pending_clear(pend);
        cell_faces(m, c, [&](int code, int other, int slot) {
            const int f = code >> 1;
            double ddg, dsrc;
            const double Sf = m->area[f];
            if (f < nIF) {
                const bool nbSide = code & 1;
                const int o = nbSide ? other : c, n = nbSide ? c : other;
                const double delta = m->dlt[f], w = m->w[f], wm = 1.0 - w;
                const double rho_f = w * rho[o] + (1.0 - w) * rho[n];
                const double dP_o = (fabs(aP[o]) > 1e-30) ? rho[o] * m->vol[o] / aP[o] : 0.0;
                const double dP_n = (fabs(aP[n]) > 1e-30) ? rho[n] * m->vol[n] / aP[n] : 0.0;
                const double dP_f = w * dP_o + (1.0 - w) * dP_n;
                const double coeff = dP_f * Sf / delta;
                const double ufx = M->Ux[o] * w + M->Ux[n] * wm;
                const double ufy = M->Uy[o] * w + M->Uy[n] * wm;
                const double ufz = M->Uz[o] * w + M->Uz[n] * wm;
                const double mFlux = rho_f * (ufx * m->nx[f] + ufy * m->ny[f] + ufz * m->nz[f]) * Sf;
                if (!nbSide) pending_put(pend, M, slot, f, -coeff, -coeff);
                ddg = coeff;
                dsrc = nbSide ? mFlux : -mFlux;
            } else {
                const int b = f - nIF;
                const double mf = rho[c] * (M->Ubx[b] * m->nx[f] + M->Uby[b] * m->ny[f] + M->Ubz[b] * m->nz[f]) * Sf;
                dsrc = -mf;
                ddg = 0.0;
                if (m->bFlags[b] & RSG_BF_OUTLET) {
                    const double dP_o = (fabs(aP[c]) > 1e-30) ? rho[c] * m->vol[c] / aP[c] : 0.0;
                    ddg = dP_o * Sf / m->dlt[f];
                }
            }
            dg += ddg; src += dsrc;
        });
        pending_flush(pend, M);
        M->diag[c] = dg;
        M->src[c] = src;
        M->pp[c] = 0.0;
    }
}

// assembleEnergy: rho Cp upwind convection, conductivity mu Cp/Pr +
// rho nuT Cp/Pr_T, alphaT relaxation; Told keeps T; the velocity change
// norm's maxima over the iteration so far (|U|, |U - U_start|).
__global__ __launch_bounds__(kBlock) void kc_asm_e(RsgMemberDev* M, RsgCompDev* Q, double* part) {
    Ctx C = gctx(M);
    const RsgMeshDev* m = C.m;
    const int nIF = C.nIF;
    const double Cp = comp_cp(Q);
    const double a = Q->alphaT;
    const double* rho = Q->rho;
    const double* mu = Q->mu;
    const double* nuT = M->nuT;
    double mx[2] = {1e-30, 0.0};
    COMP_CELLS(c) {
        double dg = 0.0, src = 0.0;
        PendingLdu pend;
        pending_clear(pend);
        cell_faces(m, c, [&](int code, int other, int slot) {
            const int f = code >> 1;
            double ddg, dsrc = 0.0;
            const double Sf = m->area[f], delta = m->dlt[f];
            if (f < nIF) {
                const bool nbSide = code & 1;
                const int o = nbSide ? other : c, n = nbSide ? c : other;
                const double w = m->w[f], wm = 1.0 - w;
                const double rho_f = w * rho[o] + (1.0 - w) * rho[n];
                const double mu_f = w * mu[o] + (1.0 - w) * mu[n];
                const double nuT_f = w * nuT[o] + (1.0 - w) * nuT[n];
                const double Df = comp_lambda(Q, mu_f, rho_f * nuT_f) * Sf / delta;
                const double ufx = M->Ux[o] * w + M->Ux[n] * wm;
                const double ufy = M->Uy[o] * w + M->Uy[n] * wm;
                const double ufz = M->Uz[o] * w + M->Uz[n] * wm;
                const double mFlux = rho_f * Cp * (ufx * m->nx[f] + ufy * m->ny[f] + ufz * m->nz[f]) * Sf;
                const double cPos = dmax(mFlux, 0.0), cNeg = dmax(-mFlux, 0.0);
                const double up = -(Df + cNeg), lo = -(Df + cPos);
                if (!nbSide) pending_put(pend, M, slot, f, up, lo);
                ddg = nbSide ? -up : -lo;
            } else {
                const int b = f - nIF;
                const double Db = comp_lambda(Q, mu[c], rho[c] * nuT[c]) * Sf / delta;
                const double mFlux =
                    rho[c] * Cp * (M->Ubx[b] * m->nx[f] + M->Uby[b] * m->ny[f] + M->Ubz[b] * m->nz[f]) * Sf;
                const bool out = mFlux >= 0;
                const double Tb = Q->Tb[b];
                ddg = out ? mFlux + Db : Db;
                dsrc = out ? Db * Tb : (Db - mFlux) * Tb;
            }
            dg += ddg; src += dsrc;
        });
        pending_flush(pend, M);
        const double Tc = Q->T[c];
        src += (1.0 - a) / a * dg * Tc;
        dg /= a;
        M->diag[c] = dg;
        M->src[c] = src;
        Q->Told[c] = Tc;
        const double ux = M->Ux[c], uy = M->Uy[c], uz = M->Uz[c];
        const double mag = sqrt(ux * ux + uy * uy + uz * uz);
        const double dx = ux - Q->Ux0[c], dy = uy - Q->Uy0[c], dz = uz - Q->Uz0[c];
        mx[0] = dmax(mx[0], mag);
        mx[1] = dmax(mx[1], sqrt(dx * dx + dy * dy + dz * dz));
    }
    comp_block_max<2>(mx, part, CQ_UVAL);
}

// T = max(T, 1) after the energy solve, and the temperature change norm's maxima
__global__ __launch_bounds__(kBlock) void kc_tpost(RsgMemberDev* M, RsgCompDev* Q, double* part) {
    Ctx C = gctx(M);
    double mx[2] = {1e-30, 0.0};
    COMP_CELLS(c) {
        const double T = dmax(Q->T[c], 1.0);
        Q->T[c] = T;
        mx[0] = dmax(mx[0], fabs(T));
        mx[1] = dmax(mx[1], fabs(T - Q->Told[c]));
    }
    comp_block_max<2>(mx, part, CQ_TVAL);
}

// assembleKEquation (isOmega 0) / assembleOmegaEquation (1): rho-weighted
// convection, mu + rho sigma nuT diffusion, rho-weighted SST sources.
__global__ __launch_bounds__(kBlock) void kc_asm_turb(RsgMemberDev* M, RsgCompDev* Q, int isOmega) {
    Ctx C = gctx(M);
    const RsgMeshDev* m = C.m;
    const int nIF = C.nIF;
    const double s1 = isOmega ? C.sw1 : C.sk1, s2 = isOmega ? C.sw2 : C.sk2;
    const double* phi = isOmega ? M->om : M->k;
    const double* phib = isOmega ? M->omb : M->kb;
    const double a = isOmega ? C.s->aOm : C.s->aK;
    const double* rho = Q->rho;
    const double* mu = Q->mu;
    const double* nuT = M->nuT;
    COMP_CELLS(c) {
        double dg = 0.0, src = 0.0;
        PendingLdu pend;
        pending_clear(pend);
        cell_faces(m, c, [&](int code, int other, int slot) {
            const int f = code >> 1;
            double ddg, dsrc = 0.0;
            const double Sf = m->area[f], delta = m->dlt[f];
            if (f < nIF) {
                const bool nbSide = code & 1;
                const int o = nbSide ? other : c, n = nbSide ? c : other;
                const double w = m->w[f], wm = 1.0 - w;
                const double F1f = w * M->F1[o] + (1.0 - w) * M->F1[n];
                const double sg = blend(s1, s2, F1f);
                const double rho_f = w * rho[o] + (1.0 - w) * rho[n];
                const double mu_f = w * mu[o] + (1.0 - w) * mu[n];
                const double nuT_f = w * nuT[o] + (1.0 - w) * nuT[n];
                const double Df = (mu_f + rho_f * sg * nuT_f) * Sf / delta;
                const double ufx = M->Ux[o] * w + M->Ux[n] * wm;
                const double ufy = M->Uy[o] * w + M->Uy[n] * wm;
                const double ufz = M->Uz[o] * w + M->Uz[n] * wm;
                const double mFlux = rho_f * (ufx * m->nx[f] + ufy * m->ny[f] + ufz * m->nz[f]) * Sf;
                const double cPos = dmax(mFlux, 0.0), cNeg = dmax(-mFlux, 0.0);
                const double up = -(Df + cNeg), lo = -(Df + cPos);
                if (!nbSide) pending_put(pend, M, slot, f, up, lo);
                ddg = nbSide ? -up : -lo;
            } else {
                const int b = f - nIF;
                const double sg = blend(s1, s2, M->F1[c]);
                const double Db = (mu[c] + rho[c] * sg * nuT[c]) * Sf / delta;
                const double mFlux = rho[c] * (M->Ubx[b] * m->nx[f] + M->Uby[b] * m->ny[f] + M->Ubz[b] * m->nz[f]) * Sf;
                const bool out = mFlux >= 0;
                ddg = out ? mFlux + Db : Db;
                dsrc = out ? Db * phib[b] : (Db - mFlux) * phib[b];
            }
            dg += ddg; src += dsrc;
        });
        pending_flush(pend, M);
        const double vol = m->vol[c];
        const double rhoC = rho[c];
        if (!isOmega) {
            src += rhoC * M->Pk[c] * vol;
            const double dest = rhoC * C.bs * dmax(M->om[c], 1e-20);
            dg += dest * vol;
        } else {
            const double F1 = M->F1[c];
            const double omC = dmax(M->om[c], 1e-20);
            const double alphaB = blend(C.al1, C.al2, F1);
            src += rhoC * alphaB * sst_prod_om(C, M->nuT[c], M->Smag[c], M->k[c], M->om[c]) * vol;
            const double betaB = blend(C.b1, C.b2, F1);
            dg += rhoC * betaB * omC * vol;
            src += rhoC * (1.0 - F1) * M->CDkw[c] * vol;
        }
        src += (1.0 - a) / a * dg * phi[c];
        dg /= a;
        M->diag[c] = dg;
        M->src[c] = src;
        if (isOmega) M->omOld[c] = phi[c]; else M->kOld[c] = phi[c];
    }
}

// omega boundary values with the owner cell's kinematic viscosity mu / rho
__global__ __launch_bounds__(kBlock) void kc_ombc(RsgMemberDev* M, RsgCompDev* Q, double beta1) {
    Ctx C = gctx(M);
    const RsgMeshDev* m = C.m;
    const int nIF = C.nIF;
    COMP_BFACES(b) {
        const int ty = C.bc->omT[b], f = nIF + b, o = m->own[f];
        if (ty == RSGBC_KEEP) continue;
        if (ty == RSGBC_WALL_KOMEGA) {
            const double y1 = m->dlt[f];
            const double nuO = Q->mu[o] / dmax(Q->rho[o], 1e-30);
            M->omb[b] = 60.0 * nuO / (beta1 * y1 * y1);
        } else if (ty == RSGBC_DIRICHLET) {
            M->omb[b] = C.bc->omV[b];
        } else {
            M->omb[b] = M->om[o];
        }
    }
}

// wall-cell omega re-pin with the local viscosity, and the turbulence change
// norms' maxima
__global__ __launch_bounds__(kBlock) void kc_pin(RsgMemberDev* M, RsgCompDev* Q, double* part) {
    Ctx C = gctx(M);
    double mx[4] = {1e-30, 1e-30, 0.0, 0.0};
    COMP_CELLS(c) {
        C.nu = Q->mu[c] / dmax(Q->rho[c], 1e-30);
        pin_cell(C, c, mx);
    }
    comp_block_max<4>(mx, part, CQ_KVAL);
}

// stateIsValid (every field finite, T, rho, p and p_thermo positive) and max |p|
__global__ __launch_bounds__(kBlock) void kc_valid(RsgMemberDev* M, RsgCompDev* Q, double* part) {
    Ctx C = gctx(M);
    double mx[2] = {0.0, 0.0};
    COMP_CELLS(c) {
        const double p = M->p[c], rho = Q->rho[c], k = M->k[c], T = Q->T[c];
        const double pTh = p - (2.0 / 3.0) * rho * dmax(k, 0.0);
        const bool finite = isfinite(M->Ux[c]) && isfinite(M->Uy[c]) && isfinite(M->Uz[c]) && isfinite(p) &&
                            isfinite(T) && isfinite(rho) && isfinite(k) && isfinite(M->om[c]) && isfinite(pTh);
        if (!finite || T <= 0.0 || rho <= 0.0 || p <= 0.0 || pTh <= 0.0) mx[1] = 1.0;
        mx[0] = dmax(mx[0], fabs(p));
    }
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

// A finished solve's export (CompressibleForwardModel.cpp:72-94): p becomes
// the thermodynamic pressure p_mech - (2/3) rho max(k, 0) (pMech keeps the
// mechanical one; boundary values stay mechanical), and the member's nu the
// effective mu(T_0) / rho_0 the observables use.
__global__ __launch_bounds__(kBlock) void kc_export(RsgMemberDev* M, RsgCompDev* Q, int nC) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x, bs = gridDim.x * blockDim.x;
    for (int c = tid; c < nC; c += bs) {
        const double p = M->p[c];
        Q->pMech[c] = p;
        M->p[c] = p - (2.0 / 3.0) * Q->rho[c] * dmax(M->k[c], 0.0);
    }
    if (tid == 0) M->nu = comp_visc(Q, Q->T[0]) / dmax(Q->rho[0], 1e-30);
}

#undef COMP_CELLS
#undef COMP_BFACES

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
