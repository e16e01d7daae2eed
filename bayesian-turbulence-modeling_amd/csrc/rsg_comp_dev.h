// rsg_comp_dev.h -- the phases of the compressible pressure-based SIMPLE
// iteration as device functions over the cells / boundary faces of a Ctx:
// shared by the kernel-per-phase launches (rsg_comp.h, one thread per cell
// over the grid) and the one-workgroup member kernel (comp_iteration below,
// a workgroup's threads striding over the member's cells).  A phase that
// reduces returns its per-thread maxima; the caller combines them (block
// partials on the grid, a workgroup reduction in the member kernel).
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
// Included at the end of rsg_member.h, inside namespace rsg.

#define COMP_CELLS(c) for (int c = C.t.tid; c < C.nC; c += C.t.bs)
#define COMP_BFACES(b) for (int b = C.t.tid; b < C.nB; b += C.t.bs)

// Sutherland viscosity (IdealGasEOS.hpp:18-22)
RSG_HD double comp_visc(const RsgCompDev* Q, double T) {
    const double ratio = T / Q->Tref;
    return Q->muRef * ratio * sqrt(ratio) * (Q->Tref + Q->S) / (T + Q->S);
}
RSG_HD double comp_cp(const RsgCompDev* Q) { return Q->gamma * Q->R / (Q->gamma - 1.0); }
// conductivity mu Cp / Pr + muT Cp / Pr_T (IdealGasEOS.hpp:25-27)
RSG_HD double comp_lambda(const RsgCompDev* Q, double mu, double muT) {
    const double Cp = comp_cp(Q);
    return mu * Cp / Q->Pr + muT * Cp / Q->PrT;
}
// rho = p_mech / (R max(T, 1) + (2/3) max(k, 0)) (updateDensity)
RSG_D double comp_rho(const RsgCompDev* Q, double p, double T, double k) {
    return p / (Q->R * dmax(T, 1.0) + (2.0 / 3.0) * dmax(k, 0.0));
}
// outlet faces whose prescribed (thermodynamic) pressure is mechanicalized
RSG_D bool comp_outlet_p(const Ctx& C, int b) {
    const int ty = C.bc->pT[b];
    return (C.m->bFlags[b] & RSG_BF_OUTLET) && (ty == RSGBC_OUTLET_PRESSURE || ty == RSGBC_DIRICHLET);
}

// initUniform's fields beyond the incompressible ones (T, rho, mu, T boundary)
RSG_D void comp_init_fields(Ctx& C, RsgCompDev* Q) {
    init_uniform_fields(C);
    COMP_CELLS(c) { Q->T[c] = Q->Tinit; Q->rho[c] = Q->rho0; Q->mu[c] = Q->mu0; }
    COMP_BFACES(b) Q->Tb[b] = Q->Tinit;
}

// updateDensity (+ updateViscosity and the iteration's starting velocity when
// `start`), the temperature BC when `tbc`, the mechanicalized outlet pressure
// when `mech` (its owner density evaluated like the cell's)
RSG_D void comp_state(Ctx& C, RsgCompDev* Q, int start, int tbc, int mech) {
    RsgMemberDev* M = C.M;
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
RSG_D void comp_bnd(Ctx& C, RsgCompDev* Q, int tbc, int mech) {
    RsgMemberDev* M = C.M;
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

// the outlet's mechanicalized pressure after p_bc (correctPressure)
RSG_D void comp_outlet_mech(Ctx& C, RsgCompDev* Q) {
    RsgMemberDev* M = C.M;
    const int nIF = C.nIF;
    COMP_BFACES(b) {
        if (comp_outlet_p(C, b)) {
            const int o = C.m->own[nIF + b];
            M->pb[b] = C.bc->pV[b] + (2.0 / 3.0) * Q->rho[o] * dmax(M->k[o], 0.0);
        }
    }
}

// SST closure with the local kinematic viscosity mu / rho (or the velocity
// gradients alone) and grad p (mechanical); C.nu is restored.  The closure's
// k gradient takes zero boundary values: the reference hands computeFields a
// fresh copy of k with its cell values only (CompressibleSIMPLESolver.cpp:
// 682-693), so the cross-diffusion term of the inlet and outlet cells sees
// k = 0 on their boundary faces.
RSG_D void comp_closure(Ctx& C, RsgCompDev* Q, int turbUpdate, int established) {
    RsgMemberDev* M = C.M;
    const double nu0 = C.nu;
    if (turbUpdate) {
        const double nuTMin = established ? 0.0 : 0.1 * (Q->mu[0] / dmax(Q->rho[0], 1e-30));
        COMP_CELLS(c) {
            C.nu = Q->mu[c] / dmax(Q->rho[c], 1e-30);
            sst_closure_cell(C, c, nuTMin, Q->zeroB);
        }
    } else {
        grad_u(C, 0, 3);
    }
    grad_scalar(C, M->p, M->pb, M->gp);
    C.nu = nu0;
}

// muEff div U per cell (current velocity gradients)
RSG_D void comp_divu(Ctx& C, RsgCompDev* Q) {
    RsgMemberDev* M = C.M;
    COMP_CELLS(c) {
        const double me = Q->mu[c] + Q->rho[c] * M->nuT[c];
        Q->mdu[c] = me * (M->g[0][c] + M->g[4][c] + M->g[8][c]);
    }
}

// grad(muEff div U), boundary values zero
RSG_D void comp_gtr(Ctx& C, RsgCompDev* Q) {
    grad_scalar(C, Q->mdu, Q->zeroB, Q->gt);
}

// assembleMomentum for one component: density-weighted upwind convection,
// central diffusion with muEff = mu + rho nuT (molecular mu at resolved
// walls), -grad p V, the transpose stress (wall coefficient: molecular mu)
// and -(2/3) grad(muEff div U) V, under-relaxation; aPx keeps the x diagonal
RSG_D void comp_asm_mom(Ctx& C, RsgCompDev* Q, int comp) {
    RsgMemberDev* M = C.M;
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
RSG_D void comp_asm_pc(Ctx& C, RsgCompDev* Q) {
    RsgMemberDev* M = C.M;
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
// rho nuT Cp/Pr_T, alphaT relaxation; Told keeps T; mx gets the velocity
// change norm's running maxima (|U|, |U - U_start|)
RSG_D void comp_asm_e(Ctx& C, RsgCompDev* Q, double (&mx)[2]) {
    RsgMemberDev* M = C.M;
    const RsgMeshDev* m = C.m;
    const int nIF = C.nIF;
    const double Cp = comp_cp(Q);
    const double a = Q->alphaT;
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
}

// T = max(T, 1) after the energy solve; mx: the temperature change norm's maxima
RSG_D void comp_tpost(Ctx& C, RsgCompDev* Q, double (&mx)[2]) {
    RsgMemberDev* M = C.M;
    COMP_CELLS(c) {
        const double T = dmax(Q->T[c], 1.0);
        Q->T[c] = T;
        mx[0] = dmax(mx[0], fabs(T));
        mx[1] = dmax(mx[1], fabs(T - Q->Told[c]));
    }
}

// assembleKEquation (isOmega 0) / assembleOmegaEquation (1): rho-weighted
// convection, mu + rho sigma nuT diffusion, rho-weighted SST sources
RSG_D void comp_asm_turb(Ctx& C, RsgCompDev* Q, int isOmega) {
    RsgMemberDev* M = C.M;
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
RSG_D void comp_ombc(Ctx& C, RsgCompDev* Q, double beta1) {
    RsgMemberDev* M = C.M;
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

// wall-cell omega re-pin with the local viscosity; mx: the turbulence change
// norms' maxima (|k|, |omega|, |k - kOld|, |omega - omegaOld|); C.nu is restored
RSG_D void comp_pin(Ctx& C, RsgCompDev* Q, double (&mx)[4]) {
    const double nu0 = C.nu;
    COMP_CELLS(c) {
        C.nu = Q->mu[c] / dmax(Q->rho[c], 1e-30);
        pin_cell(C, c, mx);
    }
    C.nu = nu0;
}

// stateIsValid (every field finite, T, rho, p and p_thermo positive): mx[1]
// becomes 1 on an invalid cell; mx[0]: max |p|
RSG_D void comp_valid(Ctx& C, RsgCompDev* Q, double (&mx)[2]) {
    RsgMemberDev* M = C.M;
    COMP_CELLS(c) {
        const double p = M->p[c], rho = Q->rho[c], k = M->k[c], T = Q->T[c];
        const double pTh = p - (2.0 / 3.0) * rho * dmax(k, 0.0);
        const bool finite = isfinite(M->Ux[c]) && isfinite(M->Uy[c]) && isfinite(M->Uz[c]) && isfinite(p) &&
                            isfinite(T) && isfinite(rho) && isfinite(k) && isfinite(M->om[c]) && isfinite(pTh);
        if (!finite || T <= 0.0 || rho <= 0.0 || p <= 0.0 || pTh <= 0.0) mx[1] = 1.0;
        mx[0] = dmax(mx[0], fabs(p));
    }
}

RSG_HD bool comp_finite(const SolveRes& r) { return isfinite(r.initialRes) && isfinite(r.finalRes); }

// One outer iteration of CompressibleSIMPLESolver::solve (:686-998) on one
// workgroup: the phases above separated by workgroup barriers, the member
// path's FP64 linear solvers, every maximum reduced over the workgroup.
template <int CMAX>
RSG_D void comp_iteration(Ctx& C, Scal& S, RsgCompDev* Q) {
    Thr& t = C.t; const RsgMeshDev* m = C.m; RsgMemberDev* M = C.M; const RsgSettingsDev* st = C.s;
    const int nIF = C.nIF;
    const int iter = S.iter;
    const bool turbActive = iter >= st->turbStart;
    const bool turbUpdate = turbActive && ((iter - st->turbStart) % st->turbInterval == 0);
    // 1. thermodynamic state
    comp_state(C, Q, 1, 0, 1);
    RSG_SYNC();
    // 2. SST closure with the local viscosity, grad p
    if (turbUpdate && !S.turbEstablished && iter >= st->turbStart + st->nutFloorIters) S.turbEstablished = 1;
    comp_closure(C, Q, turbUpdate ? 1 : 0, S.turbEstablished);
    RSG_SYNC();
    // 3. momentum
    SolveRes rU0 = {0, 0.0, 0.0, 0}, rU1 = rU0, rU2 = rU0;
    for (int comp = 0; comp < 3; ++comp) {
        if (comp == 2 && M->planarZ) break;
        if (comp > 0) { grad_u(C, comp - 1, comp); RSG_SYNC(); }
        comp_divu(C, Q);
        RSG_SYNC();
        comp_gtr(C, Q);
        RSG_SYNC();
        comp_asm_mom(C, Q, comp);
        RSG_SYNC();
        const SolveRes r = lsolve<CMAX, 0>(C, st->mSolver, ucomp(M, comp), false, false);
        if (comp == 0) rU0 = r; else if (comp == 1) rU1 = r; else rU2 = r;
        RSG_SYNC();
    }
    vel_bc(C);
    RSG_SYNC();
    // 4. pressure correction, corrections
    comp_asm_pc(C, Q);
    RSG_SYNC();
    bool amgReady = false;
    if (st->pSolver == 0) {
        const bool rebuild = S.amgBuiltIter < 0 || iter - S.amgBuiltIter >= st->amgRefresh;
        const bool coarse = rebuild || iter - S.coarseIter >= st->coarseRefresh;
        if (rebuild) {
            amg_refresh_values(C, false);
            if (t.tid == 0) amg_setup_serial(m, M->amg);
            RSG_SYNC();
            S.amgBuiltIter = iter;
        } else {
            const bool resident = res_fits(C, CMAX) && res_amg_fits(C);
            if (coarse || !resident) amg_refresh_values(C, coarse);
        }
        if (coarse) {
            amg_coarse_inverse(C);
            S.coarseIter = iter;
        }
        amgReady = true;
    }
    const SolveRes rP = lsolve<CMAX, 0>(C, st->pSolver, M->pp, amgReady, false);
    RSG_SYNC();
    COMP_BFACES(b) M->ppb[b] = (m->bFlags[b] & RSG_BF_OUTLET) ? 0.0 : M->pp[m->own[nIF + b]];
    RSG_SYNC();
    grad_scalar(C, M->pp, M->ppb, M->gp);
    RSG_SYNC();
    COMP_CELLS(c) correct_cell(C, c);
    RSG_SYNC();
    vel_bc(C);
    p_bc(C);
    comp_outlet_mech(C, Q);   // the same face-to-thread mapping as p_bc
    RSG_SYNC();
    // 5. energy
    double mu2[2] = {1e-30, 0.0};
    comp_asm_e(C, Q, mu2);
    bmax(t, mu2);
    const SolveRes rT = lsolve<CMAX, 0>(C, 2, Q->T, false, false);
    RSG_SYNC();
    double mt[2] = {1e-30, 0.0};
    comp_tpost(C, Q, mt);
    bmax(t, mt);
    comp_state(C, Q, 0, 1, 1);
    RSG_SYNC();
    // 6. turbulence
    SolveRes rK = {0, 0.0, 0.0, 0}, rOm = {0, 0.0, 0.0, 0};
    if (turbUpdate) {
        comp_asm_turb(C, Q, 0);
        RSG_SYNC();
        rK = lsolve<CMAX, 0>(C, st->tSolver, M->k, false, false);
        RSG_SYNC();
        COMP_CELLS(c) M->k[c] = dmax(st->kMin, dmin(st->kMax, M->k[c]));
        RSG_SYNC();
        k_bc(C);
        RSG_SYNC();
        comp_asm_turb(C, Q, 1);
        RSG_SYNC();
        rOm = lsolve<CMAX, 0>(C, st->tSolver, M->om, false, false);
        RSG_SYNC();
        COMP_CELLS(c) M->om[c] = dmax(st->omMin, dmin(1e15, M->om[c]));
        RSG_SYNC();
        comp_ombc(C, Q, C.b1);
        RSG_SYNC();
        double mk[4] = {1e-30, 1e-30, 0.0, 0.0};
        comp_pin(C, Q, mk);
        bmax(t, mk);
        S.lastK = mk[2] / mk[0];
        S.lastOm = mk[3] / mk[1];
    }
    // 7. state validity and max |p|
    double mv[2] = {0.0, 0.0};
    comp_valid(C, Q, mv);
    bmax(t, mv);
    const double uChange = mu2[1] / mu2[0], tChange = mt[1] / mt[0];
    const double eUx = rU0.initialRes, eUy = rU1.initialRes, eP = rP.initialRes;
    const double eK = turbActive ? S.lastK : 0.0, eOm = turbActive ? S.lastOm : 0.0;
    if (t.tid == 0) {
        RsgScalars* sc = M->sc;
        sc->innerIters[0] += rU0.iterations; sc->innerIters[1] += rU1.iterations;
        sc->innerIters[2] += rU2.iterations; sc->innerIters[3] += rP.iterations;
        sc->innerIters[4] += rK.iterations; sc->innerIters[5] += rOm.iterations;
        Q->energyIters += rT.iterations;
        if (st->recordHistory && M->hist && iter < M->histCap) {
            double* h = M->hist + 6 * (long)iter;
            h[0] = eUx; h[1] = eUy; h[2] = eP; h[3] = tChange; h[4] = eK; h[5] = eOm;
            sc->histLen = iter + 1;
        }
    }
    S.iter = iter + 1;
    bool diverged = !isfinite(eUx) || !isfinite(eUy) || !isfinite(eP) || !isfinite(tChange) || !isfinite(eK) ||
                    !isfinite(eOm) || !isfinite(uChange) || !comp_finite(rU0) || !comp_finite(rU1) ||
                    !comp_finite(rU2) || !comp_finite(rP) || !comp_finite(rT) ||
                    (turbUpdate && (!comp_finite(rK) || !comp_finite(rOm))) || mv[1] != 0.0;
    if (!diverged) diverged = mv[0] > st->divLimit;
    if (diverged) {
        S.status = RSGM_DIVERGED; S.finalIter = iter + 1;
    } else {
        double maxRes = dmax(dmax(dmax(dmax(eUx, eUy), eP), uChange), tChange);
        if (turbActive) maxRes = dmax(dmax(maxRes, S.lastK), S.lastOm);
        const bool turbScheduled = st->turbStart < st->maxIter;
        if (maxRes < st->tol && iter > 0 && (!turbScheduled || (turbActive && S.turbEstablished))) {
            // the accepted state's density from the final k, re-validated
            comp_state(C, Q, 0, 0, 1);
            RSG_SYNC();
            double mv2[2] = {0.0, 0.0};
            comp_valid(C, Q, mv2);
            bmax(t, mv2);
            S.status = mv2[1] != 0.0 ? RSGM_DIVERGED : RSGM_CONVERGED;
            S.finalIter = iter + 1;
        } else if (S.iter >= st->maxIter) {
            S.status = RSGM_MAXITER; S.finalIter = st->maxIter;
        }
    }
    RSG_SYNC();
}

// A finished solve's export (CompressibleForwardModel.cpp:72-94): p becomes
// the thermodynamic pressure (pMech keeps the mechanical one; boundary values
// stay mechanical).  Returns the effective viscosity mu(T_0) / rho_0 the
// observables use.
RSG_D double comp_export(Ctx& C, RsgCompDev* Q) {
    RsgMemberDev* M = C.M;
    COMP_CELLS(c) {
        const double p = M->p[c];
        Q->pMech[c] = p;
        M->p[c] = p - (2.0 / 3.0) * Q->rho[c] * dmax(M->k[c], 0.0);
    }
    return comp_visc(Q, Q->T[0]) / dmax(Q->rho[0], 1e-30);
}

// run_member for a compressible member (M->comp): initUniform with the
// compressible fields and BCs, then up to M->iterBudget outer iterations
// (or M->sliceTicks of wall clock), the export and the observables when it
// finishes.
template <int CMAX>
RSG_D void run_comp_member(Thr t, RsgMemberDev* M) {
    Ctx C;
    ctx_init(C, t, M);
    RsgScalars* sc = M->sc;
    RsgCompDev* Q = M->comp;
    if (sc->status != RSGM_RUNNING) return;
    if (!sc->initialized) {
        comp_init_fields(C, Q);
        RSG_SYNC();
        vel_bc(C); p_bc(C); k_bc(C); om_bc(C, 0.075);
        RSG_SYNC();
        comp_bnd(C, Q, 1, 1);
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
        comp_iteration<CMAX>(C, S, Q);
#if !defined(RSG_EMULATE) && !defined(RSG_TEAM)
        if (M->sliceTicks > 0) {
            if (C.t.tid == 0) C.t.red[0] = (wall_clock64() - tStart >= M->sliceTicks) ? 1.0 : 0.0;
            RSG_SYNC();
            const bool stop = C.t.red[0] != 0.0;
            RSG_SYNC();
            if (stop) break;
        }
#endif
    }
    if (S.status != RSGM_RUNNING) {
        RSG_SYNC();
        const double nuEff = comp_export(C, Q);
        RSG_SYNC();
        if (C.t.tid == 0) {
            M->nu = nuEff;
            if (M->obs) observe_member(M);
        }
    }
    if (C.t.tid == 0) {
        sc->iter = S.iter; sc->status = S.status; sc->turbEstablished = S.turbEstablished;
        sc->finalIter = S.finalIter; sc->amgBuiltIter = S.amgBuiltIter; sc->coarseIter = S.coarseIter;
        sc->lastK = S.lastK; sc->lastOm = S.lastOm;
        sc->nLevels = M->amg.lvN[RSG_MAX_LEVELS];
    }
}

#undef COMP_CELLS
#undef COMP_BFACES
