// rsg_resident.h -- register/LDS-resident Krylov solvers and V-cycle for
// members whose cells fit in the workgroup's registers.  Included by
// rsg_member.h after the global-memory solvers, whose helpers it uses.
//
// Thread t owns cells c = t + j*bs, j < CPT.  For the duration of one linear
// solve each thread keeps, in registers, its cells' matrix rows (diagonal and
// the internal-face coefficients in ascending face order, from the mesh's ELL
// image) and every Krylov vector.  The only cross-thread traffic is the
// vector a matrix-vector product gathers, published through one of two LDS
// buffers (double-buffered, so each product costs one workgroup barrier) and
// the block reductions.  Per cell the arithmetic, and per thread the order of
// partial sums, are those of the global-memory solvers above (the row product
// adds diag*x first and then the face terms in ascending face order, as
// ldu_row does), so the results are bit-identical to them.
//
// The AMG V-cycle keeps level 0 in registers and levels >= 1 in LDS vectors;
// the coarse CSR values stay in global memory and the coarsest dense inverse
// lives in LDS for the whole launch.
#pragma once

namespace rsg {

// Matrix rows of the owned cells.
template <int CPT> struct ResRows {
    double dg[CPT];
    double a[CPT][RSG_ELL_K];
    int nb[CPT][RSG_ELL_K];
    int cnt[CPT];
};

RSG_D double* res_buf(Thr& t) {
    double* b = t.xsel ? t.xb1 : t.xb0;
    t.xsel ^= 1;
    return b;
}

template <int CPT> RSG_D void res_load_rows(const Ctx& C, ResRows<CPT>& R) {
    const Thr& t = C.t; const RsgMeshDev* m = C.m; const RsgMemberDev* M = C.M; const int nC = C.nC;
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int c = t.tid + j * t.bs;
        R.dg[j] = 0.0; R.cnt[j] = 0;
#pragma unroll
        for (int k = 0; k < RSG_ELL_K; ++k) { R.a[j][k] = 0.0; R.nb[j][k] = 0; }
        if (c < nC) {
            R.dg[j] = M->diag[c];
            R.cnt[j] = m->ellCnt[c];
#pragma unroll
            for (int k = 0; k < RSG_ELL_K; ++k) {
                const int code = m->ellFace[k * nC + c];
                R.nb[j][k] = m->ellNb[k * nC + c];
                if (code >= 0) R.a[j][k] = (code & 1) ? M->lower[code >> 1] : M->upper[code >> 1];
            }
        }
    }
}

// y = (A x)_c for owned cell j: diag first, then the face terms in face order.
template <int CPT> RSG_D double res_row(const ResRows<CPT>& R, int j, double xc, const double* xs) {
    double y = R.dg[j] * xc;
#pragma unroll
    for (int k = 0; k < RSG_ELL_K; ++k)
        if (k < R.cnt[j]) y += R.a[j][k] * xs[R.nb[j][k]];
    return y;
}

template <int CPT> RSG_D void res_put(const Thr& t, int nC, const double (&v)[CPT], double* xs) {
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int c = t.tid + j * t.bs;
        if (c < nC) xs[c] = v[j];
    }
}

template <int CPT> RSG_D void res_get(const Thr& t, int nC, const double* g, double (&v)[CPT]) {
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int c = t.tid + j * t.bs;
        v[j] = (c < nC) ? g[c] : 0.0;
    }
}

template <int CPT> RSG_D void res_store(const Thr& t, int nC, const double (&v)[CPT], double* g) {
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int c = t.tid + j * t.bs;
        if (c < nC) g[c] = v[j];
    }
}

// BiCGSTAB (the recurrence of bicgstab() above, LinearSolver.hpp:193-245).
template <int CPT> RSG_D SolveRes bicgstab_res(Ctx& C, double* xg, int maxIter, double tol) {
    Thr& t = C.t; const int nC = C.nC; const double* b = C.M->src;
    ResRows<CPT> R;
    res_load_rows(C, R);
    double x[CPT], r[CPT], rh[CPT], p[CPT], v[CPT], s[CPT], tt[CPT], ph[CPT], sh[CPT];
    res_get(t, nC, xg, x);
    double* xs = res_buf(t);
    res_put(t, nC, x, xs);
    RSG_SYNC();
    SolveRes res = {0, 0.0, 0.0, 0};
    double acc[1] = {0.0};
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int c = t.tid + j * t.bs;
        r[j] = rh[j] = p[j] = v[j] = s[j] = tt[j] = ph[j] = sh[j] = 0.0;
        if (c < nC) {
            const double rc = b[c] - res_row(R, j, x[j], xs);
            r[j] = rc; rh[j] = rc;
            acc[0] += rc * rc;
        }
    }
    bsum(t, acc);
    const double r0 = sqrt(acc[0]);
    res.initialRes = r0;
    if (r0 < 1e-30) { res.converged = 1; return res; }
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    double rhoN = acc[0];
    bool updated = false;
    for (int it = 0; it < maxIter; ++it) {
        if (fabs(rhoN) < 1e-30) break;
        const double beta = (rhoN / (rho + 1e-30)) * (alpha / (omega + 1e-30));
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const double pc = r[j] + beta * (p[j] - omega * v[j]);
            p[j] = pc;
            ph[j] = jac(pc, R.dg[j]);
        }
        xs = res_buf(t);
        res_put(t, nC, ph, xs);
        RSG_SYNC();
        double a1[1] = {0.0};
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int c = t.tid + j * t.bs;
            if (c < nC) {
                const double vc = res_row(R, j, ph[j], xs);
                v[j] = vc;
                a1[0] += rh[j] * vc;
            }
        }
        bsum(t, a1);
        const double rv = a1[0];
        if (fabs(rv) < 1e-30) break;
        alpha = rhoN / rv;
        double a2[1] = {0.0};
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int c = t.tid + j * t.bs;
            const double sc = r[j] - alpha * v[j];
            s[j] = sc;
            sh[j] = jac(sc, R.dg[j]);
            if (c < nC) a2[0] += sc * sc;
        }
        xs = res_buf(t);
        res_put(t, nC, sh, xs);       // published by the reduction's barrier
        bsum(t, a2);
        const double sn = sqrt(a2[0]);
        if (sn / r0 < tol) {
#pragma unroll
            for (int j = 0; j < CPT; ++j) x[j] += alpha * ph[j];
            res.iterations = it + 1; res.finalRes = sn / r0; res.initialRes = r0; res.converged = 1;
            updated = true;
            break;
        }
        double a3[2] = {0.0, 0.0};
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int c = t.tid + j * t.bs;
            if (c < nC) {
                const double tc = res_row(R, j, sh[j], xs);
                tt[j] = tc;
                a3[0] += tc * tc;
                a3[1] += tc * s[j];
            }
        }
        bsum(t, a3);
        omega = (a3[0] > 1e-30) ? a3[1] / a3[0] : 0.0;
        double a4[2] = {0.0, 0.0};
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int c = t.tid + j * t.bs;
            x[j] += alpha * ph[j] + omega * sh[j];
            const double rc = s[j] - omega * tt[j];
            r[j] = rc;
            if (c < nC) { a4[0] += rc * rc; a4[1] += rh[j] * rc; }
        }
        updated = true;
        bsum(t, a4);
        const double rn = sqrt(a4[0]);
        res.iterations = it + 1;
        res.finalRes = rn / r0;
        if (res.finalRes < tol) { res.converged = 1; break; }
        if (fabs(omega) < 1e-30) break;
        rho = rhoN;
        rhoN = a4[1];
    }
    if (updated) res_store(t, nC, x, xg);
    return res;
}

// Damped-Jacobi sweep on level 0: xo = xi + wJ (b - A xi) / d (xi gathered from xs).
template <int CPT> RSG_D void res_jacobi0(const Thr& t, int nC, const ResRows<CPT>& R, const double (&b)[CPT],
                                          const double (&xi)[CPT], const double* xs, double (&xo)[CPT]) {
    const double wJ = 2.0 / 3.0;
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int c = t.tid + j * t.bs;
        double o = xi[j];
        if (c < nC) {
            const double s = b[j] - res_row(R, j, xi[j], xs);
            const double d = R.dg[j];
            o = (fabs(d) > 1e-30) ? xi[j] + wJ * s / d : xi[j];
        }
        xo[j] = o;
    }
}

// Damped-Jacobi sweep on a coarse level held in LDS (CSR from global memory).
RSG_D void res_jacobiL(const Thr& t, int n, const int* rp, const int* ci, const double* v, const double* dg,
                       const double* b, const double* xin, double* xout, bool xzero) {
    const double wJ = 2.0 / 3.0;
    for (int i = t.tid; i < n; i += t.bs) {
        double s = b[i];
        if (!xzero)
            for (int e = rp[i]; e < rp[i + 1]; ++e) s -= v[e] * xin[ci[e]];
        const double d = dg[i];
        const double x0 = xzero ? 0.0 : xin[i];
        xout[i] = (fabs(d) > 1e-30) ? x0 + wJ * s / d : x0;
    }
}

// Where the V-cycle reads the coarse ELL images from.  When the kernel
// provides LDS room (one member per CU), the images of the deepest levels
// that fit are copied there at the start of each solve (levels are appended
// in order, so they form one contiguous tail of the global arrays); the
// other levels read global memory.
#ifdef RSG_EMULATE
static int rsg_emu_staged_first = 0;
#endif
// LDS copies hold column and row indices as 16-bit integers (indices within a
// level are below RSG_RES_MAXN, and the -1 pad stays -1), so the deepest
// levels of the largest resident members fit entirely.
static_assert(RSG_RES_MAXN < 32768 && RSG_RES_MAXC < 32768, "16-bit LDS image indices");
struct CoarseImg {
    const int* gCol; const double* gVal; const int* gAgg;       // global images
    const short* sCol; const double* sVal; const short* sAgg;   // LDS copies
    int eFirst, eBase, aFirst, aBase;                            // first staged level, its offset
};

// An index array of one level, in global memory (int) or LDS (short).
struct IdxRef {
    const int* g; const short* s;
    RSG_DM int operator[](int i) const { return s ? (int)s[i] : g[i]; }
};

RSG_D IdxRef img_col(const CoarseImg& G, const RsgAmgDev& A, int l) {
    IdxRef r;
    if (l >= G.eFirst) { r.g = nullptr; r.s = G.sCol + (A.lvE[l] - G.eBase); }
    else { r.g = G.gCol + A.lvE[l]; r.s = nullptr; }
    return r;
}
RSG_D const double* img_val(const CoarseImg& G, const RsgAmgDev& A, int l) {
    return l >= G.eFirst ? G.sVal + (A.lvE[l] - G.eBase) : G.gVal + A.lvE[l];
}
RSG_D IdxRef img_agg(const CoarseImg& G, const RsgAmgDev& A, int l) {
    IdxRef r;
    if (l >= G.aFirst) { r.g = nullptr; r.s = G.sAgg + (A.lvAE[l] - G.aBase); }
    else { r.g = G.gAgg + A.lvAE[l]; r.s = nullptr; }
    return r;
}

RSG_D CoarseImg coarse_images(Thr& t, const RsgAmgDev& A) {
    CoarseImg g;
    g.gCol = A.eCol; g.gVal = A.eVal; g.gAgg = A.aCol;
    g.sCol = t.ceCol; g.sVal = t.ceVal; g.sAgg = t.caCol;
    g.eFirst = g.aFirst = RSG_MAX_LEVELS + 1;   // nothing staged
    g.eBase = g.aBase = 0;
    if (!t.ceVal) return g;
    const int L = A.lvN[RSG_MAX_LEVELS];
    for (int l = L - 1; l >= 1 && A.lvK[l] > 0 && A.used[0] - A.lvE[l] <= t.ceCap; --l) g.eFirst = l;
    for (int l = L - 1; l >= 1 && A.lvAK[l] > 0 && A.used[1] - A.lvAE[l] <= t.caCap; --l) g.aFirst = l;
#ifdef RSG_EMULATE
    rsg_emu_staged_first = g.eFirst;   // emulation build only: observed by the tests
#endif
    if (g.eFirst <= RSG_MAX_LEVELS) {
        g.eBase = A.lvE[g.eFirst];
        for (int i = t.tid; i < A.used[0] - g.eBase; i += t.bs) {
            t.ceVal[i] = A.eVal[g.eBase + i];
            t.ceCol[i] = (short)A.eCol[g.eBase + i];
        }
    }
    if (g.aFirst <= RSG_MAX_LEVELS) {
        g.aBase = A.lvAE[g.aFirst];
        for (int i = t.tid; i < A.used[1] - g.aBase; i += t.bs) t.caCol[i] = (short)A.aCol[g.aBase + i];
    }
    RSG_SYNC();
    return g;
}

// The same sweep on a level's ELL image: entries in CSR order, so the sums
// round exactly like res_jacobiL's; column -1 pads.
RSG_D void res_jacobiE(const Thr& t, int n, int K, IdxRef ec, const double* ev, const double* dg,
                       const double* b, const double* xin, double* xout, bool xzero) {
    const double wJ = 2.0 / 3.0;
    for (int i = t.tid; i < n; i += t.bs) {
        double s = b[i];
        if (!xzero) {
#pragma unroll 4
            for (int k = 0; k < K; ++k) {
                const int col = ec[k * n + i];
                const double a = ev[k * n + i];
                if (col >= 0) s -= a * xin[col];
            }
        }
        const double d = dg[i];
        const double x0 = xzero ? 0.0 : xin[i];
        xout[i] = (fabs(d) > 1e-30) ? x0 + wJ * s / d : x0;
    }
}

// Coarse-level sweep and restriction through the ELL images when present.
RSG_D void res_sweepL(const Thr& t, const RsgAmgDev& A, const CoarseImg& G, int l, const double* b,
                      const double* xin, double* xout, bool xzero) {
    const int n = A.lvN[l];
    const double* dg = A.dg + A.lvRow[l];
    if (A.lvK[l] > 0)
        res_jacobiE(t, n, A.lvK[l], img_col(G, A, l), img_val(G, A, l), dg, b, xin, xout, xzero);
    else
        res_jacobiL(t, n, A.rp + A.lvRp[l], A.ci + A.lvNz[l], A.v + A.lvNz[l], dg, b, xin, xout, xzero);
}

// b of level lc = sum over each aggregate's fine rows (restriction order) of r.
RSG_D void res_restrict(const Thr& t, const RsgAmgDev& A, const CoarseImg& G, int lc, const double* r,
                        double* bc) {
    const int nc = A.lvN[lc];
    if (A.lvAK[lc] > 0) {
        const int AK = A.lvAK[lc];
        const IdxRef ac = img_agg(G, A, lc);
        for (int I = t.tid; I < nc; I += t.bs) {
            double s = 0.0;
#pragma unroll 4
            for (int q = 0; q < AK; ++q) {
                const int i = ac[q * nc + I];
                if (i >= 0) s += r[i];
            }
            bc[I] = s;
        }
    } else {
        const int* aP = A.aggPtr + A.lvRp[lc];
        const int* aI = A.aggIdx + A.lvRow[lc - 1];
        for (int I = t.tid; I < nc; I += t.bs) {
            double s = 0.0;
            for (int q = aP[I]; q < aP[I + 1]; ++q) s += r[aI[q]];
            bc[I] = s;
        }
    }
}

// Residual row i of level l at x: b_i - sum of the row entries times x in
// CSR order (the ELL image holds the entries in that order).
RSG_D double res_row_residual(const RsgAmgDev& A, const CoarseImg& G, int l, int i, const double* b,
                              const double* x) {
    double s = 0.0;
    if (A.lvK[l] > 0) {
        const int n = A.lvN[l], K = A.lvK[l];
        const IdxRef ec = img_col(G, A, l);
        const double* ev = img_val(G, A, l);
#pragma unroll 4
        for (int k = 0; k < K; ++k) {
            const int col = ec[k * n + i];
            const double a = ev[k * n + i];
            if (col >= 0) s += a * x[col];
        }
    } else {
        const int* rp = A.rp + A.lvRp[l]; const int* ci = A.ci + A.lvNz[l]; const double* v = A.v + A.lvNz[l];
        for (int e = rp[i]; e < rp[i + 1]; ++e) s += v[e] * x[ci[e]];
    }
    return b[i] - s;
}

// b of level l+1 straight from x of level l: each coarse row sums the
// residuals of its aggregate's fine rows (restriction order) -- a residual
// pass, a barrier and res_restrict with the same roundings, one barrier fewer.
RSG_D void res_restrict_residual(const Thr& t, const RsgAmgDev& A, const CoarseImg& G, int l, const double* b,
                                 const double* x, double* bc) {
    const int lc = l + 1, nc = A.lvN[lc];
    if (A.lvAK[lc] > 0) {
        const int AK = A.lvAK[lc];
        const IdxRef ac = img_agg(G, A, lc);
        for (int I = t.tid; I < nc; I += t.bs) {
            double s = 0.0;
            for (int q = 0; q < AK; ++q) {
                const int i = ac[q * nc + I];
                if (i >= 0) s += res_row_residual(A, G, l, i, b, x);
            }
            bc[I] = s;
        }
    } else {
        const int* aP = A.aggPtr + A.lvRp[lc];
        const int* aI = A.aggIdx + A.lvRow[l];
        for (int I = t.tid; I < nc; I += t.bs) {
            double s = 0.0;
            for (int q = aP[I]; q < aP[I + 1]; ++q) s += res_row_residual(A, G, l, aI[q], b, x);
            bc[I] = s;
        }
    }
}

// First post-smoothing sweep of level l with the prolongation folded in:
// every x value is read as x[j] + xc[map[j]] (the prolonged value, rounded as
// the in-place update x[j] += xc[map[j]] rounds it), so the separate
// prolongation pass and its barrier go away.  xout = xp + wJ (b - A xp) / d.
RSG_D void res_sweep_prolong(const Thr& t, const RsgAmgDev& A, const CoarseImg& G, int l, const double* b,
                             const double* x, const double* xc, const int* map, double* xout) {
    const double wJ = 2.0 / 3.0;
    const int n = A.lvN[l];
    const double* dg = A.dg + A.lvRow[l];
    if (A.lvK[l] > 0) {
        const int K = A.lvK[l];
        const IdxRef ec = img_col(G, A, l);
        const double* ev = img_val(G, A, l);
        for (int i = t.tid; i < n; i += t.bs) {
            double s = b[i];
#pragma unroll 4
            for (int k = 0; k < K; ++k) {
                const int col = ec[k * n + i];
                const double a = ev[k * n + i];
                if (col >= 0) s -= a * (x[col] + xc[map[col]]);
            }
            const double d = dg[i];
            const double x0 = x[i] + xc[map[i]];
            xout[i] = (fabs(d) > 1e-30) ? x0 + wJ * s / d : x0;
        }
    } else {
        const int* rp = A.rp + A.lvRp[l]; const int* ci = A.ci + A.lvNz[l]; const double* v = A.v + A.lvNz[l];
        for (int i = t.tid; i < n; i += t.bs) {
            double s = b[i];
            for (int e = rp[i]; e < rp[i + 1]; ++e) s -= v[e] * (x[ci[e]] + xc[map[ci[e]]]);
            const double d = dg[i];
            const double x0 = x[i] + xc[map[i]];
            xout[i] = (fabs(d) > 1e-30) ? x0 + wJ * s / d : x0;
        }
    }
}

// z = V-cycle(rin): the cycle of amg_vcycle() with level 0 in registers.
template <int CPT> RSG_D void vcycle_res(Ctx& C, const ResRows<CPT>& R, const CoarseImg& G,
                                         const double (&rin)[CPT], double (&z)[CPT]) {
    Thr& t = C.t; RsgMemberDev* M = C.M; RsgAmgDev& A = M->amg; const int nC = C.nC;
    const int L = A.lvN[RSG_MAX_LEVELS];
    const double wJ = 2.0 / 3.0;
    const int base = A.lvRow[1];                  // first coarse row in the level arrays
    // Coarse x (rows base..) lives in the LDS coarse-vector area; coarse b in
    // the global level arrays (each b row is written and read by the thread
    // that owns it, except at the coarsest level, which reads it after a
    // barrier); the two exchange buffers, named explicitly here, carry level
    // 0's gathered vectors and, while the coarse levels run, the coarse r.
    // Every write to a buffer follows a barrier that comes after its last reads.
    double* cx = t.cvec;
    double* cb = A.b + base;
    double* B0 = t.xb0;
    double* B1 = t.xb1;
    double* cr = B1;
    double x0[CPT], tmp[CPT];
    if (L == 1) {
        // level 0 is the coarsest: dense inverse (or Jacobi sweeps) on the whole vector
        double* xs = res_buf(t);
        res_put(t, nC, rin, xs);
        RSG_SYNC();
        if (M->sc->coarseDense) {
            const int DN = RSG_COARSE_DENSE;
#pragma unroll
            for (int j = 0; j < CPT; ++j) {
                const int c = t.tid + j * t.bs;
                double s = 0.0;
                if (c < nC)
                    for (int q = 0; q < nC; ++q) s += t.lds[c * DN + q] * xs[q];
                z[j] = s;
            }
        } else {
#pragma unroll
            for (int j = 0; j < CPT; ++j) z[j] = (fabs(R.dg[j]) > 1e-30) ? wJ * rin[j] / R.dg[j] : 0.0;
            for (int sweep = 0; sweep < 50; ++sweep) {
                xs = res_buf(t);
                res_put(t, nC, z, xs);
                RSG_SYNC();
                res_jacobi0(t, nC, R, rin, z, xs, tmp);
#pragma unroll
                for (int j = 0; j < CPT; ++j) z[j] = tmp[j];
            }
        }
        return;
    }
    // level 0 down: two sweeps from zero, residual, restriction
#pragma unroll
    for (int j = 0; j < CPT; ++j) x0[j] = (fabs(R.dg[j]) > 1e-30) ? wJ * rin[j] / R.dg[j] : 0.0;
    res_put(t, nC, x0, B0);
    RSG_SYNC();
    res_jacobi0(t, nC, R, rin, x0, B0, tmp);
    res_put(t, nC, tmp, B1);
    RSG_SYNC();
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int c = t.tid + j * t.bs;
        x0[j] = tmp[j];
        if (c < nC) B0[c] = rin[j] - res_row(R, j, tmp[j], B1);
    }
    RSG_SYNC();
    res_restrict(t, A, G, 1, B0, cb + (A.lvRow[1] - base));
    RSG_TMARK(19);
    // coarse levels down (row I of a level is owned by thread I % bs throughout)
    for (int l = 1; l + 1 < L; ++l) {
        const int off = A.lvRow[l] - base;
        double* x = cx + off; double* bb = cb + off; double* r = cr + off;
        res_sweepL(t, A, G, l, bb, x, r, true);
        RSG_SYNC();
        RSG_TMARK(l == 1 ? 24 : 29);
        res_sweepL(t, A, G, l, bb, r, x, false);
        RSG_SYNC();
        RSG_TMARK(l == 1 ? 25 : 30);
        res_restrict_residual(t, A, G, l, bb, x, cb + (A.lvRow[l + 1] - base));
        RSG_TMARK(l == 1 ? 26 : 31);
    }
    RSG_SYNC();
    RSG_TMARK(20);
    {   // coarsest level
        const int l = L - 1, n = A.lvN[l];
        const int off = A.lvRow[l] - base;
        double* x = cx + off; double* bb = cb + off;
        if (M->sc->coarseDense) {
            const int DN = RSG_COARSE_DENSE;
            for (int i = t.tid; i < n; i += t.bs) {
                double s = 0.0;
                for (int j = 0; j < n; ++j) s += t.lds[i * DN + j] * bb[j];
                x[i] = s;
            }
        } else {
            const int* rp = A.rp + A.lvRp[l]; const int* ci = A.ci + A.lvNz[l];
            const double* v = A.v + A.lvNz[l]; const double* dg = A.dg + A.lvRow[l];
            double* r = cr + off;
            res_jacobiL(t, n, rp, ci, v, dg, bb, x, x, true);
            RSG_SYNC();
            for (int s = 0; s < 25; ++s) {
                res_jacobiL(t, n, rp, ci, v, dg, bb, x, r, false);
                RSG_SYNC();
                res_jacobiL(t, n, rp, ci, v, dg, bb, r, x, false);
                RSG_SYNC();
            }
        }
        RSG_SYNC();
    }
    RSG_TMARK(21);
    // coarse levels up
    for (int l = L - 2; l >= 1; --l) {
        const int n = A.lvN[l];
        const int off = A.lvRow[l] - base;
        double* x = cx + off; double* bb = cb + off; double* r = cr + off;
        const int* map = A.map + A.lvRow[l];
        const double* xc = cx + (A.lvRow[l + 1] - base);
        (void)n;
        res_sweep_prolong(t, A, G, l, bb, x, xc, map, r);
        RSG_SYNC();
        RSG_TMARK(l == 1 ? 27 : 14);
        res_sweepL(t, A, G, l, bb, r, x, false);
        RSG_SYNC();
        RSG_TMARK(l == 1 ? 28 : 15);
    }
    RSG_TMARK(22);
    // level 0 up: prolongation and two sweeps
    {
        const int* map = A.map + A.lvRow[0];
        const double* xc = cx + (A.lvRow[1] - base);
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int c = t.tid + j * t.bs;
            if (c < nC) x0[j] += xc[map[c]];
        }
    }
    res_put(t, nC, x0, B0);
    RSG_SYNC();
    res_jacobi0(t, nC, R, rin, x0, B0, tmp);
    res_put(t, nC, tmp, B1);
    RSG_SYNC();
    res_jacobi0(t, nC, R, rin, tmp, B1, z);
    RSG_TMARK(23);
}

// PCG with Jacobi or V-cycle preconditioning (the recurrence of pcg() above).
template <int CPT> RSG_D SolveRes pcg_res(Ctx& C, double* xg, int maxIter, double tol, bool useAmg) {
    Thr& t = C.t; const int nC = C.nC; const double* b = C.M->src;
    ResRows<CPT> R;
    res_load_rows(C, R);
    double x[CPT], r[CPT], z[CPT], p[CPT], q[CPT];
    res_get(t, nC, xg, x);
    double* xs = res_buf(t);
    res_put(t, nC, x, xs);
    RSG_SYNC();
    SolveRes res = {0, 0.0, 0.0, 0};
    double acc[1] = {0.0};
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int c = t.tid + j * t.bs;
        r[j] = 0.0;
        if (c < nC) {
            const double rc = b[c] - res_row(R, j, x[j], xs);
            r[j] = rc;
            acc[0] += rc * rc;
        }
    }
    bsum(t, acc);
    const double r0 = sqrt(acc[0]);
    res.initialRes = r0;
    if (r0 < 1e-30) { res.converged = 1; return res; }
    double rzv[1] = {0.0};
    CoarseImg G;
    if (useAmg) G = coarse_images(t, C.M->amg);
    if (useAmg) {
        vcycle_res(C, R, G, r, z);
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int c = t.tid + j * t.bs;
            p[j] = z[j];
            if (c < nC) rzv[0] += r[j] * z[j];
        }
    } else {
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int c = t.tid + j * t.bs;
            const double zc = jac(r[j], R.dg[j]);
            z[j] = zc; p[j] = zc;
            if (c < nC) rzv[0] += r[j] * zc;
        }
    }
    bsum(t, rzv);
    double rz = rzv[0];
    bool updated = false;
    for (int it = 0; it < maxIter; ++it) {
        xs = res_buf(t);
        res_put(t, nC, p, xs);
        RSG_SYNC();
        double a1[1] = {0.0};
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int c = t.tid + j * t.bs;
            q[j] = 0.0;
            if (c < nC) {
                const double qc = res_row(R, j, p[j], xs);
                q[j] = qc;
                a1[0] += p[j] * qc;
            }
        }
        bsum(t, a1);
        RSG_TMARK(16);
        const double pAp = a1[0];
        if (fabs(pAp) < 1e-30) break;
        const double alpha = rz / pAp;
        double a2[1] = {0.0};
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int c = t.tid + j * t.bs;
            x[j] += alpha * p[j];
            const double rc = r[j] - alpha * q[j];
            r[j] = rc;
            if (c < nC) a2[0] += rc * rc;
        }
        updated = true;
        bsum(t, a2);
        RSG_TMARK(17);
        const double rn = sqrt(a2[0]);
        res.iterations = it + 1;
        res.finalRes = rn / r0;
        if (res.finalRes < tol) { res.converged = 1; break; }
        double a3[1] = {0.0};
        if (useAmg) {
            vcycle_res(C, R, G, r, z);
#pragma unroll
            for (int j = 0; j < CPT; ++j) {
                const int c = t.tid + j * t.bs;
                if (c < nC) a3[0] += r[j] * z[j];
            }
        } else {
#pragma unroll
            for (int j = 0; j < CPT; ++j) {
                const int c = t.tid + j * t.bs;
                const double zc = jac(r[j], R.dg[j]);
                z[j] = zc;
                if (c < nC) a3[0] += r[j] * zc;
            }
        }
        bsum(t, a3);
        RSG_TMARK(18);
        const double rzNew = a3[0];
        const double beta = rzNew / (rz + 1e-30);
#pragma unroll
        for (int j = 0; j < CPT; ++j) p[j] = z[j] + beta * p[j];
        rz = rzNew;
    }
    if (updated) res_store(t, nC, x, xg);
    return res;
}

// Whether the resident path applies to this member's system.
RSG_D bool res_fits(const Ctx& C, int cptMax) {
    return C.m->ellK > 0 && C.nC <= cptMax * C.t.bs && C.nC <= RSG_RES_MAXN && C.t.xb0 != nullptr;
}

// Whether the V-cycle's coarse vectors fit the LDS area.
RSG_D bool res_amg_fits(const Ctx& C) {
    const RsgAmgDev& A = C.M->amg;
    const int L = A.lvN[RSG_MAX_LEVELS];
    if (L <= 1) return true;
    const int rows = A.lvRow[L - 1] + A.lvN[L - 1] - A.lvRow[1];
    return rows <= C.t.cvecCap && rows <= RSG_RES_MAXN;   // coarse x, b in cvec; coarse r in xb1
}

template <int CPT> RSG_D SolveRes lsolve_cpt(Ctx& C, int code, double* x, bool amg, int maxIter, double tol) {
    if (code == 2) return bicgstab_res<CPT>(C, x, maxIter, tol);
    return pcg_res<CPT>(C, x, maxIter, tol, amg);
}

// Cells per thread rounded up to an instantiated width (CMAX: the widest one
// the kernel is built with; it bounds the kernel's register allocation).
template <int CMAX> RSG_D SolveRes lsolve_res(Ctx& C, int code, double* x, bool amg, int maxIter, double tol) {
#ifdef RSG_EMULATE
    return lsolve_cpt<RSG_RES_MAXN>(C, code, x, amg, maxIter, tol);
#else
    const int cpt = (C.nC + C.t.bs - 1) / C.t.bs;
    if constexpr (CMAX >= 2) { if (cpt == 2) return lsolve_cpt<2>(C, code, x, amg, maxIter, tol); }
    if constexpr (CMAX >= 3) { if (cpt == 3) return lsolve_cpt<3>(C, code, x, amg, maxIter, tol); }
    if constexpr (CMAX >= 4) { if (cpt == 4) return lsolve_cpt<4>(C, code, x, amg, maxIter, tol); }
    if constexpr (CMAX >= 6) { if (cpt == 5 || cpt == 6) return lsolve_cpt<6>(C, code, x, amg, maxIter, tol); }
    return lsolve_cpt<1>(C, code, x, amg, maxIter, tol);
#endif
}

}  // namespace rsg
