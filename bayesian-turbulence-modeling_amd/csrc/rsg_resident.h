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
//
// Every solver body is a template on its arithmetic type T: double for FP64
// members, float for the inner solves of mixed-precision members (matrix rows
// stay FP64 in registers and are rounded to FP32 as they are used; the
// exchange buffers and coarse-level vectors hold floats; reductions
// accumulate in FP32).  lsolve_mixed_res wraps the FP32 solves in FP64
// iterative refinement.
#pragma once

namespace rsg {

// Matrix rows of the owned cells.
template <int CPT> struct ResRows {
    double dg[CPT];
    double ex[CPT];          // row sum dg + sum of the off-diagonals (FP64), for the FP32 row form
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
        double e = R.dg[j];
#pragma unroll
        for (int k = 0; k < RSG_ELL_K; ++k) e += R.a[j][k];   // padding slots hold 0
        R.ex[j] = e;
    }
}

// y = (A x)_c for owned cell j: diag first, then the face terms in face order.
// The FP32 form is the same product written as ex x_c + sum a (x_nb - x_c),
// with ex the row sum taken in FP64: rounding the coefficients to FP32 then
// perturbs each row by eps32 |a| |x_nb - x_c| instead of eps32 |diag| |x_c|,
// so the near-null space of the nearly singular p' Laplacian (constant
// vectors) survives the rounding and the FP32 solves do not lose the smooth
// error components that FP64 refinement would otherwise have to recover.
template <class T, int CPT> RSG_D T res_row(const ResRows<CPT>& R, int j, T xc, const T* xs) {
    if constexpr (sizeof(T) == sizeof(float)) {
        T y = (T)R.ex[j] * xc;
#pragma unroll
        for (int k = 0; k < RSG_ELL_K; ++k)
            if (k < R.cnt[j]) y += (T)R.a[j][k] * (xs[R.nb[j][k]] - xc);
        return y;
    } else {
        T y = (T)R.dg[j] * xc;
#pragma unroll
        for (int k = 0; k < RSG_ELL_K; ++k)
            if (k < R.cnt[j]) y += (T)R.a[j][k] * xs[R.nb[j][k]];
        return y;
    }
}

template <class T, int CPT> RSG_D void res_put(const Thr& t, int nC, const T (&v)[CPT], T* xs) {
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

// An exchange buffer viewed as T (a double buffer holds twice as many floats).
template <class T> RSG_D T* res_bufT(Thr& t) { return reinterpret_cast<T*>(res_buf(t)); }

// BiCGSTAB (the recurrence of bicgstab() above, LinearSolver.hpp:193-245)
// from the residual r at x (rr = r.r block-summed, held by every thread):
// x and r in/out in registers; `updated` reports whether x changed.
template <int CPT, class T>
RSG_D SolveRes bicgstab_res_from(Ctx& C, const ResRows<CPT>& R, T (&x)[CPT], T (&r)[CPT], T rr, int maxIter,
                                 double tol, bool& updated) {
    Thr& t = C.t; const int nC = C.nC;
    T rh[CPT], p[CPT], v[CPT], s[CPT], tt[CPT], ph[CPT], sh[CPT];
    SolveRes res = {0, 0.0, 0.0, 0};
    updated = false;
    const T r0 = sqrt(rr);
    res.initialRes = r0;
    if (r0 < T(1e-30)) { res.converged = 1; return res; }
#pragma unroll
    for (int j = 0; j < CPT; ++j) { rh[j] = r[j]; p[j] = v[j] = s[j] = tt[j] = ph[j] = sh[j] = T(0); }
    T rho = T(1), alpha = T(1), omega = T(1);
    T rhoN = rr;
    T* xs;
    for (int it = 0; it < maxIter; ++it) {
        if (fabs(rhoN) < T(1e-30)) break;
        const T beta = (rhoN / (rho + T(1e-30))) * (alpha / (omega + T(1e-30)));
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const T pc = r[j] + beta * (p[j] - omega * v[j]);
            p[j] = pc;
            ph[j] = jac<T>(pc, (T)R.dg[j]);
        }
        xs = res_bufT<T>(t);
        res_put<T>(t, nC, ph, xs);
        RSG_SYNC();
        T a1[1] = {T(0)};
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int c = t.tid + j * t.bs;
            if (c < nC) {
                const T vc = res_row<T>(R, j, ph[j], xs);
                v[j] = vc;
                a1[0] += rh[j] * vc;
            }
        }
        bsum(t, a1);
        const T rv = a1[0];
        if (fabs(rv) < T(1e-30)) break;
        alpha = rhoN / rv;
        T a2[1] = {T(0)};
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int c = t.tid + j * t.bs;
            const T sc = r[j] - alpha * v[j];
            s[j] = sc;
            sh[j] = jac<T>(sc, (T)R.dg[j]);
            if (c < nC) a2[0] += sc * sc;
        }
        xs = res_bufT<T>(t);
        res_put<T>(t, nC, sh, xs);       // published by the reduction's barrier
        bsum(t, a2);
        const T sn = sqrt(a2[0]);
        if (sn / r0 < tol) {
#pragma unroll
            for (int j = 0; j < CPT; ++j) x[j] += alpha * ph[j];
            res.iterations = it + 1; res.finalRes = sn / r0; res.initialRes = r0; res.converged = 1;
            updated = true;
            break;
        }
        T a3[2] = {T(0), T(0)};
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int c = t.tid + j * t.bs;
            if (c < nC) {
                const T tc = res_row<T>(R, j, sh[j], xs);
                tt[j] = tc;
                a3[0] += tc * tc;
                a3[1] += tc * s[j];
            }
        }
        bsum(t, a3);
        omega = (a3[0] > T(1e-30)) ? a3[1] / a3[0] : T(0);
        T a4[2] = {T(0), T(0)};
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int c = t.tid + j * t.bs;
            x[j] += alpha * ph[j] + omega * sh[j];
            const T rc = s[j] - omega * tt[j];
            r[j] = rc;
            if (c < nC) { a4[0] += rc * rc; a4[1] += rh[j] * rc; }
        }
        updated = true;
        bsum(t, a4);
        const T rn = sqrt(a4[0]);
        res.iterations = it + 1;
        res.finalRes = rn / r0;
        if (res.finalRes < tol) { res.converged = 1; break; }
        if (fabs(omega) < T(1e-30)) break;
        rho = rhoN;
        rhoN = a4[1];
    }
    return res;
}

// FP64 defect of the owned cells: r = b - A x (x published through an
// exchange buffer), returns r.r block-summed.
template <int CPT> RSG_D double res_defect(Ctx& C, const ResRows<CPT>& R, const double (&x)[CPT],
                                           double (&r)[CPT]) {
    Thr& t = C.t; const int nC = C.nC; const double* b = C.M->src;
    double* xs = res_buf(t);
    res_put<double>(t, nC, x, xs);
    RSG_SYNC();
    double acc[1] = {0.0};
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int c = t.tid + j * t.bs;
        r[j] = 0.0;
        if (c < nC) {
            const double rc = b[c] - res_row<double>(R, j, x[j], xs);
            r[j] = rc;
            acc[0] += rc * rc;
        }
    }
    bsum(t, acc);
    return acc[0];
}

template <int CPT> RSG_D SolveRes bicgstab_res(Ctx& C, double* xg, int maxIter, double tol) {
    Thr& t = C.t; const int nC = C.nC;
    ResRows<CPT> R;
    res_load_rows(C, R);
    double x[CPT], r[CPT];
    res_get(t, nC, xg, x);
    const double rr = res_defect(C, R, x, r);
    bool updated;
    const SolveRes res = bicgstab_res_from<CPT, double>(C, R, x, r, rr, maxIter, tol, updated);
    if (updated) res_store(t, nC, x, xg);
    return res;
}

// Damped-Jacobi sweep on level 0: xo = xi + wJ (b - A xi) / d (xi gathered from xs).
template <int CPT, class T>
RSG_D void res_jacobi0(const Thr& t, int nC, const ResRows<CPT>& R, const T (&b)[CPT], const T (&xi)[CPT],
                       const T* xs, T (&xo)[CPT]) {
    const T wJ = T(2.0 / 3.0);
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int c = t.tid + j * t.bs;
        T o = xi[j];
        if (c < nC) {
            const T s = b[j] - res_row<T>(R, j, xi[j], xs);
            const T d = (T)R.dg[j];
            o = (fabs(d) > T(1e-30)) ? xi[j] + wJ * s / d : xi[j];
        }
        xo[j] = o;
    }
}

// Damped-Jacobi sweep on a coarse level held in LDS (CSR from global memory).
template <class T>
RSG_D void res_jacobiL(const Thr& t, int n, const int* rp, const int* ci, const double* v, const double* dg,
                       const T* b, const T* xin, T* xout, bool xzero) {
    const T wJ = T(2.0 / 3.0);
    for (int i = t.tid; i < n; i += t.bs) {
        T s = b[i];
        if (!xzero)
            for (int e = rp[i]; e < rp[i + 1]; ++e) s -= (T)v[e] * xin[ci[e]];
        const T d = (T)dg[i];
        const T x0 = xzero ? T(0) : xin[i];
        xout[i] = (fabs(d) > T(1e-30)) ? x0 + wJ * s / d : x0;
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
template <class T>
RSG_D void res_jacobiE(const Thr& t, int n, int K, IdxRef ec, const double* ev, const double* dg,
                       const T* b, const T* xin, T* xout, bool xzero) {
    const T wJ = T(2.0 / 3.0);
    for (int i = t.tid; i < n; i += t.bs) {
        T s = b[i];
        if (!xzero) {
#pragma unroll 4
            for (int k = 0; k < K; ++k) {
                const int col = ec[k * n + i];
                const T a = (T)ev[k * n + i];
                if (col >= 0) s -= a * xin[col];
            }
        }
        const T d = (T)dg[i];
        const T x0 = xzero ? T(0) : xin[i];
        xout[i] = (fabs(d) > T(1e-30)) ? x0 + wJ * s / d : x0;
    }
}

// Coarse-level sweep and restriction through the ELL images when present.
template <class T>
RSG_D void res_sweepL(const Thr& t, const RsgAmgDev& A, const CoarseImg& G, int l, const T* b, const T* xin,
                      T* xout, bool xzero) {
    const int n = A.lvN[l];
    const double* dg = A.dg + A.lvRow[l];
    if (A.lvK[l] > 0)
        res_jacobiE<T>(t, n, A.lvK[l], img_col(G, A, l), img_val(G, A, l), dg, b, xin, xout, xzero);
    else
        res_jacobiL<T>(t, n, A.rp + A.lvRp[l], A.ci + A.lvNz[l], A.v + A.lvNz[l], dg, b, xin, xout, xzero);
}

// b of level lc = sum over each aggregate's fine rows (restriction order) of r.
template <class T>
RSG_D void res_restrict(const Thr& t, const RsgAmgDev& A, const CoarseImg& G, int lc, const T* r, T* bc) {
    const int nc = A.lvN[lc];
    if (A.lvAK[lc] > 0) {
        const int AK = A.lvAK[lc];
        const IdxRef ac = img_agg(G, A, lc);
        for (int I = t.tid; I < nc; I += t.bs) {
            T s = T(0);
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
            T s = T(0);
            for (int q = aP[I]; q < aP[I + 1]; ++q) s += r[aI[q]];
            bc[I] = s;
        }
    }
}

// Residual row i of level l at x: b_i - sum of the row entries times x in
// CSR order (the ELL image holds the entries in that order).
template <class T>
RSG_D T res_row_residual(const RsgAmgDev& A, const CoarseImg& G, int l, int i, const T* b, const T* x) {
    T s = T(0);
    if (A.lvK[l] > 0) {
        const int n = A.lvN[l], K = A.lvK[l];
        const IdxRef ec = img_col(G, A, l);
        const double* ev = img_val(G, A, l);
#pragma unroll 4
        for (int k = 0; k < K; ++k) {
            const int col = ec[k * n + i];
            const T a = (T)ev[k * n + i];
            if (col >= 0) s += a * x[col];
        }
    } else {
        const int* rp = A.rp + A.lvRp[l]; const int* ci = A.ci + A.lvNz[l]; const double* v = A.v + A.lvNz[l];
        for (int e = rp[i]; e < rp[i + 1]; ++e) s += (T)v[e] * x[ci[e]];
    }
    return b[i] - s;
}

// b of level l+1 straight from x of level l: each coarse row sums the
// residuals of its aggregate's fine rows (restriction order) -- a residual
// pass, a barrier and res_restrict with the same roundings, one barrier fewer.
template <class T>
RSG_D void res_restrict_residual(const Thr& t, const RsgAmgDev& A, const CoarseImg& G, int l, const T* b,
                                 const T* x, T* bc) {
    const int lc = l + 1, nc = A.lvN[lc];
    if (A.lvAK[lc] > 0) {
        const int AK = A.lvAK[lc];
        const IdxRef ac = img_agg(G, A, lc);
        for (int I = t.tid; I < nc; I += t.bs) {
            T s = T(0);
            for (int q = 0; q < AK; ++q) {
                const int i = ac[q * nc + I];
                if (i >= 0) s += res_row_residual<T>(A, G, l, i, b, x);
            }
            bc[I] = s;
        }
    } else {
        const int* aP = A.aggPtr + A.lvRp[lc];
        const int* aI = A.aggIdx + A.lvRow[l];
        for (int I = t.tid; I < nc; I += t.bs) {
            T s = T(0);
            for (int q = aP[I]; q < aP[I + 1]; ++q) s += res_row_residual<T>(A, G, l, aI[q], b, x);
            bc[I] = s;
        }
    }
}

// First post-smoothing sweep of level l with the prolongation folded in:
// every x value is read as x[j] + xc[map[j]] (the prolonged value, rounded as
// the in-place update x[j] += xc[map[j]] rounds it), so the separate
// prolongation pass and its barrier go away.  xout = xp + wJ (b - A xp) / d.
template <class T>
RSG_D void res_sweep_prolong(const Thr& t, const RsgAmgDev& A, const CoarseImg& G, int l, const T* b, const T* x,
                             const T* xc, const int* map, T* xout) {
    const T wJ = T(2.0 / 3.0);
    const int n = A.lvN[l];
    const double* dg = A.dg + A.lvRow[l];
    if (A.lvK[l] > 0) {
        const int K = A.lvK[l];
        const IdxRef ec = img_col(G, A, l);
        const double* ev = img_val(G, A, l);
        for (int i = t.tid; i < n; i += t.bs) {
            T s = b[i];
#pragma unroll 4
            for (int k = 0; k < K; ++k) {
                const int col = ec[k * n + i];
                const T a = (T)ev[k * n + i];
                if (col >= 0) s -= a * (x[col] + xc[map[col]]);
            }
            const T d = (T)dg[i];
            const T x0 = x[i] + xc[map[i]];
            xout[i] = (fabs(d) > T(1e-30)) ? x0 + wJ * s / d : x0;
        }
    } else {
        const int* rp = A.rp + A.lvRp[l]; const int* ci = A.ci + A.lvNz[l]; const double* v = A.v + A.lvNz[l];
        for (int i = t.tid; i < n; i += t.bs) {
            T s = b[i];
            for (int e = rp[i]; e < rp[i + 1]; ++e) s -= (T)v[e] * (x[ci[e]] + xc[map[ci[e]]]);
            const T d = (T)dg[i];
            const T x0 = x[i] + xc[map[i]];
            xout[i] = (fabs(d) > T(1e-30)) ? x0 + wJ * s / d : x0;
        }
    }
}

// z = V-cycle(rin): the cycle of amg_vcycle() with level 0 in registers.
template <int CPT, class T>
RSG_D void vcycle_res(Ctx& C, const ResRows<CPT>& R, const CoarseImg& G, const T (&rin)[CPT], T (&z)[CPT]) {
    Thr& t = C.t; RsgMemberDev* M = C.M; RsgAmgDev& A = M->amg; const int nC = C.nC;
    const int L = A.lvN[RSG_MAX_LEVELS];
    const T wJ = T(2.0 / 3.0);
    const int base = A.lvRow[1];                  // first coarse row in the level arrays
    // Coarse x (rows base..) lives in the LDS coarse-vector area; coarse b in
    // the global level arrays (each b row is written and read by the thread
    // that owns it, except at the coarsest level, which reads it after a
    // barrier); the two exchange buffers, named explicitly here, carry level
    // 0's gathered vectors and, while the coarse levels run, the coarse r.
    // Every write to a buffer follows a barrier that comes after its last reads.
    T* cx = reinterpret_cast<T*>(t.cvec);
    T* cb = reinterpret_cast<T*>(A.b) + base;
    T* B0 = reinterpret_cast<T*>(t.xb0);
    T* B1 = reinterpret_cast<T*>(t.xb1);
    T* cr = B1;
    T x0[CPT], tmp[CPT];
    if (L == 1) {
        // level 0 is the coarsest: dense inverse (FP64) or Jacobi sweeps on the whole vector
        T* xs = res_bufT<T>(t);
        res_put<T>(t, nC, rin, xs);
        RSG_SYNC();
        if (M->sc->coarseDense) {
            const int DN = RSG_COARSE_DENSE;
#pragma unroll
            for (int j = 0; j < CPT; ++j) {
                const int c = t.tid + j * t.bs;
                double s = 0.0;
                if (c < nC)
                    for (int q = 0; q < nC; ++q) s += t.lds[c * DN + q] * (double)xs[q];
                z[j] = (T)s;
            }
        } else {
#pragma unroll
            for (int j = 0; j < CPT; ++j) z[j] = (fabs((T)R.dg[j]) > T(1e-30)) ? wJ * rin[j] / (T)R.dg[j] : T(0);
            for (int sweep = 0; sweep < 50; ++sweep) {
                xs = res_bufT<T>(t);
                res_put<T>(t, nC, z, xs);
                RSG_SYNC();
                res_jacobi0<CPT, T>(t, nC, R, rin, z, xs, tmp);
#pragma unroll
                for (int j = 0; j < CPT; ++j) z[j] = tmp[j];
            }
        }
        return;
    }
    // level 0 down: two sweeps from zero, residual, restriction
#pragma unroll
    for (int j = 0; j < CPT; ++j) x0[j] = (fabs((T)R.dg[j]) > T(1e-30)) ? wJ * rin[j] / (T)R.dg[j] : T(0);
    res_put<T>(t, nC, x0, B0);
    RSG_SYNC();
    res_jacobi0<CPT, T>(t, nC, R, rin, x0, B0, tmp);
    res_put<T>(t, nC, tmp, B1);
    RSG_SYNC();
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int c = t.tid + j * t.bs;
        x0[j] = tmp[j];
        if (c < nC) B0[c] = rin[j] - res_row<T>(R, j, tmp[j], B1);
    }
    RSG_SYNC();
    res_restrict<T>(t, A, G, 1, B0, cb + (A.lvRow[1] - base));
    RSG_TMARK(19);
    // coarse levels down (row I of a level is owned by thread I % bs throughout)
    for (int l = 1; l + 1 < L; ++l) {
        const int off = A.lvRow[l] - base;
        T* x = cx + off; T* bb = cb + off; T* r = cr + off;
        res_sweepL<T>(t, A, G, l, bb, x, r, true);
        RSG_SYNC();
        RSG_TMARK(l == 1 ? 24 : 29);
        res_sweepL<T>(t, A, G, l, bb, r, x, false);
        RSG_SYNC();
        RSG_TMARK(l == 1 ? 25 : 30);
        res_restrict_residual<T>(t, A, G, l, bb, x, cb + (A.lvRow[l + 1] - base));
        RSG_TMARK(l == 1 ? 26 : 31);
    }
    RSG_SYNC();
    RSG_TMARK(20);
    {   // coarsest level (dense inverse applied in FP64)
        const int l = L - 1, n = A.lvN[l];
        const int off = A.lvRow[l] - base;
        T* x = cx + off; T* bb = cb + off;
        if (M->sc->coarseDense) {
            const int DN = RSG_COARSE_DENSE;
            for (int i = t.tid; i < n; i += t.bs) {
                double s = 0.0;
                for (int j = 0; j < n; ++j) s += t.lds[i * DN + j] * (double)bb[j];
                x[i] = (T)s;
            }
        } else {
            const int* rp = A.rp + A.lvRp[l]; const int* ci = A.ci + A.lvNz[l];
            const double* v = A.v + A.lvNz[l]; const double* dg = A.dg + A.lvRow[l];
            T* r = cr + off;
            res_jacobiL<T>(t, n, rp, ci, v, dg, bb, x, x, true);
            RSG_SYNC();
            for (int s = 0; s < 25; ++s) {
                res_jacobiL<T>(t, n, rp, ci, v, dg, bb, x, r, false);
                RSG_SYNC();
                res_jacobiL<T>(t, n, rp, ci, v, dg, bb, r, x, false);
                RSG_SYNC();
            }
        }
        RSG_SYNC();
    }
    RSG_TMARK(21);
    // coarse levels up
    for (int l = L - 2; l >= 1; --l) {
        const int off = A.lvRow[l] - base;
        T* x = cx + off; T* bb = cb + off; T* r = cr + off;
        const int* map = A.map + A.lvRow[l];
        const T* xc = cx + (A.lvRow[l + 1] - base);
        res_sweep_prolong<T>(t, A, G, l, bb, x, xc, map, r);
        RSG_SYNC();
        RSG_TMARK(l == 1 ? 27 : 14);
        res_sweepL<T>(t, A, G, l, bb, r, x, false);
        RSG_SYNC();
        RSG_TMARK(l == 1 ? 28 : 15);
    }
    RSG_TMARK(22);
    // level 0 up: prolongation and two sweeps
    {
        const int* map = A.map + A.lvRow[0];
        const T* xc = cx + (A.lvRow[1] - base);
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int c = t.tid + j * t.bs;
            if (c < nC) x0[j] += xc[map[c]];
        }
    }
    res_put<T>(t, nC, x0, B0);
    RSG_SYNC();
    res_jacobi0<CPT, T>(t, nC, R, rin, x0, B0, tmp);
    res_put<T>(t, nC, tmp, B1);
    RSG_SYNC();
    res_jacobi0<CPT, T>(t, nC, R, rin, tmp, B1, z);
    RSG_TMARK(23);
}

// PCG with Jacobi or V-cycle preconditioning (the recurrence of pcg() above)
// from the residual r at x (rr = r.r block-summed): x, r in/out in registers.
template <int CPT, class T>
RSG_D SolveRes pcg_res_from(Ctx& C, const ResRows<CPT>& R, const CoarseImg& G, T (&x)[CPT], T (&r)[CPT], T rr,
                            int maxIter, double tol, bool useAmg, bool& updated) {
    Thr& t = C.t; const int nC = C.nC;
    T z[CPT], p[CPT], q[CPT];
    SolveRes res = {0, 0.0, 0.0, 0};
    updated = false;
    const T r0 = sqrt(rr);
    res.initialRes = r0;
    if (r0 < T(1e-30)) { res.converged = 1; return res; }
    T rzv[1] = {T(0)};
    if (useAmg) {
        vcycle_res<CPT, T>(C, R, G, r, z);
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
            const T zc = jac<T>(r[j], (T)R.dg[j]);
            z[j] = zc; p[j] = zc;
            if (c < nC) rzv[0] += r[j] * zc;
        }
    }
    bsum(t, rzv);
    T rz = rzv[0];
    for (int it = 0; it < maxIter; ++it) {
        T* xs = res_bufT<T>(t);
        res_put<T>(t, nC, p, xs);
        RSG_SYNC();
        T a1[1] = {T(0)};
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int c = t.tid + j * t.bs;
            q[j] = T(0);
            if (c < nC) {
                const T qc = res_row<T>(R, j, p[j], xs);
                q[j] = qc;
                a1[0] += p[j] * qc;
            }
        }
        bsum(t, a1);
        RSG_TMARK(16);
        const T pAp = a1[0];
        if (fabs(pAp) < T(1e-30)) break;
        const T alpha = rz / pAp;
        T a2[1] = {T(0)};
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int c = t.tid + j * t.bs;
            x[j] += alpha * p[j];
            const T rc = r[j] - alpha * q[j];
            r[j] = rc;
            if (c < nC) a2[0] += rc * rc;
        }
        updated = true;
        bsum(t, a2);
        RSG_TMARK(17);
        const T rn = sqrt(a2[0]);
        res.iterations = it + 1;
        res.finalRes = rn / r0;
        if (res.finalRes < tol) { res.converged = 1; break; }
        T a3[1] = {T(0)};
        if (useAmg) {
            vcycle_res<CPT, T>(C, R, G, r, z);
#pragma unroll
            for (int j = 0; j < CPT; ++j) {
                const int c = t.tid + j * t.bs;
                if (c < nC) a3[0] += r[j] * z[j];
            }
        } else {
#pragma unroll
            for (int j = 0; j < CPT; ++j) {
                const int c = t.tid + j * t.bs;
                const T zc = jac<T>(r[j], (T)R.dg[j]);
                z[j] = zc;
                if (c < nC) a3[0] += r[j] * zc;
            }
        }
        bsum(t, a3);
        RSG_TMARK(18);
        const T rzNew = a3[0];
        const T beta = rzNew / (rz + T(1e-30));
#pragma unroll
        for (int j = 0; j < CPT; ++j) p[j] = z[j] + beta * p[j];
        rz = rzNew;
    }
    return res;
}

template <int CPT> RSG_D SolveRes pcg_res(Ctx& C, double* xg, int maxIter, double tol, bool useAmg) {
    Thr& t = C.t; const int nC = C.nC;
    ResRows<CPT> R;
    res_load_rows(C, R);
    double x[CPT], r[CPT];
    res_get(t, nC, xg, x);
    const double rr = res_defect(C, R, x, r);
    CoarseImg G;
    // the images are staged only when the solve has work to do (as before)
    if (useAmg && !(sqrt(rr) < 1e-30)) G = coarse_images(t, C.M->amg);
    bool updated;
    const SolveRes res = pcg_res_from<CPT, double>(C, R, G, x, r, rr, maxIter, tol, useAmg, updated);
    if (updated) res_store(t, nC, x, xg);
    return res;
}

// Mixed precision on the resident path: the FP64 defect r = b - A x, an FP32
// inner solve of A d = r from d = 0 (same solver as the FP64 member would
// use) asked for the reduction still missing, x += d in FP64, and the FP64
// defect again -- iterative refinement until the true FP64 residual is below
// tol relative to the first defect, at most RSG_MIXED_SWEEPS sweeps and
// maxIter inner iterations in all.  initialRes is the FP64 first defect, as
// in the FP64 solvers, so the outer iteration's residual accounting holds.
template <int CPT> RSG_D SolveRes lsolve_mixed_res(Ctx& C, int code, double* xg, bool amg, int maxIter,
                                                   double tol) {
    Thr& t = C.t; const int nC = C.nC;
    ResRows<CPT> R;
    res_load_rows(C, R);
    double x[CPT], r[CPT];
    res_get(t, nC, xg, x);
    double rn = sqrt(res_defect(C, R, x, r));
    const double r0 = rn;
    SolveRes res = {0, 0.0, r0, 0};
    if (r0 < 1e-30) { res.converged = 1; return res; }
    CoarseImg G;
    if (code == 0 && amg) G = coarse_images(t, C.M->amg);
    bool changed = false;
    for (int sweep = 0; sweep < RSG_MIXED_SWEEPS && res.iterations < maxIter; ++sweep) {
        // the FP32 solve sees the defect scaled to unit norm (a scalar, so the
        // system itself stays unscaled): the unscaled systems' residuals reach
        // 1e-20 and below, whose squares FP32 cannot represent
        float d[CPT], rf[CPT];
        const double inv = 1.0 / rn;
#pragma unroll
        for (int j = 0; j < CPT; ++j) { d[j] = 0.0f; rf[j] = (float)(r[j] * inv); }
        const float rr = 1.0f;
        const double tolIn = tol * r0 / rn;
        bool updated;
        const SolveRes in =
            code == 2 ? bicgstab_res_from<CPT, float>(C, R, d, rf, rr, maxIter - res.iterations, tolIn, updated)
                      : pcg_res_from<CPT, float>(C, R, G, d, rf, rr, maxIter - res.iterations, tolIn,
                                                 code == 0 && amg, updated);
        res.iterations += in.iterations;
        if (!updated) break;
#pragma unroll
        for (int j = 0; j < CPT; ++j) x[j] += rn * (double)d[j];
        changed = true;
        const double rnNew = sqrt(res_defect(C, R, x, r));
        res.finalRes = rnNew / r0;
        if (res.finalRes < tol) { res.converged = 1; break; }
        if (!(rnNew < rn)) break;   // the FP32 solve stalls: stop refining
        rn = rnNew;
    }
    if (changed) res_store(t, nC, x, xg);
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

// allowMixed false: this solve runs in FP64 even for a mixed member (the
// omega equation, see simple_iteration).
template <int CPT, int PREC> RSG_D SolveRes lsolve_cpt(Ctx& C, int code, double* x, bool amg, int maxIter,
                                                       double tol, bool allowMixed) {
    if constexpr (PREC != 0) {
        if (allowMixed && is_mixed<PREC>(C)) return lsolve_mixed_res<CPT>(C, code, x, amg, maxIter, tol);
    }
    if (code == 2) return bicgstab_res<CPT>(C, x, maxIter, tol);
    return pcg_res<CPT>(C, x, maxIter, tol, amg);
}

// Cells per thread rounded up to an instantiated width (CMAX: the widest one
// the kernel is built with; it bounds the kernel's register allocation).
template <int CMAX, int PREC = -1>
RSG_D SolveRes lsolve_res(Ctx& C, int code, double* x, bool amg, int maxIter, double tol, bool allowMixed = true) {
#ifdef RSG_EMULATE
    return lsolve_cpt<RSG_RES_MAXN, PREC>(C, code, x, amg, maxIter, tol, allowMixed);
#else
    const int cpt = (C.nC + C.t.bs - 1) / C.t.bs;
    if constexpr (CMAX >= 2) { if (cpt == 2) return lsolve_cpt<2, PREC>(C, code, x, amg, maxIter, tol, allowMixed); }
    if constexpr (CMAX >= 3) { if (cpt == 3) return lsolve_cpt<3, PREC>(C, code, x, amg, maxIter, tol, allowMixed); }
    if constexpr (CMAX >= 4) { if (cpt == 4) return lsolve_cpt<4, PREC>(C, code, x, amg, maxIter, tol, allowMixed); }
    if constexpr (CMAX >= 6) { if (cpt == 5 || cpt == 6) return lsolve_cpt<6, PREC>(C, code, x, amg, maxIter, tol, allowMixed); }
    return lsolve_cpt<1, PREC>(C, code, x, amg, maxIter, tol, allowMixed);
#endif
}

}  // namespace rsg
