// rsg_grid.h -- grid-wide PCG / AMG-PCG for one large LDU system (every CU
// works on the one system), FP64 or FP32-inner with FP64 refinement.
//
// The member solver (rsg_member.h) gives each ensemble member one workgroup;
// a single large system (10^5..10^6 rows, the stand-alone pressure solve)
// instead spreads every phase over the whole GPU: one kernel per sparse
// operation, rows one per thread, deterministic two-pass dot products (block
// partials, then one block summing them in block order).
//
// Algorithms (reference: core/include/LinearSolver.hpp):
//   lduToCSR            :344-404  level-0 CSR, columns ascending per row
//   buildAggregates     :407-443  the greedy pass, bit for bit (rsg::amg_aggregate)
//   galerkinProduct     :446-485  coarse entry (I,J) = sum of fine entries in
//                                 ascending fine row then entry
//   PCGSolver / AMG     :147-186, :602-653  PCG recurrence, tol relative to r0
// The V-cycle smooths with damped Jacobi (omega 2/3, V(2,2)) and solves the
// coarsest level (<= 64 rows) with a dense inverse, as the member V-cycle does;
// a larger coarsest level (where the greedy aggregation stalls, up to
// RSG_DENSE_MAX rows) is inverted on the FP64 matrix cores (rsg_dense.h) and
// applied as one matrix-vector product per V-cycle.
// Kernels stay one phase each, CSR, one row per thread: on a 261,600-row
// system with 17 levels, folding the residual into the restriction and the
// prolongation into the next sweep, ELL images, and a single-workgroup kernel
// for the small levels were each measured slower (longer dependent-load
// chains per thread, padded rows, one CU's latency) than these plain launches.
//
// The hierarchy is built on the host (the greedy aggregation is sequential);
// every iteration runs on the device.  Included by rsg_capi.hip.
#pragma once
#include <chrono>
#include <type_traits>
#include <vector>

#include "rsg_dense.h"

namespace rsg {
namespace grid {

constexpr int kBlock = 256;

// ------------------------------------------------------------------ kernels
// Row products.  ex == null: y_i = sum_e v[e] x[ci[e]] (FP64 levels).  With
// ex (FP32 levels): y_i = ex_i x_i + sum_e v[e] (x[ci[e]] - x_i), where ex_i
// is the FP64 row sum rounded once and the diagonal entry's v is 0 -- the
// constants (the near-null space of a pinned or periodic pressure Laplacian)
// stay in the near-null space after the coefficients are rounded to FP32.
template <typename T>
__device__ __forceinline__ T row_ax(int i, const int* __restrict__ rp, const int* __restrict__ ci,
                                    const T* __restrict__ v, const T* __restrict__ ex, const T* x) {
    T s = 0;
    if (ex) {
        const T xi = x[i];
        s = ex[i] * xi;
        for (int e = rp[i]; e < rp[i + 1]; ++e) s += v[e] * (x[ci[e]] - xi);
    } else {
        for (int e = rp[i]; e < rp[i + 1]; ++e) s += v[e] * x[ci[e]];
    }
    return s;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_spmv(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                  const T* __restrict__ v, const T* __restrict__ ex,
                                                  const T* __restrict__ x, T* __restrict__ y) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    y[i] = row_ax<T>(i, rp, ci, v, ex, x);
}

// r = b - A x
template <typename T>
__global__ __launch_bounds__(kBlock) void k_residual(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                      const T* __restrict__ v, const T* __restrict__ ex,
                                                      const T* __restrict__ b, const T* __restrict__ x,
                                                      T* __restrict__ r) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    r[i] = b[i] - row_ax<T>(i, rp, ci, v, ex, x);
}

// xo = xi + w (b - A xi) / d  (xi == null: from zero)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_jacobi(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                    const T* __restrict__ v, const T* __restrict__ ex,
                                                    const T* __restrict__ dg, const T* __restrict__ b, const T* xi,
                                                    T* xo) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const T wJ = T(2) / T(3);
    T s = b[i];
    T x0 = 0;
    if (xi) {
        x0 = xi[i];
        if (ex) {
            s -= ex[i] * x0;
            for (int e = rp[i]; e < rp[i + 1]; ++e) s -= v[e] * (xi[ci[e]] - x0);
        } else {
            for (int e = rp[i]; e < rp[i + 1]; ++e) s -= v[e] * xi[ci[e]];
        }
    }
    const T d = dg[i];
    xo[i] = (fabs((double)d) > 1e-30) ? x0 + wJ * s / d : x0;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_restrict(int nc, const int* __restrict__ aP, const int* __restrict__ aI,
                                                      const T* __restrict__ r, T* __restrict__ bc) {
    const int I = blockIdx.x * kBlock + threadIdx.x;
    if (I >= nc) return;
    T s = 0;
    for (int q = aP[I]; q < aP[I + 1]; ++q) s += r[aI[q]];
    bc[I] = s;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_prolong(int n, const int* __restrict__ map, const T* __restrict__ xc,
                                                     T* __restrict__ x) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    x[i] += xc[map[i]];
}

// x = C^{-1} b on the coarsest level (n <= 64), one block
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dense(int n, const T* __restrict__ cinv, const T* __restrict__ b,
                                                   T* __restrict__ x) {
    const int i = threadIdx.x;
    if (i >= n) return;
    T s = 0;
    for (int j = 0; j < n; ++j) s += cinv[i * 64 + j] * b[j];
    x[i] = s;
}

// per-block partial sums of up to two dot products, accumulated in FP64
template <typename T>
__device__ inline double block_sum(double v, double* sh) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int q = 0; q < kBlock / 64; ++q) s += sh[q];
    return s;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot2(int n, const T* a, const T* b, const T* c, const T* d,
                                                  double* part) {
    __shared__ double sh[2][kBlock / 64];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double s0 = 0.0, s1 = 0.0;
    if (i < n) {
        s0 = (double)a[i] * (double)b[i];
        if (c) s1 = (double)c[i] * (double)d[i];
    }
    const double t0 = block_sum<T>(s0, sh[0]);
    const double t1 = block_sum<T>(s1, sh[1]);
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = t0; part[2 * blockIdx.x + 1] = t1; }
}

// sums the block partials in block order into out[0] (and out[1] when two)
static __global__ __launch_bounds__(kBlock) void k_finish(int nb, const double* part, double* out, int two) {
    __shared__ double sh[2][kBlock / 64];
    double s0 = 0.0, s1 = 0.0;
    // each thread sums a contiguous stripe, stripes combined in thread order
    const int per = (nb + kBlock - 1) / kBlock;
    for (int q = threadIdx.x * per; q < nb && q < (threadIdx.x + 1) * per; ++q) {
        s0 += part[2 * q];
        s1 += part[2 * q + 1];
    }
    // ordered combination: wave butterflies are deterministic, waves in order
    const double t0 = block_sum<double>(s0, sh[0]);
    const double t1 = block_sum<double>(s1, sh[1]);
    if (threadIdx.x == 0) { out[0] = t0; if (two) out[1] = t1; }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(int n, T* x, T v) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] = v;
}

template <typename Td, typename Ts>
__global__ __launch_bounds__(kBlock) void k_convert(int n, const Ts* s, Td* d) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) d[i] = (Td)s[i];
}

// Device scalars of the PCG loop: sc[0] rz, sc[1] pAp, sc[2] rr, sc[3] rzNew.
// x += alpha p, r -= alpha q with alpha = rz / pAp; no update when |pAp| < 1e-30
// (the loop then stops, as the reference's does before updating).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_update(int n, const double* sc, T* x, T* r, const T* p,
                                                        const T* q) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n || !(fabs(sc[1]) >= 1e-30)) return;
    const T alpha = (T)(sc[0] / sc[1]);
    x[i] += alpha * p[i];
    r[i] -= alpha * q[i];
}

// p = z + beta p with beta = rzNew / (rz + 1e-30)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(int n, const double* sc, const T* z, T* p) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const T beta = (T)(sc[3] / (sc[0] + 1e-30));
    p[i] = z[i] + beta * p[i];
}

// rz = rzNew (one thread, after every block of k_pcg_dir has read rz)
static __global__ void k_pcg_shift(double* sc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) sc[0] = sc[3];
}

// z = r / d (Jacobi), zero divisor guard as the reference's jacobiPrecond
template <typename T>
__global__ __launch_bounds__(kBlock) void k_jac(int n, const T* dg, const T* r, T* z) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    z[i] = (fabs((double)dg[i]) > 1e-30) ? r[i] / dg[i] : r[i];
}

// e (FP64) += d (any), for the refinement update
template <typename T>
__global__ __launch_bounds__(kBlock) void k_add(int n, const T* d, double* x) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] += (double)d[i];
}

// ------------------------------------------------------------------ host side
struct Level {
    int n = 0, nnz = 0;
    std::vector<int> rp, ci;
    std::vector<double> v, dg;
    std::vector<int> map;            // fine row -> aggregate (levels < L-1)
    std::vector<int> aP, aI;         // aggregate -> fine rows of the finer level (levels >= 1)
};

// lduToCSR (LinearSolver.hpp:344-404): columns ascending per row
inline void ldu_to_csr(int n, int nif, const int* own, const int* nbr, const double* diag, const double* upper,
                       const double* lower, Level& L) {
    std::vector<int> cnt(n + 1, 1);
    for (int f = 0; f < nif; ++f) { cnt[own[f] + 1]++; cnt[nbr[f] + 1]++; }
    L.n = n;
    L.rp.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) L.rp[i + 1] = L.rp[i] + cnt[i + 1];
    L.nnz = L.rp[n];
    L.ci.assign(L.nnz, 0); L.v.assign(L.nnz, 0.0); L.dg.assign(diag, diag + n);
    std::vector<int> fill(L.rp.begin(), L.rp.end() - 1);
    for (int i = 0; i < n; ++i) { L.ci[fill[i]] = i; L.v[fill[i]] = diag[i]; fill[i]++; }
    for (int f = 0; f < nif; ++f) {
        L.ci[fill[own[f]]] = nbr[f]; L.v[fill[own[f]]] = upper[f]; fill[own[f]]++;
        L.ci[fill[nbr[f]]] = own[f]; L.v[fill[nbr[f]]] = lower[f]; fill[nbr[f]]++;
    }
    for (int i = 0; i < n; ++i) {   // insertion sort by column (rows are short)
        for (int a = L.rp[i] + 1; a < L.rp[i + 1]; ++a) {
            const int c = L.ci[a]; const double val = L.v[a];
            int b = a - 1;
            while (b >= L.rp[i] && L.ci[b] > c) { L.ci[b + 1] = L.ci[b]; L.v[b + 1] = L.v[b]; --b; }
            L.ci[b + 1] = c; L.v[b + 1] = val;
        }
    }
}

// The Galerkin level `c` of `f` under the aggregate map f.map (nc aggregates).
inline void galerkin(const Level& f, int nc, Level& c);

// One coarsening step: aggregates of `f` and the Galerkin level `c`.
inline bool coarsen(Level& f, Level& c) {
    f.map.assign(f.n, -1);
    const int nc = amg_aggregate(f.n, f.rp.data(), f.ci.data(), f.v.data(), f.dg.data(), f.map.data());
    if (nc == 0 || nc >= f.n) { f.map.clear(); return false; }
    galerkin(f, nc, c);
    return true;
}

// Pairwise matching (deterministic, row order): each unmatched row joins its
// unmatched neighbour of largest |a_ij|; a row left without a partner then
// joins the aggregate of its neighbour of largest |a_ij| (it stays alone only
// without neighbours).  Aggregates are numbered in order of their first row.
// Returns the count.
inline int pair_aggregate(const Level& f, std::vector<int>& map) {
    std::vector<int> tmp(f.n, -1);
    std::vector<char> single(f.n, 0);
    int nt = 0;
    for (int i = 0; i < f.n; ++i) {
        if (tmp[i] >= 0) continue;
        int best = -1;
        double bv = 0.0;
        for (int e = f.rp[i]; e < f.rp[i + 1]; ++e) {
            const int j = f.ci[e];
            if (j == i || tmp[j] >= 0) continue;
            const double a = std::fabs(f.v[e]);
            if (a > bv) { bv = a; best = j; }
        }
        tmp[i] = nt;
        if (best >= 0) tmp[best] = nt; else single[i] = 1;
        ++nt;
    }
    for (int i = 0; i < f.n; ++i) {
        if (!single[i]) continue;
        int best = -1;
        double bv = 0.0;
        for (int e = f.rp[i]; e < f.rp[i + 1]; ++e) {
            const int j = f.ci[e];
            if (j == i) continue;
            const double a = std::fabs(f.v[e]);
            if (a > bv) { bv = a; best = j; }
        }
        if (best >= 0) tmp[i] = tmp[best];
    }
    std::vector<int> id(nt, -1);
    map.assign(f.n, -1);
    int nc = 0;
    for (int i = 0; i < f.n; ++i) {
        if (id[tmp[i]] < 0) id[tmp[i]] = nc++;
        map[i] = id[tmp[i]];
    }
    return nc;
}

// Coarsening for levels where the reference's greedy aggregation stalls
// (fewer than a quarter of the rows merged): two rounds of pairwise matching,
// the second on the Galerkin image of the first, composed into one map
// (aggregates of up to four rows).  Not the reference's aggregation: used by
// the large-member path only, below the levels where the reference's pass
// still coarsens.
inline bool coarsen_pairwise(Level& f, Level& c) {
    std::vector<int> m1, m2;
    const int n1 = pair_aggregate(f, m1);
    if (n1 >= f.n) return false;
    Level mid;
    f.map = m1;
    galerkin(f, n1, mid);
    const int n2 = pair_aggregate(mid, m2);
    for (int i = 0; i < f.n; ++i) f.map[i] = m2[m1[i]];
    galerkin(f, n2, c);
    return true;
}

inline void galerkin(const Level& f, int nc, Level& c) {
    c.n = nc;
    c.aP.assign(nc + 1, 0);
    for (int i = 0; i < f.n; ++i) c.aP[f.map[i] + 1]++;
    for (int I = 0; I < nc; ++I) c.aP[I + 1] += c.aP[I];
    c.aI.assign(f.n, 0);
    std::vector<int> fill(c.aP.begin(), c.aP.end() - 1);
    for (int i = 0; i < f.n; ++i) c.aI[fill[f.map[i]]++] = i;
    // coarse pattern row by row; entry (I,J) sums fine entries in ascending
    // fine row (aggregate order), then ascending fine entry
    std::vector<int> mark(nc, -1), pos(nc, 0), cols;
    c.rp.assign(nc + 1, 0); c.ci.clear(); c.v.clear(); c.dg.assign(nc, 0.0);
    for (int I = 0; I < nc; ++I) {
        cols.clear();
        for (int q = c.aP[I]; q < c.aP[I + 1]; ++q) {
            const int i = c.aI[q];
            for (int e = f.rp[i]; e < f.rp[i + 1]; ++e) {
                const int J = f.map[f.ci[e]];
                if (mark[J] != I) { mark[J] = I; cols.push_back(J); }
            }
        }
        std::sort(cols.begin(), cols.end());
        const int base = (int)c.ci.size();
        for (size_t a = 0; a < cols.size(); ++a) { pos[cols[a]] = base + (int)a; c.ci.push_back(cols[a]); c.v.push_back(0.0); }
        for (int q = c.aP[I]; q < c.aP[I + 1]; ++q) {
            const int i = c.aI[q];
            for (int e = f.rp[i]; e < f.rp[i + 1]; ++e) c.v[pos[f.map[f.ci[e]]]] += f.v[e];
        }
        c.rp[I + 1] = (int)c.ci.size();
        for (int a = c.rp[I]; a < c.rp[I + 1]; ++a) if (c.ci[a] == I) c.dg[I] = c.v[a];
    }
    c.nnz = c.rp[nc];
}

// AMGSolver::setup (LinearSolver.hpp:516-550): level 0 from the LDU arrays,
// then coarsening while the level has more than RSG_MIN_COARSE rows, at most
// `maxCoarsenings` times, stopping early when aggregation no longer reduces.
inline void build_hierarchy(int n, int nif, const int* own, const int* nbr, const double* diag,
                            const double* upper, const double* lower, int maxCoarsenings, std::vector<Level>& H) {
    H.assign(1, Level());
    ldu_to_csr(n, nif, own, nbr, diag, upper, lower, H[0]);
    for (int l = 0; l < maxCoarsenings; ++l) {
        if (H.back().n <= RSG_MIN_COARSE) break;
        Level c;
        if (!coarsen(H.back(), c)) break;
        H.push_back(std::move(c));
    }
}

// Gauss-Jordan inverse of the coarsest level (no pivoting: SPD M-matrix image),
// the same elimination as the member solver's amg_coarse_inverse.
inline bool dense_inverse(const Level& L, std::vector<double>& G) {
    const int n = L.n, DN = 64;
    G.assign(DN * DN, 0.0);
    for (int i = 0; i < n; ++i)
        for (int e = L.rp[i]; e < L.rp[i + 1]; ++e) G[i * DN + L.ci[e]] += L.v[e];
    for (int k = 0; k < n; ++k) {
        const double piv = G[k * DN + k];
        if (!(std::fabs(piv) > 1e-300)) return false;
        for (int j = 0; j < n; ++j) G[k * DN + j] = (j == k ? 1.0 : G[k * DN + j]) / piv;
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            const double fct = G[i * DN + k];
            for (int j = 0; j < n; ++j) G[i * DN + j] = (j == k ? 0.0 : G[i * DN + j]) - fct * G[k * DN + j];
        }
    }
    return true;
}

// Device image of one level in precision T.
template <typename T> struct DevLevel {
    int n = 0, nnz = 0;
    int *rp = nullptr, *ci = nullptr, *map = nullptr, *aP = nullptr, *aI = nullptr;
    T *v = nullptr, *dg = nullptr;
    T* ex = nullptr;       // FP32 levels: row sums (row_ax's form); null on FP64 levels
    T *x = nullptr, *b = nullptr, *r = nullptr;
};

struct DevAlloc {
    std::vector<void*> ptrs;
    hipError_t err = hipSuccess;
    template <typename T> T* take(size_t n) {
        void* p = nullptr;
        if (err == hipSuccess) err = hipMalloc(&p, sizeof(T) * (n ? n : 1));
        if (err == hipSuccess) ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    ~DevAlloc() { for (void* p : ptrs) (void)hipFree(p); }
};

template <typename T> struct Hierarchy {
    std::vector<DevLevel<T>> lv;
    T* cinv = nullptr;
    bool dense = false;
    const double* dinv = nullptr;   // FP64 inverse of a large coarsest level (row stride dinvN)
    int dinvN = 0;
};

template <typename T>
hipError_t upload(const std::vector<Level>& H, const std::vector<double>& G, bool dense, DevAlloc& A,
                  Hierarchy<T>& D, hipStream_t s) {
    D.lv.resize(H.size());
    std::vector<T> tmp;
    auto put = [&](void* dst, const void* src, size_t bytes) {
        if (A.err == hipSuccess && bytes) A.err = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    };
    auto putT = [&](T* dst, const std::vector<double>& src) {
        tmp.assign(src.begin(), src.end());
        put(dst, tmp.data(), sizeof(T) * tmp.size());
        if (A.err == hipSuccess) A.err = hipStreamSynchronize(s);   // tmp is reused
    };
    for (size_t l = 0; l < H.size(); ++l) {
        const Level& L = H[l];
        DevLevel<T>& d = D.lv[l];
        d.n = L.n; d.nnz = L.nnz;
        d.rp = A.take<int>(L.n + 1); d.ci = A.take<int>(L.nnz);
        d.v = A.take<T>(L.nnz); d.dg = A.take<T>(L.n);
        d.x = A.take<T>(L.n); d.b = A.take<T>(L.n); d.r = A.take<T>(L.n);
        put(d.rp, L.rp.data(), 4 * L.rp.size()); put(d.ci, L.ci.data(), 4 * L.ci.size());
        if (std::is_same<T, float>::value) {
            // FP32 level in row_ax's form: off-diagonal values, FP64 row sums
            std::vector<double> off(L.v), rs(L.n, 0.0);
            for (int i = 0; i < L.n; ++i)
                for (int e = L.rp[i]; e < L.rp[i + 1]; ++e) {
                    rs[i] += L.v[e];
                    if (L.ci[e] == i) off[e] = 0.0;
                }
            d.ex = A.take<T>(L.n);
            putT(d.v, off); putT(d.ex, rs);
        } else {
            putT(d.v, L.v);
        }
        putT(d.dg, L.dg);
        if (!L.map.empty()) { d.map = A.take<int>(L.n); put(d.map, L.map.data(), 4 * L.map.size()); }
        if (!L.aP.empty()) {
            d.aP = A.take<int>(L.n + 1); d.aI = A.take<int>(L.aI.size());
            put(d.aP, L.aP.data(), 4 * L.aP.size()); put(d.aI, L.aI.data(), 4 * L.aI.size());
        }
    }
    D.dense = dense;
    if (dense) { D.cinv = A.take<T>(64 * 64); putT(D.cinv, G); }
    if (A.err == hipSuccess) A.err = hipStreamSynchronize(s);
    return A.err;
}

inline int blocks(int n) { return (n + kBlock - 1) / kBlock; }

// Device workspace of the dense coarsest-level inverse (rsg_dense.h): the
// padded N x N image (the inverse after dense_invert), the two 64 x N panels,
// two pivot blocks and the pivot flag.
struct DenseWork {
    int n = 0, N = 0;
    double *M = nullptr, *R = nullptr, *CT = nullptr, *P = nullptr;
    int* flag = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool take(int rows, DevAlloc& A) {
        n = rows; N = dense::padded(rows);
        M = A.take<double>((size_t)N * N);
        R = A.take<double>((size_t)dense::kNB * N); CT = A.take<double>((size_t)N * dense::kNB);
        P = A.take<double>(2 * dense::kNB * dense::kNB);
        flag = A.take<int>(1);
        if (A.err == hipSuccess) A.err = hipEventCreate(&e0);
        if (A.err == hipSuccess) A.err = hipEventCreate(&e1);
        return A.err == hipSuccess;
    }
    void release() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        e0 = e1 = nullptr;
    }
};

// W.M = inverse of the n-row CSR matrix (rp, ci, v: device arrays) on stream
// s; *ok is false when a pivot was not positive (the caller then smooths the
// level); *ms gets the inversion's device time.  Waits for the stream.
inline hipError_t dense_invert(DenseWork& W, const int* rp, const int* ci, const double* v, hipStream_t s, bool* ok,
                               double* ms) {
    *ok = false; *ms = 0.0;
    const size_t total = (size_t)W.N * W.N;
    const int g = (int)std::min<size_t>((total + dense::kThreads - 1) / dense::kThreads, 8192);
    hipLaunchKernelGGL(dense::k_dense_clear, dim3(g), dim3(dense::kThreads), 0, s, W.n, W.N, W.M);
    hipLaunchKernelGGL(dense::k_dense_scatter, dim3((W.n + dense::kThreads - 1) / dense::kThreads),
                       dim3(dense::kThreads), 0, s, W.n, W.N, rp, ci, v, W.M);
    hipError_t e = hipEventRecord(W.e0, s);
    if (e == hipSuccess) dense::invert(W.N, W.M, W.R, W.CT, W.P, W.flag, s);
    if (e == hipSuccess) e = hipEventRecord(W.e1, s);
    int hflag = 1;
    if (e == hipSuccess) e = hipMemcpyAsync(&hflag, W.flag, sizeof hflag, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, W.e0, W.e1);
    *ms = t;
    *ok = e == hipSuccess && hflag == 0;
    return e;
}

// Explicit FP64 inverse of level L (host CSR) on the device: *inv (row stride
// *N = L.n rounded up to 64) lives in A.  *ok as for dense_invert.
inline hipError_t dense_inverse_device(const Level& L, DevAlloc& A, hipStream_t s, double** inv, int* N, bool* ok,
                                       double* ms) {
    *ok = false; *inv = nullptr; *ms = 0.0;
    const int n = L.n;
    int* rp = A.take<int>(n + 1); int* ci = A.take<int>(L.nnz); double* v = A.take<double>(L.nnz);
    DenseWork W;
    if (!W.take(n, A)) { W.release(); return A.err; }
    *N = W.N;
    hipError_t e = hipMemcpyAsync(rp, L.rp.data(), 4 * (size_t)(n + 1), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(ci, L.ci.data(), 4 * (size_t)L.nnz, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(v, L.v.data(), 8 * (size_t)L.nnz, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = dense_invert(W, rp, ci, v, s, ok, ms);
    if (*ok) *inv = W.M;
    W.release();
    return e;
}

// z = V-cycle(r) on hierarchy D (level 0 vectors: rin, z; level matrices in D)
template <typename T>
void vcycle(const Hierarchy<T>& D, const T* rin, T* z, hipStream_t s) {
    const int L = (int)D.lv.size();
    if (L == 1) {
        const DevLevel<T>& d = D.lv[0];
        if (D.dense) { hipLaunchKernelGGL(k_dense<T>, dim3(1), dim3(kBlock), 0, s, d.n, D.cinv, rin, z); return; }
        if (D.dinv) {
            hipLaunchKernelGGL(dense::k_dense_apply<T>, dim3(dense::apply_blocks(d.n)), dim3(dense::kThreads), 0, s, d.n,
                               D.dinvN, D.dinv, rin, z);
            return;
        }
        hipLaunchKernelGGL(k_jacobi<T>, dim3(blocks(d.n)), dim3(kBlock), 0, s, d.n, d.rp, d.ci, d.v, d.ex, d.dg, rin,
                           (const T*)nullptr, z);
        return;
    }
    for (int l = 0; l + 1 < L; ++l) {
        const DevLevel<T>& d = D.lv[l];
        const T* b = (l == 0) ? rin : d.b;
        const dim3 g(blocks(d.n)), bl(kBlock);
        hipLaunchKernelGGL(k_jacobi<T>, g, bl, 0, s, d.n, d.rp, d.ci, d.v, d.ex, d.dg, b, (const T*)nullptr, d.r);
        hipLaunchKernelGGL(k_jacobi<T>, g, bl, 0, s, d.n, d.rp, d.ci, d.v, d.ex, d.dg, b, (const T*)d.r, d.x);
        hipLaunchKernelGGL(k_residual<T>, g, bl, 0, s, d.n, d.rp, d.ci, d.v, d.ex, b, (const T*)d.x, d.r);
        const DevLevel<T>& c = D.lv[l + 1];
        hipLaunchKernelGGL(k_restrict<T>, dim3(blocks(c.n)), bl, 0, s, c.n, c.aP, c.aI, (const T*)d.r, c.b);
    }
    {
        const DevLevel<T>& c = D.lv[L - 1];
        if (D.dense) {
            hipLaunchKernelGGL(k_dense<T>, dim3(1), dim3(kBlock), 0, s, c.n, D.cinv, (const T*)c.b, c.x);
        } else if (D.dinv) {
            hipLaunchKernelGGL(dense::k_dense_apply<T>, dim3(dense::apply_blocks(c.n)), dim3(dense::kThreads), 0, s, c.n,
                               D.dinvN, D.dinv, (const T*)c.b, c.x);
        } else {
            const dim3 g(blocks(c.n)), bl(kBlock);
            hipLaunchKernelGGL(k_jacobi<T>, g, bl, 0, s, c.n, c.rp, c.ci, c.v, c.ex, c.dg, (const T*)c.b, (const T*)nullptr, c.x);
            for (int q = 0; q < 25; ++q) {
                hipLaunchKernelGGL(k_jacobi<T>, g, bl, 0, s, c.n, c.rp, c.ci, c.v, c.ex, c.dg, (const T*)c.b, (const T*)c.x, c.r);
                hipLaunchKernelGGL(k_jacobi<T>, g, bl, 0, s, c.n, c.rp, c.ci, c.v, c.ex, c.dg, (const T*)c.b, (const T*)c.r, c.x);
            }
        }
    }
    for (int l = L - 2; l >= 0; --l) {
        const DevLevel<T>& d = D.lv[l];
        const T* b = (l == 0) ? rin : d.b;
        const dim3 g(blocks(d.n)), bl(kBlock);
        hipLaunchKernelGGL(k_prolong<T>, g, bl, 0, s, d.n, d.map, (const T*)D.lv[l + 1].x, d.x);
        hipLaunchKernelGGL(k_jacobi<T>, g, bl, 0, s, d.n, d.rp, d.ci, d.v, d.ex, d.dg, b, (const T*)d.x, d.r);
        hipLaunchKernelGGL(k_jacobi<T>, g, bl, 0, s, d.n, d.rp, d.ci, d.v, d.ex, d.dg, b, (const T*)d.r, l == 0 ? z : d.x);
    }
}

// hipGraphs of whole V-cycles.  A V-cycle's launch sequence and every
// argument are fixed by the hierarchy (level arrays, coarsest inverse) and the
// two level-0 vectors, so it is captured once on the stream and replayed with
// one hipGraphLaunch afterwards (no memset nodes: every kernel writes its
// outputs).  The owner clears the cache whenever the hierarchy is rebuilt
// (new allocations may reuse old addresses).
struct VcycleGraphs {
    struct Entry {
        const void *l0, *rin, *z, *dinv;
        bool dense;
        size_t levels;
        hipGraphExec_t exec;
    };
    std::vector<Entry> e;
    bool enabled = true;
    void clear() {
        for (Entry& x : e) (void)hipGraphExecDestroy(x.exec);
        e.clear();
    }
    ~VcycleGraphs() { clear(); }
};

template <typename T> void vcycle(const Hierarchy<T>& D, const T* rin, T* z, hipStream_t s);

// z = V-cycle(r): replayed from G's graph of this (hierarchy, rin, z), captured
// on first use; launched directly when G is null or disabled.
template <typename T>
hipError_t vcycle_run(const Hierarchy<T>& D, const T* rin, T* z, hipStream_t s, VcycleGraphs* G) {
    if (!G || !G->enabled) { vcycle<T>(D, rin, z, s); return hipSuccess; }
    const void* l0 = D.lv.empty() ? nullptr : (const void*)D.lv[0].v;
    for (const VcycleGraphs::Entry& x : G->e)
        if (x.l0 == l0 && x.rin == rin && x.z == z && x.dinv == D.dinv && x.dense == D.dense &&
            x.levels == D.lv.size())
            return hipGraphLaunch(x.exec, s);
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    vcycle<T>(D, rin, z, s);
    e = hipStreamEndCapture(s, &g);
    if (e == hipSuccess) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) return e;
    G->e.push_back({l0, rin, z, D.dinv, D.dense, D.lv.size(), ex});
    return hipGraphLaunch(ex, s);
}

struct Scratch {
    double* part = nullptr;   // [2 * blocks]
    double* dots = nullptr;   // [8] device scalars
    double* hdots = nullptr;  // pinned host mirror [8]
};

// out[0..1] = (a.b, c.d) on the device; copied to the host mirror when `read`
template <typename T>
void dot2(int n, const T* a, const T* b, const T* c, const T* d, const Scratch& S, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_dot2<T>, dim3(blocks(n)), dim3(kBlock), 0, s, n, a, b, c, d, S.part);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kBlock), 0, s, blocks(n), (const double*)S.part, out,
                       c != nullptr ? 1 : 0);
}

struct PcgResult { int iterations = 0; double finalRes = 0, initialRes = 0; int converged = 0; };

// PCG (PCGSolver::solve LinearSolver.hpp:147-186, AMGSolver::solve :602-653):
// x in/out, tol relative to ||b - A x0||.  M: V-cycle (amg) or Jacobi.
template <typename T>
hipError_t pcg(const Hierarchy<T>& H, const T* b, T* x, T* r, T* z, T* p, T* q, int maxIter, double tol, bool amg,
               const Scratch& S, hipStream_t s, PcgResult& res, VcycleGraphs* G = nullptr) {
    const DevLevel<T>& A0 = H.lv[0];
    const int n = A0.n;
    const dim3 g(blocks(n)), bl(kBlock);
    double* dv = S.dots;          // [0] rz, [1] pAp, [2] rr, [3] rzNew
    hipLaunchKernelGGL(k_residual<T>, g, bl, 0, s, n, A0.rp, A0.ci, A0.v, A0.ex, b, (const T*)x, r);
    dot2<T>(n, r, r, (const T*)nullptr, (const T*)nullptr, S, dv + 2, s);
    hipError_t e = hipMemcpyAsync(S.hdots, dv, 4 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    const double r0 = std::sqrt(S.hdots[2]);
    res.initialRes = r0;
    if (r0 < 1e-30) { res.converged = 1; return hipSuccess; }
    if (amg) {
        if ((e = vcycle_run<T>(H, r, z, s, G)) != hipSuccess) return e;
    }
    else hipLaunchKernelGGL(k_jac<T>, g, bl, 0, s, n, (const T*)A0.dg, (const T*)r, z);
    hipLaunchKernelGGL((k_convert<T, T>), g, bl, 0, s, n, (const T*)z, p);
    dot2<T>(n, r, z, (const T*)nullptr, (const T*)nullptr, S, dv + 0, s);          // rz
    for (int it = 0; it < maxIter; ++it) {
        hipLaunchKernelGGL(k_spmv<T>, g, bl, 0, s, n, A0.rp, A0.ci, A0.v, A0.ex, (const T*)p, q);
        dot2<T>(n, p, q, (const T*)nullptr, (const T*)nullptr, S, dv + 1, s);      // pAp
        hipLaunchKernelGGL(k_pcg_update<T>, g, bl, 0, s, n, (const double*)dv, x, r, (const T*)p, (const T*)q);
        dot2<T>(n, r, r, (const T*)nullptr, (const T*)nullptr, S, dv + 2, s);      // rr
        e = hipMemcpyAsync(S.hdots, dv, 4 * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        if (!(std::fabs(S.hdots[1]) >= 1e-30)) break;                            // no update was made
        res.iterations = it + 1;
        res.finalRes = std::sqrt(S.hdots[2]) / r0;
        if (res.finalRes < tol) { res.converged = 1; break; }
        if (amg) {
            if ((e = vcycle_run<T>(H, r, z, s, G)) != hipSuccess) return e;
        }
        else hipLaunchKernelGGL(k_jac<T>, g, bl, 0, s, n, (const T*)A0.dg, (const T*)r, z);
        dot2<T>(n, r, z, (const T*)nullptr, (const T*)nullptr, S, dv + 3, s);      // rzNew
        hipLaunchKernelGGL(k_pcg_dir<T>, g, bl, 0, s, n, (const double*)dv, (const T*)z, p);
        hipLaunchKernelGGL(k_pcg_shift, dim3(1), dim3(64), 0, s, dv);
    }
    return hipGetLastError();
}

}  // namespace grid
}  // namespace rsg
