// rsg_dense.h -- explicit inverse of the AMG coarsest level on the FP64
// matrix cores (v_mfma_f64_16x16x4f64), and its application per V-cycle.
//
// The reference solves its coarsest level (<= 50 rows) with 50 Gauss-Seidel
// sweeps (core/include/LinearSolver.hpp:508-512, 557-562).  On the large
// pressure systems the reference's greedy aggregation stalls thousands of rows
// above that size (7,779 rows on the 261,600-cell backward-facing-step system),
// where sweeps no longer solve the level and dominate the V-cycle.  Here the
// coarsest level is inverted once per hierarchy set-up and every V-cycle
// applies the inverse as one matrix-vector product (x = C^-1 b): a dense
// solve, bandwidth-bound at n^2 doubles per application.
//
// Inversion: blocked Gauss-Jordan without pivoting (the coarsest Galerkin
// level of an SPD M-matrix is SPD, so every pivot block is an SPD Schur
// complement block), block width 64, on the matrix padded to N = 64 k rows
// with identity rows.  After the blocks S = {0..K-1} have been eliminated the
// matrix holds  [[A_SS^-1, A_SS^-1 A_SU], [-A_US A_SS^-1, Schur complement]]:
// tile (J, I) is +-tile (I, J)^T (minus across the S/U split), so only the
// tiles I >= J are stored and updated -- N^3 flops instead of 2 N^3.  Step K:
//   Y_J  = M_KJ    (from tile (K, J) for J < K, tile (J, K)^T for J > K)
//   C_J  = M_JK    (= -Y_J^T for J < K, Y_J^T for J > K)
//   R_J  = P Y_J,  R_K = P,  P = M_KK^-1                    k_gj_panel (MFMA)
//   M_IJ <- [J != K] M_IJ - C_I R_J   (I >= J, I != K),  M_KJ <- R_J (J <= K)
//                                                           k_gj_update (MFMA)
// The pivot block of step K + 1 is updated first, by workgroup 0 of step K's
// update, which then inverts it (Gauss-Jordan in LDS) while
// the other workgroups update the rest; after the last step k_gj_mirror
// copies the lower tiles to the upper ones.  A pivot that is not positive or
// not finite sets a flag the host reads after the last step (the caller then
// smooths the level instead).
//
// MFMA lane layout of v_mfma_f64_16x16x4f64 (one wave): lane l supplies
// A[l % 16][l / 16] of the 16 x 4 operand and B[l / 16][l % 16] of the 4 x 16
// operand; its four results are D[l / 16 + 4 i][l % 16], i = 0..3 (checked on
// the device against a host product).
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace rsg {
namespace dense {
// internal linkage: the header is included by more than one translation unit
namespace {

constexpr int kNB = 64;          // block width (pivot block, tile edge)
constexpr int kThreads = 256;    // four waves; each owns a 32 x 32 quadrant of a tile
constexpr int kLd = kNB + 1;     // LDS row stride in doubles (spreads the banks)

typedef double v4d __attribute__((ext_vector_type(4)));

inline int padded(int n) { return (n + kNB - 1) / kNB * kNB; }

// ------------------------------------------------------------- assembly
// A (N x N, row-major) = 0 with identity on the padding rows.
__global__ __launch_bounds__(kThreads) void k_dense_clear(int n, int N, double* __restrict__ A) {
    const size_t total = (size_t)N * N;
    for (size_t q = (size_t)blockIdx.x * kThreads + threadIdx.x; q < total; q += (size_t)gridDim.x * kThreads) {
        const int i = (int)(q / N), j = (int)(q % N);
        A[q] = (i == j && i >= n) ? 1.0 : 0.0;
    }
}

// rows of a CSR matrix (n rows) scattered into the dense image, one thread per row
__global__ __launch_bounds__(kThreads) void k_dense_scatter(int n, int N, const int* __restrict__ rp,
                                                            const int* __restrict__ ci, const double* __restrict__ v,
                                                            double* __restrict__ A) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    for (int e = rp[i]; e < rp[i + 1]; ++e) A[(size_t)i * N + ci[e]] += v[e];
}

// --------------------------------------------------------- pivot block
// P = G^-1 (64 x 64 in LDS, overwritten; P row-major in global memory) by
// Gauss-Jordan without pivoting, all threads of the workgroup (kThreads):
// thread t owns column t % 64 of rows t / 64 + 4 q.  Two barriers per step:
// every thread reads the pivot and its factors, then updates its entries.
// flag[0] = 1 if a pivot is not positive or not finite.
__device__ inline void block_gj(double (*G)[kLd], double* __restrict__ P, int* __restrict__ flag) {
    constexpr int RPT = kNB / (kThreads / kNB);     // rows per thread
    const int t = threadIdx.x, j = t % kNB, r0 = t / kNB;
    bool bad = false;
    for (int k = 0; k < kNB; ++k) {
        double piv = G[k][k];
        if (!(piv > 0.0) || !std::isfinite(piv)) { bad = true; piv = 1.0; }
        double fct[RPT];
#pragma unroll
        for (int q = 0; q < RPT; ++q) fct[q] = G[r0 + 4 * q][k];
        const double rowk = (j == k ? 1.0 : G[k][j]) / piv;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int i = r0 + 4 * q;
            G[i][j] = (i == k) ? rowk : (j == k ? 0.0 : G[i][j]) - fct[q] * rowk;
        }
        __syncthreads();
    }
    for (int q = t; q < kNB * kNB; q += kThreads) P[q] = G[q / kNB][q % kNB];
    if (bad && t == 0) flag[0] = 1;
}

// P = inverse of diagonal tile (K, K).
__global__ __launch_bounds__(kThreads) void k_gj_pivot(int N, int K, const double* __restrict__ A,
                                                       double* __restrict__ P, int* __restrict__ flag) {
    __shared__ double G[kNB][kLd];
    const size_t base = (size_t)K * kNB * N + (size_t)K * kNB;
    for (int q = threadIdx.x; q < kNB * kNB; q += kThreads)
        G[q / kNB][q % kNB] = A[base + (size_t)(q / kNB) * N + q % kNB];
    __syncthreads();
    block_gj(G, P, flag);
}

// ------------------------------------------------------------ MFMA tile
// acc (a wave's 32 x 32 quadrant at (wr, wc) of a 64 x 64 tile) += X Y with
// X = 64 x 64 held transposed in LDS (Xt[k][row]) and Y = 64 x 64 (Y[k][col]).
__device__ inline void mfma_tile(const double (*Xt)[kLd], const double (*Y)[kLd], int wr, int wc, int lane,
                                 v4d acc[2][2]) {
    const int r = lane % 16, kq = lane / 16;
#pragma unroll 4
    for (int k0 = 0; k0 < kNB; k0 += 4) {
        const double a0 = Xt[k0 + kq][wr * 32 + r];
        const double a1 = Xt[k0 + kq][wr * 32 + 16 + r];
        const double b0 = Y[k0 + kq][wc * 32 + r];
        const double b1 = Y[k0 + kq][wc * 32 + 16 + r];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
}

// lower-triangle tile t (row-major over I >= J) -> (I, J)
__device__ inline void tri_tile(long t, int& I, int& J) {
    int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((long)i * (i + 1) / 2 > t) --i;
    while ((long)(i + 1) * (i + 2) / 2 <= t) ++i;
    I = i;
    J = (int)(t - (long)i * (i + 1) / 2);
}

// ----------------------------------------------------------- panels
// Workgroup J: Y_J = M_KJ from the stored lower tile, CT[:, J] = C_J^T
// (64 x N, CT[k][J 64 + i] = M_JK[i][k]) and R[:, J] = P Y_J (R_K = P).
__global__ __launch_bounds__(kThreads) void k_gj_panel(int N, int K, const double* __restrict__ A,
                                                       const double* __restrict__ P, double* __restrict__ R,
                                                       double* __restrict__ CT) {
    __shared__ double Pt[kNB][kLd];
    __shared__ double Y[kNB][kLd];
    const int J = blockIdx.x, t = threadIdx.x;
    if (J == K) {
        for (int q = t; q < kNB * kNB; q += kThreads) {
            R[(size_t)(q / kNB) * N + K * kNB + q % kNB] = P[q];
            CT[(size_t)(q / kNB) * N + K * kNB + q % kNB] = 0.0;
        }
        return;
    }
    const double sign = J < K ? -1.0 : 1.0;
    for (int q = t; q < kNB * kNB; q += kThreads) {
        const int i = q / kNB, k = q % kNB;
        Pt[k][i] = P[q];                                                          // P[i][k]
        if (J < K) Y[i][k] = A[(size_t)(K * kNB + i) * N + J * kNB + k];         // tile (K, J)
        else Y[k][i] = A[(size_t)(J * kNB + i) * N + K * kNB + k];               // tile (J, K)^T
    }
    __syncthreads();
    for (int q = t; q < kNB * kNB; q += kThreads)
        CT[(size_t)(q / kNB) * N + J * kNB + q % kNB] = sign * Y[q / kNB][q % kNB];
    const int w = t / 64, lane = t % 64, wr = w / 2, wc = w % 2;
    v4d acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
    mfma_tile(Pt, Y, wr, wc, lane, acc);
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int i = 0; i < 4; ++i) {
                const int row = wr * 32 + a * 16 + lane / 16 + 4 * i, col = wc * 32 + b * 16 + lane % 16;
                R[(size_t)row * N + J * kNB + col] = acc[a][b][i];
            }
}

// ------------------------------------------------------- trailing update
// One workgroup per stored tile (I >= J).  Tile row K: M_KJ <- R_J.  Others:
// M_IJ <- [J != K] M_IJ - C_I R_J, operands read straight from CT and R in
// the MFMA lane layout (both coalesced: 16 consecutive doubles per k).
// Workgroup 0 takes tile (K + 1, K + 1) and then inverts it into Pn.
__global__ __launch_bounds__(kThreads) void k_gj_update(int N, int K, double* __restrict__ A,
                                                        const double* __restrict__ R, const double* __restrict__ CT,
                                                        double* __restrict__ Pn, int* __restrict__ flag) {
    __shared__ double G[kNB][kLd];
    const int nt = N / kNB;
    const bool fuse = K + 1 < nt;
    int I, J;
    if (fuse && blockIdx.x == 0) {
        I = J = K + 1;
    } else {
        long t = fuse ? (long)blockIdx.x - 1 : (long)blockIdx.x;
        if (fuse && t >= (long)(K + 1) * (K + 2) / 2 + (K + 1)) ++t;             // skip tile (K+1, K+1)
        tri_tile(t, I, J);
    }
    const int tid = threadIdx.x;
    if (I == K) {
        for (int q = tid; q < kNB * kNB; q += kThreads) {
            const int i = q / kNB, j = q % kNB;
            A[(size_t)(K * kNB + i) * N + J * kNB + j] = R[(size_t)i * N + J * kNB + j];
        }
        return;
    }
    const int w = tid / 64, lane = tid % 64, wr = w / 2, wc = w % 2;
    const int r = lane % 16, kq = lane / 16;
    const bool keep = J != K;
    // the old tile values first (their loads overlap the products)
    double old[2][2][4];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int i = 0; i < 4; ++i) {
                const int row = wr * 32 + a * 16 + kq + 4 * i, col = wc * 32 + b * 16 + r;
                old[a][b][i] = keep ? A[(size_t)(I * kNB + row) * N + J * kNB + col] : 0.0;
            }
    v4d acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
    const double* cI = CT + (size_t)I * kNB + wr * 32 + r;
    const double* rJ = R + (size_t)J * kNB + wc * 32 + r;
#pragma unroll
    for (int k0 = 0; k0 < kNB; k0 += 4) {
        const size_t o = (size_t)(k0 + kq) * N;
        const double a0 = cI[o], a1 = cI[o + 16], b0 = rJ[o], b1 = rJ[o + 16];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    const bool pivot = fuse && blockIdx.x == 0;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int i = 0; i < 4; ++i) {
                const int row = wr * 32 + a * 16 + kq + 4 * i, col = wc * 32 + b * 16 + r;
                const double v = old[a][b][i] - acc[a][b][i];
                A[(size_t)(I * kNB + row) * N + J * kNB + col] = v;
                if (pivot) G[row][col] = v;
            }
    if (!pivot) return;
    __syncthreads();
    block_gj(G, Pn, flag);
}

// Upper tiles from the lower ones (the inverse is symmetric): tile (I, J),
// I < J, = tile (J, I)^T.  One workgroup per upper tile.
__global__ __launch_bounds__(kThreads) void k_gj_mirror(int N, double* __restrict__ A) {
    __shared__ double T[kNB][kLd];
    int J, I;
    tri_tile(blockIdx.x, J, I);          // strictly-lower index space of (J - 1, I)
    ++J;
    for (int q = threadIdx.x; q < kNB * kNB; q += kThreads)
        T[q % kNB][q / kNB] = A[(size_t)(J * kNB + q / kNB) * N + I * kNB + q % kNB];
    __syncthreads();
    for (int q = threadIdx.x; q < kNB * kNB; q += kThreads)
        A[(size_t)(I * kNB + q / kNB) * N + J * kNB + q % kNB] = T[q / kNB][q % kNB];
}

// ------------------------------------------------------------- apply
// x = Ainv b on the first n rows (Ainv's padding columns are zero for those
// rows).  One wave per pair of rows: every lane reads two adjacent columns of
// both rows per 128-column chunk (16-byte loads, two chunks in flight), then
// the lanes' partial sums are combined by an xor butterfly -- a fixed order,
// so the product is the same on every call.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_dense_apply(int n, int N, const double* __restrict__ Ainv,
                                                          const T* __restrict__ b, T* __restrict__ x) {
    const int wave = blockIdx.x * (kThreads / 64) + threadIdx.x / 64, lane = threadIdx.x % 64;
    const int r0 = 2 * wave;
    if (r0 >= n) return;
    const int r1 = (r0 + 1 < n) ? r0 + 1 : r0;
    const double2* a0 = reinterpret_cast<const double2*>(Ainv + (size_t)r0 * N);
    const double2* a1 = reinterpret_cast<const double2*>(Ainv + (size_t)r1 * N);
    double s0 = 0.0, s1 = 0.0, t0 = 0.0, t1 = 0.0;
    int c = 0;
    for (; c + 256 <= n; c += 256) {
        const int j = c + 2 * lane, k = j + 128;
        const double bj0 = (double)b[j], bj1 = (double)b[j + 1], bk0 = (double)b[k], bk1 = (double)b[k + 1];
        const double2 p = a0[j / 2], q = a1[j / 2], u = a0[k / 2], v = a1[k / 2];
        s0 += p.x * bj0 + p.y * bj1;
        s1 += q.x * bj0 + q.y * bj1;
        t0 += u.x * bk0 + u.y * bk1;
        t1 += v.x * bk0 + v.y * bk1;
    }
    for (int j = c + lane; j < n; j += 64) {
        const double bj = (double)b[j];
        s0 += Ainv[(size_t)r0 * N + j] * bj;
        s1 += Ainv[(size_t)r1 * N + j] * bj;
    }
    s0 += t0;
    s1 += t1;
    for (int off = 32; off >= 1; off >>= 1) {
        s0 += __shfl_xor(s0, off, 64);
        s1 += __shfl_xor(s1, off, 64);
    }
    if (lane == 0) {
        x[r0] = (T)s0;
        if (r1 != r0) x[r1] = (T)s1;
    }
}

// grid size of k_dense_apply for n rows
inline int apply_blocks(int n) { return (n + 2 * (kThreads / 64) - 1) / (2 * (kThreads / 64)); }

__global__ void k_flag_clear(int* flag) { flag[0] = 0; }

// Inverts the dense SPD image A (N x N, padded) in place on stream s; R and CT
// (64 x N each), P (2 x 64 x 64) and flag (one int, cleared here) are device
// workspace.  The caller reads flag after the stream has finished.
inline void invert(int N, double* A, double* R, double* CT, double* P, int* flag, hipStream_t s) {
    const int nt = N / kNB;
    const int lower = nt * (nt + 1) / 2;
    hipLaunchKernelGGL(k_flag_clear, dim3(1), dim3(1), 0, s, flag);
    hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(kThreads), 0, s, N, 0, (const double*)A, P, flag);
    for (int K = 0; K < nt; ++K) {
        double* Pk = P + (K & 1) * kNB * kNB;
        double* Pn = P + ((K + 1) & 1) * kNB * kNB;
        hipLaunchKernelGGL(k_gj_panel, dim3(nt), dim3(kThreads), 0, s, N, K, (const double*)A, (const double*)Pk,
                           R, CT);
        hipLaunchKernelGGL(k_gj_update, dim3(lower), dim3(kThreads), 0, s, N, K, A, (const double*)R,
                           (const double*)CT, Pn, flag);
    }
    if (nt > 1) hipLaunchKernelGGL(k_gj_mirror, dim3(nt * (nt - 1) / 2), dim3(kThreads), 0, s, N, A);
}

}  // namespace
}  // namespace dense
}  // namespace rsg
