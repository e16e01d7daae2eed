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
// with identity rows.  Step K (pivot block rows/cols [64K, 64K + 64)):
//   P      = A_KK^-1                         k_gj_pivot  (one workgroup, LDS)
//   R_J    = P A_KJ (J != K),  R_K = P        k_gj_panel  (MFMA, one workgroup per column tile)
//   C_I    = A_IK                             (copied by k_gj_panel)
//   A_IJ  <- [J != K] A_IJ - C_I R_J (I != K),  A_KJ <- R_J
//                                             k_gj_update (MFMA, one workgroup per 64 x 64 tile)
// which leaves A^-1 in place after N/64 steps: 2 N^3 flops, all but O(N^2 64)
// of them on MFMA.  A pivot that is not positive or not finite sets a flag the
// host reads after the last step (the caller then smooths the level instead).
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
// P = A_KK^-1 by Gauss-Jordan in LDS (no pivoting).  flag[0] is set to 1 if a
// pivot is not positive or not finite.
__global__ __launch_bounds__(kThreads) void k_gj_pivot(int N, int K, const double* __restrict__ A,
                                                       double* __restrict__ P, int* __restrict__ flag) {
    __shared__ double G[kNB][kLd];
    __shared__ double piv;
    const int t = threadIdx.x;
    const size_t base = (size_t)K * kNB * N + (size_t)K * kNB;
    for (int q = t; q < kNB * kNB; q += kThreads) G[q / kNB][q % kNB] = A[base + (size_t)(q / kNB) * N + q % kNB];
    __syncthreads();
    for (int k = 0; k < kNB; ++k) {
        if (t == 0) {
            piv = G[k][k];
            if (!(piv > 0.0) || !std::isfinite(piv)) { flag[0] = 1; piv = 1.0; }
        }
        __syncthreads();
        // column k of the other rows, read before row k is scaled
        double fct[kNB / (kThreads / kNB)];
        const int j = t % kNB, r0 = t / kNB;          // thread: column j, rows r0, r0 + 4, ...
        for (int q = 0; q < kNB / (kThreads / kNB); ++q) fct[q] = G[r0 + 4 * q][k];
        const double rowk = (j == k ? 1.0 : G[k][j]) / piv;
        __syncthreads();
        for (int q = 0; q < kNB / (kThreads / kNB); ++q) {
            const int i = r0 + 4 * q;
            if (i == k) G[i][j] = rowk;
            else G[i][j] = (j == k ? 0.0 : G[i][j]) - fct[q] * rowk;
        }
        __syncthreads();
    }
    for (int q = t; q < kNB * kNB; q += kThreads) P[q] = G[q / kNB][q % kNB];
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

// ----------------------------------------------------------- row panel
// Workgroup J (column tile J): R[:, J] = P A_KJ (J != K) or P (J == K), and
// C[J rows, :] = A[J rows, K cols].  R is 64 x N, C is N x 64, row-major.
__global__ __launch_bounds__(kThreads) void k_gj_panel(int N, int K, const double* __restrict__ A,
                                                       const double* __restrict__ P, double* __restrict__ R,
                                                       double* __restrict__ C) {
    __shared__ double Pt[kNB][kLd];
    __shared__ double Y[kNB][kLd];
    const int J = blockIdx.x, t = threadIdx.x;
    // C tile: rows of tile J, columns of the pivot block
    for (int q = t; q < kNB * kNB; q += kThreads) {
        const int i = q / kNB, k = q % kNB;
        C[(size_t)(J * kNB + i) * kNB + k] = A[(size_t)(J * kNB + i) * N + K * kNB + k];
    }
    if (J == K) {
        for (int q = t; q < kNB * kNB; q += kThreads) R[(size_t)(q / kNB) * N + K * kNB + q % kNB] = P[q];
        return;
    }
    for (int q = t; q < kNB * kNB; q += kThreads) {
        const int i = q / kNB, k = q % kNB;
        Pt[k][i] = P[q];                                              // P[i][k]
        Y[i][k] = A[(size_t)(K * kNB + i) * N + J * kNB + k];         // A_KJ[i][k]
    }
    __syncthreads();
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
// Workgroup (I, J): A_IJ <- [J != K] A_IJ - C_I R_J for I != K; A_KJ <- R_J.
__global__ __launch_bounds__(kThreads) void k_gj_update(int N, int K, double* __restrict__ A,
                                                        const double* __restrict__ R, const double* __restrict__ C) {
    __shared__ double Xt[kNB][kLd];
    __shared__ double Y[kNB][kLd];
    const int nt = N / kNB;
    const int I = blockIdx.x / nt, J = blockIdx.x % nt, t = threadIdx.x;
    if (I == K) {
        for (int q = t; q < kNB * kNB; q += kThreads) {
            const int i = q / kNB, j = q % kNB;
            A[(size_t)(K * kNB + i) * N + J * kNB + j] = R[(size_t)i * N + J * kNB + j];
        }
        return;
    }
    for (int q = t; q < kNB * kNB; q += kThreads) {
        const int i = q / kNB, k = q % kNB;
        Xt[k][i] = C[(size_t)(I * kNB + i) * kNB + k];               // C_I[i][k]
        Y[i][k] = R[(size_t)i * N + J * kNB + k];                     // R_J[i][k]
    }
    __syncthreads();
    const int w = t / 64, lane = t % 64, wr = w / 2, wc = w % 2;
    v4d acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
    mfma_tile(Xt, Y, wr, wc, lane, acc);
    const bool keep = J != K;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int i = 0; i < 4; ++i) {
                const int row = wr * 32 + a * 16 + lane / 16 + 4 * i, col = wc * 32 + b * 16 + lane % 16;
                double* dst = A + (size_t)(I * kNB + row) * N + J * kNB + col;
                *dst = (keep ? *dst : 0.0) - acc[a][b][i];
            }
}

// ------------------------------------------------------------- apply
// x = Ainv b on the first n rows, one wave per row, FP64 sums in a fixed
// order (lane-strided partial sums, then an xor butterfly).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_dense_apply(int n, int N, const double* __restrict__ Ainv,
                                                          const T* __restrict__ b, T* __restrict__ x) {
    const int row = blockIdx.x * (kThreads / 64) + threadIdx.x / 64, lane = threadIdx.x % 64;
    if (row >= n) return;
    const double* a = Ainv + (size_t)row * N;
    double s = 0.0;
    for (int j = lane; j < n; j += 64) s += a[j] * (double)b[j];
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) x[row] = (T)s;
}

__global__ void k_flag_clear(int* flag) { flag[0] = 0; }

// Inverts the dense SPD image A (N x N, padded) in place on stream s; R
// (64 x N), C (N x 64), P (64 x 64) and flag (one int, cleared here) are
// device workspace.  The caller reads flag after the stream has finished.
inline void invert(int N, double* A, double* R, double* C, double* P, int* flag, hipStream_t s) {
    const int nt = N / kNB;
    hipLaunchKernelGGL(k_flag_clear, dim3(1), dim3(1), 0, s, flag);
    for (int K = 0; K < nt; ++K) {
        hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(kThreads), 0, s, N, K, (const double*)A, P, flag);
        hipLaunchKernelGGL(k_gj_panel, dim3(nt), dim3(kThreads), 0, s, N, K, (const double*)A, (const double*)P, R, C);
        hipLaunchKernelGGL(k_gj_update, dim3(nt * nt), dim3(kThreads), 0, s, N, K, A, (const double*)R,
                           (const double*)C);
    }
}

}  // namespace
}  // namespace dense
}  // namespace rsg
