// rsg_geom.h -- mesh geometry on the device: the wall distance of every cell.
//
// Mesh::computeWallDistance (core/src/Mesh.cpp:345-373): for every cell the
// minimum Euclidean distance from its centre to any wall-face centre, brute
// force over all wall faces.  One thread per cell pair (two cells per
// thread), wall-face centres staged through LDS in tiles shared by the whole
// workgroup (every lane reads the same LDS word: a broadcast).  The distance
// is the reference's expression, Vec3::norm of (cell - wall) = sqrt(x*x +
// y*y + z*z) evaluated left to right with no fused multiply-add (the library
// is built with -ffp-contract=off); the square root is taken once, of the
// minimum square, which equals the minimum of the roots because a correctly
// rounded sqrt is monotone.  The result is bit-identical to the host loop.
//
// The running minimum is v_min_f64 (fmin), which selects exactly what the
// reference's `if (d < minD) minD = d` selects: a NaN distance keeps the old
// minimum in both, the minimum itself is never NaN, and squares are >= +0.
//
// Cost: 9 FP64 operations per (cell, wall face) pair and nothing from HBM but
// the centres once, so the kernel is bound by FP64 VALU issue (256 CUs x 64
// FP64 lanes per clock), not by memory.
#pragma once

#include <hip/hip_runtime.h>

namespace rsg {
namespace geom {

constexpr int kWallBlock = 256;
constexpr int kWallCellsPerThread = 2;
constexpr int kWallTile = 1024;

__global__ __launch_bounds__(kWallBlock) void k_wall_distance(int nC, const double* __restrict__ cc, int nW,
                                                              const double* __restrict__ wp,
                                                              double* __restrict__ out) {
    __shared__ double sx[kWallTile], sy[kWallTile], sz[kWallTile];
    const int base = blockIdx.x * (kWallBlock * kWallCellsPerThread) + threadIdx.x;
    double cx[kWallCellsPerThread], cy[kWallCellsPerThread], cz[kWallCellsPerThread], m[kWallCellsPerThread];
#pragma unroll
    for (int i = 0; i < kWallCellsPerThread; ++i) {
        const int c = base + i * kWallBlock;
        cx[i] = c < nC ? cc[3 * (size_t)c] : 0.0;
        cy[i] = c < nC ? cc[3 * (size_t)c + 1] : 0.0;
        cz[i] = c < nC ? cc[3 * (size_t)c + 2] : 0.0;
        m[i] = 1.7976931348623157e308;                     // std::numeric_limits<double>::max()
    }
    for (int t0 = 0; t0 < nW; t0 += kWallTile) {
        const int nt = min(kWallTile, nW - t0);
        __syncthreads();
        for (int j = threadIdx.x; j < nt; j += kWallBlock) {
            const size_t w = 3 * (size_t)(t0 + j);
            sx[j] = wp[w]; sy[j] = wp[w + 1]; sz[j] = wp[w + 2];
        }
        __syncthreads();
        for (int j = 0; j < nt; ++j) {
            const double wx = sx[j], wy = sy[j], wz = sz[j];
#pragma unroll
            for (int i = 0; i < kWallCellsPerThread; ++i) {
                const double dx = cx[i] - wx, dy = cy[i] - wy, dz = cz[i] - wz;
                const double d2 = dx * dx + dy * dy + dz * dz;
                m[i] = fmin(d2, m[i]);     // = (d2 < m ? d2 : m): no NaN, m never NaN, d2 >= +0
            }
        }
    }
#pragma unroll
    for (int i = 0; i < kWallCellsPerThread; ++i) {
        const int c = base + i * kWallBlock;
        if (c < nC) out[c] = sqrt(m[i]);
    }
}

inline int wall_blocks(int nC) {
    const int per = kWallBlock * kWallCellsPerThread;
    return (nC + per - 1) / per;
}

}  // namespace geom
}  // namespace rsg
