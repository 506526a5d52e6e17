// GoldenC stencils on row-major [nx][ny] images, shared by the parallel
// kernels (kernels.hip) and the reference-order kernels (ordered.hip).
#pragma once

#include "common.hpp"

namespace mi {

// a[i][j], or `pad` outside the image.
template <typename T>
__device__ __forceinline__ T gc_ld(const T *a, int i, int j, int nx, int ny, T pad) {
    return (i < 0 || j < 0 || i >= nx || j >= ny) ? pad : a[(int64_t)i * ny + j];
}

// Image-gradient regulariser at pixel (i, j), f zero padded and v one
// padded, in the expression order of origin/GoldenC/src/main.c:114-124.
template <typename T>
__device__ __forceinline__ double gc_reg(const T *u, const T *v, int i, int j, int nx, int ny, double eps) {
    const double z = (double)u[(int64_t)i * ny + j];
    const double zd = (double)gc_ld(u, i + 1, j, nx, ny, T(0)), zr = (double)gc_ld(u, i, j + 1, nx, ny, T(0));
    const double zu = (double)gc_ld(u, i - 1, j, nx, ny, T(0)), zl = (double)gc_ld(u, i, j - 1, nx, ny, T(0));
    const double p = (double)v[(int64_t)i * ny + j];
    const double pu = (double)gc_ld(v, i - 1, j, nx, ny, T(1)), pl = (double)gc_ld(v, i, j - 1, nx, ny, T(1));
    return +(p * p) * (zd - z) + (p * p) * (zr - z) - (pu * pu) * (z - zu) - (pl * pl) * (z - zl) +
           0.01 * (eps * eps) * (zd + zr - 4 * z + zu + zl);
}

}  // namespace mi
