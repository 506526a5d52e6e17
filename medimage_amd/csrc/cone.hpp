// 3D cone-beam projectors of the industry3d geometry (flat panel, circular
// orbit): host side of cone.hip.
#pragma once

#include "runtime.hpp"

namespace mi {

// Geometry as the kernels read it (industry3d Parameter, Cone beam).
struct ConeDev {
    int nx, ny, nz, nviews, nrows, ncols, halfdet;
    float sample, pixel, sod, sdd;
};

// Per-geometry device tables: cosf / sinf of each view's angle (computed on
// the host with the C library, as the reference computes them) and, per
// detector row, w = dstZ - srcZ of the row's rays and 1 / w (double).
struct ConeTables {
    mi_cone_geometry g{};
    ConeDev d{};
    DevBuf<float> cs;      // [2 * nviews]: cos, sin interleaved
    DevBuf<double> rows;   // [2 * nrows]: w, 1 / w
    int64_t nnz = -1;      // traced entries of all rays (-1: not counted yet)
};

void validate_cone(const mi_cone_geometry &g);
// The context's tables for g (rebuilt when g differs from the cached one).
ConeTables &cone_tables(mi_context *ctx, const mi_cone_geometry &g);

void cone_trace(hipStream_t st, const ConeTables &t, int view, int row, int col, int64_t *ind, float *len,
                int cap, int *numb);
void cone_forward(hipStream_t st, const ConeTables &t, const float *vol, float *sino);
void cone_backproject(hipStream_t st, const ConeTables &t, const float *sino, float *vol);

// Traced entries of all rays (counted on the device once per geometry).
int64_t cone_nnz(hipStream_t st, ConeTables &t);
// Algorithmic bytes of one projection (voxel reads of every traced entry,
// sinogram written / read once, volume written / read once).
double cone_fp_bytes(const ConeTables &t);
double cone_bp_bytes(const ConeTables &t);

}  // namespace mi
