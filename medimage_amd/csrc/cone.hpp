// 3D cone-beam projectors of the industry3d geometry (flat panel, circular
// orbit): host side of cone.hip.
#pragma once

#include "runtime.hpp"

namespace mi {

// Geometry as the kernels read it (industry3d Parameter, Cone beam).
struct ConeDev {
    int nx, ny, nz, nviews, nrows, ncols, halfdet;
    float sample, pixel, sod, sdd;
    int parallel;      // BEAM "Parallel"
    float halfsize;    // HALFSIZE = NX / 2.0 (float), the parallel beam's centre on every axis
};

// Per-geometry device tables: cosf / sinf of each view's angle (computed on
// the host with the C library, as the reference computes them) and, per
// detector row, w = dstZ - srcZ of the row's rays and 1 / w (double; for the
// parallel beam the rows' constant z instead).
struct ConeTables {
    mi_cone_geometry g{};
    ConeDev d{};
    DevBuf<float> cs;      // [2 * nviews]: cos, sin interleaved
    DevBuf<double> rows;   // [2 * nrows]: w, 1 / w
    DevBuf<int> all;       // 0 .. nviews - 1 (the view list of a full projection)
    DevBuf<double2> vinv;  // [nviews * ncols]: 1 / (dx - sx), 1 / (dy - sy) of the column's rays
    DevBuf<float> vl2;     // [nviews * ncols]: (dx - sx)^2 + (dy - sy)^2
    DevBuf<float4> vsrc;   // [nviews]: source point (cone beam)
    DevBuf<float2> vpos;   // [nviews * ncols]: (x, y) of the column's rays' start (parallel beam)
    int64_t nnz = -1;      // traced entries of all rays (-1: not counted yet)
};

void validate_cone(const mi_cone_geometry &g);
// The context's tables for g (rebuilt when g differs from the cached one).
ConeTables &cone_tables(mi_context *ctx, const mi_cone_geometry &g);

void cone_trace(hipStream_t st, const ConeTables &t, int view, int row, int col, int64_t *ind, float *len,
                int cap, int *numb);
void cone_forward(hipStream_t st, const ConeTables &t, const float *vol, float *sino);
void cone_backproject(hipStream_t st, const ConeTables &t, const float *sino, float *vol);

// Forward projection over the device view list views[0 .. nv):
//   CFP_WRITE: out = A vol;
//   CFP_RESID: out = meas - A vol (float), part[i] = sum of out^2 over the
//              rays of workgroup i;
//   CFP_SQ:    part[i] = sum of (A vol)^2, nothing written.
// Returns the number of partial sums written (part needs that many).
enum { CFP_WRITE = 0, CFP_RESID = 1, CFP_SQ = 2 };
size_t cone_fp_views(hipStream_t st, const ConeTables &t, const int *views, int nv, int mode, const float *vol,
                     const float *meas, float *out, double *part);
// Gather back projection over the view list (overwrites vol).
// device_scratch >= 0: carry the sums in that device's fixed scratch array
// (no allocation, no wait; volumes up to cone_scratch_voxels()) instead of
// a buffer allocated for the call.
void cone_bp_views(hipStream_t st, const ConeTables &t, const int *views, int nv, const float *sino, float *vol,
                   int device_scratch = -1);
// Looks up the device address of the scratch array of `device` (current).
void cone_scratch_init(int device);
int64_t cone_scratch_voxels();

// Ordered-subset AT reconstruction of the volume (cone_at.hip): sino, u and
// v (initial values in, results out) on the host, or on the device when dev.
void cone_at_solve(mi_context *ctx, const mi_cone_geometry &g, const mi_at_params &p, const float *sino,
                   float *u, float *v, mi_report *rep, bool dev = false);

// Traced entries of all rays (counted on the device once per geometry).
int64_t cone_nnz(hipStream_t st, ConeTables &t);
// Algorithmic bytes of one projection (voxel reads of every traced entry,
// sinogram written / read once, volume written / read once).
double cone_fp_bytes(const ConeTables &t);
double cone_bp_bytes(const ConeTables &t);

}  // namespace mi
