// Device ray tracers.  Each walks one ray and hands every (pixel, length)
// pair to an emit functor; the count pass and the fill pass of the system
// matrix use the same walk, so the two passes cannot disagree.
//
// The arithmetic is written in the reference's own types and operation
// order and the whole library is compiled with -ffp-contract=off, which
// keeps the index/length sequences bit-identical to a strict-IEEE build of
// the reference tracers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace mi {

// ---------------------------------------------------------------------------
// Parallel beam, incremental tracer of msbeam/CPU_OpenMP/src/proj.c:67-211
// (= blockrecon/src/wray.cpp:23-167), posit() of proj.c:54-65 with its sin/cos
// supplied per view by the host (libm, in the reference's type).  CT is the
// type of the tolerance constants: double for msbeam (proto.h:16-17), float
// for blockrecon (config.h:35-36); mixed expressions promote as they do
// there.
template <typename CT, typename Emit>
__device__ __forceinline__ void trace_parallel(float s_view, float c_view, int nr, int nray, int n, Emit &&emit) {
    const CT F_TOL = (CT)1e-8;
    const CT GRID_TOL = (CT)1e-2;
    float px, py, cf, sf, rcos, tanp, rtan, c, fp, yc;
    int ix, iy, incx, incy, idx, flip, swapxy;
    bool up;

    {
        float off = (float)(nr - nray / 2) * 1;
        px = -s_view * off;
        py = c_view * off;
        cf = c_view;
        sf = s_view;
    }
    if (sf < -F_TOL) { sf = -sf; cf = -cf; }
    flip = 0;
    if (sf * cf < -F_TOL * F_TOL) { flip = 1; cf = -cf; }
    if (flip > 0) px = -px;
    if ((CT)fabsf(sf) <= F_TOL) { if (sf < 0) sf = -F_TOL * F_TOL; else sf = F_TOL * F_TOL; }
    if ((CT)fabsf(cf) <= F_TOL) { if (cf < 0) cf = -F_TOL * F_TOL; else cf = F_TOL * F_TOL; }
    if ((CT)fabsf(cf) > F_TOL) tanp = sf / cf; else tanp = 1 / (F_TOL * F_TOL);
    if ((CT)fabsf(sf) > F_TOL) rtan = cf / sf; else rtan = 1 / (F_TOL * F_TOL);

    swapxy = 0;
    if (tanp > 1 + F_TOL) {
        swapxy = 1;
        tanp = 1 / tanp;
        rtan = 1 / rtan;
        float t = py; py = px; px = t;
    }
    // The reference's double sqrt rounded to float (double sqrt is correctly
    // rounded on the device; HIP's __fsqrt_rn is not: it carries a 3-ulp
    // accuracy annotation and lowers to the bare hardware sqrt).
    rcos = (float)sqrt((double)(1 + tanp * tanp));
    py = (float)((double)py + n / 2.0);
    px = (float)((double)px + n / 2.0);
    c = py - px * tanp;

    if ((c - n) >= GRID_TOL) return;
    if (c < -GRID_TOL) {
        c = -c * rtan;
        if ((c - n) >= GRID_TOL) return;
        ix = (int)c;
        iy = 0;
        fp = ((float)(ix + 1)) - c;
        yc = fp * tanp;
        yc = (float)((double)yc + -1.0);
    } else {
        ix = 0;
        iy = (int)c;
        fp = ((float)(iy + 1)) - c;
        yc = -fp;
        yc += tanp;
    }
    up = (yc >= 0);

    if (swapxy > 0) {
        idx = iy + ix * n; incx = n; incy = 1;
        if (flip > 0) { idx = n - 1 - iy + ix * n; incx = n; incy = -1; }
    } else {
        idx = iy * n + ix; incx = 1; incy = n;
        if (flip > 0) { idx = n - 1 - ix + iy * n; incx = -1; incy = n; }
    }

    for (;;) {
        if (up) {
            if (iy + 1 >= n || ix < 0) return;
            float w = rcos * (1 - yc * rtan);
            emit(idx, (double)w);
            iy++; idx += incy;
            if (ix + 1 >= n) return;
            w = rcos * (yc * rtan);
            emit(idx, (double)w);
            ix++; idx += incx;
            yc = (float)((double)yc + ((double)tanp - 1.0));
        } else {
            if (iy >= n || ix + 1 >= n || ix < 0) return;
            emit(idx, (double)rcos);
            ix++; idx += incx;
            yc += tanp;
        }
        up = (yc >= 0);
    }
}

// ---------------------------------------------------------------------------
// Parametric Siddon in double, helical/rebin/src/tracing.cpp:22-147.  Image
// spans [0, nx*vx] x [0, ny*vy]; pixel (x, y) -> x*ny + y.  The walk stops
// 5e-2 before the exit (tracing.cpp:118).  The grid test inside the loop only
// triggers for rays through a grid corner, where the reference would index
// outside its arrays.
__device__ __forceinline__ double sd_cross(double p, double s, double d, double L) { return L * (p - s) / (d - s); }

// Walk state of the parametric Siddon tracer: lam is the current crossing
// parameter, ax / ay the next x / y crossings, (cx, cy) the current pixel.
struct SiddonState {
    double lam, ax, ay, stx, sty, hi;
    int cx, cy, sgx, sgy;
};

// Entry computation of the walk; false when the ray misses the image.
__device__ __forceinline__ bool siddon_init(double sx, double sy, double dx, double dy, int nx, int ny, double vx,
                                            double vy, SiddonState &st) {
    double rx = dx - sx, ry = dy - sy;
    double L = sqrt(rx * rx + ry * ry);
    double lo = 0.0, hi = L, a0, a1, t, ax, ay, lam;
    int sgx = (rx > 0) ? 1 : -1, sgy = (ry > 0) ? 1 : -1;
    double arx = fabs(rx), ary = fabs(ry);
    double stx = (arx > 1.e-4) ? (L / arx) * vx : 1.e6;
    double sty = (ary > 1.e-4) ? (L / ary) * vy : 1.e6;
    int entry = 0, cx, cy;

    a0 = sd_cross(0, sx, dx, L);
    a1 = sd_cross(nx * vx, sx, dx, L);
    t = (a0 < a1) ? a0 : a1;
    lo = (lo > t) ? lo : t;
    if (lo == t) entry = 1;
    t = (a0 > a1) ? a0 : a1;
    hi = (hi < t) ? hi : t;

    a0 = sd_cross(0, sy, dy, L);
    a1 = sd_cross(ny * vy, sy, dy, L);
    t = (a0 < a1) ? a0 : a1;
    lo = (lo > t) ? lo : t;
    if (lo == t) entry = 2;
    t = (a0 > a1) ? a0 : a1;
    hi = (hi < t) ? hi : t;

    if (lo >= hi) return false;
    lam = lo;
    if (entry == 1) {
        cx = (sgx == 1) ? 0 : nx - 1;
        ax = lam + stx;
        cy = (sy + lam * ry / L) / vy;
        double py = cy * vy;
        ay = (ary < 1.e-4) ? 1.e6 : sd_cross(py + (sgy > 0) * vy, sy, dy, L);
    } else {
        cy = (sgy == 1) ? 0 : ny - 1;
        ay = lam + sty;
        cx = (sx + lam * rx / L) / vx;
        double px = cx * vx;
        ax = (arx < 1.e-4) ? 1.e6 : sd_cross(px + (sgx > 0) * vx, sx, dx, L);
    }
    st = SiddonState{lam, ax, ay, stx, sty, hi, cx, cy, sgx, sgy};
    return true;
}

// One crossing of the walk: the current pixel and its intersection length.
// Same operations, in the same order, as the loop of the reference tracer.
__device__ __forceinline__ void siddon_step(SiddonState &st, int &cx, int &cy, double &len) {
    cx = st.cx;
    cy = st.cy;
    if (st.ax <= st.ay) {
        len = st.ax - st.lam;
        st.lam = st.ax;
        st.ax += st.stx;
        st.cx += st.sgx;
    } else {
        len = st.ay - st.lam;
        st.lam = st.ay;
        st.ay += st.sty;
        st.cy += st.sgy;
    }
}

template <typename Emit>
__device__ __forceinline__ void trace_siddon2d(double sx, double sy, double dx, double dy,
                                               int nx, int ny, double vx, double vy, Emit &&emit) {
    SiddonState st;
    if (!siddon_init(sx, sy, dx, dy, nx, ny, vx, vy, st)) return;
    while (st.lam < st.hi - 5.e-2) {
        if (st.cx < 0 || st.cx >= nx || st.cy < 0 || st.cy >= ny) return;
        int cx, cy;
        double len;
        siddon_step(st, cx, cy, len);
        emit(cx * ny + cy, len);
    }
}

// Flat-detector fan geometry: industry3d/src/tracing.cpp:51-81 (cone(), on
// the mid-plane) with rotate_2d of industry3d/src/utility.cpp:82-98, in
// double and millimetres, image corner at the origin (as
// helical/rebin/src/ms2d.cpp:172-183 places its arc-detector rays).
struct FanGeom {
    int nx, ny, nbins;
    double vx, vy, bin_mm, sod, sdd, det_offset;
    int arc;                  // 1: equiangular detector (arc_endpoints)
};

__device__ __forceinline__ void fan_rotate(double &x, double &y, double c, double s) {
    double r0 = 0.0, r1 = 0.0;
    r0 += c * x; r1 += s * x;
    r0 += -s * y; r1 += c * y;
    x = r0; y = r1;
}

// Source and detector points of fan ray (view with cos c / sin s, bin),
// shifted to the image-corner origin.
__device__ __forceinline__ void fan_endpoints(const FanGeom &g, double c, double s, int bin, double &sx, double &sy,
                                              double &dx, double &dy) {
    double x0 = 0.0, y0 = -g.sod;
    double x1 = ((double)(bin - g.nbins / 2) + 0.5 + g.det_offset) * g.bin_mm;
    double y1 = g.sdd - g.sod;
    double cx = g.nx * g.vx / 2.0, cy = g.ny * g.vy / 2.0;
    fan_rotate(x0, y0, c, s);
    fan_rotate(x1, y1, c, s);
    sx = x0 + cx;
    sy = y0 + cy;
    dx = x1 + cx;
    dy = y1 + cy;
}

// Arc-detector fan geometry, helical/rebin/src/ms2d.cpp:172-183: source at
// angle (cos c, sin s) on the circle of radius sod around the image centre,
// detector element at angle (dc, ds) seen from the source, distance sdd
// (the image-centre term uses nx for both coordinates, as the reference
// does; arc images are square).
__device__ __forceinline__ void arc_endpoints(const FanGeom &g, double c, double s, double dc, double ds,
                                              double &sx, double &sy, double &dx, double &dy) {
    sx = g.sod * c + g.nx * g.vx / 2.0;
    sy = g.sod * s + g.nx * g.vy / 2.0;
    dx = g.sdd * (dc * c - ds * s) + sx;
    dy = g.sdd * (ds * c + dc * s) + sy;
}

// Endpoints of fan ray (view with cos c / sin s, bin) of either detector;
// dcos / dsin: per-bin tables of the arc detector (unused when flat).
__device__ __forceinline__ void ray_endpoints(const FanGeom &g, double c, double s, const double *dcos,
                                              const double *dsin, int bin, double &sx, double &sy, double &dx,
                                              double &dy) {
    if (g.arc) arc_endpoints(g, c, s, dcos[bin], dsin[bin], sx, sy, dx, dy);
    else fan_endpoints(g, c, s, bin, sx, sy, dx, dy);
}

template <typename Emit>
__device__ __forceinline__ void trace_fan(const FanGeom &g, double c, double s, const double *dcos,
                                          const double *dsin, int bin, Emit &&emit) {
    double sx, sy, dx, dy;
    ray_endpoints(g, c, s, dcos, dsin, bin, sx, sy, dx, dy);
    trace_siddon2d(sx, sy, dx, dy, g.nx, g.ny, g.vx, g.vy, emit);
}

}  // namespace mi
