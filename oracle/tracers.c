/*
 * Oracle tracers and projectors (test infrastructure only; see oracle.h).
 * Built with -O2 -ffp-contract=off -fno-fast-math: every float/double
 * operation below rounds as written, which is what makes the index/length
 * sequences comparable bit for bit with the GPU tracers.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* The tracer body is instantiated twice (wray_body.inc):
 *  flavour 0, msbeam: double constants and double sin/cos,
 *             msbeam/CPU_OpenMP/src/proto.h:16-19, proj.c:55-56;
 *  flavour 1, blockrecon: float constants (blockrecon/src/config.h:35-38)
 *             and the float overloads of sin/cos/fabs selected by
 *             `using namespace std` (blockrecon/src/wray.cpp:1-12).
 * blockrecon/phantom/std_g.dat was written by flavour 1. */
#define WR_NAME wray_msbeam
#define WR_CT double
#define WR_SIN(x) sin(x)
#define WR_COS(x) cos(x)
#define WR_FABS(x) fabs(x)
#define WR_PI_V 3.14159265359
#include "wray_body.inc"
#undef WR_NAME
#undef WR_CT
#undef WR_SIN
#undef WR_COS
#undef WR_FABS
#undef WR_PI_V

#define WR_NAME wray_block
#define WR_CT float
#define WR_SIN(x) sinf(x)
#define WR_COS(x) cosf(x)
#define WR_FABS(x) fabsf(x)
#define WR_PI_V 3.14159265359f
#include "wray_body.inc"

void or_wray(int np, int nr, int nproj, int nray, int n, int flavour,
             int *line, float *weight, int *numb, float *snorm)
{
    if (flavour == 1) wray_block(np, nr, nproj, nray, n, line, weight, numb, snorm);
    else wray_msbeam(np, nr, nproj, nray, n, line, weight, numb, snorm);
}

void or_wray_table(int nproj, int flavour, float *sin_t, float *cos_t)
{
    for (int np = 0; np < nproj; ++np) {
        if (flavour == 1) {
            const float pi = 3.14159265359f;
            sin_t[np] = sinf(pi / nproj * np);
            cos_t[np] = cosf(pi / nproj * np);
        } else {
            sin_t[np] = sin(3.14159265359 / nproj * np);
            cos_t[np] = cos(3.14159265359 / nproj * np);
        }
    }
}

void or_par_fp_f32(float *g, const float *f, int nproj, int nray, int n, int flavour)
{
    int *line = (int *)malloc(sizeof(int) * 2 * n);
    float *wt = (float *)malloc(sizeof(float) * 2 * n);
    for (int np = 0; np < nproj; ++np)
        for (int nr = 0; nr < nray; ++nr) {
            int numb; float sn; float acc = 0;
            or_wray(np, nr, nproj, nray, n, flavour, line, wt, &numb, &sn);
            for (int i = 0; i < numb; ++i) acc += f[line[i]] * wt[i];
            g[np * nray + nr] = acc;
        }
    free(line); free(wt);
}

void or_par_bp_f32(float *f, const float *g, int nproj, int nray, int n, int flavour)
{
    int *line = (int *)malloc(sizeof(int) * 2 * n);
    float *wt = (float *)malloc(sizeof(float) * 2 * n);
    memset(f, 0, sizeof(float) * (size_t)n * n);
    for (int np = 0; np < nproj; ++np)
        for (int nr = 0; nr < nray; ++nr) {
            int numb; float sn;
            or_wray(np, nr, nproj, nray, n, flavour, line, wt, &numb, &sn);
            for (int i = 0; i < numb; ++i) f[line[i]] += g[np * nray + nr] * wt[i];
        }
    free(line); free(wt);
}

/* ------------------------------------------------------------------------ */
/* 2D parametric Siddon, helical/rebin/src/tracing.cpp:22-147.              */

#define SD_MAX(a, b) (((a) > (b)) ? (a) : (b))
#define SD_MIN(a, b) (((a) < (b)) ? (a) : (b))

/* Ray parameter (mm from the source) where the ray meets coordinate p. */
static double sd_cross(double p, double s, double d, double L) { return L * (p - s) / (d - s); }

int or_siddon2d(double sx, double sy, double dx, double dy,
                int nx, int ny, double vx, double vy,
                int *ind, double *wgt, int cap)
{
    double rx = dx - sx, ry = dy - sy;
    double L = sqrt(rx * rx + ry * ry);
    double lo = 0.0, hi = L, a0, a1, t, ax, ay, lam;
    int sgx = (rx > 0) ? 1 : -1, sgy = (ry > 0) ? 1 : -1;
    double arx = fabs(rx), ary = fabs(ry);
    double stx = (arx > 1.e-4) ? (L / arx) * vx : 1.e6;
    double sty = (ary > 1.e-4) ? (L / ary) * vy : 1.e6;
    int entry = 0, cx, cy, numb = 0;

    a0 = sd_cross(0, sx, dx, L);
    a1 = sd_cross(nx * vx, sx, dx, L);
    t = SD_MIN(a0, a1);
    lo = SD_MAX(lo, t);
    if (lo == t) entry = 1;
    t = SD_MAX(a0, a1);
    hi = SD_MIN(hi, t);

    a0 = sd_cross(0, sy, dy, L);
    a1 = sd_cross(ny * vy, sy, dy, L);
    t = SD_MIN(a0, a1);
    lo = SD_MAX(lo, t);
    if (lo == t) entry = 2;
    t = SD_MAX(a0, a1);
    hi = SD_MIN(hi, t);

    if (lo >= hi) return 0;
    lam = lo;
    if (entry == 1) {
        double py;
        cx = (sgx == 1) ? 0 : nx - 1;
        ax = lam + stx;
        cy = (sy + lam * ry / L) / vy;
        py = cy * vy;
        ay = (ary < 1.e-4) ? 1.e6 : sd_cross(py + (sgy > 0) * vy, sy, dy, L);
    } else {
        double px;
        cy = (sgy == 1) ? 0 : ny - 1;
        ay = lam + sty;
        cx = (sx + lam * rx / L) / vx;
        px = cx * vx;
        ax = (arx < 1.e-4) ? 1.e6 : sd_cross(px + (sgx > 0) * vx, sx, dx, L);
    }

    while (lam < hi - 5.e-2) {
        if (cx < 0 || cx >= nx || cy < 0 || cy >= ny || numb >= cap) break;
        if (ax <= ay) {
            ind[numb] = cx * ny + cy;
            wgt[numb] = ax - lam;
            ++numb;
            lam = ax; ax += stx; cx += sgx;
        } else {
            ind[numb] = cx * ny + cy;
            wgt[numb] = ay - lam;
            ++numb;
            lam = ay; ay += sty; cy += sgy;
        }
    }
    return numb;
}

void or_fan_tables(const or_fan_geom *g, double *cos_t, double *sin_t)
{
    for (int a = 0; a < g->nviews; ++a) {
        if (g->arc) {
            /* helical/rebin/src/ms2d.cpp:44-46 (angle in degrees, from the
             * slice's table when there is one) */
            double angle = g->view_deg ? g->view_deg[a] : g->span_deg * (double)a / (double)g->nviews;
            cos_t[a] = cos(angle / 180.0 * 3.14159265358979323846);
            sin_t[a] = sin(angle / 180.0 * 3.14159265358979323846);
            continue;
        }
        const double span = g->span_deg * (3.14159265358979323846 / 180.0);
        double th = span * (double)a / (double)g->nviews;
        cos_t[a] = cos(th);
        sin_t[a] = sin(th);
    }
}

/* Arc detector element `bin` (j = bin - nbins/2) seen from the source,
 * helical/rebin/src/ms2d.cpp:53-57, and the ray endpoints of
 * ms2d.cpp:172-183 (source on the circle of radius SOURCE_TO_ISO around the
 * image centre NX*v/2, the detector SOURCE_TO_DET away from the source). */
static void arc_endpoints(const or_fan_geom *g, double c, double s, int bin,
                          double *sx, double *sy, double *dx, double *dy)
{
    const double pi = 3.14159265358979323846;
    double lpd = g->bin_mm * pi / 180.;
    int j = bin - g->nbins / 2;
    double dc = cos((j + g->det_offset) * lpd + pi);
    double ds = sin((j + g->det_offset) * lpd + pi);
    double x0 = g->sod * c + g->nx * g->vx / 2.0;
    double y0 = g->sod * s + g->nx * g->vy / 2.0;
    *sx = x0;
    *sy = y0;
    *dx = g->sdd * (dc * c - ds * s) + x0;
    *dy = g->sdd * (ds * c + dc * s) + y0;
}

/* rotate_2d of industry3d/src/utility.cpp:82-98 in double: the 2x2 product
 * accumulated from zero, column by column. */
static void fan_rotate(double *x, double *y, double c, double s)
{
    double r0 = 0.0, r1 = 0.0;
    r0 += c * *x; r1 += s * *x;
    r0 += -s * *y; r1 += c * *y;
    *x = r0; *y = r1;
}

void or_fan_endpoints(const or_fan_geom *g, const double *cos_t, const double *sin_t,
                      int view, int bin, double *sx, double *sy, double *dx, double *dy)
{
    double c = cos_t[view], s = sin_t[view];
    if (g->arc) {
        arc_endpoints(g, c, s, bin, sx, sy, dx, dy);
        return;
    }
    double x0 = 0.0, y0 = -g->sod;
    double x1 = ((double)(bin - g->nbins / 2) + 0.5 + g->det_offset) * g->bin_mm;
    double y1 = g->sdd - g->sod;
    double cx = g->nx * g->vx / 2.0, cy = g->ny * g->vy / 2.0;
    fan_rotate(&x0, &y0, c, s);
    fan_rotate(&x1, &y1, c, s);
    *sx = x0 + cx; *sy = y0 + cy;
    *dx = x1 + cx; *dy = y1 + cy;
}

int or_fan_ray(const or_fan_geom *g, const double *cos_t, const double *sin_t,
               int view, int bin, int *ind, double *wgt, int cap)
{
    double sx, sy, dx, dy;
    or_fan_endpoints(g, cos_t, sin_t, view, bin, &sx, &sy, &dx, &dy);
    return or_siddon2d(sx, sy, dx, dy, g->nx, g->ny, g->vx, g->vy, ind, wgt, cap);
}

static int fan_cap(const or_fan_geom *g) { return 2 * (g->nx + g->ny) + 8; }

void or_fan_fp(const or_fan_geom *g, const int *views, int nv, const double *img, double *sino)
{
    int cap = fan_cap(g);
    int *ind = (int *)malloc(sizeof(int) * cap);
    double *wgt = (double *)malloc(sizeof(double) * cap);
    double *ct = (double *)malloc(sizeof(double) * g->nviews);
    double *st = (double *)malloc(sizeof(double) * g->nviews);
    or_fan_tables(g, ct, st);
    for (int k = 0; k < nv; ++k)
        for (int b = 0; b < g->nbins; ++b) {
            int n = or_fan_ray(g, ct, st, views[k], b, ind, wgt, cap);
            double acc = 0.0;
            for (int i = 0; i < n; ++i) acc += img[ind[i]] * wgt[i];
            sino[(size_t)views[k] * g->nbins + b] = acc;
        }
    free(ind); free(wgt); free(ct); free(st);
}

void or_fan_bp(const or_fan_geom *g, const int *views, int nv, const double *sino, double *img)
{
    int cap = fan_cap(g);
    int *ind = (int *)malloc(sizeof(int) * cap);
    double *wgt = (double *)malloc(sizeof(double) * cap);
    double *ct = (double *)malloc(sizeof(double) * g->nviews);
    double *st = (double *)malloc(sizeof(double) * g->nviews);
    or_fan_tables(g, ct, st);
    memset(img, 0, sizeof(double) * (size_t)g->nx * g->ny);
    for (int k = 0; k < nv; ++k)
        for (int b = 0; b < g->nbins; ++b) {
            int n = or_fan_ray(g, ct, st, views[k], b, ind, wgt, cap);
            double r = sino[(size_t)views[k] * g->nbins + b];
            for (int i = 0; i < n; ++i) img[ind[i]] += r * wgt[i];
        }
    free(ind); free(wgt); free(ct); free(st);
}

int64_t or_fan_nnz(const or_fan_geom *g)
{
    int cap = fan_cap(g);
    int *ind = (int *)malloc(sizeof(int) * cap);
    double *wgt = (double *)malloc(sizeof(double) * cap);
    double *ct = (double *)malloc(sizeof(double) * g->nviews);
    double *st = (double *)malloc(sizeof(double) * g->nviews);
    int64_t total = 0;
    or_fan_tables(g, ct, st);
    for (int a = 0; a < g->nviews; ++a)
        for (int b = 0; b < g->nbins; ++b) total += or_fan_ray(g, ct, st, a, b, ind, wgt, cap);
    free(ind); free(wgt); free(ct); free(st);
    return total;
}

/* ------------------------------------------------------------------------ */
/* GoldenC symmetric tiled DDA, origin/GoldenC/src/proj.c:14-397.           */
/* One traversal per (angle 1..45, tile, ray 0..383) feeds up to eight      */
/* (view, bin) pairs by the 90-degree rotations and the mirror of the grid. */

#define GC_N OR_GC_N
#define GC_NB OR_GC_BINS
#define GC_MID 383
#define GC_TILE 128
#define GC_TILES_X 4
#define GC_NTILES 16

/* Visit callback: pixel (r, c) of ray (view, bin) with weight wt. */
typedef void (*gc_visit)(void *ctx, int view, int bin, int r, int c, double wt);

static void gc_emit8(gc_visit f, void *ctx, int al, int ray, int ix, int iy, int dx, int dy, double wt)
{
    const int W = GC_N;
    f(ctx, al, ray, iy + dy, ix + dx, wt);
    f(ctx, 90 + al, ray, W - 1 - ix - dx, iy + dy, wt);
    if (ray != GC_MID) {
        f(ctx, al, GC_NB - 1 - ray, W - 1 - iy - dy, W - 1 - ix - dx, wt);
        f(ctx, 90 + al, GC_NB - 1 - ray, ix + dx, W - 1 - iy - dy, wt);
    }
    if (al != 45) {
        f(ctx, 90 - al, ray, ix + dx, iy + dy, wt);
        f(ctx, 180 - al, ray, W - 1 - iy - dy, ix + dx, wt);
        if (ray != GC_MID) {
            f(ctx, 90 - al, GC_NB - 1 - ray, W - 1 - ix - dx, W - 1 - iy - dy, wt);
            f(ctx, 180 - al, GC_NB - 1 - ray, iy + dy, W - 1 - ix - dx, wt);
        }
    }
}

/* Walk of proj.c:38-203; view-90 column offset differs between A (W/2+ind,
 * proj.c:32) and AStar (W/2-1+ind, proj.c:225): adjoint selects the latter. */
static void gc_walk(gc_visit f, void *ctx, int adjoint)
{
    const int W = GC_N, H = GC_N;
    for (int ray = 0; ray < GC_NB; ++ray) {
        int k = ray - (GC_NB / 2);
        if (k > (H / 2 - 1) || k < (1 - H / 2)) continue;
        int row = H / 2 - 1 + k;
        for (int c = 0; c < W; ++c) f(ctx, 0, ray, row, c, 1.0);
    }
    for (int ray = 0; ray < GC_NB; ++ray) {
        int k = ray - (GC_NB / 2);
        if (k > (W / 2 - 1) || k < (1 - W / 2)) continue;
        int col = adjoint ? (W / 2 - 1 + k) : (W / 2 + k);
        for (int r = 0; r < H; ++r) f(ctx, 90, ray, r, col, 1.0);
    }
    for (int al = 1; al < 46; ++al) {
        double sf = sin(3.141593 / 180 * al), cf = cos(3.141593 / 180 * al);
        double hx = ((double)1.0) / sf, hy = ((double)1.0) / cf;
        double tanp = sf / cf, rtan = cf / sf;
        for (int tile = 0; tile < GC_NTILES; ++tile)
            for (int ray = 0; ray < GC_MID + 1; ++ray) {
                int k = ray - (GC_NB / 2);
                int dx = GC_TILE * (GC_TILES_X - 1 - (tile % GC_TILES_X));
                int dy = GC_TILE * (tile / GC_TILES_X);
                double c, fp, w, yc, xc;
                int ix, iy, incx, step;
                c = cf * k + sf * k * tanp;
                c = (W / 2 - dy) - tanp * (W / 2 - GC_TILE * (tile % GC_TILES_X)) - c;
                if (c < 0) {
                    c = -c * rtan;
                    ix = (int)c; iy = 0;
                    fp = ((double)(ix + 1)) - c;
                    ix = GC_TILE - ix - 1; incx = -1;
                    w = hy * fp; yc = fp * tanp; xc = -fp;
                    step = 2;
                } else {
                    ix = GC_TILE - 1; incx = -1;
                    iy = (int)c;
                    fp = ((double)(iy + 1)) - c;
                    yc = -fp; w = hx * fp; xc = fp * rtan;
                    step = 1;
                }
                while (ix >= 0 && ix < GC_TILE && iy >= 0 && iy < GC_TILE) {
                    if (step == 1) {
                        xc -= (double)1.0;
                        yc += tanp;
                        if (xc > 0) {
                            gc_emit8(f, ctx, al, ray, ix, iy, dx, dy, hy);
                            w -= hy; ix += incx;
                        } else {
                            gc_emit8(f, ctx, al, ray, ix, iy, dx, dy, w);
                            w = hy - w; iy++; step = 2;
                        }
                    } else {
                        yc -= 1.0;
                        xc += rtan;
                        if (yc > 0) {
                            w -= hx;
                            gc_emit8(f, ctx, al, ray, ix, iy, dx, dy, hx);
                            iy++;
                        } else {
                            gc_emit8(f, ctx, al, ray, ix, iy, dx, dy, w);
                            w = hx - w; ix += incx; step = 1;
                        }
                    }
                }
            }
    }
}

typedef struct { double *sino; const double *img; } gc_fp_ctx;
static void gc_fp_visit(void *p, int view, int bin, int r, int c, double wt)
{
    gc_fp_ctx *x = (gc_fp_ctx *)p;
    x->sino[view * GC_NB + bin] += x->img[r * GC_N + c] * wt;
}
typedef struct { double *img; const double *sino; } gc_bp_ctx;
static void gc_bp_visit(void *p, int view, int bin, int r, int c, double wt)
{
    gc_bp_ctx *x = (gc_bp_ctx *)p;
    x->img[r * GC_N + c] += x->sino[view * GC_NB + bin] * wt;
}

void or_gc_fp(double *proj, const double *orig)
{
    gc_fp_ctx x = { proj, orig };
    memset(proj, 0, sizeof(double) * OR_GC_VIEWS * GC_NB);
    gc_walk(gc_fp_visit, &x, 0);
}

void or_gc_bp(double *back, const double *proj)
{
    gc_bp_ctx x = { back, proj };
    memset(back, 0, sizeof(double) * GC_N * GC_N);
    gc_walk(gc_bp_visit, &x, 1);
}

/* ------------------------------------------------------------------------ */
/* Generic operator over a view list.                                       */

static void wr_fp_d(const or_op *op, const int *views, int nv, const double *img, double *sino)
{
    int n = op->h;
    int *line = (int *)malloc(sizeof(int) * 2 * n);
    float *wt = (float *)malloc(sizeof(float) * 2 * n);
    for (int k = 0; k < nv; ++k)
        for (int b = 0; b < op->nbins; ++b) {
            int numb; float sn; double acc = 0.0;
            or_wray(views[k], b, op->nviews, op->nbins, n, op->flavour, line, wt, &numb, &sn);
            for (int i = 0; i < numb; ++i) acc += img[line[i]] * (double)wt[i];
            sino[(size_t)views[k] * op->nbins + b] = acc;
        }
    free(line); free(wt);
}

static void wr_bp_d(const or_op *op, const int *views, int nv, const double *sino, double *img)
{
    int n = op->h;
    int *line = (int *)malloc(sizeof(int) * 2 * n);
    float *wt = (float *)malloc(sizeof(float) * 2 * n);
    memset(img, 0, sizeof(double) * (size_t)n * n);
    for (int k = 0; k < nv; ++k)
        for (int b = 0; b < op->nbins; ++b) {
            int numb; float sn;
            double r = sino[(size_t)views[k] * op->nbins + b];
            or_wray(views[k], b, op->nviews, op->nbins, n, op->flavour, line, wt, &numb, &sn);
            for (int i = 0; i < numb; ++i) img[line[i]] += r * (double)wt[i];
        }
    free(line); free(wt);
}

void or_op_fp(const or_op *op, const int *views, int nv, const double *img, double *sino)
{
    if (op->kind == 0) {
        or_fan_fp(&op->fan, views, nv, img, sino);
    } else if (op->kind == 1) {
        double *full = (double *)malloc(sizeof(double) * OR_GC_VIEWS * GC_NB);
        or_gc_fp(full, img);
        for (int k = 0; k < nv; ++k)
            memcpy(sino + (size_t)views[k] * GC_NB, full + (size_t)views[k] * GC_NB, sizeof(double) * GC_NB);
        free(full);
    } else {
        wr_fp_d(op, views, nv, img, sino);
    }
}

void or_op_bp(const or_op *op, const int *views, int nv, const double *sino, double *img)
{
    if (op->kind == 0) {
        or_fan_bp(&op->fan, views, nv, sino, img);
    } else if (op->kind == 1) {
        double *masked = (double *)calloc((size_t)OR_GC_VIEWS * GC_NB, sizeof(double));
        for (int k = 0; k < nv; ++k)
            memcpy(masked + (size_t)views[k] * GC_NB, sino + (size_t)views[k] * GC_NB, sizeof(double) * GC_NB);
        or_gc_bp(img, masked);
        free(masked);
    } else {
        wr_bp_d(op, views, nv, sino, img);
    }
}

/* ------------------------------------------------------------------------ */
/* Whole-geometry CSR (rays in [view][bin] order) for comparisons.          */

int64_t or_fan_csr(const or_fan_geom *g, int64_t *rowptr, int *ind, double *len, int64_t cap)
{
    int rc = fan_cap(g);
    int *ix = (int *)malloc(sizeof(int) * rc);
    double *w = (double *)malloc(sizeof(double) * rc);
    double *ct = (double *)malloc(sizeof(double) * g->nviews);
    double *st = (double *)malloc(sizeof(double) * g->nviews);
    int64_t k = 0;
    or_fan_tables(g, ct, st);
    rowptr[0] = 0;
    for (int a = 0; a < g->nviews; ++a)
        for (int b = 0; b < g->nbins; ++b) {
            int n = or_fan_ray(g, ct, st, a, b, ix, w, rc);
            for (int i = 0; i < n; ++i, ++k)
                if (k < cap) { ind[k] = ix[i]; len[k] = w[i]; }
            rowptr[(int64_t)a * g->nbins + b + 1] = k;
        }
    free(ix); free(w); free(ct); free(st);
    return k;
}

int64_t or_wray_csr(int nproj, int nray, int n, int flavour, int64_t *rowptr, int *ind, float *len, int64_t cap)
{
    int *line = (int *)malloc(sizeof(int) * 2 * n);
    float *wt = (float *)malloc(sizeof(float) * 2 * n);
    int64_t k = 0;
    rowptr[0] = 0;
    for (int a = 0; a < nproj; ++a)
        for (int b = 0; b < nray; ++b) {
            int numb; float sn;
            or_wray(a, b, nproj, nray, n, flavour, line, wt, &numb, &sn);
            for (int i = 0; i < numb; ++i, ++k)
                if (k < cap) { ind[k] = line[i]; len[k] = wt[i]; }
            rowptr[(int64_t)a * nray + b + 1] = k;
        }
    free(line); free(wt);
    return k;
}
