/*
 * Oracle of the 3D cone-beam path (test infrastructure only; see oracle.h).
 *
 * or_cone_ray / or_cone_trace restate industry3d's cone() and forward_proj
 * (industry3d/src/tracing.cpp:51-81, 119-307, with rotate_2d and sqr of
 * industry3d/src/utility.cpp:82-98) in float, the reference's type, and in
 * its operation order; this file is built with -ffp-contract=off so each
 * operation rounds as written.  The C++ source calls the float overloads
 * of cos, sin, sqrt and fabs (`using namespace std` with float arguments),
 * hence cosf / sinf / sqrtf / fabsf here.  Mixed float/double expressions
 * (the 1.e-4 and 1.e6 constants, the N/2.0 shifts, the 5.e-2 stop test)
 * are evaluated in double as C++ promotes them.
 *
 * The walk stops when its current voxel leaves the volume.  The reference
 * would read and write outside (or, through the flat index, at an aliased
 * voxel of) its arrays there; it happens only for rays that enter through
 * an edge or a corner of the volume.
 *
 * or_cone_exact lists the exact (double) intersection lengths of the same
 * ray (the float end points of cone(), widened) with every voxel it
 * crosses: the merged plane crossings of the three axes.  It defines the
 * back projector the device applies in gather form.
 *
 * Pins: the reference's own tracing.cpp needs boost through utility.cpp and
 * is not compiled here, so or_cone_trace is pinned by closed forms
 * (axis-aligned rays: unit lengths, one voxel per plane; sum of lengths =
 * clipped chord minus the dropped final piece; entry count = plane
 * crossings + 1; exact-rational enumeration of the voxel sequence on small
 * grids), and or_cone_exact by a per-voxel slab clip of every voxel.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define CONE_PI 3.1415926535897932385

/* theta of view a: (float)alpha / NPROJ * 2 * PI (tracing.cpp:58), the last
 * product in double, narrowed to the float parameter of rotate_2d. */
static float cone_theta(const or_cone_geom *g, int a) {
    float t = (float)a / (float)g->nproj;
    t = t * 2.0f;
    return (float)((double)t * CONE_PI);
}

void or_cone_tables(const or_cone_geom *g, float *cos_t, float *sin_t) {
    for (int a = 0; a < g->nproj; ++a) {
        const float th = cone_theta(g, a);
        cos_t[a] = cosf(th);
        sin_t[a] = sinf(th);
    }
}

/* rotate_2d (utility.cpp:82-98): d = R x accumulated from zero, k outer. */
static void rot2(float *x, float *y, float c, float s) {
    float d0 = 0.0f, d1 = 0.0f;
    d0 += c * *x;
    d1 += s * *x;
    d0 += -s * *y;
    d1 += c * *y;
    *x = d0;
    *y = d1;
}

/* parallel(), industry3d/src/tracing.cpp:9-49: source and detector points
 * share their x (panel column) and z (panel row); every coordinate is
 * shifted by HALFSIZE = NX / 2.0 held as a float (ct3d.cpp:88-93), in float. */
static void par_ray(const or_cone_geom *g, float c, float s, int row, int col, float *src, float *dst) {
    const int halfdet = g->ndx / 2;
    const float half = (float)(g->nx / 2.0);
    float sX = (float)((double)col + 0.5 - halfdet);
    sX *= g->pixel;
    float sY = -g->sod;
    float sZ = (float)((double)(halfdet - row) - 0.5);
    sZ *= g->pixel;
    float dX = (float)((double)col + 0.5 - halfdet);
    dX *= g->pixel;
    float dY = g->sdd - g->sod;
    float dZ = (float)((double)(halfdet - row) - 0.5);
    dZ *= g->pixel;
    rot2(&sX, &sY, c, s);
    rot2(&dX, &dY, c, s);
    sX /= g->sample; sY /= g->sample; sZ /= g->sample;
    dX /= g->sample; dY /= g->sample; dZ /= g->sample;
    src[0] = sX + half; src[1] = sY + half; src[2] = sZ + half;
    dst[0] = dX + half; dst[1] = dY + half; dst[2] = dZ + half;
}

void or_cone_ray(const or_cone_geom *g, float c, float s, int row, int col, float *src, float *dst) {
    if (g->parallel) {
        par_ray(g, c, s, row, col, src, dst);
        return;
    }
    const int halfdet = g->ndx / 2;
    float sX = 0.0f, sY = -g->sod, sZ = 0.0f;
    rot2(&sX, &sY, c, s);
    float dX = (float)((double)(col - halfdet) + 0.5);
    dX *= g->pixel;
    float dY = g->sdd - g->sod;
    float dZ = (float)((double)(halfdet - row) - 0.5);
    dZ *= g->pixel;
    rot2(&dX, &dY, c, s);
    sX /= g->sample; sZ /= g->sample; sY /= g->sample;
    dX /= g->sample; dY /= g->sample; dZ /= g->sample;
    src[0] = (float)((double)sX + g->nx / 2.0);
    src[1] = (float)((double)sY + g->ny / 2.0);
    src[2] = (float)((double)sZ + g->nz / 2.0);
    dst[0] = (float)((double)dX + g->nx / 2.0);
    dst[1] = (float)((double)dY + g->ny / 2.0);
    dst[2] = (float)((double)dZ + g->nz / 2.0);
}

/* LAMBDA_X(i, s, d, L) = L * ((float)i - s) / (d - s), tracing.cpp:83-85. */
static float lam(float i, float s, float d, float L) { return L * (i - s) / (d - s); }

int or_cone_walk(const or_cone_geom *g, const float *S, const float *D, int64_t *ind, float *wgt, int cap) {
    const int NX = g->nx, NY = g->ny, NZ = g->nz;
    const float rx = D[0] - S[0], ry = D[1] - S[1], rz = D[2] - S[2];
    const float L = sqrtf(rx * rx + ry * ry + rz * rz);
    float lmin = 0.0f, lmax = L, l0, lN, t;
    const int sgx = (rx > 0) ? 1 : -1, sgy = (ry > 0) ? 1 : -1, sgz = (rz > 0) ? 1 : -1;
    const float ax = fabsf(rx), ay = fabsf(ry), az = fabsf(rz);
    const float lenx = ((double)ax > 1.e-4) ? (L / ax) : 1.e6f;
    const float leny = ((double)ay > 1.e-4) ? (L / ay) : 1.e6f;
    const float lenz = ((double)az > 1.e-4) ? (L / az) : 1.e6f;
    int index = 0;

    l0 = lam(0.0f, S[0], D[0], L); lN = lam((float)NX, S[0], D[0], L);
    t = (l0 < lN) ? l0 : lN; lmin = (lmin > t) ? lmin : t; if (lmin == t) index = 1;
    t = (l0 > lN) ? l0 : lN; lmax = (lmax < t) ? lmax : t;
    l0 = lam(0.0f, S[1], D[1], L); lN = lam((float)NY, S[1], D[1], L);
    t = (l0 < lN) ? l0 : lN; lmin = (lmin > t) ? lmin : t; if (lmin == t) index = 2;
    t = (l0 > lN) ? l0 : lN; lmax = (lmax < t) ? lmax : t;
    l0 = lam(0.0f, S[2], D[2], L); lN = lam((float)NZ, S[2], D[2], L);
    t = (l0 < lN) ? l0 : lN; lmin = (lmin > t) ? lmin : t; if (lmin == t) index = 3;
    t = (l0 > lN) ? l0 : lN; lmax = (lmax < t) ? lmax : t;
    if (lmin >= lmax) return 0;

    float lam0 = lmin, lx, ly, lz;
    int vx, vy, vz;
    if (index == 1) {
        vx = (sgx == 1) ? 0 : NX - 1;
        lx = lam0 + lenx;
        vy = (int)(S[1] + lam0 * ry / L);
        ly = ((double)ay < 1.e-4) ? 1.e6f : lam((float)vy + (float)(sgy > 0), S[1], D[1], L);
        vz = (int)(S[2] + lam0 * rz / L);
        lz = ((double)az < 1.e-4) ? 1.e6f : lam((float)vz + (float)(sgz > 0), S[2], D[2], L);
    } else if (index == 2) {
        vy = (sgy == 1) ? 0 : NY - 1;
        ly = lam0 + leny;
        vx = (int)(S[0] + lam0 * rx / L);
        lx = ((double)ax < 1.e-4) ? 1.e6f : lam((float)vx + (float)(sgx > 0), S[0], D[0], L);
        vz = (int)(S[2] + lam0 * rz / L);
        lz = ((double)az < 1.e-4) ? 1.e6f : lam((float)vz + (float)(sgz > 0), S[2], D[2], L);
    } else {
        vz = (sgz == 1) ? 0 : NZ - 1;
        lz = lam0 + lenz;
        vx = (int)(S[0] + lam0 * rx / L);
        lx = ((double)ax < 1.e-4) ? 1.e6f : lam((float)vx + (float)(sgx > 0), S[0], D[0], L);
        vy = (int)(S[1] + lam0 * ry / L);
        ly = ((double)ay < 1.e-4) ? 1.e6f : lam((float)vy + (float)(sgy > 0), S[1], D[1], L);
    }
    int numb = 0;
    while ((double)lam0 < (double)lmax - 5.e-2) {
        if (vx < 0 || vx >= NX || vy < 0 || vy >= NY || vz < 0 || vz >= NZ) break;
        const int64_t idx = (int64_t)vz * NX * NY + (int64_t)vx * NY + vy;
        float w;
        if (lx <= ly && lx <= lz) {
            w = lx - lam0; lam0 = lx; lx += lenx; vx += sgx;
        } else if (ly <= lz) {
            w = ly - lam0; lam0 = ly; ly += leny; vy += sgy;
        } else {
            w = lz - lam0; lam0 = lz; lz += lenz; vz += sgz;
        }
        if (numb < cap) { ind[numb] = idx; wgt[numb] = w; }
        ++numb;
    }
    return numb;
}

int or_cone_trace(const or_cone_geom *g, int view, int row, int col, int64_t *ind, float *wgt, int cap) {
    float S[3], D[3];
    const float th = cone_theta(g, view);
    or_cone_ray(g, cosf(th), sinf(th), row, col, S, D);
    return or_cone_walk(g, S, D, ind, wgt, cap);
}

void or_cone_fp(const or_cone_geom *g, const float *vol, double *sino) {
    const int cap = 4 * (g->nx + g->ny + g->nz) + 8;
    int64_t *ind = (int64_t *)malloc(sizeof(int64_t) * cap);
    float *wgt = (float *)malloc(sizeof(float) * cap);
    float *ct = (float *)malloc(sizeof(float) * g->nproj), *st = (float *)malloc(sizeof(float) * g->nproj);
    or_cone_tables(g, ct, st);
    for (int a = 0; a < g->nproj; ++a)
        for (int r = 0; r < g->ndx; ++r)
            for (int c = 0; c < g->ndy; ++c) {
                float S[3], D[3];
                or_cone_ray(g, ct[a], st[a], r, c, S, D);
                const int n = or_cone_walk(g, S, D, ind, wgt, cap);
                double acc = 0.0;
                for (int i = 0; i < n && i < cap; ++i) acc += (double)vol[ind[i]] * (double)wgt[i];
                sino[((int64_t)a * g->ndx + r) * g->ndy + c] = acc;
            }
    free(ind); free(wgt); free(ct); free(st);
}

static int cmp_dbl(const void *a, const void *b) {
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

int or_cone_exact(const or_cone_geom *g, const float *Sf, const float *Df, int64_t *ind, double *len, int cap) {
    const int N[3] = {g->nx, g->ny, g->nz};
    double S[3], r[3];
    for (int a = 0; a < 3; ++a) { S[a] = Sf[a]; r[a] = (double)Df[a] - (double)Sf[a]; }
    const double L = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    double lo = 0.0, hi = 1.0;
    for (int a = 0; a < 3; ++a) {
        if (r[a] == 0.0) {
            if (!(S[a] >= 0.0 && S[a] < N[a])) return 0;
            continue;
        }
        double t0 = (0.0 - S[a]) / r[a], t1 = ((double)N[a] - S[a]) / r[a];
        if (t0 > t1) { double q = t0; t0 = t1; t1 = q; }
        if (t0 > lo) lo = t0;
        if (t1 < hi) hi = t1;
    }
    if (!(lo < hi)) return 0;
    int nt = 0;
    const int tcap = N[0] + N[1] + N[2] + 8;
    double *ts = (double *)malloc(sizeof(double) * tcap);
    ts[nt++] = lo;
    for (int a = 0; a < 3; ++a) {
        if (r[a] == 0.0) continue;
        for (int k = 0; k <= N[a]; ++k) {
            const double t = ((double)k - S[a]) / r[a];
            if (t > lo && t < hi) ts[nt++] = t;
        }
    }
    ts[nt++] = hi;
    qsort(ts, nt, sizeof(double), cmp_dbl);
    int n = 0;
    for (int i = 0; i + 1 < nt; ++i) {
        const double t1 = ts[i], t2 = ts[i + 1];
        if (!(t2 > t1)) continue;
        const double tm = 0.5 * (t1 + t2);
        int v[3];
        for (int a = 0; a < 3; ++a) {
            v[a] = (int)floor(S[a] + tm * r[a]);
            if (v[a] < 0) v[a] = 0;
            if (v[a] >= N[a]) v[a] = N[a] - 1;
        }
        if (n < cap) {
            ind[n] = (int64_t)v[2] * g->nx * g->ny + (int64_t)v[0] * g->ny + v[1];
            len[n] = (t2 - t1) * L;
        }
        ++n;
    }
    free(ts);
    return n;
}

void or_cone_bp(const or_cone_geom *g, const double *sino, double *vol) {
    const int cap = 4 * (g->nx + g->ny + g->nz) + 8;
    int64_t *ind = (int64_t *)malloc(sizeof(int64_t) * cap);
    double *len = (double *)malloc(sizeof(double) * cap);
    float *ct = (float *)malloc(sizeof(float) * g->nproj), *st = (float *)malloc(sizeof(float) * g->nproj);
    or_cone_tables(g, ct, st);
    memset(vol, 0, sizeof(double) * (size_t)g->nx * g->ny * g->nz);
    for (int a = 0; a < g->nproj; ++a)
        for (int r = 0; r < g->ndx; ++r)
            for (int c = 0; c < g->ndy; ++c) {
                float S[3], D[3];
                or_cone_ray(g, ct[a], st[a], r, c, S, D);
                const int n = or_cone_exact(g, S, D, ind, len, cap);
                const double y = sino[((int64_t)a * g->ndx + r) * g->ndy + c];
                for (int i = 0; i < n && i < cap; ++i) vol[ind[i]] += y * len[i];
            }
    free(ind); free(len); free(ct); free(st);
}

double or_cone_clip(const float *Sf, const float *Df, int x, int y, int z) {
    const double S[3] = {Sf[0], Sf[1], Sf[2]};
    const double r[3] = {(double)Df[0] - Sf[0], (double)Df[1] - Sf[1], (double)Df[2] - Sf[2]};
    const int v[3] = {x, y, z};
    double lo = 0.0, hi = 1.0;
    for (int a = 0; a < 3; ++a) {
        if (r[a] == 0.0) {
            if (!(S[a] >= v[a] && S[a] < v[a] + 1)) return 0.0;
            continue;
        }
        double t0 = ((double)v[a] - S[a]) / r[a], t1 = ((double)v[a] + 1.0 - S[a]) / r[a];
        if (t0 > t1) { double q = t0; t0 = t1; t1 = q; }
        if (t0 > lo) lo = t0;
        if (t1 < hi) hi = t1;
    }
    if (!(hi > lo)) return 0.0;
    return (hi - lo) * sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
}

/* ---------------------------------------------------------------------------
 * Ordered-subset AT iteration on the cone-beam volume (the 3D form the
 * device runs).  Regulariser over the 26-neighbourhood: the 13 offsets k
 * with a positive first non-zero component, weight 1 / |k|^2, every pair
 * (x, x + k) with at least one voxel inside the volume, f zero-padded and v
 * one-padded (the padding of origin/GoldenC/src/main.c:112-125):
 *   R(f; v) = sum w_k (v(x)^2 + eta) (f(x + k) - f(x))^2,  eta = 0.01 eps^2.
 * u step: d = S A_s^T (g_s - A_s u) - alpha/2 dR/df, exact line search
 * c = |d|^2 / (S |A_s d|^2 + alpha R(d; v)), u = clamp(u + c d).
 * v step: AT energy alpha sum v^2 n + beta eps Q(v) + beta/(4 eps) |1 - v|^2
 * with n(x) = sum_k w_k (f(x + k) - f(x))^2 and Q the same pair sum of v
 * with unit weights; iter_edge steepest-descent steps with exact line
 * search, v clamped to [0, 1].  This is the 2D ordered-subset form
 * (or_os_at) with GoldenC's 4-neighbour stencils widened to 26 neighbours.
 * ------------------------------------------------------------------------- */
static const int K13[13][3] = {{0, 0, 1},  {0, 1, -1}, {0, 1, 0},  {0, 1, 1},  {1, -1, -1},
                               {1, -1, 0}, {1, -1, 1}, {1, 0, -1}, {1, 0, 0},  {1, 0, 1},
                               {1, 1, -1}, {1, 1, 0},  {1, 1, 1}};   /* (dz, dx, dy) */

static double kw(int k) {
    return 1.0 / (double)(K13[k][0] * K13[k][0] + K13[k][1] * K13[k][1] + K13[k][2] * K13[k][2]);
}

static int inside(int z, int x, int y, int nz, int nx, int ny) {
    return z >= 0 && z < nz && x >= 0 && x < nx && y >= 0 && y < ny;
}

static double zp3(const double *f, int z, int x, int y, int nz, int nx, int ny) {
    return inside(z, x, y, nz, nx, ny) ? f[((size_t)z * nx + x) * ny + y] : 0.0;
}

static double op3(const double *v, int z, int x, int y, int nz, int nx, int ny) {
    return inside(z, x, y, nz, nx, ny) ? v[((size_t)z * nx + x) * ny + y] : 1.0;
}

void or_cone_ugrad(double *d, const double *f, const double *v, double alpha, double eps, int nz, int nx, int ny) {
    const double eta = 0.01 * eps * eps;
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) {
                const double f0 = f[((size_t)z * nx + x) * ny + y], v0 = v[((size_t)z * nx + x) * ny + y];
                double acc = 0.0;
                for (int k = 0; k < 13; ++k) {
                    const int a = K13[k][0], b = K13[k][1], c = K13[k][2];
                    const double fp = zp3(f, z + a, x + b, y + c, nz, nx, ny);
                    const double fm = zp3(f, z - a, x - b, y - c, nz, nx, ny);
                    const double vm = op3(v, z - a, x - b, y - c, nz, nx, ny);
                    acc += kw(k) * ((v0 * v0 + eta) * (fp - f0) - (vm * vm + eta) * (f0 - fm));
                }
                d[((size_t)z * nx + x) * ny + y] += alpha * acc;
            }
}

/* Pair sum of p (zero outside) with weights (v(x)^2 + eta) when wv, else 1. */
static double pair_sum(const double *p, const double *v, double eta, int wv, int nz, int nx, int ny) {
    double acc = 0.0;
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) {
                const double p0 = p[((size_t)z * nx + x) * ny + y];
                const double v0 = v ? v[((size_t)z * nx + x) * ny + y] : 1.0;
                const double s0 = wv ? v0 * v0 + eta : 1.0, so = wv ? 1.0 + eta : 1.0;
                for (int k = 0; k < 13; ++k) {
                    const int a = K13[k][0], b = K13[k][1], c = K13[k][2];
                    const double q = zp3(p, z + a, x + b, y + c, nz, nx, ny) - p0;
                    acc += kw(k) * s0 * q * q;
                    if (!inside(z - a, x - b, y - c, nz, nx, ny)) acc += kw(k) * so * p0 * p0;
                }
            }
    return acc;
}

double or_cone_nV(const double *p, const double *v, double eps, int nz, int nx, int ny) {
    return pair_sum(p, v, 0.01 * eps * eps, 1, nz, nx, ny);
}

void or_cone_edge_step(double *v, const double *f, double alpha, double beta, double eps, int iters, int nz,
                       int nx, int ny) {
    const size_t nv = (size_t)nz * nx * ny;
    double *dv = (double *)malloc(sizeof(double) * nv), *nn = (double *)malloc(sizeof(double) * nv);
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) {
                const double f0 = f[((size_t)z * nx + x) * ny + y];
                double acc = 0.0;
                for (int k = 0; k < 13; ++k) {
                    const double q = zp3(f, z + K13[k][0], x + K13[k][1], y + K13[k][2], nz, nx, ny) - f0;
                    acc += kw(k) * q * q;
                }
                nn[((size_t)z * nx + x) * ny + y] = acc;
            }
    for (int it = 0; it < iters; ++it) {
        double num = 0.0, dn = 0.0;
        for (int z = 0; z < nz; ++z)
            for (int x = 0; x < nx; ++x)
                for (int y = 0; y < ny; ++y) {
                    const size_t i = ((size_t)z * nx + x) * ny + y;
                    const double v0 = v[i];
                    double lap = 0.0;
                    for (int k = 0; k < 13; ++k) {
                        const int a = K13[k][0], b = K13[k][1], c = K13[k][2];
                        lap += kw(k) * ((op3(v, z + a, x + b, y + c, nz, nx, ny) - v0) -
                                        (v0 - op3(v, z - a, x - b, y - c, nz, nx, ny)));
                    }
                    dv[i] = -alpha * v0 * nn[i] + beta * eps * lap - beta / (4.0 * eps) * (v0 - 1.0);
                    num += dv[i] * dv[i];
                    dn += nn[i] * dv[i] * dv[i];
                }
        const double den = alpha * dn + beta * eps * pair_sum(dv, NULL, 0.0, 0, nz, nx, ny) +
                           beta / (4.0 * eps) * num;
        const double c = den > 0 ? num / den : 0.0;
        for (size_t i = 0; i < nv; ++i) {
            double t = v[i] + c * dv[i];
            if (t < 0.0) t = 0.0;
            if (t > 1.0) t = 1.0;
            v[i] = t;
        }
    }
    free(dv);
    free(nn);
}

void or_cone_fp_views(const or_cone_geom *g, const int *views, int nv, const double *vol, double *sino) {
    const int cap = 4 * (g->nx + g->ny + g->nz) + 8;
    int64_t *ind = (int64_t *)malloc(sizeof(int64_t) * cap);
    float *wgt = (float *)malloc(sizeof(float) * cap);
    float *ct = (float *)malloc(sizeof(float) * g->nproj), *st = (float *)malloc(sizeof(float) * g->nproj);
    or_cone_tables(g, ct, st);
    for (int q = 0; q < nv; ++q) {
        const int a = views[q];
        for (int r = 0; r < g->ndx; ++r)
            for (int c = 0; c < g->ndy; ++c) {
                float S[3], D[3];
                or_cone_ray(g, ct[a], st[a], r, c, S, D);
                const int n = or_cone_walk(g, S, D, ind, wgt, cap);
                double acc = 0.0;
                for (int i = 0; i < n && i < cap; ++i) acc += vol[ind[i]] * (double)wgt[i];
                sino[((int64_t)a * g->ndx + r) * g->ndy + c] = acc;
            }
    }
    free(ind); free(wgt); free(ct); free(st);
}

void or_cone_bp_views(const or_cone_geom *g, const int *views, int nv, const double *sino, double *vol) {
    const int cap = 4 * (g->nx + g->ny + g->nz) + 8;
    int64_t *ind = (int64_t *)malloc(sizeof(int64_t) * cap);
    double *len = (double *)malloc(sizeof(double) * cap);
    float *ct = (float *)malloc(sizeof(float) * g->nproj), *st = (float *)malloc(sizeof(float) * g->nproj);
    or_cone_tables(g, ct, st);
    memset(vol, 0, sizeof(double) * (size_t)g->nx * g->ny * g->nz);
    for (int q = 0; q < nv; ++q) {
        const int a = views[q];
        for (int r = 0; r < g->ndx; ++r)
            for (int c = 0; c < g->ndy; ++c) {
                float S[3], D[3];
                or_cone_ray(g, ct[a], st[a], r, c, S, D);
                const int n = or_cone_exact(g, S, D, ind, len, cap);
                const double y = sino[((int64_t)a * g->ndx + r) * g->ndy + c];
                for (int i = 0; i < n && i < cap; ++i) vol[ind[i]] += y * len[i];
            }
    }
    free(ind); free(len); free(ct); free(st);
}

int or_cone_os_at(const or_cone_geom *g, const double *sino, double *u, double *v, const or_at_params *p,
                  double *df_hist) {
    const int nz = g->nz, nx = g->nx, ny = g->ny, S = p->n_subsets;
    const size_t ni = (size_t)nz * nx * ny, ns = (size_t)g->nproj * g->ndx * g->ndy, pv = (size_t)g->ndx * g->ndy;
    int *views = (int *)malloc(sizeof(int) * g->nproj);
    double *q = (double *)calloc(ns, sizeof(double));
    double *b = (double *)malloc(sizeof(double) * ni);
    double *d = (double *)malloc(sizeof(double) * ni);
    double df_prev = 0;
    int it;
    for (it = 1; it <= p->max_iters; ++it) {
        double df = 0;
        for (int s = 0; s < S; ++s) {
            const int nv = or_subset_views(g->nproj, S, s, views);
            double num = 0, ad2 = 0;
            or_cone_fp_views(g, views, nv, u, q);
            for (int k = 0; k < nv; ++k)
                for (size_t j = 0; j < pv; ++j) {
                    const size_t i = (size_t)views[k] * pv + j;
                    q[i] = sino[i] - q[i];
                    df += q[i] * q[i];
                }
            or_cone_bp_views(g, views, nv, q, b);
            for (size_t i = 0; i < ni; ++i) d[i] = S * b[i];
            or_cone_ugrad(d, u, v, p->alpha, p->eps, nz, nx, ny);
            or_cone_fp_views(g, views, nv, d, q);
            for (int k = 0; k < nv; ++k)
                for (size_t j = 0; j < pv; ++j) {
                    const double t = q[(size_t)views[k] * pv + j];
                    ad2 += t * t;
                }
            for (size_t i = 0; i < ni; ++i) num += d[i] * d[i];
            const double den = S * ad2 + p->alpha * or_cone_nV(d, v, p->eps, nz, nx, ny);
            const double c = (den > 0) ? num / den : 0.0;
            for (size_t i = 0; i < ni; ++i) {
                double x = u[i] + c * d[i];
                if (x < p->lower) x = p->lower;
                if (x > p->upper) x = p->upper;
                u[i] = x;
            }
        }
        if (p->iter_edge > 0) or_cone_edge_step(v, u, p->alpha, p->beta, p->eps, p->iter_edge, nz, nx, ny);
        if (df_hist) df_hist[it - 1] = df;
        if (it > 1 && fabs(df - df_prev) < p->rel_tol * df_prev) break;
        df_prev = df;
    }
    free(views); free(q); free(b); free(d);
    return (it > p->max_iters) ? p->max_iters : it;
}
