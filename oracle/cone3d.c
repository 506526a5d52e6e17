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

void or_cone_ray(const or_cone_geom *g, float c, float s, int row, int col, float *src, float *dst) {
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
