/*
 * Oracle AT solvers (test infrastructure only; see oracle.h).
 * GoldenC image/edge steps, median filter, MumfordShah alternation, the
 * ordered-subset AT iteration and the serial msbeam per-ray driver.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double dsq(double x) { return x * x; }

/* Zero-padded image value and one-padded edge value at (i, j). */
static double zpad(const double *a, int i, int j, int h, int w)
{
    return (i < 0 || j < 0 || i >= h || j >= w) ? 0.0 : a[(size_t)i * w + j];
}
static double opad(const double *a, int i, int j, int h, int w)
{
    return (i < 0 || j < 0 || i >= h || j >= w) ? 1.0 : a[(size_t)i * w + j];
}

/* origin/GoldenC/src/main.c:112-125 (same expression at :226-239). */
void or_gc_ugrad_reg(double *d, const double *f, const double *v,
                     double alpha, double eps, int h, int w)
{
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            double z = zpad(f, i, j, h, w);
            double zd = zpad(f, i + 1, j, h, w), zr = zpad(f, i, j + 1, h, w);
            double zu = zpad(f, i - 1, j, h, w), zl = zpad(f, i, j - 1, h, w);
            double p = opad(v, i, j, h, w), pu = opad(v, i - 1, j, h, w), pl = opad(v, i, j - 1, h, w);
            d[(size_t)i * w + j] += alpha * (
                +dsq(p) * (zd - z)
                + dsq(p) * (zr - z)
                - dsq(pu) * (z - zu)
                - dsq(pl) * (z - zl)
                + 0.01 * dsq(eps) * (zd + zr - 4 * z + zu + zl));
        }
}

/* main.c:152-158. */
double or_gc_nV(const double *p, const double *v, double eps, int h, int w)
{
    double acc = 0;
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            double z = zpad(p, i, j, h, w);
            acc += (dsq(opad(v, i, j, h, w)) + 0.01 * dsq(eps)) *
                   (dsq(z - zpad(p, i - 1, j, h, w)) + dsq(z - zpad(p, i, j - 1, h, w)));
        }
    return acc;
}

/* main.c:135-139: interior pixels only, added onto acc. */
static double gc_energy_add(double acc, const double *f, const double *v,
                            double alpha, double eps, int h, int w)
{
    for (int i = 1; i <= h - 2; ++i)
        for (int j = 1; j <= w - 2; ++j) {
            double z = f[(size_t)i * w + j];
            acc += alpha * (dsq(v[(size_t)i * w + j]) + 0.01 * dsq(eps)) *
                   (dsq(z - f[(size_t)(i - 1) * w + j]) + dsq(z - f[(size_t)i * w + j - 1]));
        }
    return acc;
}

double or_gc_energy_reg(const double *f, const double *v, double alpha, double eps, int h, int w)
{
    return gc_energy_add(0.0, f, v, alpha, eps, h, w);
}

/* main.c:366-390.  Any exact selection of the 5th of 9 values gives the
 * reference's result; the exchange order below is the reference's. */
void or_medfilt2(double *f, int h, int w)
{
    double *src = (double *)malloc(sizeof(double) * (size_t)h * w);
    memcpy(src, f, sizeof(double) * (size_t)h * w);
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            double t[9];
            int k = 0;
            t[k++] = zpad(src, i, j, h, w);
            t[k++] = zpad(src, i - 1, j - 1, h, w);
            t[k++] = zpad(src, i - 1, j, h, w);
            t[k++] = zpad(src, i - 1, j + 1, h, w);
            t[k++] = zpad(src, i, j + 1, h, w);
            t[k++] = zpad(src, i + 1, j + 1, h, w);
            t[k++] = zpad(src, i + 1, j, h, w);
            t[k++] = zpad(src, i + 1, j - 1, h, w);
            t[k++] = zpad(src, i, j - 1, h, w);
            for (int a = 0; a < 9; ++a)
                for (int b = a + 1; b < 9; ++b)
                    if (t[a] > t[b]) { double x = t[a]; t[a] = t[b]; t[b] = x; }
            f[(size_t)i * w + j] = t[4];
        }
    free(src);
}

static int *all_views(int n)
{
    int *v = (int *)malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) v[i] = i;
    return v;
}

/* minimize_in_IMAGE_variable, origin/GoldenC/src/main.c:81-257. */
int or_gc_image_step(const or_op *op, double *out, const double *g, const double *f,
                     const double *v, double alpha, double eps, int iter_image, int *why)
{
    const int h = op->h, w = op->w;
    const size_t ni = (size_t)h * w, ns = (size_t)op->nviews * op->nbins;
    int *views = all_views(op->nviews);
    double *upd = (double *)malloc(sizeof(double) * ni);
    double *d = (double *)malloc(sizeof(double) * ni);
    double *p = (double *)malloc(sizeof(double) * ni);
    double *tmp = (double *)malloc(sizeof(double) * ni);
    double *Af = (double *)malloc(sizeof(double) * ns);
    double *r = (double *)malloc(sizeof(double) * ns);
    double *q = (double *)malloc(sizeof(double) * ns);
    double *Ad = (double *)malloc(sizeof(double) * ns);
    double at_old, at_new;
    int k, cadence = (int)(iter_image / 2 + 0.0001);
    int reason = 0, ran = 0;

    memcpy(upd, f, sizeof(double) * ni);
    or_op_fp(op, views, op->nviews, upd, Af);
    for (size_t i = 0; i < ns; ++i) r[i] = g[i] - Af[i];
    or_op_bp(op, views, op->nviews, r, d);
    or_gc_ugrad_reg(d, upd, v, alpha, eps, h, w);
    at_old = 0;
    for (size_t i = 0; i < ns; ++i) at_old += dsq(Af[i] - g[i]);
    at_old = gc_energy_add(at_old, upd, v, alpha, eps, h, w);
    memcpy(p, d, sizeof(double) * ni);

    for (k = 1; k <= iter_image; ++k) {
        double nv, pd, nap, c1, maxcp, adap, c2;
        ran = k;
        or_op_fp(op, views, op->nviews, p, q);
        nv = or_gc_nV(p, v, eps, h, w);
        pd = 0;
        for (size_t i = 0; i < ni; ++i) pd += p[i] * d[i];
        nap = 0;
        for (size_t i = 0; i < ns; ++i) nap += dsq(q[i]);
        c1 = pd / (nap + alpha * nv);
        memcpy(tmp, upd, sizeof(double) * ni);
        maxcp = 0;
        for (size_t i = 0; i < ni; ++i) {
            double cp;
            upd[i] += c1 * p[i];
            cp = fabs(c1 * p[i]);
            if (maxcp < cp) maxcp = cp;
        }
        /* The reference takes k % 0 when iter_image < 2 (undefined); no
         * median is applied then. */
        if (cadence > 0 && k % cadence == 0) or_medfilt2(upd, h, w);
        if (k == iter_image) { reason = 0; break; }
        if (maxcp < 1e-5) { reason = 1; break; }
        or_op_fp(op, views, op->nviews, upd, Af);
        at_new = 0;
        for (size_t i = 0; i < ns; ++i) at_new += dsq(Af[i] - g[i]);
        at_new = gc_energy_add(at_new, upd, v, alpha, eps, h, w);
        if (1.1 * at_old < at_new && 1 < k) {
            memcpy(upd, tmp, sizeof(double) * ni);
            reason = 2;
            break;
        }
        at_old = at_new;
        for (size_t i = 0; i < ns; ++i) r[i] -= c1 * q[i];
        or_op_bp(op, views, op->nviews, r, d);
        or_gc_ugrad_reg(d, upd, v, alpha, eps, h, w);
        or_op_fp(op, views, op->nviews, d, Ad);
        adap = 0;
        for (size_t i = 0; i < ns; ++i) adap += Ad[i] * q[i];
        c2 = -adap / nap;
        for (size_t i = 0; i < ni; ++i) p[i] = d[i] + c2 * p[i];
    }
    memcpy(out, upd, sizeof(double) * ni);
    if (why) *why = reason;
    free(views); free(upd); free(d); free(p); free(tmp);
    free(Af); free(r); free(q); free(Ad);
    return ran;
}

/* minimize_in_EDGE_variable, origin/GoldenC/src/main.c:259-360. */
int or_gc_edge_step(double *out, const double *f, const double *v, double alpha,
                    double beta, double eps, int iter_edge, int h, int w)
{
    const size_t ni = (size_t)h * w;
    double *upd = (double *)malloc(sizeof(double) * ni);
    double *tmp = (double *)malloc(sizeof(double) * ni);
    double *n = (double *)malloc(sizeof(double) * ni);
    double *d = (double *)malloc(sizeof(double) * ni);
    int k, ran = 0;
    memcpy(upd, v, sizeof(double) * ni);
    for (k = 1; k <= iter_edge; ++k) {
        double nnd = 0, nd = 0, nd2 = 0, c, maxcd = 0;
        ran = k;
        memcpy(tmp, upd, sizeof(double) * ni);
        for (int i = 0; i < h; ++i)
            for (int j = 0; j < w; ++j) {
                double dxf = 0.5 * (zpad(f, i + 1, j, h, w) - zpad(f, i - 1, j, h, w));
                double dyf = 0.5 * (zpad(f, i, j + 1, h, w) - zpad(f, i, j - 1, h, w));
                double t = opad(tmp, i, j, h, w);
                double lap, gr;
                n[(size_t)i * w + j] = dsq(dxf) + dsq(dyf);
                lap = 0.25 * (opad(tmp, i + 1, j, h, w) + opad(tmp, i, j + 1, h, w)
                              - 4 * t + opad(tmp, i - 1, j, h, w) + opad(tmp, i, j - 1, h, w));
                gr = alpha * t * n[(size_t)i * w + j] - beta * eps * lap + beta * (t - 1) / (4 * eps);
                d[(size_t)i * w + j] = -gr;
            }
        for (int i = 0; i < h; ++i)
            for (int j = 0; j < w; ++j)
                nnd = nnd + dsq(0.5 * (zpad(d, i + 1, j, h, w) - zpad(d, i - 1, j, h, w)))
                          + dsq(0.5 * (zpad(d, i, j + 1, h, w) - zpad(d, i, j - 1, h, w)));
        for (size_t i = 0; i < ni; ++i) nd += dsq(d[i]);
        for (size_t i = 0; i < ni; ++i) nd2 += n[i] * dsq(d[i]);
        c = nd / (alpha * nd2 + beta * eps * nnd + (beta / (4 * eps)) * nd);
        for (size_t i = 0; i < ni; ++i) {
            double cd;
            upd[i] = tmp[i] + c * d[i];
            cd = fabs(c * d[i]);
            if (cd > maxcd) maxcd = cd;
        }
        for (size_t i = 0; i < ni; ++i) {
            if (upd[i] > 1) upd[i] = 1;
            if (upd[i] < 0) upd[i] = 0;
        }
        if (k == iter_edge) break;
        if (maxcd < 1e-5) break;
    }
    memcpy(out, upd, sizeof(double) * ni);
    free(upd); free(tmp); free(n); free(d);
    return ran;
}

/* MumfordShah, origin/GoldenC/src/main.c:35-49. */
void or_gc_mumford_shah(const or_op *op, const double *g, double *f, double *v,
                        double alpha, double beta, double eps,
                        int iter_alt, int iter_image, int iter_edge)
{
    const size_t ni = (size_t)op->h * op->w;
    for (size_t i = 0; i < ni; ++i) { f[i] = 0.0; v[i] = 1.0; }
    for (int i = 1; i <= iter_alt; ++i) {
        or_gc_image_step(op, f, g, f, v, alpha, eps, iter_image, NULL);
        or_gc_edge_step(v, f, v, alpha, beta, eps, iter_edge, op->h, op->w);
    }
}

/* ------------------------------------------------------------------------ */
/* Ordered subsets.                                                         */

void or_subset_order(int S, int *order)
{
    int bits = 0, k = 0;
    while ((1 << bits) < S) ++bits;
    for (int i = 0; i < (1 << bits); ++i) {
        int r = 0;
        for (int b = 0; b < bits; ++b)
            if (i & (1 << b)) r |= 1 << (bits - 1 - b);
        if (r < S) order[k++] = r;
    }
}

int or_subset_views(int nviews, int S, int s, int *views)
{
    int *order = (int *)malloc(sizeof(int) * (S > 0 ? S : 1));
    int n = 0;
    or_subset_order(S, order);
    for (int v = 0; v < nviews; ++v)
        if (v % S == order[s]) views[n++] = v;
    free(order);
    return n;
}

double or_os_step(const or_op *op, const double *g, double *u, const double *v,
                  const or_at_params *p, int s, double *d_out, double *df)
{
    const int h = op->h, w = op->w, S = p->n_subsets, nb = op->nbins;
    const size_t ni = (size_t)h * w, ns = (size_t)op->nviews * nb;
    int *views = (int *)malloc(sizeof(int) * op->nviews);
    double *q = (double *)malloc(sizeof(double) * ns);
    double *b = (double *)malloc(sizeof(double) * ni);
    double *d = d_out ? d_out : (double *)malloc(sizeof(double) * ni);
    const int nv = or_subset_views(op->nviews, S, s, views);
    double num = 0, ad2 = 0, den, c, fid = 0;
    double *acc = df ? df : &fid;    /* added in view order, as the pass sums it */

    or_op_fp(op, views, nv, u, q);
    for (int k = 0; k < nv; ++k)
        for (int j = 0; j < nb; ++j) {
            size_t i = (size_t)views[k] * nb + j;
            q[i] = g[i] - q[i];
            *acc += dsq(q[i]);
        }
    or_op_bp(op, views, nv, q, b);
    for (size_t i = 0; i < ni; ++i) d[i] = S * b[i];
    or_gc_ugrad_reg(d, u, v, p->alpha, p->eps, h, w);
    or_op_fp(op, views, nv, d, q);
    for (int k = 0; k < nv; ++k)
        for (int j = 0; j < nb; ++j) ad2 += dsq(q[(size_t)views[k] * nb + j]);
    for (size_t i = 0; i < ni; ++i) num += dsq(d[i]);
    den = S * ad2 + p->alpha * or_gc_nV(d, v, p->eps, h, w);
    c = (den > 0) ? num / den : 0.0;
    for (size_t i = 0; i < ni; ++i) {
        double x = u[i] + c * d[i];
        if (x < p->lower) x = p->lower;
        if (x > p->upper) x = p->upper;
        u[i] = x;
    }
    free(views); free(q); free(b);
    if (!d_out) free(d);
    return c;
}

int or_os_at(const or_op *op, const double *g, double *u, double *v,
             const or_at_params *p, double *df_hist)
{
    const int h = op->h, w = op->w, S = p->n_subsets;
    double df_prev = 0;
    int it;

    for (it = 1; it <= p->max_iters; ++it) {
        double df = 0;
        for (int s = 0; s < S; ++s) or_os_step(op, g, u, v, p, s, NULL, &df);
        if (p->median_every > 0 && it % p->median_every == 0) or_medfilt2(u, h, w);
        if (p->iter_edge > 0) or_gc_edge_step(v, u, v, p->alpha, p->beta, p->eps, p->iter_edge, h, w);
        if (df_hist) df_hist[it - 1] = df;
        if (it > 1 && fabs(df - df_prev) < p->rel_tol * df_prev) break;
        df_prev = df;
    }
    return (it > p->max_iters) ? p->max_iters : it;
}

/* ------------------------------------------------------------------------ */
/* msbeam per-ray driver, msbeam/CPU_OpenMP/src/main.c:213-301, 352-394.    */

static float fsq(float x) { return x * x; }

void or_msbeam_solve(const float *g, float *f, float *v, int nproj, int nray, int n,
                     double alpha, double beta, double eps, int iters, int flavour)
{
    int *line = (int *)malloc(sizeof(int) * 2 * n);
    float *wt = (float *)malloc(sizeof(float) * 2 * n);
    float *dd = (float *)malloc(sizeof(float) * 2 * n);
    /* ALPHA is an int literal (used in float expressions); BETA and EPSILON
     * are double literals (msbeam/CPU_OpenMP/src/proto.h:8-10). */
    const float A = (float)alpha;
    const double B = beta, E = eps;
    float lambda = 0.001;
    for (int i = 0; i < n * n; ++i) { f[i] = 0.; v[i] = 1.; }
    for (int it = 1; it <= iters; ++it) {
        for (int np = 0; np < nproj; ++np)
            for (int nr = 0; nr < nray; ++nr) {
                int numb; float sn;
                float Af = 0.;
                or_wray(np, nr, nproj, nray, n, flavour, line, wt, &numb, &sn);
                /* minIMAGE, main.c:213-261 */
                for (int i = 0; i < numb; ++i) Af += f[line[i]] * wt[i];
                Af -= g[np * nray + nr];
                for (int i = 0; i < numb; ++i) {
                    int ind = line[i], x = ind / n, y = ind % n;
                    float tmp = 0., lap = 0.;
                    if (x + 1 < n) tmp += fsq(v[ind]) * (f[ind + n] - f[ind]);
                    else tmp += fsq(v[ind]) * (0 - f[ind]);
                    if (y + 1 < n) tmp += fsq(v[ind]) * (f[ind + 1] - f[ind]);
                    else tmp += fsq(v[ind]) * (0 - f[ind]);
                    if (x - 1 >= 0) tmp -= fsq(v[ind - n]) * (f[ind] - f[ind - n]);
                    else tmp -= (f[ind] - 0);
                    if (y - 1 >= 0) tmp -= fsq(v[ind - 1]) * (f[ind] - f[ind - 1]);
                    else tmp -= (f[ind] - 0);
                    if (x + 1 < n) lap += f[ind + n];
                    if (y + 1 < n) lap += f[ind + 1];
                    if (x - 1 >= 0) lap += f[ind - n];
                    if (y - 1 >= 0) lap += f[ind - 1];
                    lap -= 4 * f[ind];
                    dd[i] = -Af * wt[i] + A * (tmp + fsq((float)E) * lap);
                }
                for (int i = 0; i < numb; ++i) {
                    int ind = line[i];
                    f[ind] += lambda * dd[i];
                    if (f[ind] < 0) f[ind] = 0;
                }
                /* minEDGE, main.c:263-301 */
                for (int i = 0; i < numb; ++i) {
                    int ind = line[i], x = ind / n, y = ind % n;
                    float a = 0., b = 0., c = 0.;
                    if (x - 1 >= 0) a += fsq(f[ind] - f[ind - n]);
                    else a += fsq(f[ind] - 0);
                    if (y - 1 >= 0) a += fsq(f[ind] - f[ind - 1]);
                    else a += fsq(f[ind] - 0);
                    a *= v[ind];
                    b = v[ind] - 1;
                    if (x + 1 < n) c += v[ind + n];
                    if (y + 1 < n) c += v[ind + 1];
                    if (x - 1 >= 0) c += v[ind - n];
                    if (y - 1 >= 0) c += v[ind - 1];
                    c -= 4 * v[ind];
                    dd[i] = -A * a - B / (4 * E) * b + B * E * c;
                }
                for (int i = 0; i < numb; ++i) {
                    int ind = line[i];
                    v[ind] += lambda * dd[i];
                    if (v[ind] < 0) v[ind] = 0;
                    if (v[ind] > 1) v[ind] = 1;
                }
            }
        lambda = lambda / (1 + 500 * lambda);
    }
    free(line); free(wt); free(dd);
}
