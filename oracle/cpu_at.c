/* CPU baseline of the ordered-subset AT iteration (test and benchmark
 * infrastructure: only the benchmark's CPU leg and the tests use it).
 *
 * A multi-threaded restatement of the same iteration as or_os_at
 * (solvers.c), in the style of the reference's OpenMP programs: every
 * projection re-traces its rays (msbeam/CPU_OpenMP/src/proj.c:14-52 calls
 * wray per ray and view), the bulk is FP32 (msbeam's float images and
 * sinograms), the sums that set step sizes are double.  Rays are split over
 * OpenMP threads; the back projection scatters into one image per thread,
 * reduced in thread order afterwards.  The regulariser, nV, median and edge
 * step are the oracle's FP64 routines on converted copies (O(N^2) work).
 * Only fan geometries (op->kind == 0). */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

static int fan_cap_c(const or_fan_geom *g) { return 2 * (g->nx + g->ny) + 8; }

/* sino rows of `views` <- A img (FP32 bulk, per-ray double accumulation of
 * float products is not used: float sums as msbeam's A does). */
static void fp_f32(const or_fan_geom *g, const double *ct, const double *st, const int *views, int nv,
                   const float *img, float *sino)
{
    const int nb = g->nbins, cap = fan_cap_c(g);
#pragma omp parallel
    {
        int *ind = (int *)malloc(sizeof(int) * cap);
        double *w = (double *)malloc(sizeof(double) * cap);
#pragma omp for schedule(dynamic, 16)
        for (long r = 0; r < (long)nv * nb; ++r) {
            const int k = (int)(r / nb), b = (int)(r % nb);
            const int n = or_fan_ray(g, ct, st, views[k], b, ind, w, cap);
            float acc = 0.f;
            for (int i = 0; i < n; ++i) acc += img[ind[i]] * (float)w[i];
            sino[(size_t)views[k] * nb + b] = acc;
        }
        free(ind);
        free(w);
    }
}

/* img <- A^T of the listed rows (per-thread images, reduced in order). */
static void bp_f32(const or_fan_geom *g, const double *ct, const double *st, const int *views, int nv,
                   const float *sino, float *img, float *scratch, int nthreads)
{
    const int nb = g->nbins, cap = fan_cap_c(g);
    const size_t ni = (size_t)g->nx * g->ny;
    memset(scratch, 0, sizeof(float) * ni * nthreads);
#pragma omp parallel num_threads(nthreads)
    {
        float *mine = scratch + ni * omp_get_thread_num();
        int *ind = (int *)malloc(sizeof(int) * cap);
        double *w = (double *)malloc(sizeof(double) * cap);
#pragma omp for schedule(dynamic, 16)
        for (long r = 0; r < (long)nv * nb; ++r) {
            const int k = (int)(r / nb), b = (int)(r % nb);
            const int n = or_fan_ray(g, ct, st, views[k], b, ind, w, cap);
            const float q = sino[(size_t)views[k] * nb + b];
            for (int i = 0; i < n; ++i) mine[ind[i]] += q * (float)w[i];
        }
        free(ind);
        free(w);
    }
#pragma omp parallel for num_threads(nthreads)
    for (long i = 0; i < (long)ni; ++i) {
        float s = 0.f;
        for (int t = 0; t < nthreads; ++t) s += scratch[(size_t)t * ni + i];
        img[i] = s;
    }
}

int cpu_os_at_f32(const or_op *op, const float *g, float *u, float *v, const or_at_params *p, int nthreads)
{
    if (op->kind != 0) return -1;
    const or_fan_geom *fg = &op->fan;
    const int h = op->h, w = op->w, S = p->n_subsets, nb = op->nbins;
    const size_t ni = (size_t)h * w, ns = (size_t)op->nviews * nb;
    omp_set_num_threads(nthreads);
    int *views = (int *)malloc(sizeof(int) * op->nviews);
    double *ct = (double *)malloc(sizeof(double) * op->nviews), *st = (double *)malloc(sizeof(double) * op->nviews);
    float *q = (float *)malloc(sizeof(float) * ns), *b = (float *)malloc(sizeof(float) * ni);
    float *scratch = (float *)malloc(sizeof(float) * ni * nthreads);
    double *ud = (double *)malloc(sizeof(double) * ni), *vd = (double *)malloc(sizeof(double) * ni);
    double *dd = (double *)malloc(sizeof(double) * ni);
    float *df32 = (float *)malloc(sizeof(float) * ni);
    double df_prev = 0;
    int it;
    or_fan_tables(fg, ct, st);
    for (it = 1; it <= p->max_iters; ++it) {
        double df = 0;
        for (size_t i = 0; i < ni; ++i) vd[i] = v[i];
        for (int s = 0; s < S; ++s) {
            const int nv = or_subset_views(op->nviews, S, s, views);
            double num = 0, ad2 = 0, den, c;
            fp_f32(fg, ct, st, views, nv, u, q);
            for (int k = 0; k < nv; ++k)
                for (int j = 0; j < nb; ++j) {
                    const size_t i = (size_t)views[k] * nb + j;
                    q[i] = g[i] - q[i];
                    df += (double)q[i] * q[i];
                }
            bp_f32(fg, ct, st, views, nv, q, b, scratch, nthreads);
            for (size_t i = 0; i < ni; ++i) {
                dd[i] = (double)S * b[i];
                ud[i] = u[i];
            }
            or_gc_ugrad_reg(dd, ud, vd, p->alpha, p->eps, h, w);
            for (size_t i = 0; i < ni; ++i) df32[i] = (float)dd[i];
            fp_f32(fg, ct, st, views, nv, df32, q);
            for (int k = 0; k < nv; ++k)
                for (int j = 0; j < nb; ++j) {
                    const double x = q[(size_t)views[k] * nb + j];
                    ad2 += x * x;
                }
            for (size_t i = 0; i < ni; ++i) num += (double)df32[i] * df32[i];
            for (size_t i = 0; i < ni; ++i) dd[i] = df32[i];
            den = S * ad2 + p->alpha * or_gc_nV(dd, vd, p->eps, h, w);
            c = (den > 0) ? num / den : 0.0;
            for (size_t i = 0; i < ni; ++i) {
                double x = u[i] + c * df32[i];
                if (x < p->lower) x = p->lower;
                if (x > p->upper) x = p->upper;
                u[i] = (float)x;
            }
        }
        for (size_t i = 0; i < ni; ++i) ud[i] = u[i];
        if (p->median_every > 0 && it % p->median_every == 0) {
            or_medfilt2(ud, h, w);
            for (size_t i = 0; i < ni; ++i) u[i] = (float)ud[i];
        }
        if (p->iter_edge > 0) {
            or_gc_edge_step(vd, ud, vd, p->alpha, p->beta, p->eps, p->iter_edge, h, w);
            for (size_t i = 0; i < ni; ++i) v[i] = (float)vd[i];
        }
        if (it > 1 && fabs(df - df_prev) < p->rel_tol * df_prev) break;
        df_prev = df;
    }
    free(views); free(ct); free(st); free(q); free(b); free(scratch); free(ud); free(vd); free(dd); free(df32);
    return (it > p->max_iters) ? p->max_iters : it;
}
