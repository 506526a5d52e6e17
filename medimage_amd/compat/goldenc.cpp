// libmedimage_goldenc.so: GoldenC's call surface (include/medimage_goldenc.h)
// over libmedimage_amd's C ABI.  Host glue only: every projection, CG and
// SD step runs on the device in the reference's summation order.
#include "medimage_goldenc.h"
#include "medimage_amd.h"

#include <cstdio>
#include <cstdlib>

double init_f[MI_GC_IMAGE][MI_GC_IMAGE];
double f[MI_GC_IMAGE][MI_GC_IMAGE], v[MI_GC_IMAGE][MI_GC_IMAGE];
double gDelta[MI_GC_VIEWS][MI_GC_BINS];

namespace {

// origin/GoldenC/src/config.h:10-17
constexpr int kIterAlt = 10, kIterImage = 10, kIterEdge = 10;

mi_context *g_ctx = nullptr;

void check(int rc, const char *what) {
    if (rc == MI_OK) return;
    std::fprintf(stderr, "medimage_goldenc: %s failed (%d): %s\n", what, rc, mi_last_error());
    std::exit(1);
}

void release() {
    if (g_ctx) mi_context_destroy(g_ctx);
    g_ctx = nullptr;
}

mi_context *ctx() {
    if (!g_ctx) {
        check(mi_context_create(0, &g_ctx), "mi_context_create");
        std::atexit(release);
    }
    return g_ctx;
}

mi_geometry geometry() {
    mi_geometry g{};
    g.beam = MI_BEAM_GOLDENC;
    g.nx = g.ny = MI_GC_IMAGE;
    g.pixel_mm = 1.0;
    g.n_views = MI_GC_VIEWS;
    g.span_deg = 180.0;
    g.n_bins = MI_GC_BINS;
    g.bin_mm = 1.0;
    return g;
}

}  // namespace

extern "C" {

void A(double proj[MI_GC_VIEWS][MI_GC_BINS], double orig[MI_GC_IMAGE][MI_GC_IMAGE]) {
    const mi_geometry g = geometry();
    check(mi_forward_ordered(ctx(), &g, MI_FP64, &orig[0][0], &proj[0][0]), "A");
}

void AStar(double back[MI_GC_IMAGE][MI_GC_IMAGE], double proj[MI_GC_VIEWS][MI_GC_BINS]) {
    const mi_geometry g = geometry();
    check(mi_backproject_ordered(ctx(), &g, MI_FP64, &proj[0][0], &back[0][0]), "AStar");
}

void MumfordShah(double alpha, double beta, double epsilon) {
    const mi_geometry g = geometry();
    mi_at_params p{};
    p.alpha = alpha;
    p.beta = beta;
    p.epsilon = epsilon;
    p.iter_alt = kIterAlt;
    p.iter_image = kIterImage;
    p.iter_edge = kIterEdge;
    p.precision = MI_FP64;
    p.order = MI_ORDER_REFERENCE;
    mi_report rep{};
    check(mi_mumford_shah(ctx(), &g, &p, &gDelta[0][0], &f[0][0], &v[0][0], &rep), "MumfordShah");
}

void minimize_in_IMAGE_variable(double out[MI_GC_IMAGE][MI_GC_IMAGE], double g[MI_GC_VIEWS][MI_GC_BINS],
                                double fin[MI_GC_IMAGE][MI_GC_IMAGE], double vin[MI_GC_IMAGE][MI_GC_IMAGE],
                                double alpha, double epsilon, int IterImage) {
    const mi_geometry geo = geometry();
    check(mi_image_step_f64(ctx(), &geo, &out[0][0], &g[0][0], &fin[0][0], &vin[0][0], alpha, epsilon, IterImage),
          "minimize_in_IMAGE_variable");
}

void minimize_in_EDGE_variable(double out[MI_GC_IMAGE][MI_GC_IMAGE], double fin[MI_GC_IMAGE][MI_GC_IMAGE],
                               double vin[MI_GC_IMAGE][MI_GC_IMAGE], double alpha, double beta, double epsilon,
                               int IterEdge) {
    check(mi_edge_step_f64(ctx(), MI_GC_IMAGE, MI_GC_IMAGE, &out[0][0], &fin[0][0], &vin[0][0], alpha, beta,
                           epsilon, IterEdge),
          "minimize_in_EDGE_variable");
}

void medfilt2(double img[MI_GC_IMAGE][MI_GC_IMAGE]) {
    check(mi_medfilt2_f64(ctx(), MI_GC_IMAGE, MI_GC_IMAGE, &img[0][0]), "medfilt2");
}

}  // extern "C"
