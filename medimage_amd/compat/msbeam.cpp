// libmedimage_msbeam.so: msbeam's A / AStar (include/medimage_msbeam.h)
// over libmedimage_amd's C ABI; the projections run on the device in
// msbeam's summation order and type.
#include "medimage_msbeam.h"
#include "medimage_amd.h"

#include <cstdio>
#include <cstdlib>

namespace {

mi_context *g_ctx = nullptr;

void check(int rc, const char *what) {
    if (rc == MI_OK) return;
    std::fprintf(stderr, "medimage_msbeam: %s failed (%d): %s\n", what, rc, mi_last_error());
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

// msbeam/CPU_OpenMP/src/proto.h:4-6; posit()'s double pi (flavour 0).
mi_geometry geometry() {
    mi_geometry g{};
    g.beam = MI_BEAM_PARALLEL;
    g.nx = g.ny = MI_MSBEAM_IMGSIZE;
    g.pixel_mm = 1.0;
    g.n_views = MI_MSBEAM_NPROJ;
    g.span_deg = 180.0;
    g.n_bins = MI_MSBEAM_NRAY;
    g.bin_mm = 1.0;
    g.flavour = 0;
    return g;
}

}  // namespace

extern "C" {

void A(float *g, float *f) {
    const mi_geometry geo = geometry();
    check(mi_forward_ordered(ctx(), &geo, MI_FP32, f, g), "A");
}

void AStar(float *f, float *g) {
    const mi_geometry geo = geometry();
    check(mi_backproject_ordered(ctx(), &geo, MI_FP32, g, f), "AStar");
}

}  // extern "C"
