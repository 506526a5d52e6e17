/*
 * libmedimage_msbeam.so: msbeam's projector pair with its own signatures
 * (msbeam/CPU_OpenMP/src/proj.h:7-8), served on the GPU through
 * libmedimage_amd's C ABI, for callers linking msbeam's proj.c today.
 *
 *   A(g, f)      g[NPROJ * NRAY] = projections of f[IMGSIZE * IMGSIZE]
 *                (msbeam/CPU_OpenMP/src/proj.c:14-32)
 *   AStar(f, g)  f = back projection of g (proj.c:34-52)
 *
 * Sizes are msbeam's proto.h (IMGSIZE 512, NPROJ 180, NRAY 768), the
 * tracer is wray with msbeam's double constants (MI_BEAM_PARALLEL, flavour
 * 0), and each value's float products are added in msbeam's loop order, so
 * the results equal msbeam's bit for bit.  Like the reference, the
 * functions return nothing; a failure prints the library's message to
 * stderr and exits with status 1.  Device 0, one context per process.
 */
#ifndef MEDIMAGE_MSBEAM_H
#define MEDIMAGE_MSBEAM_H

#ifdef __cplusplus
extern "C" {
#endif

#define MI_MSBEAM_IMGSIZE 512
#define MI_MSBEAM_NPROJ 180
#define MI_MSBEAM_NRAY 768

void A(float *g, float *f);
void AStar(float *f, float *g);

#ifdef __cplusplus
}
#endif
#endif
