/*
 * libmedimage_goldenc.so: the GoldenC program's own call surface, served on
 * the GPU through libmedimage_amd's C ABI, so that a GoldenC-style main
 * (read a phantom into init_f, A(gDelta, init_f), MumfordShah(ALPHA, BETA,
 * EPSILON), write f and v) links against it in place of proj.c and the
 * solver functions of main.c.
 *
 *   A, AStar                    origin/GoldenC/src/proj.h:6-7, proj.c:14-397
 *   MumfordShah                 origin/GoldenC/src/main.c:35-49
 *   minimize_in_IMAGE_variable  main.c:25-27, 81-257
 *   minimize_in_EDGE_variable   main.c:28-30, 259-360
 *   medfilt2                    main.c:31, 366-390
 *   init_f, f, v, gDelta        main.c:8-10 (MumfordShah works on f, v and
 *                               gDelta, as the reference's does)
 *
 * Sizes and iteration counts are GoldenC's config.h (512 x 512 image, 180
 * views of 767 bins, 10 alternations of 10 CG and 10 SD steps).  Every
 * result equals the reference program's bit for bit (FP64, the reference's
 * summation order, MI_ORDER_REFERENCE).  Like the reference, the functions
 * return nothing; a device or runtime failure prints the library's message
 * to stderr and exits with status 1 (as industry3d/src/ct3d.cpp:26-37 does
 * on bad input).  Device 0, one context for the process, created on first
 * use; not reentrant (the reference keeps its state in globals too).
 */
#ifndef MEDIMAGE_GOLDENC_H
#define MEDIMAGE_GOLDENC_H

#ifdef __cplusplus
extern "C" {
#endif

#define MI_GC_IMAGE 512   /* imageHight = imageWidth */
#define MI_GC_VIEWS 180   /* nAlpha                  */
#define MI_GC_BINS 767    /* projWidth               */

extern double init_f[MI_GC_IMAGE][MI_GC_IMAGE];
extern double f[MI_GC_IMAGE][MI_GC_IMAGE], v[MI_GC_IMAGE][MI_GC_IMAGE];
extern double gDelta[MI_GC_VIEWS][MI_GC_BINS];

void A(double proj[MI_GC_VIEWS][MI_GC_BINS], double orig[MI_GC_IMAGE][MI_GC_IMAGE]);
void AStar(double back[MI_GC_IMAGE][MI_GC_IMAGE], double proj[MI_GC_VIEWS][MI_GC_BINS]);
void MumfordShah(double alpha, double beta, double epsilon);
void minimize_in_IMAGE_variable(double out[MI_GC_IMAGE][MI_GC_IMAGE], double g[MI_GC_VIEWS][MI_GC_BINS],
                                double f[MI_GC_IMAGE][MI_GC_IMAGE], double v[MI_GC_IMAGE][MI_GC_IMAGE],
                                double alpha, double epsilon, int IterImage);
void minimize_in_EDGE_variable(double out[MI_GC_IMAGE][MI_GC_IMAGE], double f[MI_GC_IMAGE][MI_GC_IMAGE],
                               double v[MI_GC_IMAGE][MI_GC_IMAGE], double alpha, double beta, double epsilon,
                               int IterEdge);
void medfilt2(double f[MI_GC_IMAGE][MI_GC_IMAGE]);

#ifdef __cplusplus
}
#endif
#endif
