/* An msbeam-style caller of the projector pair (msbeam/CPU_OpenMP/src/proj.h:7-8)
 * written against include/medimage_msbeam.h: g = A f, then b = AStar g.
 *
 * usage: msbeam_caller in.bin out.bin   (in: f as IMGSIZE^2 floats;
 *                                        out: g, then b) */
#include <stdio.h>
#include <stdlib.h>

#include "medimage_msbeam.h"

int main(int argc, char **argv) {
    const size_t npix = (size_t)MI_MSBEAM_IMGSIZE * MI_MSBEAM_IMGSIZE;
    const size_t nrow = (size_t)MI_MSBEAM_NPROJ * MI_MSBEAM_NRAY;
    float *f, *g, *b;
    FILE *fp;
    if (argc != 3) {
        fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
        return 2;
    }
    f = malloc(npix * sizeof(float));
    b = malloc(npix * sizeof(float));
    g = malloc(nrow * sizeof(float));
    if (!f || !b || !g) return 2;
    fp = fopen(argv[1], "rb");
    if (!fp || fread(f, sizeof(float), npix, fp) != npix) return 2;
    fclose(fp);

    A(g, f);
    AStar(b, g);

    fp = fopen(argv[2], "wb");
    if (!fp) return 2;
    fwrite(g, sizeof(float), nrow, fp);
    fwrite(b, sizeof(float), npix, fp);
    fclose(fp);
    free(f);
    free(g);
    free(b);
    printf("msbeam_caller: done\n");
    return 0;
}
