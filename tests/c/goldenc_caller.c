/* A GoldenC-style driver, the calling pattern of origin/GoldenC/src/main.c:51-80
 * written against include/medimage_goldenc.h: read the phantom (text, %lf)
 * into init_f, A(gDelta, init_f), MumfordShah(ALPHA, BETA, EPSILON) of
 * config.h:6-8, then write gDelta, f and v.  The outputs are written as raw
 * doubles (the reference prints them with %f) so that a test can compare
 * them bit for bit.
 *
 * usage: goldenc_caller phantom.txt out.bin   (out: gDelta, f, v) */
#include <stdio.h>

#include "medimage_goldenc.h"

int main(int argc, char **argv) {
    int i, j;
    FILE *fin, *fout;
    if (argc != 3) {
        fprintf(stderr, "usage: %s phantom.txt out.bin\n", argv[0]);
        return 2;
    }
    fin = fopen(argv[1], "r");
    if (!fin) return 2;
    for (i = 0; i < MI_GC_IMAGE; ++i)
        for (j = 0; j < MI_GC_IMAGE; ++j)
            if (fscanf(fin, "%lf", &init_f[i][j]) != 1) return 2;
    fclose(fin);

    A(gDelta, init_f);
    MumfordShah(1000, 0.002, 0.0001);

    fout = fopen(argv[2], "wb");
    if (!fout) return 2;
    fwrite(gDelta, sizeof(double), (size_t)MI_GC_VIEWS * MI_GC_BINS, fout);
    fwrite(f, sizeof(double), (size_t)MI_GC_IMAGE * MI_GC_IMAGE, fout);
    fwrite(v, sizeof(double), (size_t)MI_GC_IMAGE * MI_GC_IMAGE, fout);
    fclose(fout);
    printf("goldenc_caller: done\n");
    return 0;
}
