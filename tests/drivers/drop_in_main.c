/*
 * The reference's program (HIP/mhd.cpp:39-64: init, then cuda_main on host
 * arrays) written against include/mhd_api.h, as a maintainer would swap the
 * library in: read the fixture with the C reader's order, run the loop on
 * the device, write result_u1.dat / result_b1.dat (Heterogeneous_MHD/C/
 * mhd.c:126-163) into the given directory, and the dt sequence to dts.bin.
 *
 *   drop_in_main <fixture.dat> <n> <nsteps> <outdir>
 *
 * Exit status 0 on success; every library error is printed with its message.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "mhd_api.h"

int main(int argc, char **argv)
{
    if (argc != 5) {
        fprintf(stderr, "usage: %s <fixture.dat> <n> <nsteps> <outdir>\n", argv[0]);
        return 2;
    }
    int nx = atoi(argv[2]), ny = nx, nz = nx;
    const int nsteps = atoi(argv[3]);
    const size_t n3 = (size_t)nx * ny * nz;
    float *u = malloc(sizeof(float) * 5 * n3), *b = malloc(sizeof(float) * 3 * n3);
    float *dts = malloc(sizeof(float) * (nsteps > 0 ? nsteps : 1));
    if (!u || !b || !dts) return 3;
    /* init(u, b, &nx, &ny, &nz) -> */
    if (mhd_init(argv[1], 0, u, b, &nx, &ny, &nz)) {
        fprintf(stderr, "mhd_init: %s\n", mhd_last_error());
        return 1;
    }
    /* cuda_main(u, b, &nx, &ny, &nz) -> */
    int done = 0;
    if (mhd_main(u, b, &nx, &ny, &nz, nsteps, INFINITY, dts, nsteps, &done)) {
        fprintf(stderr, "mhd_main: %s\n", mhd_last_error());
        return 1;
    }
    char pu[4096], pb[4096], pd[4096];
    snprintf(pu, sizeof pu, "%s/result_u1.dat", argv[4]);
    snprintf(pb, sizeof pb, "%s/result_b1.dat", argv[4]);
    snprintf(pd, sizeof pd, "%s/dts.bin", argv[4]);
    if (mhd_write_result(pu, pb, u, b, &nx, &ny, &nz)) {
        fprintf(stderr, "mhd_write_result: %s\n", mhd_last_error());
        return 1;
    }
    FILE *fd = fopen(pd, "wb");
    if (!fd || fwrite(dts, sizeof(float), (size_t)done, fd) != (size_t)done) return 1;
    fclose(fd);
    printf("drop_in_main: %d steps of %d^3 on the device, results in %s\n", done, nx, argv[4]);
    free(u);
    free(b);
    free(dts);
    return 0;
}
