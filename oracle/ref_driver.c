/*
 * Calls into the reference's own C functions (Heterogeneous_MHD/C, compiled
 * from /root/reference by oracle/Makefile into oracle/_ref/libmhd_ref.so).
 * TEST INFRASTRUCTURE ONLY: used to pin the oracle restatement and as the
 * "reference" CPU baseline.  This file is ours; it restates nothing and only
 * sequences the reference's entry points the way its main() does
 * (Heterogeneous_MHD/C/mhd.c:60-119), with the double buffer update_u/b and
 * copy_matrix after every op.
 */
#include <stdlib.h>
#include <string.h>

void fluidx(float *u, float *b, int *nx, int *ny, int *nz, float *dt, float *uu, float *ub, int iter);
void copy_matrix(float *u, float *b, int *nx, int *ny, int *nz, float *uu, float *ub);
void advectbyzx(float *u, float *b, int *nx, int *ny, int *nz, float *dt, float *uu, float *ub);
void transpose12(float *u, float *b, float *uu, float *ub, int *nx, int *ny, int *nz, int iter);
void transpose13(float *u, float *b, float *uu, float *ub, int *nx, int *ny, int *nz);
float calcfl(float *u, float *b, int *nx, int *ny, int *nz);

static float *g_uu, *g_ub;
static size_t g_cells;

static void ensure(size_t cells)
{
    if (cells > g_cells) {
        free(g_uu);
        free(g_ub);
        g_uu = (float *)malloc(sizeof(float) * 5 * cells);
        g_ub = (float *)malloc(sizeof(float) * 3 * cells);
        g_cells = cells;
    }
}

static void sync(float *u, float *b, int nx, int ny, int nz)
{
    copy_matrix(u, b, &nx, &ny, &nz, g_uu, g_ub);
}

/* Single ops, each preceded by the copy that mhd.c performs after the
 * previous op, so that update_u/b hold the current state. */
void ref_fluidx(float *u, float *b, int nx, int ny, int nz, float dt)
{
    ensure((size_t)nx * ny * nz);
    sync(u, b, nx, ny, nz);
    fluidx(u, b, &nx, &ny, &nz, &dt, g_uu, g_ub, 1);
}

void ref_advectbyzx(float *u, float *b, int nx, int ny, int nz, float dt)
{
    ensure((size_t)nx * ny * nz);
    sync(u, b, nx, ny, nz);
    advectbyzx(u, b, &nx, &ny, &nz, &dt, g_uu, g_ub);
}

/* transpose12/13 read update_u/b and write u/b in the transposed frame. */
void ref_transpose12(float *u, float *b, int nx, int ny, int nz)
{
    ensure((size_t)nx * ny * nz);
    sync(u, b, nx, ny, nz);
    transpose12(u, b, g_uu, g_ub, &nx, &ny, &nz, 1);
}

void ref_transpose13(float *u, float *b, int nx, int ny, int nz)
{
    ensure((size_t)nx * ny * nz);
    sync(u, b, nx, ny, nz);
    transpose13(u, b, g_uu, g_ub, &nx, &ny, &nz);
}

float ref_calcfl(float *u, float *b, int nx, int ny, int nz)
{
    return calcfl(u, b, &nx, &ny, &nz);
}

/* nsteps iterations of the loop body of mhd.c:61-119; dts[s] gets each dt. */
void ref_run(float *u, float *b, int nx, int ny, int nz, int nsteps, float tf, float *dts)
{
    ensure((size_t)nx * ny * nz);
    float t = 0;
    int iter = 0;
    for (int s = 0; s < nsteps; s++) {
        iter++;
        float dt = 0.9 * calcfl(u, b, &nx, &ny, &nz);
        if (dt > (tf - t) / 2) dt = (tf - t) / 2;
        t = t + 2 * dt;
        if (dts) dts[s] = dt;
        sync(u, b, nx, ny, nz);
        fluidx(u, b, &nx, &ny, &nz, &dt, g_uu, g_ub, iter);
        sync(u, b, nx, ny, nz);
        advectbyzx(u, b, &nx, &ny, &nz, &dt, g_uu, g_ub);
        sync(u, b, nx, ny, nz);
        transpose12(u, b, g_uu, g_ub, &nx, &ny, &nz, iter);
        sync(u, b, ny, nx, nz);
        fluidx(u, b, &ny, &nx, &nz, &dt, g_uu, g_ub, iter);
        sync(u, b, ny, nx, nz);
        advectbyzx(u, b, &ny, &nx, &nz, &dt, g_uu, g_ub);
        sync(u, b, ny, nx, nz);
        transpose13(u, b, g_uu, g_ub, &ny, &nx, &nz);
        sync(u, b, nz, nx, ny);
        fluidx(u, b, &nz, &nx, &ny, &dt, g_uu, g_ub, iter);
        sync(u, b, nz, nx, ny);
        advectbyzx(u, b, &nz, &nx, &ny, &dt, g_uu, g_ub);
        sync(u, b, nz, nx, ny);
        advectbyzx(u, b, &nz, &nx, &ny, &dt, g_uu, g_ub);
        sync(u, b, nz, nx, ny);
        fluidx(u, b, &nz, &nx, &ny, &dt, g_uu, g_ub, iter);
        sync(u, b, nz, nx, ny);
        transpose13(u, b, g_uu, g_ub, &nz, &nx, &ny);
        sync(u, b, ny, nx, nz);
        advectbyzx(u, b, &ny, &nx, &nz, &dt, g_uu, g_ub);
        sync(u, b, ny, nx, nz);
        fluidx(u, b, &ny, &nx, &nz, &dt, g_uu, g_ub, iter);
        sync(u, b, ny, nx, nz);
        transpose12(u, b, g_uu, g_ub, &ny, &nx, &nz, iter);
        sync(u, b, nx, ny, nz);
        advectbyzx(u, b, &nx, &ny, &nz, &dt, g_uu, g_ub);
        sync(u, b, nx, ny, nz);
        fluidx(u, b, &nx, &ny, &nz, &dt, g_uu, g_ub, iter);
    }
}
