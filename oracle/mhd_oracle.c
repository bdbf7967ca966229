/*
 * The oracle: a CPU restatement of the reference's relaxing-TVD MHD timestep
 * (Heterogeneous_MHD/C), built as liboracle_mhd.so.  TEST INFRASTRUCTURE
 * ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it to check results against; the product library never links or calls it.
 *
 * Parity pins (each function is checked by tests/test_oracle.py):
 *   - against the reference's own C sources compiled into oracle/_ref
 *     (bit-exact, every op and whole steps, on the reference's fixtures
 *     data/source_init_alfvenlinear_{16,32}cube.dat and seeded states),
 *   - against committed golden vectors written from oracle/_ref
 *     (tests/golden/), so the pin holds where /root/reference is absent,
 *   - against closed forms and invariants: transposes are permutations and
 *     involutions, uniform states are fixed points, fluidx conserves
 *     mass/momentum/energy and advectbyzx keeps div B, the linear Alfven wave
 *     travels at v_A = 1, and an FP64 instantiation agrees with FP32 to
 *     rounding.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define R float
#define SFX _f32
#include "mhd_oracle_impl.h"
#undef R
#undef SFX

#define R double
#define SFX _f64
#include "mhd_oracle_impl.h"
#undef R
#undef SFX

/* Threads of the oracle's OpenMP pencil loops (OMP_NUM_THREADS or the
 * runtime's default), for reporting the CPU timing beside the GPU one. */
int mhd_oracle_threads(void) { return omp_get_max_threads(); }

/* Initial-state reader of Heterogeneous_MHD/C/init.c:33-54: the file holds
 * 8 floats per cell (u1..u5, b1..b3); the reader's loop runs i outermost and
 * k innermost, so the file's fastest axis lands on the array's k. */
void mhd_oracle_read_ref_order(const float *rec, float *u, float *b, int nx, int ny, int nz)
{
    size_t r = 0;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++, r++) {
                const size_t s = ((size_t)k * ny + j) * nx + i;
                for (int c = 0; c < 5; c++) u[5 * s + c] = rec[8 * r + c];
                for (int c = 0; c < 3; c++) b[3 * s + c] = rec[8 * r + 5 + c];
            }
}

/* FP64 conservation monitors of a state in the x frame: total mass, total
 * energy, total momentum (3), and max |div B| with B on the left faces. */
void mhd_oracle_monitors_f32(const float *u, const float *b, int nx, int ny, int nz, double *out6)
{
    double m = 0, e = 0, px = 0, py = 0, pz = 0, dmax = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const size_t s = ((size_t)k * ny + j) * nx + i;
                const size_t sip = ((size_t)k * ny + j) * nx + (i + 1) % nx;
                const size_t sjp = ((size_t)k * ny + (j + 1) % ny) * nx + i;
                const size_t skp = ((size_t)((k + 1) % nz) * ny + j) * nx + i;
                m += u[5 * s];
                px += u[5 * s + 1];
                py += u[5 * s + 2];
                pz += u[5 * s + 3];
                e += u[5 * s + 4];
                const double d = ((double)b[3 * sip] - b[3 * s]) + ((double)b[3 * sjp + 1] - b[3 * s + 1]) +
                                 ((double)b[3 * skp + 2] - b[3 * s + 2]);
                if (fabs(d) > dmax) dmax = fabs(d);
            }
    out6[0] = m;
    out6[1] = e;
    out6[2] = px;
    out6[3] = py;
    out6[4] = pz;
    out6[5] = dmax;
}
