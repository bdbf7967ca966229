/*
 * CPU restatement of the relaxing-TVD MHD timestep: TEST INFRASTRUCTURE ONLY.
 *
 * This file is included twice by mhd_oracle.c, once with R=float (suffix
 * _f32, the reference's own precision and mixed FP32/FP64 evaluation order)
 * and once with R=double (suffix _f64, the FP64 oracle used to state the
 * FP32 error of each workload).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load the resulting library, and only as the
 * checker: nothing in the product path calls it.
 *
 * Layout: Fortran order, component fastest.  u[((k*ny+j)*nx+i)*5+c],
 * b[((k*ny+j)*nx+i)*3+c], periodic in every direction
 * (Heterogeneous_MHD/C/array_definition.h:12-14).
 *
 * Evaluation order follows the reference C exactly: float operands promoted
 * to double wherever the reference has a double literal or calls fabs/sqrt,
 * and the helper `a*X + b*Y + c` of Heterogeneous_MHD/C/includefunc_spu.h:73-94
 * evaluated as ((a*X)+(b*Y))+c.  Build with -ffp-contract=off -fno-fast-math.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

/* a*X + b*Y + c, Heterogeneous_MHD/C/includefunc_spu.h:73-94. */
static inline R FN(abxyc)(R a, R x, R bb, R y, R cc) { return a * x + bb * y + cc; }

/* van Leer harmonic limiter, Heterogeneous_MHD/C/tvd1.c:71-81 and tvdb.c:72-76:
 * the test is a product in R, the value 2*p*m/(p+m) is evaluated in double
 * with (p+m) summed in R, then rounded to R. */
static inline R FN(vanleer)(R p, R m)
{
    R d = (R)0.0;
    if (p * m > 0) d = (R)(2.0 * (double)p * (double)m / (double)(p + m));
    return d;
}

/* mhdflux: MHD x-flux and freezing speed of one pencil.
 * Heterogeneous_MHD/C/mhdflux.c:9-59 (find_max :61-72). */
static void FN(mhdflux)(R *v, R *cout, const R *u, const R *b, int n)
{
    const R gamma = (R)(5.0 / 3.0);
    R c = (R)0.0;
    for (int i = 0; i < n; i++) {
        const R u1 = u[5 * i + 0], u2 = u[5 * i + 1], u3 = u[5 * i + 2], u4 = u[5 * i + 3],
                u5 = u[5 * i + 4];
        const R b1 = b[3 * i + 0], b2 = b[3 * i + 1], b3 = b[3 * i + 2];
        const R vx = u2 / u1;                                   /* mhdflux.c:23 */
        const R S = u2 * u2 + u3 * u3 + u4 * u4;
        const R B = b1 * b1 + b2 * b2 + b3 * b3;
        const R q = S / u1;
        /* mhdflux.c:28: (u5 - S/u1/2.0)*(gamma-1.0) + (2.0-gamma)*B/2.0, in double */
        const R ps = (R)(((double)u5 - (double)q / 2.0) * ((double)gamma - 1.0) +
                         (2.0 - (double)gamma) * (double)B / 2.0);
        v[5 * i + 0] = u2;                                      /* mhdflux.c:33-37 */
        v[5 * i + 1] = u2 * vx + ps - b1 * b1;
        v[5 * i + 2] = u3 * vx - b2 * b1;
        v[5 * i + 3] = u4 * vx - b3 * b1;
        v[5 * i + 4] = (u5 + ps) * vx - b1 * (b1 * u2 + b2 * u3 + b3 * u4) / u1;
        const R p = (R)((double)ps - (double)B / 2.0);          /* mhdflux.c:41 */
        /* mhdflux.c:45: fabs(vx) + sqrt(fabs((B + gamma*p)/u1)), in double */
        const R ci = (R)(fabs((double)vx) + sqrt(fabs((double)((B + gamma * p) / u1))));
        if (ci > c) c = ci;                                     /* find_max: strict >, from 0 */
    }
    if (c > 0)
        for (int e = 0; e < 5 * n; e++) v[e] = v[e] / c;        /* mhdflux.c:48-57 */
    *cout = c;
}

/* tvd1: relaxing TVD, second-order Runge-Kutta, on one pencil in place.
 * Heterogeneous_MHD/C/tvd1.c:8-118. */
static void FN(tvd1)(R *u, const R *b, int n, R dt)
{
    R *v = (R *)malloc(sizeof(R) * 5 * n * 7);
    R *wr = v + 5 * n, *wl = wr + 5 * n, *fr = wl + 5 * n, *fl = fr + 5 * n, *flux = fl + 5 * n,
      *u1 = flux + 5 * n;
    R c;
    const R h = (R)(-1.0 * (double)dt / 2.0); /* tvd1.c:49 coefficient */
    const R mdt = (R)(-1.0 * (double)dt);     /* tvd1.c:116 coefficient */

    /* first half step: tvd1.c:34-49 */
    FN(mhdflux)(v, &c, u, b, n);
    for (int e = 0; e < 5 * n; e++) {
        wr[e] = FN(abxyc)(1, u[e], 1, v[e], 0);
        wl[e] = FN(abxyc)(1, u[e], -1, v[e], 0);
        fr[e] = FN(abxyc)(c, wr[e], 0, wr[e], 0);
    }
    for (int i = 0; i < n; i++) {
        const int ip = (i + 1) % n;
        for (int m = 0; m < 5; m++) fl[5 * i + m] = FN(abxyc)(c, wl[5 * ip + m], 0, wl[5 * ip + m], 0);
    }
    for (int e = 0; e < 5 * n; e++) flux[e] = FN(abxyc)((R)0.5, fr[e], (R)-0.5, fl[e], 0);
    for (int i = 0; i < n; i++) {
        const int im = (i + n - 1) % n;
        for (int m = 0; m < 5; m++) {
            const R d = FN(abxyc)(1, flux[5 * i + m], -1, flux[5 * im + m], 0);
            u1[5 * i + m] = FN(abxyc)(1, u[5 * i + m], h, d, 0);
        }
    }

    /* full step with van Leer limited fluxes: tvd1.c:52-116 */
    FN(mhdflux)(v, &c, u1, b, n);
    for (int e = 0; e < 5 * n; e++) {
        wr[e] = FN(abxyc)(1, u1[e], 1, v[e], 0);
        wl[e] = FN(abxyc)(1, u1[e], -1, v[e], 0);
        fr[e] = FN(abxyc)(c, wr[e], 0, wr[e], 0);
    }
    for (int i = 0; i < n; i++) {
        const int ip = (i + 1) % n;
        for (int m = 0; m < 5; m++) fl[5 * i + m] = FN(abxyc)(c, wl[5 * ip + m], 0, wl[5 * ip + m], 0);
    }
    for (int i = 0; i < n; i++) {
        const int ip = (i + 1) % n, im = (i + n - 1) % n;
        for (int m = 0; m < 5; m++) {
            const R dfrp = FN(abxyc)((R)0.5, fr[5 * ip + m], (R)-0.5, fr[5 * i + m], 0);
            const R dfrm = FN(abxyc)((R)0.5, fr[5 * i + m], (R)-0.5, fr[5 * im + m], 0);
            const R dflp = FN(abxyc)((R)0.5, fl[5 * i + m], (R)-0.5, fl[5 * ip + m], 0);
            const R dflm = FN(abxyc)((R)0.5, fl[5 * im + m], (R)-0.5, fl[5 * i + m], 0);
            const R dfr = FN(vanleer)(dfrp, dfrm);
            const R dfl = FN(vanleer)(dflp, dflm);
            const R t1 = FN(abxyc)(1, fr[5 * i + m], -1, fl[5 * i + m], 0);
            const R t2 = FN(abxyc)(1, dfr, -1, dfl, 0);
            flux[5 * i + m] = FN(abxyc)((R)0.5, t1, (R)0.5, t2, 0);
        }
    }
    for (int i = 0; i < n; i++) {
        const int im = (i + n - 1) % n;
        for (int m = 0; m < 5; m++) {
            const R d = FN(abxyc)(1, flux[5 * i + m], -1, flux[5 * im + m], 0);
            wr[5 * i + m] = FN(abxyc)(1, u[5 * i + m], mdt, d, 0);
        }
    }
    memcpy(u, wr, sizeof(R) * 5 * n);
    free(v);
}

/* fluidx: per (j,k) pencil, face B averaged to the cell centre, then tvd1.
 * Heterogeneous_MHD/C/fluidx.c:19-102.  Each pencil reads only its own u and
 * b at (j,k), (jp,k), (j,kp), so the update is done in place. */
void FN(mhd_oracle_fluidx)(R *u, const R *b, int nx, int ny, int nz, R dt)
{
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++) {
        R *pu = (R *)malloc(sizeof(R) * 8 * nx);
        R *b3x = pu + 5 * nx;
        for (int j = 0; j < ny; j++) {
            const int jp = (j + 1) % ny, kp = (k + 1) % nz;
            const size_t row = ((size_t)k * ny + j) * nx;
            const size_t rowjp = ((size_t)k * ny + jp) * nx;
            const size_t rowkp = ((size_t)kp * ny + j) * nx;
            memcpy(pu, u + 5 * row, sizeof(R) * 5 * nx);
            for (int i = 0; i < nx; i++) {
                const int ip = (i + 1) % nx;
                /* fluidx.c:71 halves b, :73-78 adds the halved b1 at i+1,
                 * :79-80 add b2(jp)/2 and b3(kp)/2 */
                b3x[3 * i + 0] = (R)0.5 * b[3 * (row + i) + 0] + (R)0.5 * b[3 * (row + ip) + 0];
                b3x[3 * i + 1] = (R)0.5 * b[3 * (row + i) + 1] + b[3 * (rowjp + i) + 1] / (R)2;
                b3x[3 * i + 2] = (R)0.5 * b[3 * (row + i) + 2] + b[3 * (rowkp + i) + 2] / (R)2;
            }
            FN(tvd1)(pu, b3x, nx, dt);
            memcpy(u + 5 * row, pu, sizeof(R) * 5 * nx);
        }
        free(pu);
    }
}

/* tvdb: 1-D TVD upwind advection of a face field; flux on the right face.
 * Heterogeneous_MHD/C/tvdb.c:10-91.  b is updated in place, flux returned. */
static void FN(tvdb)(R *flux, R *b, const R *vg, int n, R dt)
{
    R *vh = (R *)malloc(sizeof(R) * 4 * n);
    R *bv = vh + n, *flux1 = bv + n, *b1 = flux1 + n;
    for (int i = 0; i < n; i++) {
        const int ip = (i + 1) % n;
        vh[i] = FN(abxyc)((R)0.5, vg[i], (R)0.5, vg[ip], 0);   /* tvdb.c:26-27 */
        bv[i] = b[i] * vg[i];                                   /* tvdb.c:31-34 */
    }
    for (int i = 0; i < n; i++) flux1[i] = (vh[i] > 0) ? bv[i] : bv[(i + 1) % n]; /* :35-46 */
    for (int i = 0; i < n; i++) {
        const int im = (i + n - 1) % n;
        /* tvdb.c:51: b - (flux1-flux1[i-1])*dt/2.0, the halving and subtraction in double */
        b1[i] = (R)((double)b[i] - (double)((flux1[i] - flux1[im]) * dt) / 2.0);
    }
    for (int i = 0; i < n; i++) {                               /* tvdb.c:54-78 */
        const int ip = (i + 1) % n, ipp = (ip + 1) % n, im = (i + n - 1) % n;
        R w, wp, wm;
        if (vh[i] > 0) {
            w = vg[i] * b1[i];
            wp = (R)((double)(vg[ip] * b1[ip] - w) / 2.0);
            wm = (R)((double)(w - vg[im] * b1[im]) / 2.0);
        } else {
            w = vg[ip] * b1[ip];
            wp = (R)((double)(w - vg[ipp] * b1[ipp]) / 2.0);
            wm = (R)((double)(vg[i] * b1[i] - w) / 2.0);
        }
        const R dw = FN(vanleer)(wm, wp);
        flux[i] = (w + dw) * dt;
    }
    for (int i = 0; i < n; i++) b1[i] = b[i] - (flux[i] - flux[(i + n - 1) % n]); /* :80-88 */
    memcpy(b, b1, sizeof(R) * n);
    free(vh);
}

/* advectbyzx: constrained-transport advection of b2 along the pencil with the
 * velocity averaged to the j-face (A), then of b3 with the velocity averaged
 * to the k-face (B), each with the EMF correction of b1.
 * Heterogeneous_MHD/C/advectbyzx.c:14-23, advectbyzxA1.c:18-63,
 * advectbyzxA2.c:9-28, advectbyzxB1.c:19-62, advectbyzxB2.c:9-28.
 * b1 ends as (((b1 - FA_j) + FA_{j+1}) - FB_k) + FB_{k+1}, the reference's
 * rounding order. */
void FN(mhd_oracle_advectbyzx)(const R *u, R *b, int nx, int ny, int nz, R dt)
{
    const size_t ncell = (size_t)nx * ny * nz;
    R *tmp = (R *)malloc(sizeof(R) * ncell);
    for (int pass = 0; pass < 2; pass++) {
        /* pass 0 = A (neighbour jm, field b2), pass 1 = B (neighbour km, field b3) */
        const int fc = pass == 0 ? 1 : 2;
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nz; k++) {
            R *vx = (R *)malloc(sizeof(R) * 4 * nx);
            R *vs = vx + nx, *bx = vs + nx, *fx = bx + nx;
            for (int j = 0; j < ny; j++) {
                const int jm = pass == 0 ? (j + ny - 1) % ny : j;
                const int km = pass == 0 ? k : (k + nz - 1) % nz;
                const size_t row = ((size_t)k * ny + j) * nx;
                const size_t rowm = ((size_t)km * ny + jm) * nx;
                for (int i = 0; i < nx; i++)   /* A1.c:34 / B1.c:34 */
                    vx[i] = (u[5 * (rowm + i) + 1] + u[5 * (row + i) + 1]) /
                            (u[5 * (rowm + i) + 0] + u[5 * (row + i) + 0]);
                for (int i = 0; i < nx; i++) { /* A1.c:41: (vx[i-1]+vx[i+1] + 2.0*vx)/4.0 */
                    const int ip = (i + 1) % nx, im = (i + nx - 1) % nx;
                    vs[i] = (R)(((double)(vx[im] + vx[ip]) + 2.0 * (double)vx[i]) / 4.0);
                    bx[i] = b[3 * (row + i) + fc];
                }
                FN(tvdb)(fx, bx, vs, nx, dt);
                for (int i = 0; i < nx; i++) { /* A1.c:52-56 */
                    const R f = fx[(i + nx - 1) % nx];
                    b[3 * (row + i) + fc] = bx[i];
                    b[3 * (row + i) + 0] = b[3 * (row + i) + 0] - f;
                    tmp[row + i] = f;
                }
            }
            free(vx);
        }
        /* A2.c:16-26 / B2.c:16-26: b1 at the minus neighbour gains the flux */
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++) {
                const int jm = pass == 0 ? (j + ny - 1) % ny : j;
                const int km = pass == 0 ? k : (k + nz - 1) % nz;
                const size_t row = ((size_t)k * ny + j) * nx;
                const size_t rowm = ((size_t)km * ny + jm) * nx;
                for (int i = 0; i < nx; i++) b[3 * (rowm + i)] = b[3 * (rowm + i)] + tmp[row + i];
            }
    }
    free(tmp);
}

/* transpose12: (i,j,k) -> (j,i,k), u components (1,3,2,4,5), b (2,1,3).
 * Heterogeneous_MHD/C/transpose12.c:7-47.  out is (ny,nx,nz). */
void FN(mhd_oracle_transpose12)(R *uo, R *bo, const R *u, const R *b, int nx, int ny, int nz)
{
    static const int pu[5] = {0, 2, 1, 3, 4}, pb[3] = {1, 0, 2};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const size_t s = ((size_t)k * ny + j) * nx + i;
                const size_t d = ((size_t)k * nx + i) * ny + j;
                for (int c = 0; c < 5; c++) uo[5 * d + c] = u[5 * s + pu[c]];
                for (int c = 0; c < 3; c++) bo[3 * d + c] = b[3 * s + pb[c]];
            }
}

/* transpose13: (i,j,k) -> (k,j,i), u components (1,4,3,2,5), b (3,2,1).
 * Heterogeneous_MHD/C/transpose13.c:9-62.  out is (nz,ny,nx).  The reference
 * only acts when nx == nz (else it prints and leaves the state alone); this
 * restatement is the same permutation for any shape. */
void FN(mhd_oracle_transpose13)(R *uo, R *bo, const R *u, const R *b, int nx, int ny, int nz)
{
    static const int pu[5] = {0, 3, 2, 1, 4}, pb[3] = {2, 1, 0};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const size_t s = ((size_t)k * ny + j) * nx + i;
                const size_t d = ((size_t)i * ny + j) * nz + k;
                for (int c = 0; c < 5; c++) uo[5 * d + c] = u[5 * s + pu[c]];
                for (int c = 0; c < 3; c++) bo[3 * d + c] = b[3 * s + pb[c]];
            }
}

/* calcfl: 1/max over cells of max|v| + sqrt(|2B^2 + gamma p|/rho), with the
 * face values of B at (i,j,k) (Heterogeneous_MHD/C/calcfl.c:13-65; the cell
 * averages bx,by,bz computed at :38-43 are unused there).  avg_b != 0 selects
 * those averages of calcfl.c:38-43 as B^2 (the choice of HIP/cu/cuda_cfl.cu:
 * 100-115 and OpenMP/mhd.f90:112-120, here with the C path's arithmetic).
 * The max is the reference's running c = max_2num(c, t) from c = 0 in
 * (k,j,i) order (calcfl.c:26-57, :93-103), NaN semantics included: a NaN t
 * makes c NaN and the next t replaces it.  Each plane k runs that recursion
 * from its first cell, recording whether it met a NaN; the planes are then
 * chained in order (a plane with a NaN sets c to its own tail value, any
 * other plane p gives max_2num(c, max of p)), which is the sequential result
 * whatever the thread count.  Returns 1/c computed in double and rounded to
 * R (calcfl.c:61). */
R FN(mhd_oracle_calcfl)(const R *u, const R *b, int nx, int ny, int nz, int avg_b)
{
    const R gamma = (R)(5.0 / 3.0);
    R *plane_c = (R *)malloc(sizeof(R) * (size_t)nz);
    int *plane_nan = (int *)malloc(sizeof(int) * (size_t)nz);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++) {
        R cl = (R)-INFINITY;
        int seen = 0;
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const size_t s = ((size_t)k * ny + j) * nx + i;
                const R u1 = u[5 * s], u2 = u[5 * s + 1], u3 = u[5 * s + 2], u4 = u[5 * s + 3], u5 = u[5 * s + 4];
                R B;
                if (avg_b) {
                    const size_t sip = ((size_t)k * ny + j) * nx + (i + 1) % nx;
                    const size_t sjp = ((size_t)k * ny + (j + 1) % ny) * nx + i;
                    const size_t skp = ((size_t)((k + 1) % nz) * ny + j) * nx + i;
                    /* calcfl.c:39-43: float sum, halved in double */
                    const R bx = (R)((double)(b[3 * s + 0] + b[3 * sip + 0]) / 2.0);
                    const R by = (R)((double)(b[3 * s + 1] + b[3 * sjp + 1]) / 2.0);
                    const R bz = (R)((double)(b[3 * s + 2] + b[3 * skp + 2]) / 2.0);
                    B = bx * bx + by * by + bz * bz;
                } else {
                    B = b[3 * s] * b[3 * s] + b[3 * s + 1] * b[3 * s + 1] + b[3 * s + 2] * b[3 * s + 2];
                }
                /* calcfl.c:45-48 */
                const R t1 = (R)fabs((double)(u2 / u1)), t2 = (R)fabs((double)(u3 / u1)),
                        t3 = (R)fabs((double)(u4 / u1));
                const R v = t1 > t2 ? (t1 > t3 ? t1 : t3) : (t2 > t3 ? t2 : t3); /* max_3num */
                const R S = u2 * u2 + u3 * u3 + u4 * u4;
                const R ps = (R)(((double)u5 - (double)(S / u1) / 2.0) * ((double)gamma - 1.0) +
                                 (2.0 - (double)gamma) * (double)B / 2.0); /* calcfl.c:50 */
                const R p = (R)((double)ps - (double)B / 2.0);             /* calcfl.c:52 */
                /* calcfl.c:54: v + sqrt(fabs(B*2.0 + gamma*p)/u1), in double */
                const R t = (R)((double)v + sqrt(fabs((double)B * 2.0 + (double)(gamma * p)) / (double)u1));
                if (t != t) seen = 1;
                cl = cl > t ? cl : t; /* max_2num */
            }
        plane_c[k] = cl;
        plane_nan[k] = seen;
    }
    R c = (R)0.0;
    if (nx > 0 && ny > 0)
        for (int k = 0; k < nz; k++) c = plane_nan[k] ? plane_c[k] : (c > plane_c[k] ? c : plane_c[k]);
    free(plane_c);
    free(plane_nan);
    return (R)(1.0 / (double)c);
}

/* One loop iteration of the reference driver, Heterogeneous_MHD/C/mhd.c:61-119:
 * dt = 0.9*cfl in double rounded to R, clamp to (tf-t)/2, t += 2*dt, then
 * the Strang-split sweeps with explicit transposes.  Returns dt. */
R FN(mhd_oracle_step)(R *u, R *b, int nx, int ny, int nz, R *t, R tf, int avg_b)
{
    const size_t ncell = (size_t)nx * ny * nz;
    R *ut = (R *)malloc(sizeof(R) * 8 * ncell);
    R *bt = ut + 5 * ncell;
    const R cfl = FN(mhd_oracle_calcfl)(u, b, nx, ny, nz, avg_b);
    R dt = (R)(0.9 * (double)cfl);
    if (dt > (tf - *t) / 2) dt = (tf - *t) / 2;
    *t = *t + 2 * dt;
#define T12(a, bb, c) do { FN(mhd_oracle_transpose12)(ut, bt, u, b, a, bb, c); \
        memcpy(u, ut, sizeof(R) * 5 * ncell); memcpy(b, bt, sizeof(R) * 3 * ncell); } while (0)
#define T13(a, bb, c) do { FN(mhd_oracle_transpose13)(ut, bt, u, b, a, bb, c); \
        memcpy(u, ut, sizeof(R) * 5 * ncell); memcpy(b, bt, sizeof(R) * 3 * ncell); } while (0)
    FN(mhd_oracle_fluidx)(u, b, nx, ny, nz, dt);
    FN(mhd_oracle_advectbyzx)(u, b, nx, ny, nz, dt);
    T12(nx, ny, nz);
    FN(mhd_oracle_fluidx)(u, b, ny, nx, nz, dt);
    FN(mhd_oracle_advectbyzx)(u, b, ny, nx, nz, dt);
    T13(ny, nx, nz);
    FN(mhd_oracle_fluidx)(u, b, nz, nx, ny, dt);
    FN(mhd_oracle_advectbyzx)(u, b, nz, nx, ny, dt);
    FN(mhd_oracle_advectbyzx)(u, b, nz, nx, ny, dt);
    FN(mhd_oracle_fluidx)(u, b, nz, nx, ny, dt);
    T13(nz, nx, ny);
    FN(mhd_oracle_advectbyzx)(u, b, ny, nx, nz, dt);
    FN(mhd_oracle_fluidx)(u, b, ny, nx, nz, dt);
    T12(ny, nx, nz);
    FN(mhd_oracle_advectbyzx)(u, b, nx, ny, nz, dt);
    FN(mhd_oracle_fluidx)(u, b, nx, ny, nz, dt);
#undef T12
#undef T13
    free(ut);
    return dt;
}

/* nsteps loop iterations; dts[s] receives each step's dt when non-NULL. */
void FN(mhd_oracle_run)(R *u, R *b, int nx, int ny, int nz, int nsteps, R tf, int avg_b, R *dts)
{
    R t = 0;
    for (int s = 0; s < nsteps; s++) {
        const R dt = FN(mhd_oracle_step)(u, b, nx, ny, nz, &t, tf, avg_b);
        if (dts) dts[s] = dt;
    }
}

#undef FN
#undef CAT
#undef CAT_
