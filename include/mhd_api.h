/*
 * mhd_api.h - C ABI of the MI355X (gfx950) relaxing-TVD MHD timestep library
 * (libmhdgpu.so).  Plain pointers and sizes only.
 *
 * Layout everywhere: the reference's Fortran-order AoS arrays
 *   u[((k*ny+j)*nx+i)*5+c]  c = rho, rho vx, rho vy, rho vz, E   (cell centred)
 *   b[((k*ny+j)*nx+i)*3+c]  c = bx, by, bz on the left faces of cell (i,j,k)
 * periodic in all three directions (HIP/array_definition.h:5-7,
 * Heterogeneous_MHD/C/array_definition.h:12-14, OpenMP/mhd.f90:35).
 *
 * Errors: every entry point returns MHD_OK (0) or a negative MHD_ERR_* code
 * and never exits; mhd_last_error() gives the message of the calling thread's
 * last failure.  The reference's wrappers return void and either exit(-1)
 * (HIP/cuda_function.h:4-12) or ignore errors (HIP/hip_function.h:5-13).
 */
#ifndef MHD_API_H
#define MHD_API_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MHD_OK 0
#define MHD_ERR_ARG (-1)      /* bad pointer or argument */
#define MHD_ERR_SHAPE (-2)    /* unsupported grid shape: pencil lengths 2..1024 (FP64 2..512), as the
                                 reference's one block of n threads; the batched and served forms need
                                 n = L*K with L <= 64, K <= 8 */
#define MHD_ERR_HIP (-3)      /* HIP runtime error (message in mhd_last_error) */
#define MHD_ERR_NOMEM (-4)    /* device allocation failed or budget exceeded */
#define MHD_ERR_STATE (-5)    /* call not valid in the object's current state */
#define MHD_ERR_IO (-6)       /* file could not be read */
#define MHD_ERR_FULL (-7)     /* request refused: queue full / over budget */
#define MHD_ERR_NONFINITE (-8)/* non-finite value found where a finite one is required */

const char *mhd_last_error(void);
/* library version string, e.g. "mhdgpu 1.0 gfx950" */
const char *mhd_version(void);

/* ---- device memory plumbing for callers without a HIP toolchain --------- */
int mhd_dev_alloc(void **ptr, size_t bytes);
int mhd_dev_free(void *ptr);
int mhd_memcpy_h2d(void *dst, const void *src, size_t bytes);
int mhd_memcpy_d2h(void *dst, const void *src, size_t bytes);
int mhd_memcpy_d2d(void *dst, const void *src, size_t bytes);
int mhd_device_sync(void);
/* device-to-device copy by a plain streaming kernel (bytes % 16 == 0), the
 * yardstick for HBM byte counters; no counterpart in the reference */
int mhd_dev_copy_kernel(void *dst, const void *src, size_t bytes);
/* free and total device memory in bytes */
int mhd_mem_info(size_t *free_bytes, size_t *total_bytes);
/* hardware property the FP32 limiter relies on (no counterpart in the
 * reference): *rcp_f64_err = max |1 - D*rcp(D)| of the FP64 reciprocal
 * instruction over every finite nonzero float-valued D (exhaustive scan,
 * a few milliseconds); the limiter's exactness needs it below 2^-22 */
int mhd_selfcheck(double *rcp_f64_err);
/* hardware property the FP32 divisions of the sweeps rely on (no
 * counterpart in the reference): mismatches[0] counts the pairs a, d in
 * [1, 2) whose one-correction quotient from the refined v_rcp_f32 differs
 * from the IEEE a/d, over every dividend and the divisor mantissas 0,
 * stride, 2*stride, ... (stride 1: all 2^46 pairs, about 30 s);
 * mismatches[1] counts the floats of exponent [-40, 40] whose v_rcp_f32 is
 * not exactly the scaled, signed value at the mantissa.  Both must be 0. */
int mhd_selfcheck_division(int stride, unsigned long long *mismatches);
/* hardware property the freezing speed of the sweeps relies on (no
 * counterpart in the reference): *mismatches counts the positive finite
 * floats x whose shortened FP64 square root (one coupled Newton step and one
 * residual correction from the reciprocal square root estimate) differs from
 * the IEEE sqrt((double)x); exhaustive, a few milliseconds.  Must be 0. */
int mhd_selfcheck_sqrt(unsigned long long *mismatches);
/* the FP32 van Leer limiter as the sweeps evaluate it (no counterpart in the
 * reference; a check of the limiter shortcuts on the device): out[i] from
 * the branching form of the tiles, out_nb[i] from the branch-free form of the
 * streamed tvdb, for the operand pairs (p[i], m[i]); host arrays of n floats.
 * Both must equal (float)(2.0*p*m/(p+m)) where the FP32 p*m > 0 and 0
 * elsewhere (tvd1.c:71-81, tvdb.c:72-76). */
int mhd_selfcheck_limiter(const float *p, const float *m, float *out, float *out_nb, int n);

/* ---- step-level drop-ins (device pointers, FP32) ------------------------
 * Each replaces one wrapper of HIP/hip_subroutine.h:4-12 (CUDA originals in
 * HIP/cu/).  Arrays are device pointers in the frame the caller passes (the
 * dims are the frame's, e.g. (ny,nx,nz) after transpose12, as in
 * HIP/cu/cuda_main.cu:105); nx, ny, nz are HOST pointers (the reference also
 * passes device copies d_nx.. for its kernels; here the library reads the
 * host values).  dt is a DEVICE pointer, so a dt computed on the device is
 * used without a round trip.  The calls are synchronous, like the reference's
 * wrappers (cudaThreadSynchronize after each launch, e.g.
 * HIP/cu/cuda_fluidx.cu:19). */

/* replaces cuda_fluidx (HIP/cu/cuda_fluidx.cu:12, HIP/cuda_fluidx.cpp:13):
 * one relaxing-TVD sweep along i of every (j,k) pencil, u updated in place */
int mhd_fluidx(float *u, const float *b, const int *nx, const int *ny, const int *nz, const float *dt);

/* replaces cuda_advectbyzx (HIP/cu/cuda_advectbyzx.cu:15, HIP/cuda_advectbyzx.cpp:16):
 * constrained-transport advection of b2 then b3 along i, with the EMF
 * correction of b1; the reference's N^3 scratch `temp` is not needed */
int mhd_advectbyzx(const float *u, float *b, const int *nx, const int *ny, const int *nz, const float *dt);

/* mhd_advectbyzx without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The input
 * copy of b lives in a fixed per-device array of 2^21 cells (128^3; 24 MiB,
 * held until the process ends); larger boxes return MHD_ERR_SHAPE.  Every
 * call on one GPU shares that array, so calls that overlap on one GPU (from
 * several host threads) race; results equal mhd_advectbyzx's. */
int mhd_advectbyzx_noalloc(const float *u, float *b, const int *nx, const int *ny, const int *nz, const float *dt);

/* replaces cuda_transpose12 (HIP/cu/cuda_transpose12.cu:12, HIP/cuda_transpose12.cpp:13):
 * (i,j,k)->(j,i,k) with u components (1,3,2,4,5) and b (2,1,3); the result
 * is written to ut,bt and copied back into u,b as in the reference, so after
 * the call u,b hold the (ny,nx,nz) frame */
int mhd_transpose12(float *ut, float *bt, float *u, float *b, const int *nx, const int *ny, const int *nz);

/* replaces cuda_transpose13 (HIP/cu/cuda_transpose13.cu:14, HIP/cuda_transpose13.cpp:15):
 * (i,j,k)->(k,j,i) with u components (1,4,3,2,5) and b (3,2,1).  The
 * reference supports only nx == nz (it prints "nz<>nx not supported" and
 * leaves the state alone otherwise, cuda_transpose13.cu:27-42); this
 * permutation is exact for any shape. */
int mhd_transpose13(float *ut, float *bt, float *u, float *b, const int *nx, const int *ny, const int *nz);

/* replaces cuda_transpose12_ (HIP/cu/cuda_transpose12.cu:69-127), the
 * Fortran-callable form: trailing underscore, every argument by reference,
 * HOST arrays; `call mhd_transpose12(ut, bt, u, b, nx, ny, nz, dt)` from
 * Fortran links to it.  u, b are copied to the device, transposed there and
 * the result copied into ut, bt (u, b unchanged; the caller may pass the
 * same arrays for ut and u).  dt is unused, as in the reference.  The
 * transpose13 form is its sibling (no Fortran form in the reference). */
int mhd_transpose12_(float *ut, float *bt, const float *u, const float *b, const int *nx, const int *ny,
                     const int *nz, const float *dt);
int mhd_transpose13_(float *ut, float *bt, const float *u, const float *b, const int *nx, const int *ny,
                     const int *nz, const float *dt);

/* replaces cuda_cfl (HIP/cu/cuda_cfl.cu:14) and calcfl (Heterogeneous_MHD/C/calcfl.c:13):
 * cfl = 1/max over cells of max|v| + sqrt(|2B^2 + gamma p|/rho), with the
 * C path's face-value B^2 (avg_b = 0, calcfl.c:45-56), or (avg_b = 1) the
 * same C-path arithmetic with B^2 of the cell-averaged field
 * ((b(i)+b(i+1))/2 per axis, the averages calcfl.c:38-43 forms and then
 * leaves unused).  The running max keeps calcfl's max_2num semantics: a NaN
 * speed resets it (calcfl.c:55).  Written to *d_cfl (device) and, when
 * h_cfl is not NULL, to *h_cfl (host).  The caller forms dt = 0.9*cfl as
 * Heterogeneous_MHD/C/mhd.c:63 does (the GPU reference instead used 1/c as
 * dt on the device, HIP/cu/cuda_cfl.cu:41-43). */
int mhd_cfl(const float *u, const float *b, const int *nx, const int *ny, const int *nz, int avg_b, float *d_cfl,
            float *h_cfl);

/* ---- program-level drop-ins ---------------------------------------------- */

/* replaces init (HIP/c_subroutine.h:4, Heterogeneous_MHD/C/init.c:11): read
 * an initial-state file of 8 float32 per cell (u1..u5, b1..b3).  order 0 is
 * the reference C reader (file's fastest axis lands on k, init.c:33-54);
 * order 1 is plain Fortran order (x fastest, OpenMP_offloadMIC/work.f90:127-132).
 * u (5 n^3) and b (3 n^3) are caller-owned host arrays; *nx,*ny,*nz give the
 * grid (the reference fixes them at compile time). */
int mhd_init(const char *path, int order, float *u, float *b, const int *nx, const int *ny, const int *nz);

/* write the result files of Heterogeneous_MHD/C/mhd.c:126-163 (records in the
 * order of the reader with order 0): 5 floats per cell to path_u, 3 to path_b */
int mhd_write_result(const char *path_u, const char *path_b, const float *u, const float *b, const int *nx,
                     const int *ny, const int *nz);

/* replaces cuda_main (HIP/cu/cuda_main.cu:14, HIP/cuda_main.cpp:17): host
 * arrays in and out, device-resident loop.  nsteps > 0 runs exactly nsteps
 * loop iterations; nsteps <= 0 runs while t < tf like the reference
 * (tf must then be finite).  dt = 0.9*cfl clamped to (tf-t)/2 as in
 * Heterogeneous_MHD/C/mhd.c:63-65; pass tf = INFINITY for no clamp.  When
 * h_dts is not NULL it receives the dt of each step (capacity max_dts).
 * A NaN or infinite dt (the state held NaN, a non-positive density or no
 * signal speed; the steps follow the reference's arithmetic regardless)
 * returns MHD_ERR_NONFINITE after the state and the dts are written back. */
int mhd_main(float *h_u, float *h_b, const int *nx, const int *ny, const int *nz, int nsteps, float tf,
             float *h_dts, int max_dts, int *steps_done);

/* ---- context API: one device-resident box ------------------------------- */
typedef struct mhd_ctx mhd_ctx;

/* precision: 4 (FP32, the reference's) or 8 (FP64).  avg_b as in mhd_cfl.
 * Allocates 48 B (FP32) or 96 B (FP64) per cell on the current device: the
 * state (32 B), the second field buffer of advectbyzx (12 B) and the
 * scratch plane of its streamed y/z passes (4 B). */
int mhd_ctx_create(mhd_ctx **ctx, int nx, int ny, int nz, int precision, int avg_b);
int mhd_ctx_destroy(mhd_ctx *ctx);
/* host arrays in the ctx's precision (float* or double*) */
int mhd_ctx_set_state(mhd_ctx *ctx, const void *h_u, const void *h_b);
int mhd_ctx_get_state(mhd_ctx *ctx, void *h_u, void *h_b);
/* tf for the dt clamp (default +inf); resets t, the dt history and the step count */
int mhd_ctx_reset_clock(mhd_ctx *ctx, double tf);
/* enqueue nsteps loop iterations on the ctx stream (a captured hipGraph
 * replayed per step; dt stays on the device).  Asynchronous. */
int mhd_ctx_step(mhd_ctx *ctx, int nsteps);
/* same, bracketed by HIP events on the ctx stream; *ms = elapsed time */
int mhd_ctx_step_timed(mhd_ctx *ctx, int nsteps, float *ms);
int mhd_ctx_sync(mhd_ctx *ctx);
/* dt history (in the ctx precision, as double) of the steps since the last
 * reset, at most max; *count = steps taken */
int mhd_ctx_dts(mhd_ctx *ctx, double *h_dts, int max, int *count);
/* t as accumulated in the ctx precision and as an FP64 sum */
int mhd_ctx_time(mhd_ctx *ctx, double *t, double *t64);
/* FP64 monitors: out[6] = total mass, total energy, momentum x, y, z, max |div B| */
int mhd_ctx_monitors(mhd_ctx *ctx, double *out6);
/* FP64 monitors recorded inside the captured step after every step since
 * the last reset (no counterpart in the reference, whose drivers print probe
 * values per step, Heterogeneous_MHD/C/mhd.c:67, HIP/cu/cuda_main.cu:86):
 * h_out[6*s .. 6*s+5] = the six values of mhd_ctx_monitors for the state
 * after step s+1, summed in a fixed order (deterministic run to run).  At
 * most max entries are copied (the device keeps the first 65536); *count =
 * steps taken. */
int mhd_ctx_monitor_history(mhd_ctx *ctx, double *h_out, int max, int *count);
/* device pointers of the state, for zero-copy callers: the context keeps the
 * state as structure of arrays, d_u = 5 planes (rho, rho vx, rho vy, rho vz,
 * E) and d_b = 3 planes (bx, by, bz), each laid out as mhd_ctx_layout says */
int mhd_ctx_device_state(mhd_ctx *ctx, void **d_u, void **d_b);
/* layout of each plane: cell (i,j,k) at i + j*sy + k*sz, planes `plane` values
 * apart (sy >= nx, sz >= sy*ny: padded rows/planes avoid power-of-two strides) */
int mhd_ctx_layout(mhd_ctx *ctx, long *sy, long *sz, long *plane);
/* kernel form of the y/z advectbyzx sweeps (no counterpart in the
 * reference; a tuning and test switch): 0 automatic (the default: LDS tiles
 * for boxes too small to stream, else the fused column stream where
 * nx <= 1024, else the two-pass stream), 1 LDS tiles, 2 two-pass column
 * stream, 3 fused column stream (MHD_ERR_ARG when nx > 1024).  Results are
 * identical in every form. */
int mhd_ctx_set_advect_form(mhd_ctx *ctx, int form);
/* time per kernel of one step, for the roofline: runs `reps` steps with
 * events around every launch; ms_out[0..12] = dt, then the 12 sweeps in
 * order (fluidx x, advect x, fluidx y, advect y, fluidx z, advect z,
 * advect z, fluidx z, advect y, fluidx y, advect x, fluidx x), averaged.
 * The CFL reduction of the next step is fused into the last sweep. */
int mhd_ctx_profile_step(mhd_ctx *ctx, int reps, float *ms_out13);

/* ---- batched form: many independent boxes of uneven sizes ----------------
 * No counterpart in the reference (one compile-time box per program,
 * Heterogeneous_MHD/C/parameter.h:6).  Each box advances with its own dt
 * exactly as a lone context would (results are bit-identical to
 * mhd_ctx_step on the same box).  One step of the batch is captured in a
 * hipGraph: under the default schedule the boxes are grouped by their
 * cells-per-lane K, and each group runs one dt launch and 12 sweep kernels
 * over all its boxes (the last sweep also forms the next step's CFL), the
 * groups concurrently.  FP32.
 * dims holds nx,ny,nz of each box. */
typedef struct mhd_batch mhd_batch;
int mhd_batch_create(mhd_batch **bt, int nbox, const int *dims, int avg_b);
int mhd_batch_destroy(mhd_batch *bt);
int mhd_batch_set_state(mhd_batch *bt, int box, const float *h_u, const float *h_b);
int mhd_batch_get_state(mhd_batch *bt, int box, float *h_u, float *h_b);
int mhd_batch_step(mhd_batch *bt, int nsteps);
int mhd_batch_step_timed(mhd_batch *bt, int nsteps, float *ms);
int mhd_batch_sync(mhd_batch *bt);
int mhd_batch_dts(mhd_batch *bt, int box, double *h_dts, int max, int *count);
/* schedule of the sweeps: blocks per CU of the persistent kernels (0 =
 * occupancy limit) and policy (0 = persistent, one queue per XCD with
 * stealing; 1 = persistent, one global queue; 2 = one kernel per sweep over
 * all boxes, one block per work item; 3 = the default: boxes grouped by K,
 * one exact-K kernel per sweep and group, the groups' chains concurrent) */
int mhd_batch_set_schedule(mhd_batch *bt, int blocks_per_cu, int policy);
/* times the candidate schedules on the batch itself (trial_steps each),
 * restores every box's state and clock, keeps the fastest */
int mhd_batch_autotune(mhd_batch *bt, int trial_steps, int *blocks_per_cu, int *policy, float *ms_per_step);

/* ---- served form: a request loop with continuous batching ---------------
 * No counterpart in the reference.  mode 0: requests are admitted in arrival
 * order while their device state fits hbm_budget bytes (0 = 90% of the free
 * device memory), at most max_active run at once (0 = 256) and, beyond the
 * first, their cells stay under max_active_cells (0 = 16 Mi, enough to keep
 * every CU busy); admitted requests stay resident and advance together
 * through the batched engine, new ones joining at the next step boundary
 * (results are bit-identical to running each request alone).  mode 1: the FIFO baseline, one request at a time
 * through the device-resident plain path.  A service thread does the GPU
 * work; submit/poll/pending/stats are thread-safe.  The caller keeps h_u,h_b
 * (input, AoS) and out_u,out_b (result) alive until the request is polled.
 * A request larger than the whole budget is refused with MHD_ERR_FULL. */
typedef struct mhd_server mhd_server;
int mhd_server_create(mhd_server **sv, int mode, size_t hbm_budget, int max_active, long long max_active_cells);
int mhd_server_submit(mhd_server *sv, int nx, int ny, int nz, int nsteps, const float *h_u, const float *h_b,
                      float *out_u, float *out_b, long long *req_id);
/* completed requests since the last poll: id, time from submit to result,
 * time spent queued before admission, return code; *n of at most max */
int mhd_server_poll(mhd_server *sv, long long *ids, double *latency_s, double *queue_s, int *rcs, int max, int *n);
int mhd_server_pending(mhd_server *sv, int *queued, int *active);
/* out[7]: zone-updates, seconds first arrival -> last completion, GPU busy
 * seconds, batched steps, work-list rebuilds, bytes resident, budget */
int mhd_server_stats(mhd_server *sv, double *out7);
/* out[4]: seconds the service thread spent uploading admitted requests,
 * rebuilding the work lists and the step graph, stepping (launch and wait)
 * and reading finished requests back (continuous batching; zeros for FIFO).
 * No counterpart in the reference (its driver has no request loop). */
int mhd_server_timing(mhd_server *sv, double *out4);
/* finishes every submitted request, then stops the service thread */
int mhd_server_destroy(mhd_server *sv);

#ifdef __cplusplus
}
#endif
#endif
