// Host-side launchers of the gfx950 timestep kernels (mhd_kernels.hip).
// All launches are stream-ordered; T is float (the reference's precision and
// evaluation order) or double (the FP64 instantiation).
#pragma once
#include "mhd_common.h"

namespace mhd {

// smallest supported cells-per-lane K with 64*K >= n and n % K == 0; 0 if none
int pick_k(int n);
// pencils per wave of the FP32 sweeps whose pencils fill a wave several at a
// time: 2 or 4 when n = (64/P)*K with K = 1 or 3 (n = 16, 32, 48, 96), else 1
int pick_p(int n);
// pick_p(n) where several pencils per wave pay for a launch over `pencils`
// pencils: where they keep pick_k(n) cells per lane, or where the launch
// still has 4096 waves (four per SIMD); else 1.  Sharing a wave at a larger K
// lengthens each lane's work, which a launch too small to fill the GPU pays
// for in latency (48^3 measured 6% slower at 4 pencils per wave, 96^3 7%
// faster at 2).
int p_pays(int n, long pencils);
Frame make_frame(int fr, int nx, int ny, int nz);

// sweeps over an x-major AoS box nx*ny*nz in frame fr (0: x, 1: y, 2: z)
template <typename T>
hipError_t launch_fluidx(T *u, const T *b, int fr, int nx, int ny, int nz, const T *dt, hipStream_t st);
// reads bin, writes bout (must be distinct buffers)
template <typename T>
hipError_t launch_advect(const T *u, const T *bin, T *bout, int fr, int nx, int ny, int nz, const T *dt,
                         hipStream_t st);

// per-cell CFL speed max into *cmax (IEEE bits, atomically max-ed)
template <typename T>
hipError_t launch_cfl(const T *u, const T *b, int nx, int ny, int nz, int avg, unsigned long long *cmax,
                      hipStream_t st);
// dt from the block maxima of a fused CFL (part) and the accumulator, NaN
// maxima resolved over the SoA state (u, b, g, avg) as the reference's
// sequential max does; clears the accumulator
template <typename T>
hipError_t launch_dt_part(unsigned long long *cmax, const T *part, int np, T *dt, T *t, double *t64, T tf, T *dts,
                          int *iter, int cap, const T *u, const T *b, const SoaGrid &g, int avg, hipStream_t st);
template <typename T> hipError_t launch_fill_zero(T *p, int n, hipStream_t st);
// *out = 1/c from the accumulator (cleared), NaN maxima resolved over the AoS state
template <typename T>
hipError_t launch_cfl_value(unsigned long long *cmax, T *out, const T *u, const T *b, int nx, int ny, int nz, int avg,
                            hipStream_t st);
template <typename T> hipError_t prepare_sweeps(int nx, int ny, int nz);
hipError_t launch_zero_u64(unsigned long long *p, int n, hipStream_t st);
// plain device copy kernel (bytes a multiple of 16)
hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st);
// n floats from src to dst by a kernel (no runtime copy engine)
hipError_t launch_copy_f32(float *dst, const float *src, long n, hipStream_t st);
// *maxbits (zeroed by the caller) = FP64 bits of max |1 - D*rcp_f64(D)| over all float-valued D
hipError_t launch_rcp_scan(unsigned long long *maxbits, hipStream_t st);
hipError_t launch_div_scan(unsigned stride, unsigned long long *mismatches, hipStream_t st);
hipError_t launch_limiter_eval(const float *p, const float *m, float *out, float *out_nb, int n, hipStream_t st);
hipError_t launch_sqrt_scan(unsigned long long *mismatches, hipStream_t st);

// which = 12 or 13: out-of-place transposes with the reference's component maps
template <typename T>
hipError_t launch_transpose(T *uo, T *bo, const T *u, const T *b, int which, int nx, int ny, int nz, hipStream_t st,
                            bool both = false);

// deterministic reduction of np monitor partial records (6 doubles each:
// mass, energy, momentum x, y, z, max |div B|) into out[6] (when not null)
// and, when hist is not null, into hist[6*(*iter-1)..] if that is below cap.
// Above 2048 records part must have room for ceil(np/256) more records
// after the np inputs (a first reduction stage writes them there).
hipError_t launch_monitor_reduce(const double *part, int np, double *out, double *hist, const int *iter, int cap,
                                 hipStream_t st);

}  // namespace mhd
