// C-ABI of libmhdgpu.so: step-level drop-ins, the program-level driver and
// the device-resident context (include/mhd_api.h).
#include "../../include/mhd_api.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "mhd_kernels.h"
#include "mhd_generic.h"
#include "mhd_internal.h"

namespace mhd {

static thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *where)
{
    return fail(MHD_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
}

// a pencil length the sweeps take in precision prec: the fast kernels'
// n = L*K (K <= 8), else the generic ones' 2 <= n <= 1024 (FP64: 512)
static bool pencil_ok(int n, int prec) { return pick_k(n) || gen_k(n, prec); }

int check_shape(int nx, int ny, int nz, int prec)
{
    if (nx < 2 || ny < 1 || nz < 1) return fail(MHD_ERR_SHAPE, "grid %dx%dx%d: need nx>=2, ny>=1, nz>=1", nx, ny, nz);
    if ((long)nx * ny * nz * 5 >= (1L << 31))
        return fail(MHD_ERR_SHAPE, "grid %dx%dx%d: 5*cells must stay below 2^31", nx, ny, nz);
    if (!pencil_ok(nx, prec))
        return fail(MHD_ERR_SHAPE, "pencil length %d unsupported (2..1024 cells; FP64 2..512)", nx);
    return MHD_OK;
}

// all three sweep directions of a box (the context path sweeps y and z too)
int check_shape3(int nx, int ny, int nz, int prec)
{
    int rc = check_shape(nx, ny, nz, prec);
    if (rc) return rc;
    if (ny < 2 || nz < 2) return fail(MHD_ERR_SHAPE, "grid %dx%dx%d: the full step needs ny, nz >= 2", nx, ny, nz);
    if (!pencil_ok(ny, prec) || !pencil_ok(nz, prec))
        return fail(MHD_ERR_SHAPE, "grid %dx%dx%d: pencil length unsupported (2..1024 cells; FP64 2..512)", nx, ny,
                    nz);
    return MHD_OK;
}

int check_shape_fast3(int nx, int ny, int nz)
{
    int rc = check_shape3(nx, ny, nz, 4);
    if (rc) return rc;
    if (!pick_k(nx) || !pick_k(ny) || !pick_k(nz))
        return fail(MHD_ERR_SHAPE,
                    "grid %dx%dx%d: the batched and served forms need every pencil length n = L*K with K <= 8 and "
                    "L <= 64 (the plain context takes any length up to 1024)",
                    nx, ny, nz);
    return MHD_OK;
}

// per-device scratch of the step-level entry points: the CFL max accumulator
// (the caller holds g_scr_mu from the lookup until the accumulator is read
// and cleared, so concurrent mhd_cfl calls do not share it)
static std::mutex g_scr_mu;
static unsigned long long *g_scr[64];

static int step_scratch(unsigned long long **p)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= 64) return fail(MHD_ERR_ARG, "device %d out of range", dev);
    if (!g_scr[dev]) {
        e = hipMalloc((void **)&g_scr[dev], 64);
        if (e != hipSuccess) return fail(MHD_ERR_NOMEM, "hipMalloc scratch: %s", hipGetErrorString(e));
        e = launch_zero_u64(g_scr[dev], 8, 0);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) return hip_fail(e, "scratch init");
    }
    *p = g_scr[dev];
    return MHD_OK;
}

// per-device copy buffer of the step-level advectbyzx (the input b it reads
// while it writes the caller's b), grown on demand
static std::mutex g_bcopy_mu;
static void *g_bcopy[64];
static size_t g_bcopy_bytes[64];

static int bcopy_scratch(size_t bytes, void **p)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= 64) return fail(MHD_ERR_ARG, "device %d out of range", dev);
    if (g_bcopy_bytes[dev] < bytes) {
        if (g_bcopy[dev]) (void)hipFree(g_bcopy[dev]);
        g_bcopy[dev] = nullptr;
        g_bcopy_bytes[dev] = 0;
        e = hipMalloc(&g_bcopy[dev], bytes);
        if (e != hipSuccess) return fail(MHD_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        g_bcopy_bytes[dev] = bytes;
    }
    *p = g_bcopy[dev];
    return MHD_OK;
}

// Fixed scratch of mhd_advectbyzx_noalloc: the input copy of b that the
// sweep reads while it writes the caller's b, for boxes of at most
// kNoAllocCells cells (3 floats each: 24 MiB per device, held until the
// process ends).  In host code `scratch` names a host-side stand-in, so the
// device address of each GPU's copy is looked up once per device and kept.
static constexpr long kNoAllocCells = 1L << 21;  // 128^3
static __device__ __align__(256) float scratch[3 * kNoAllocCells];
static std::mutex g_noalloc_mu;
static float *g_noalloc_ptr[64];

static int noalloc_scratch(float **p)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= 64) return fail(MHD_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *q = nullptr;
        e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!q) return fail(MHD_ERR_HIP, "hipGetSymbolAddress(scratch) gave a null address");
        g_noalloc_ptr[dev] = (float *)q;
    }
    *p = g_noalloc_ptr[dev];
    return MHD_OK;
}

static int sync_check(hipError_t e, const char *where)
{
    if (e != hipSuccess) return hip_fail(e, where);
    e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(e, where);
    return MHD_OK;
}

}  // namespace mhd

using namespace mhd;

extern "C" {

const char *mhd_last_error(void) { return g_err.c_str(); }
const char *mhd_version(void) { return "mhdgpu 1.0 gfx950"; }

int mhd_dev_alloc(void **ptr, size_t bytes)
{
    if (!ptr) return fail(MHD_ERR_ARG, "mhd_dev_alloc: null ptr");
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(MHD_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return MHD_OK;
}

int mhd_dev_free(void *ptr)
{
    hipError_t e = hipFree(ptr);
    return e == hipSuccess ? MHD_OK : hip_fail(e, "hipFree");
}

int mhd_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    return e == hipSuccess ? MHD_OK : hip_fail(e, "hipMemcpy H2D");
}

int mhd_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    return e == hipSuccess ? MHD_OK : hip_fail(e, "hipMemcpy D2H");
}

int mhd_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice);
    return e == hipSuccess ? MHD_OK : hip_fail(e, "hipMemcpy D2D");
}

int mhd_dev_copy_kernel(void *dst, const void *src, size_t bytes)
{
    if (!dst || !src || bytes % 16) return fail(MHD_ERR_ARG, "mhd_dev_copy_kernel: bad argument");
    return sync_check(launch_copy(dst, src, bytes, 0), "mhd_dev_copy_kernel");
}

int mhd_selfcheck(double *rcp_f64_err)
{
    if (!rcp_f64_err) return fail(MHD_ERR_ARG, "mhd_selfcheck: null argument");
    unsigned long long *d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof *d);
    if (e != hipSuccess) return hip_fail(e, "mhd_selfcheck");
    unsigned long long h = 0;
    e = hipMemset(d, 0, sizeof *d);
    if (e == hipSuccess) e = launch_rcp_scan(d, 0);
    if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof h, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "mhd_selfcheck");
    memcpy(rcp_f64_err, &h, sizeof h);
    return MHD_OK;
}

int mhd_selfcheck_division(int stride, unsigned long long *mismatches)
{
    if (!mismatches || stride < 1) return fail(MHD_ERR_ARG, "mhd_selfcheck_division: bad argument");
    unsigned long long *d = nullptr;
    hipError_t e = hipMalloc(&d, 2 * sizeof *d);
    if (e != hipSuccess) return hip_fail(e, "mhd_selfcheck_division");
    unsigned long long h[2] = {0, 0};
    e = hipMemset(d, 0, 2 * sizeof *d);
    if (e == hipSuccess) e = launch_div_scan((unsigned)stride, d, 0);
    if (e == hipSuccess) e = hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "mhd_selfcheck_division");
    mismatches[0] = h[0];
    mismatches[1] = h[1];
    return MHD_OK;
}

int mhd_selfcheck_sqrt(unsigned long long *mismatches)
{
    if (!mismatches) return fail(MHD_ERR_ARG, "mhd_selfcheck_sqrt: null argument");
    unsigned long long *d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof *d);
    if (e != hipSuccess) return hip_fail(e, "mhd_selfcheck_sqrt");
    unsigned long long h = 0;
    e = hipMemset(d, 0, sizeof *d);
    if (e == hipSuccess) e = launch_sqrt_scan(d, 0);
    if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof h, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "mhd_selfcheck_sqrt");
    *mismatches = h;
    return MHD_OK;
}

int mhd_selfcheck_limiter(const float *p, const float *m, float *out, float *out_nb, int n)
{
    if (!p || !m || !out || !out_nb || n < 0) return fail(MHD_ERR_ARG, "mhd_selfcheck_limiter: bad argument");
    if (n == 0) return MHD_OK;
    const size_t bytes = (size_t)n * sizeof(float);
    float *d = nullptr;
    hipError_t e = hipMalloc(&d, 4 * bytes);
    if (e != hipSuccess) return hip_fail(e, "mhd_selfcheck_limiter");
    e = hipMemcpy(d, p, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + n, m, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_limiter_eval(d, d + n, d + 2 * (size_t)n, d + 3 * (size_t)n, n, 0);
    if (e == hipSuccess) e = hipMemcpy(out, d + 2 * (size_t)n, bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(out_nb, d + 3 * (size_t)n, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "mhd_selfcheck_limiter");
    return MHD_OK;
}

int mhd_device_sync(void)
{
    hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? MHD_OK : hip_fail(e, "hipDeviceSynchronize");
}

int mhd_mem_info(size_t *free_bytes, size_t *total_bytes)
{
    size_t f = 0, t = 0;
    hipError_t e = hipMemGetInfo(&f, &t);
    if (e != hipSuccess) return hip_fail(e, "hipMemGetInfo");
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return MHD_OK;
}

int mhd_fluidx(float *u, const float *b, const int *nx, const int *ny, const int *nz, const float *dt)
{
    if (!u || !b || !nx || !ny || !nz || !dt) return fail(MHD_ERR_ARG, "mhd_fluidx: null argument");
    int rc = check_shape(*nx, *ny, *nz, 4);
    if (rc) return rc;
    return sync_check(launch_fluidx<float>(u, b, 0, *nx, *ny, *nz, dt, 0), "mhd_fluidx");
}

int mhd_advectbyzx(const float *u, float *b, const int *nx, const int *ny, const int *nz, const float *dt)
{
    if (!u || !b || !nx || !ny || !nz || !dt) return fail(MHD_ERR_ARG, "mhd_advectbyzx: null argument");
    int rc = check_shape(*nx, *ny, *nz, 4);
    if (rc) return rc;
    const size_t bytes = (size_t)*nx * *ny * *nz * 3 * sizeof(float);
    std::lock_guard<std::mutex> lk(g_bcopy_mu);
    void *bin = nullptr;
    rc = bcopy_scratch(bytes, &bin);
    if (rc) return rc;
    // the field's copy for the sweep to read: the 16-byte copy kernel where
    // both ends allow it (measured faster than the runtime's copy here)
    const bool vec = bytes % 16 == 0 && (uintptr_t)b % 16 == 0 && (uintptr_t)bin % 16 == 0;
    hipError_t e = vec ? launch_copy(bin, b, bytes, 0) : hipMemcpyAsync(bin, b, bytes, hipMemcpyDeviceToDevice, 0);
    if (e == hipSuccess) e = launch_advect<float>(u, (const float *)bin, b, 0, *nx, *ny, *nz, dt, 0);
    return sync_check(e, "mhd_advectbyzx");
}

int mhd_advectbyzx_noalloc(const float *u, float *b, const int *nx, const int *ny, const int *nz, const float *dt)
{
    if (!u || !b || !nx || !ny || !nz || !dt) return fail(MHD_ERR_ARG, "mhd_advectbyzx_noalloc: null argument");
    int rc = check_shape(*nx, *ny, *nz, 4);
    if (rc) return rc;
    const long cells = (long)*nx * *ny * *nz;
    if (cells > kNoAllocCells)
        return fail(MHD_ERR_SHAPE,
                    "mhd_advectbyzx_noalloc: %dx%dx%d = %ld cells exceed its fixed scratch of %ld cells; "
                    "use mhd_advectbyzx",
                    *nx, *ny, *nz, cells, kNoAllocCells);
    float *bin = nullptr;
    rc = noalloc_scratch(&bin);
    if (rc) return rc;
    hipError_t e = launch_copy_f32(bin, b, 3 * cells, 0);
    if (e == hipSuccess) e = launch_advect<float>(u, bin, b, 0, *nx, *ny, *nz, dt, 0);
    return sync_check(e, "mhd_advectbyzx_noalloc");
}

static int transpose_common(int which, float *ut, float *bt, float *u, float *b, const int *nx, const int *ny,
                            const int *nz)
{
    if (!ut || !bt || !u || !b || !nx || !ny || !nz) return fail(MHD_ERR_ARG, "mhd_transpose: null argument");
    if (*nx < 1 || *ny < 1 || *nz < 1) return fail(MHD_ERR_SHAPE, "mhd_transpose: bad grid");
    // the new frame into ut,bt and into u,b, as the reference's transpose
    // and copy back leave them (HIP/cu/cuda_transpose12.cu:33-34): in place
    // for square planes, else out of place and copied
    const hipError_t e = launch_transpose<float>(ut, bt, u, b, which, *nx, *ny, *nz, 0, true);
    return sync_check(e, "mhd_transpose");
}

int mhd_transpose12(float *ut, float *bt, float *u, float *b, const int *nx, const int *ny, const int *nz)
{
    return transpose_common(12, ut, bt, u, b, nx, ny, nz);
}

int mhd_transpose13(float *ut, float *bt, float *u, float *b, const int *nx, const int *ny, const int *nz)
{
    return transpose_common(13, ut, bt, u, b, nx, ny, nz);
}

// host arrays in, the transposed frame out into h_ut, h_bt (h_u, h_b are
// read before anything is written, so a caller may pass the same arrays)
static int transpose_host(int which, float *h_ut, float *h_bt, const float *h_u, const float *h_b, const int *nx,
                          const int *ny, const int *nz)
{
    if (!h_ut || !h_bt || !h_u || !h_b || !nx || !ny || !nz) return fail(MHD_ERR_ARG, "mhd_transpose_: null argument");
    if (*nx < 1 || *ny < 1 || *nz < 1) return fail(MHD_ERR_SHAPE, "mhd_transpose_: bad grid");
    const size_t nc = (size_t)*nx * *ny * *nz, bu = nc * 5 * sizeof(float), bb = nc * 3 * sizeof(float);
    float *d = nullptr;
    hipError_t e = hipMalloc(&d, 2 * (bu + bb));
    if (e != hipSuccess) return fail(MHD_ERR_NOMEM, "mhd_transpose_: hipMalloc: %s", hipGetErrorString(e));
    float *du = d, *db = du + 5 * nc, *dut = db + 3 * nc, *dbt = dut + 5 * nc;
    e = hipMemcpy(du, h_u, bu, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, h_b, bb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_transpose<float>(dut, dbt, du, db, which, *nx, *ny, *nz, 0);
    if (e == hipSuccess) e = hipMemcpy(h_ut, dut, bu, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(h_bt, dbt, bb, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return e == hipSuccess ? MHD_OK : hip_fail(e, "mhd_transpose_");
}

int mhd_transpose12_(float *h_ut, float *h_bt, const float *h_u, const float *h_b, const int *nx, const int *ny,
                     const int *nz, const float *h_dt)
{
    (void)h_dt;  // unused, as in the reference
    return transpose_host(12, h_ut, h_bt, h_u, h_b, nx, ny, nz);
}

int mhd_transpose13_(float *h_ut, float *h_bt, const float *h_u, const float *h_b, const int *nx, const int *ny,
                     const int *nz, const float *h_dt)
{
    (void)h_dt;
    return transpose_host(13, h_ut, h_bt, h_u, h_b, nx, ny, nz);
}

int mhd_cfl(const float *u, const float *b, const int *nx, const int *ny, const int *nz, int avg_b, float *d_cfl,
            float *h_cfl)
{
    if (!u || !b || !nx || !ny || !nz || !d_cfl) return fail(MHD_ERR_ARG, "mhd_cfl: null argument");
    if (*nx < 1 || *ny < 1 || *nz < 1) return fail(MHD_ERR_SHAPE, "mhd_cfl: bad grid");
    unsigned long long *scr = nullptr;
    std::lock_guard<std::mutex> lk(g_scr_mu);
    int rc = step_scratch(&scr);
    if (rc) return rc;
    hipError_t e = launch_cfl<float>(u, b, *nx, *ny, *nz, avg_b, scr, 0);
    if (e == hipSuccess) e = launch_cfl_value<float>(scr, d_cfl, u, b, *nx, *ny, *nz, avg_b ? 1 : 0, 0);
    rc = sync_check(e, "mhd_cfl");
    if (rc) return rc;
    if (h_cfl) return mhd_memcpy_d2h(h_cfl, d_cfl, sizeof(float));
    return MHD_OK;
}

int mhd_init(const char *path, int order, float *u, float *b, const int *nx, const int *ny, const int *nz)
{
    if (!path || !u || !b || !nx || !ny || !nz) return fail(MHD_ERR_ARG, "mhd_init: null argument");
    const int X = *nx, Y = *ny, Z = *nz;
    if (X < 1 || Y < 1 || Z < 1) return fail(MHD_ERR_SHAPE, "mhd_init: bad grid");
    FILE *fp = fopen(path, "rb");
    if (!fp) return fail(MHD_ERR_IO, "mhd_init: cannot open %s", path);
    const size_t nc = (size_t)X * Y * Z;
    std::vector<float> rec(8 * nc);
    const size_t got = fread(rec.data(), sizeof(float), rec.size(), fp);
    fclose(fp);
    if (got != rec.size()) return fail(MHD_ERR_IO, "mhd_init: %s holds %zu floats, need %zu", path, got, rec.size());
    size_t r = 0;
    if (order == 0) {  // i outer, k inner (Heterogeneous_MHD/C/init.c:33-54)
        for (int i = 0; i < X; i++)
            for (int j = 0; j < Y; j++)
                for (int k = 0; k < Z; k++, r++) {
                    const size_t s = ((size_t)k * Y + j) * X + i;
                    for (int c = 0; c < 5; c++) u[5 * s + c] = rec[8 * r + c];
                    for (int c = 0; c < 3; c++) b[3 * s + c] = rec[8 * r + 5 + c];
                }
    } else {  // x fastest
        for (size_t s = 0; s < nc; s++) {
            for (int c = 0; c < 5; c++) u[5 * s + c] = rec[8 * s + c];
            for (int c = 0; c < 3; c++) b[3 * s + c] = rec[8 * s + 5 + c];
        }
    }
    return MHD_OK;
}

int mhd_write_result(const char *path_u, const char *path_b, const float *u, const float *b, const int *nx,
                     const int *ny, const int *nz)
{
    if (!path_u || !path_b || !u || !b || !nx || !ny || !nz) return fail(MHD_ERR_ARG, "mhd_write_result: null");
    const int X = *nx, Y = *ny, Z = *nz;
    FILE *fu = fopen(path_u, "wb");
    FILE *fb = fopen(path_b, "wb");
    if (!fu || !fb) {
        if (fu) fclose(fu);
        if (fb) fclose(fb);
        return fail(MHD_ERR_IO, "mhd_write_result: cannot open output");
    }
    // Heterogeneous_MHD/C/mhd.c:130-161: first array index outermost
    for (int i = 0; i < X; i++)
        for (int j = 0; j < Y; j++)
            for (int k = 0; k < Z; k++) {
                const size_t s = ((size_t)k * Y + j) * X + i;
                fwrite(u + 5 * s, sizeof(float), 5, fu);
                fwrite(b + 3 * s, sizeof(float), 3, fb);
            }
    const int bad = ferror(fu) || ferror(fb);
    fclose(fu);
    fclose(fb);
    return bad ? fail(MHD_ERR_IO, "mhd_write_result: write failed") : MHD_OK;
}

int mhd_main(float *h_u, float *h_b, const int *nx, const int *ny, const int *nz, int nsteps, float tf,
             float *h_dts, int max_dts, int *steps_done)
{
    if (!h_u || !h_b || !nx || !ny || !nz) return fail(MHD_ERR_ARG, "mhd_main: null argument");
    if (nsteps <= 0 && !isfinite(tf)) return fail(MHD_ERR_ARG, "mhd_main: run-to-tf needs a finite tf");
    mhd_ctx *ctx = nullptr;
    int rc = mhd_ctx_create(&ctx, *nx, *ny, *nz, 4, 0);
    if (rc) return rc;
    int done = 0;
    rc = mhd_ctx_set_state(ctx, h_u, h_b);
    if (!rc) rc = mhd_ctx_reset_clock(ctx, tf);
    if (!rc) {
        if (nsteps > 0) {
            rc = mhd_ctx_step(ctx, nsteps);
            done = nsteps;
        } else {
            // while (t < tf) (Heterogeneous_MHD/C/mhd.c:125, HIP/cu/cuda_main.cu:128)
            double t = 0, t64 = 0;
            while (!rc && t < (double)tf) {
                rc = mhd_ctx_step(ctx, 1);
                if (!rc) rc = mhd_ctx_time(ctx, &t, &t64);
                done++;
            }
        }
    }
    if (!rc) rc = mhd_ctx_sync(ctx);
    if (!rc) rc = mhd_ctx_get_state(ctx, h_u, h_b);
    // the dt sequence: into h_dts, and checked (a NaN or infinite dt means the
    // state holds NaN, a non-positive density or an all-zero speed)
    int bad_step = -1;
    float bad_dt = 0.0f;
    if (!rc && done > 0) {
        std::vector<double> d(done);
        int cnt = 0;
        rc = mhd_ctx_dts(ctx, d.data(), done, &cnt);
        for (int i = 0; !rc && i < cnt && i < done; i++) {
            if (h_dts && i < max_dts) h_dts[i] = (float)d[i];
            if (bad_step < 0 && !isfinite(d[i])) {
                bad_step = i;
                bad_dt = (float)d[i];
            }
        }
    }
    if (steps_done) *steps_done = done;
    const int rc2 = mhd_ctx_destroy(ctx);
    if (!rc && !rc2 && bad_step >= 0)
        return fail(MHD_ERR_NONFINITE,
                    "mhd_main: step %d has dt = %g (the state holds NaN, a non-positive density or no signal speed); "
                    "the state and the dts were written back",
                    bad_step, (double)bad_dt);
    return rc ? rc : rc2;
}

}  // extern "C"
