// Device-resident timestep driver (the mhd_ctx_* entry points).
//
// One loop iteration of the reference (Heterogeneous_MHD/C/mhd.c:61-119,
// HIP/cu/cuda_main.cu:89-121) is 1 cfl reduction + 1 dt kernel + 12 sweeps.
// The reference transposes the whole state four times per step so that
// every sweep runs along the contiguous axis; here each sweep reads its
// pencils in place through a frame (strides plus the transposes' component
// relabelling), so the transposes and their 8 copies disappear.  dt is
// computed and consumed on the device; the step is captured once into a
// hipGraph and replayed.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../include/mhd_api.h"
#include "mhd_internal.h"
#include "mhd_soa.h"

using namespace mhd;

struct mhd_ctx {
    int nx, ny, nz, prec, avg;
    size_t ncell;              // nx*ny*nz
    SoaGrid g;                 // device layout: padded planes of g.N values
    hipStream_t st;
    void *u, *b;               // state: 5 + 3 planes of ncell values (SoA, ctx precision)
    void *b2;                  // second field buffer: advectbyzx reads one, writes the other
    void *fa;                  // scratch plane: the EMF flux of the y/z advectbyzx (pass A -> pass B)
    void *cfl_part;            // block maxima of the CFL speed fused into the step's last sweep
    int nparts;
    unsigned long long *cmax;  // CFL max accumulator
    void *dt, *t;              // device scalars in ctx precision
    double *t64;
    void *dts;                 // dt history
    int cap;
    int *iter;
    double *mon_part, *mon_out;
    int mon_nblk;
    double *mon_tile;          // per-tile monitor partials of the step's last sweep (6 per x-frame tile)
    double *mon_hist;          // FP64 monitors after each step (6 per step, kMonCap steps)
    double tf;
    hipGraph_t graph, graph_n;     // one step; kGraphSteps steps
    hipGraphExec_t exec, exec_n;
    bool captured;
    int adv_form;              // y/z advectbyzx: 0 automatic, 1 tiles, 2 two-pass stream, 3 fused stream
    hipEvent_t ev0, ev1;
};

static const int kDtCap = 1 << 20;
static const int kMonCap = 1 << 16;
// steps per replay of the long graph: one graph launch costs ~9 us of idle
// GPU between steps (measured on MI355X), paid once per kGraphSteps steps
static const int kGraphSteps = 8;

// Padding of the SoA rows and planes (in cells).  A plane of 256x256 floats
// is 256 KiB: z-sweep tiles read rows one plane apart, and at a power-of-two
// plane stride the loads of a wave pile onto the same memory channels (the
// z sweeps ran ~30% slower than the y sweeps).  Padding each plane by 256
// bytes removes that (measured on MI355X: fluidx_z 0.555 -> 0.415 ms at
// 256^3); padding the rows did not help.  MHD_PAD_Y / MHD_PAD_Z (cells)
// override the defaults; they are rounded up to the 16-byte vector width.
static int env_pad(const char *name, int dflt, int prec)
{
    const char *v = getenv(name);
    int p = v ? atoi(v) : dflt;
    const int ve = 16 / prec;
    return p < 0 ? 0 : (p + ve - 1) / ve * ve;
}
namespace mhd {
int soa_pad_y(int prec) { return env_pad("MHD_PAD_Y", 0, prec); }
int soa_pad_z(int prec) { return env_pad("MHD_PAD_Z", 256 / prec, prec); }
}  // namespace mhd

// The CFL of the state a step starts from is the max of the block maxima
// that the previous step's last sweep (fluidx x) wrote, and of the
// accumulator, which a standalone CFL fills whenever the state was set from
// outside (prime_cfl): the dt kernel takes both and clears the accumulator.
template <typename T> static hipError_t enqueue_dt(mhd_ctx *c)
{
    return launch_dt_part<T>(c->cmax, (const T *)c->cfl_part, c->nparts, (T *)c->dt, (T *)c->t, c->t64, (T)c->tf,
                             (T *)c->dts, c->iter, c->cap, (const T *)c->u, (const T *)c->b, c->g, c->avg, c->st);
}

template <typename T> static hipError_t prime_cfl(mhd_ctx *c)
{
    hipError_t e = launch_fill_zero<T>((T *)c->cfl_part, c->nparts, c->st);
    if (e == hipSuccess) e = launch_zero_u64(c->cmax, 1, c->st);
    if (e == hipSuccess) e = launch_cfl_soa<T>((const T *)c->u, (const T *)c->b, c->g, c->avg, c->cmax, c->st);
    return e;
}

static hipError_t prime_cfl_any(mhd_ctx *c) { return c->prec == 4 ? prime_cfl<float>(c) : prime_cfl<double>(c); }

// sweep s of the Strang sequence (HIP/cu/cuda_main.cu:99-121): frame and op
static const int kSweepFrame[12] = {0, 0, 1, 1, 2, 2, 2, 2, 1, 1, 0, 0};
static const int kSweepOp[12] = {0, 1, 0, 1, 0, 1, 1, 0, 1, 0, 1, 0};  // 0 fluidx, 1 advectbyzx

// The field lives in b before sweep 0; each advectbyzx moves it to the other
// buffer.  The six advectbyzx of a step bring it back to b.
static int field_buffer_before(int s)
{
    int nadv = 0;
    for (int i = 0; i < s; i++) nadv += kSweepOp[i];
    return nadv & 1;
}

template <typename T> static hipError_t enqueue_sweep(mhd_ctx *c, int s)
{
    const int fr = kSweepFrame[s];
    T *bcur = (T *)(field_buffer_before(s) ? c->b2 : c->b);
    T *bnext = (T *)(field_buffer_before(s) ? c->b : c->b2);
    const int n = fr == 0 ? c->nx : fr == 1 ? c->ny : c->nz;
    const bool generic = !pick_k(n);  // a pencil length only the generic kernels take
    if (kSweepOp[s] == 0) {
        // the last sweep (fluidx, x frame) also forms the next step's CFL; the
        // generic kernels do not, so a standalone CFL of the new state follows
        // (and a standalone monitor pass); then the FP64 monitors of the new
        // state are reduced into the step's history entry
        if (s == 11 && generic) {
            hipError_t e = launch_fluidx_soa<T>((T *)c->u, bcur, fr, c->g, (const T *)c->dt, c->st);
            if (e == hipSuccess)
                e = launch_cfl_soa<T>((const T *)c->u, (const T *)bcur, c->g, c->avg, c->cmax, c->st);
            if (e == hipSuccess)
                e = launch_monitors_soa<T>((const T *)c->u, (const T *)bcur, c->g, c->mon_tile, c->mon_nblk, nullptr,
                                           c->st);
            if (e == hipSuccess)
                e = launch_monitor_reduce(c->mon_tile, c->mon_nblk, nullptr, c->mon_hist, c->iter, kMonCap, c->st);
            return e;
        }
        if (s == 11) {
            hipError_t e = launch_fluidx_soa<T>((T *)c->u, bcur, fr, c->g, (const T *)c->dt, c->st, (T *)c->cfl_part,
                                                c->avg ? 2 : 1, c->mon_tile);
            if (e == hipSuccess)
                e = launch_monitor_reduce(c->mon_tile, c->nparts, nullptr, c->mon_hist, c->iter, kMonCap, c->st);
            return e;
        }
        return launch_fluidx_soa<T>((T *)c->u, bcur, fr, c->g, (const T *)c->dt, c->st);
    }
    // x frame, and y/z frames of small boxes or generic pencil lengths: LDS
    // tiles; else column streaming
    const bool tiles =
        generic || (c->adv_form == 0 ? !stream_pays(make_stream_box(fr, c->g)) : c->adv_form == 1);
    if (fr == 0 || tiles)
        return launch_advect_soa<T>((const T *)c->u, bcur, bnext, fr, c->g, (const T *)c->dt, c->st);
    return launch_advect_stream<T>((const T *)c->u, bcur, bnext, (T *)c->fa, fr, c->g, (const T *)c->dt, c->st,
                                   c->adv_form == 0 ? 0 : c->adv_form - 1);
}

template <typename T> static hipError_t enqueue_step(mhd_ctx *c)
{
    hipError_t e = enqueue_dt<T>(c);
    for (int s = 0; s < 12 && e == hipSuccess; s++) e = enqueue_sweep<T>(c, s);
    return e;
}

static hipError_t enqueue_step_any(mhd_ctx *c)
{
    return c->prec == 4 ? enqueue_step<float>(c) : enqueue_step<double>(c);
}

static void drop_graph(mhd_ctx *c);

// capture `steps` steps of the loop into one executable graph
static int capture_steps(mhd_ctx *c, int steps, hipGraph_t *gout, hipGraphExec_t *xout)
{
    hipError_t e = hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    hipError_t el = hipSuccess;
    for (int s = 0; s < steps && el == hipSuccess; s++) el = enqueue_step_any(c);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(c->st, &g);
    if (el != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(el, "capture of the step");
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(e, "hipGraphInstantiate");
    }
    *gout = g;
    *xout = x;
    return MHD_OK;
}

static int ensure_graph(mhd_ctx *c)
{
    if (c->captured) return MHD_OK;
    int rc = capture_steps(c, 1, &c->graph, &c->exec);
    if (!rc) rc = capture_steps(c, kGraphSteps, &c->graph_n, &c->exec_n);
    if (rc) {
        drop_graph(c);
        return rc;
    }
    c->captured = true;
    return MHD_OK;
}

// enqueue nsteps steps: replays of the long graph, then single steps
static hipError_t launch_steps(mhd_ctx *c, int nsteps)
{
    hipError_t e = hipSuccess;
    int s = 0;
    for (; s + kGraphSteps <= nsteps && e == hipSuccess; s += kGraphSteps) e = hipGraphLaunch(c->exec_n, c->st);
    for (; s < nsteps && e == hipSuccess; s++) e = hipGraphLaunch(c->exec, c->st);
    return e;
}

static void drop_graph(mhd_ctx *c)
{
    if (c->exec) (void)hipGraphExecDestroy(c->exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->exec_n) (void)hipGraphExecDestroy(c->exec_n);
    if (c->graph_n) (void)hipGraphDestroy(c->graph_n);
    c->exec = c->exec_n = nullptr;
    c->graph = c->graph_n = nullptr;
    c->captured = false;
}

extern "C" {

int mhd_ctx_create(mhd_ctx **out, int nx, int ny, int nz, int precision, int avg_b)
{
    if (!out) return fail(MHD_ERR_ARG, "mhd_ctx_create: null out");
    *out = nullptr;
    if (precision != 4 && precision != 8) return fail(MHD_ERR_ARG, "precision must be 4 or 8");
    int rc = check_shape3(nx, ny, nz, precision);
    if (rc) return rc;
    mhd_ctx *c = new mhd_ctx();
    memset(c, 0, sizeof *c);
    c->nx = nx;
    c->ny = ny;
    c->nz = nz;
    c->prec = precision;
    c->avg = avg_b ? 1 : 0;
    c->ncell = (size_t)nx * ny * nz;
    c->g = make_grid(nx, ny, nz, soa_pad_y(precision), soa_pad_z(precision));
    c->tf = INFINITY;
    c->cap = kDtCap;
    c->mon_nblk = 1024;
    const size_t P = precision;
    hipError_t e = hipStreamCreate(&c->st);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "hipStreamCreate");
    }
    const size_t NP = (size_t)c->g.N * P;
    c->nparts = precision == 4 ? fluidx_tiles_x<float>(c->g) : fluidx_tiles_x<double>(c->g);
    struct A {
        void **p;
        size_t n;
    } al[] = {{&c->u, 5 * NP},   {&c->b, 3 * NP}, {&c->b2, 3 * NP}, {&c->fa, NP},
              {&c->cfl_part, (size_t)c->nparts * P},
              {(void **)&c->cmax, 64},
              {&c->dt, 64},                 {&c->t, 64},               {(void **)&c->t64, 64},
              {&c->dts, (size_t)kDtCap * P}, {(void **)&c->iter, 64},  {(void **)&c->mon_part, 6 * 8 * 1024},
              {(void **)&c->mon_out, 64},
              // the tile records and, after them, the reduction's first stage
              {(void **)&c->mon_tile, 6 * 8 * (size_t)((c->nparts > c->mon_nblk ? c->nparts : c->mon_nblk) * 2)},
              {(void **)&c->mon_hist, 6 * 8 * (size_t)kMonCap}};
    for (auto &a : al) {
        e = hipMalloc(a.p, a.n);
        if (e != hipSuccess) {
            mhd_ctx_destroy(c);
            return fail(MHD_ERR_NOMEM, "mhd_ctx_create: hipMalloc(%zu): %s", a.n, hipGetErrorString(e));
        }
    }
    e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) e = c->prec == 4 ? prepare_soa<float>(nx, ny, nz) : prepare_soa<double>(nx, ny, nz);
    if (e != hipSuccess) {
        mhd_ctx_destroy(c);
        return hip_fail(e, "mhd_ctx_create");
    }
    rc = mhd_ctx_reset_clock(c, INFINITY);
    if (rc) {
        mhd_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return MHD_OK;
}

int mhd_ctx_destroy(mhd_ctx *c)
{
    if (!c) return MHD_OK;
    if (c->st) (void)hipStreamSynchronize(c->st);
    drop_graph(c);
    void *ps[] = {c->u,    c->b,  c->b2,   c->fa,       c->cfl_part, c->cmax,    c->dt,      c->t,
                  c->t64,  c->dts, c->iter, c->mon_part, c->mon_out,  c->mon_tile, c->mon_hist};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return MHD_OK;
}

// host AoS <-> device SoA through a temporary device buffer
static int transfer_state(mhd_ctx *c, void *h_u, void *h_b, int to_device)
{
    const size_t bu = 5 * c->ncell * c->prec, bb = 3 * c->ncell * c->prec;
    void *tmp = nullptr;
    hipError_t e = hipMalloc(&tmp, bu);
    if (e != hipSuccess) return fail(MHD_ERR_NOMEM, "state transfer buffer: %s", hipGetErrorString(e));
    struct Part {
        void *host, *dev;
        size_t bytes;
        int ncomp;
    } parts[2] = {{h_u, c->u, bu, 5}, {h_b, c->b, bb, 3}};
    for (auto &p : parts) {
        if (e != hipSuccess) break;
        if (to_device) {
            e = hipMemcpyAsync(tmp, p.host, p.bytes, hipMemcpyHostToDevice, c->st);
            if (e == hipSuccess)
                e = c->prec == 4 ? launch_aos_soa<float>((float *)p.dev, (float *)tmp, p.ncomp, c->g, 1, c->st)
                                 : launch_aos_soa<double>((double *)p.dev, (double *)tmp, p.ncomp, c->g, 1, c->st);
        } else {
            e = c->prec == 4 ? launch_aos_soa<float>((float *)p.dev, (float *)tmp, p.ncomp, c->g, 0, c->st)
                             : launch_aos_soa<double>((double *)p.dev, (double *)tmp, p.ncomp, c->g, 0, c->st);
            if (e == hipSuccess) e = hipMemcpyAsync(p.host, tmp, p.bytes, hipMemcpyDeviceToHost, c->st);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    }
    (void)hipFree(tmp);
    return e == hipSuccess ? MHD_OK : hip_fail(e, to_device ? "mhd_ctx_set_state" : "mhd_ctx_get_state");
}

int mhd_ctx_set_state(mhd_ctx *c, const void *h_u, const void *h_b)
{
    if (!c || !h_u || !h_b) return fail(MHD_ERR_ARG, "mhd_ctx_set_state: null argument");
    hipError_t e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return hip_fail(e, "mhd_ctx_set_state");
    const int rc = transfer_state(c, const_cast<void *>(h_u), const_cast<void *>(h_b), 1);
    if (rc) return rc;
    e = prime_cfl_any(c);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    return e == hipSuccess ? MHD_OK : hip_fail(e, "mhd_ctx_set_state (CFL)");
}

int mhd_ctx_get_state(mhd_ctx *c, void *h_u, void *h_b)
{
    if (!c || !h_u || !h_b) return fail(MHD_ERR_ARG, "mhd_ctx_get_state: null argument");
    return transfer_state(c, h_u, h_b, 0);
}

int mhd_ctx_set_advect_form(mhd_ctx *c, int form)
{
    if (!c) return fail(MHD_ERR_ARG, "mhd_ctx_set_advect_form: null ctx");
    if (form < 0 || form > 3) return fail(MHD_ERR_ARG, "mhd_ctx_set_advect_form: form must be 0..3");
    if (form == 3 && (!stream_fusable(make_stream_box(1, c->g)) || !stream_fusable(make_stream_box(2, c->g))))
        return fail(MHD_ERR_ARG, "mhd_ctx_set_advect_form: the fused stream needs nx <= 1024");
    const hipError_t e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return hip_fail(e, "mhd_ctx_set_advect_form");
    // the sweeps are baked into the captured graph: capture again on the next step
    drop_graph(c);
    c->adv_form = form;
    return MHD_OK;
}

int mhd_ctx_reset_clock(mhd_ctx *c, double tf)
{
    if (!c) return fail(MHD_ERR_ARG, "mhd_ctx_reset_clock: null ctx");
    hipError_t e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return hip_fail(e, "mhd_ctx_reset_clock");
    // tf is baked into the captured dt kernel: capture again on the next step
    drop_graph(c);
    c->tf = tf;
    e = hipMemsetAsync(c->cmax, 0, 64, c->st);
    if (e == hipSuccess) e = hipMemsetAsync(c->t, 0, 64, c->st);
    if (e == hipSuccess) e = hipMemsetAsync(c->dt, 0, 64, c->st);
    if (e == hipSuccess) e = hipMemsetAsync(c->t64, 0, 64, c->st);
    if (e == hipSuccess) e = hipMemsetAsync(c->iter, 0, 64, c->st);
    if (e == hipSuccess) e = prime_cfl_any(c);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    return e == hipSuccess ? MHD_OK : hip_fail(e, "mhd_ctx_reset_clock");
}

int mhd_ctx_step(mhd_ctx *c, int nsteps)
{
    if (!c || nsteps < 0) return fail(MHD_ERR_ARG, "mhd_ctx_step: bad argument");
    int rc = ensure_graph(c);
    if (rc) return rc;
    const hipError_t e = launch_steps(c, nsteps);
    return e == hipSuccess ? MHD_OK : hip_fail(e, "hipGraphLaunch");
}

int mhd_ctx_step_timed(mhd_ctx *c, int nsteps, float *ms)
{
    if (!c || !ms) return fail(MHD_ERR_ARG, "mhd_ctx_step_timed: null argument");
    int rc = ensure_graph(c);
    if (rc) return rc;
    hipError_t e = hipEventRecord(c->ev0, c->st);
    if (e == hipSuccess) e = launch_steps(c, nsteps);
    if (e == hipSuccess) e = hipEventRecord(c->ev1, c->st);
    if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, c->ev0, c->ev1);
    return e == hipSuccess ? MHD_OK : hip_fail(e, "mhd_ctx_step_timed");
}

int mhd_ctx_sync(mhd_ctx *c)
{
    if (!c) return fail(MHD_ERR_ARG, "mhd_ctx_sync: null ctx");
    hipError_t e = hipStreamSynchronize(c->st);
    return e == hipSuccess ? MHD_OK : hip_fail(e, "mhd_ctx_sync");
}

int mhd_ctx_dts(mhd_ctx *c, double *h_dts, int max, int *count)
{
    if (!c || !count) return fail(MHD_ERR_ARG, "mhd_ctx_dts: null argument");
    int it = 0;
    hipError_t e = hipStreamSynchronize(c->st);
    if (e == hipSuccess) e = hipMemcpy(&it, c->iter, sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "mhd_ctx_dts");
    *count = it;
    int n = it < max ? it : max;
    if (n > c->cap) n = c->cap;
    if (n > 0 && h_dts) {
        std::vector<unsigned char> raw((size_t)n * c->prec);
        e = hipMemcpy(raw.data(), c->dts, raw.size(), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "mhd_ctx_dts copy");
        for (int i = 0; i < n; i++)
            h_dts[i] = c->prec == 4 ? (double)((const float *)raw.data())[i] : ((const double *)raw.data())[i];
    }
    return MHD_OK;
}

int mhd_ctx_time(mhd_ctx *c, double *t, double *t64)
{
    if (!c) return fail(MHD_ERR_ARG, "mhd_ctx_time: null ctx");
    unsigned char raw[8] = {0};
    double d64 = 0;
    hipError_t e = hipStreamSynchronize(c->st);
    if (e == hipSuccess) e = hipMemcpy(raw, c->t, c->prec, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&d64, c->t64, 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "mhd_ctx_time");
    if (t) {
        if (c->prec == 4) {
            float f;
            memcpy(&f, raw, 4);
            *t = f;
        } else {
            memcpy(t, raw, 8);
        }
    }
    if (t64) *t64 = d64;
    return MHD_OK;
}

int mhd_ctx_monitors(mhd_ctx *c, double *out6)
{
    if (!c || !out6) return fail(MHD_ERR_ARG, "mhd_ctx_monitors: null argument");
    hipError_t e = c->prec == 4
                       ? launch_monitors_soa<float>((const float *)c->u, (const float *)c->b, c->g, c->mon_part,
                                                    c->mon_nblk, c->mon_out, c->st)
                       : launch_monitors_soa<double>((const double *)c->u, (const double *)c->b, c->g, c->mon_part,
                                                     c->mon_nblk, c->mon_out, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(out6, c->mon_out, 6 * sizeof(double), hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    return e == hipSuccess ? MHD_OK : hip_fail(e, "mhd_ctx_monitors");
}

int mhd_ctx_monitor_history(mhd_ctx *c, double *h_out, int max, int *count)
{
    if (!c || !count) return fail(MHD_ERR_ARG, "mhd_ctx_monitor_history: null argument");
    int it = 0;
    hipError_t e = hipStreamSynchronize(c->st);
    if (e == hipSuccess) e = hipMemcpy(&it, c->iter, sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "mhd_ctx_monitor_history");
    *count = it;
    int n = it < max ? it : max;
    if (n > kMonCap) n = kMonCap;
    if (n > 0 && h_out) {
        e = hipMemcpy(h_out, c->mon_hist, (size_t)n * 6 * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "mhd_ctx_monitor_history copy");
    }
    return MHD_OK;
}

int mhd_ctx_device_state(mhd_ctx *c, void **d_u, void **d_b)
{
    if (!c) return fail(MHD_ERR_ARG, "mhd_ctx_device_state: null ctx");
    if (d_u) *d_u = c->u;
    if (d_b) *d_b = c->b;
    return MHD_OK;
}

int mhd_ctx_layout(mhd_ctx *c, long *sy, long *sz, long *plane)
{
    if (!c) return fail(MHD_ERR_ARG, "mhd_ctx_layout: null ctx");
    if (sy) *sy = c->g.sy;
    if (sz) *sz = c->g.sz;
    if (plane) *plane = c->g.N;
    return MHD_OK;
}

int mhd_ctx_profile_step(mhd_ctx *c, int reps, float *ms13)
{
    if (!c || !ms13 || reps < 1) return fail(MHD_ERR_ARG, "mhd_ctx_profile_step: bad argument");
    hipEvent_t ev[14];
    for (int i = 0; i < 14; i++) {
        hipError_t e = hipEventCreate(&ev[i]);
        if (e != hipSuccess) {
            for (int j = 0; j < i; j++) (void)hipEventDestroy(ev[j]);
            return hip_fail(e, "hipEventCreate");
        }
    }
    double acc[13] = {0};
    hipError_t e = hipSuccess;
    for (int r = 0; r < reps && e == hipSuccess; r++) {
        e = hipEventRecord(ev[0], c->st);
        if (e == hipSuccess) e = c->prec == 4 ? enqueue_dt<float>(c) : enqueue_dt<double>(c);
        if (e == hipSuccess) e = hipEventRecord(ev[1], c->st);
        for (int s = 0; s < 12 && e == hipSuccess; s++) {
            e = c->prec == 4 ? enqueue_sweep<float>(c, s) : enqueue_sweep<double>(c, s);
            if (e == hipSuccess) e = hipEventRecord(ev[s + 2], c->st);
        }
        if (e == hipSuccess) e = hipEventSynchronize(ev[13]);
        for (int s = 0; s < 13 && e == hipSuccess; s++) {
            float ms = 0;
            e = hipEventElapsedTime(&ms, ev[s], ev[s + 1]);
            acc[s] += ms;
        }
    }
    for (int i = 0; i < 14; i++) (void)hipEventDestroy(ev[i]);
    if (e != hipSuccess) return hip_fail(e, "mhd_ctx_profile_step");
    for (int s = 0; s < 13; s++) ms13[s] = (float)(acc[s] / reps);
    return MHD_OK;
}

}  // extern "C"
