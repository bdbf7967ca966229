// Serving loop: simulation requests arrive at any time (mhd_server_submit),
// are admitted in arrival order while their device state fits an HBM budget,
// stay resident on the device, and advance together through the batched
// engine (continuous batching: a request that arrives while others run joins
// at the next step boundary, a finished one leaves).  Two threads own the GPU
// work: the service thread steps the active set on the engine's compute
// stream; the staging thread uploads admitted requests and reads finished
// ones back on the engine's transfer stream, so those copies overlap the
// steps of the others.  submit/poll are thread-safe.  Mode 1 is the FIFO
// baseline: one request at a time through the plain device-resident path.
#include <math.h>
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/mhd_api.h"
#include "mhd_batch.h"
#include "mhd_internal.h"

using namespace mhd;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    long long id;
    int nx, ny, nz, nsteps;
    const float *in_u, *in_b;
    float *out_u, *out_b;
    Clock::time_point arrived, admitted, done;
    int slot = -1, remaining = 0, rc = MHD_OK;
    size_t bytes = 0;
    long long cells() const { return (long long)nx * ny * nz; }
};

size_t device_bytes(const Request &r)
{
    // state 32 B + second field buffer 12 B + EMF scratch 4 B per cell
    // (padded planes), plus the 20 B/cell transfer buffer held while the
    // state is uploaded or read back
    const size_t c = (size_t)r.nx * r.ny * r.nz;
    return 48 * c + 20 * c + (1 << 20);
}

}  // namespace

struct mhd_server {
    int mode;  // 0 continuous batching, 1 FIFO one at a time (plain path)
    int device = 0;
    size_t budget;
    int max_active;
    long long max_cells;  // admitted work cap: enough cells to saturate the GPU
    long long active_cells = 0;
    Engine *eng = nullptr;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request *> queue;     // submitted, not admitted
    std::deque<Request *> ready;     // uploaded, joining at the next chunk
    std::deque<Request *> outgoing;  // stepped to the end, to be read back
    std::vector<Request *> active, finished;
    int staging = 0;                 // being uploaded
    int in_flight = 0;               // admitted and not finished
    bool service_done = false;
    std::thread th_stage;
    long long next_id = 1;
    bool stop = false;
    size_t resident = 0;
    // statistics
    double zone_updates = 0, busy_s = 0;
    double t_upload = 0, t_rebuild = 0, t_step = 0, t_download = 0;  // service-thread seconds
    long long steps_run = 0, rebuilds = 0;
    Clock::time_point first_arrival, last_done;
    bool any = false;
    int fatal = MHD_OK;
    char fatal_msg[256] = {0};

    bool admissible() const
    {
        if (queue.empty() || fatal) return false;
        const Request *r = queue.front();
        return in_flight < max_active && resident + r->bytes <= budget &&
               (active_cells == 0 || active_cells + r->cells() <= max_cells);
    }
    void fail_all(std::deque<Request *> &q, int rc);
    void loop_batched();
    void loop_stage();
    void loop_fifo();
};

static double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// (under mu) finish every request of q with rc; admitted ones (holding a
// slot) give back their slot and their share of the budget
void mhd_server::fail_all(std::deque<Request *> &q, int rc)
{
    for (Request *r : q) {
        if (r->slot >= 0) {
            eng->stage_out(r->slot, nullptr, nullptr);
            r->slot = -1;
            resident -= r->bytes;
            active_cells -= r->cells();
            in_flight--;
        }
        r->rc = rc;
        r->done = Clock::now();
        finished.push_back(r);
    }
    q.clear();
}

// staging thread: read finished requests back first (their latency is
// running), else upload the next admissible request
void mhd_server::loop_stage()
{
    (void)hipSetDevice(device);
    for (;;) {
        Request *out = nullptr, *in = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] {
                return !outgoing.empty() || admissible() || fatal ||
                       (service_done && queue.empty() && ready.empty() && outgoing.empty());
            });
            if (!outgoing.empty()) {
                out = outgoing.front();
                outgoing.pop_front();
            } else if (fatal) {
                // the service thread stopped on an error: nothing more will step
                fail_all(ready, fatal);
                fail_all(queue, fatal);
                if (service_done) return;
                cv.wait(lk, [&] { return service_done || !outgoing.empty(); });
                continue;
            } else if (admissible()) {
                in = queue.front();
                queue.pop_front();
                resident += in->bytes;
                active_cells += in->cells();
                in_flight++;
                staging++;
            } else {
                return;  // service done and nothing left
            }
        }
        const auto t0 = Clock::now();
        if (out) {
            const int rc = eng->stage_out(out->slot, out->out_u, out->out_b);
            std::lock_guard<std::mutex> lk(mu);
            out->rc = rc;
            out->slot = -1;
            out->done = Clock::now();
            resident -= out->bytes;
            active_cells -= out->cells();
            in_flight--;
            last_done = out->done;
            finished.push_back(out);
            t_download += secs(t0, out->done);
        } else {
            in->admitted = t0;
            int slot = -1;
            const int rc = eng->stage_in(in->nx, in->ny, in->nz, in->in_u, in->in_b, &slot);
            std::lock_guard<std::mutex> lk(mu);
            staging--;
            t_upload += secs(t0, Clock::now());
            if (rc) {
                in->rc = rc;
                in->done = Clock::now();
                resident -= in->bytes;
                active_cells -= in->cells();
                in_flight--;
                finished.push_back(in);
            } else {
                in->slot = slot;
                in->remaining = in->nsteps;
                ready.push_back(in);
            }
        }
        cv.notify_all();
    }
}

// service thread: join the uploaded requests, step the active set in chunks,
// hand the finished ones to the staging thread
void mhd_server::loop_batched()
{
    (void)hipSetDevice(device);
    double est_step_ms = 1.0;  // running estimate of one batched step
    for (;;) {
        bool changed = false;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] {
                return !ready.empty() || !active.empty() ||
                       (stop && queue.empty() && staging == 0 && ready.empty());
            });
            if (active.empty() && ready.empty()) {
                service_done = true;
                cv.notify_all();
                return;
            }
            while (!ready.empty()) {
                active.push_back(ready.front());
                ready.pop_front();
                changed = true;
            }
        }
        if (changed) {
            std::vector<int> slots;
            for (Request *r : active) slots.push_back(r->slot);
            eng->set_active(slots);
            rebuilds++;
        }
        const auto tb = Clock::now();
        int rc = eng->prepare();
        t_rebuild += secs(tb, Clock::now());
        // steps until the next completion, capped so that arrivals join
        // within ~10 ms of GPU time
        int k = 1 << 30;
        for (Request *r : active) k = std::min(k, r->remaining);
        const int cap = std::max(1, (int)(10.0 / std::max(est_step_ms, 1e-3)));
        k = std::max(1, std::min(k, cap));
        const auto t0 = Clock::now();
        if (!rc) rc = eng->step(k);
        if (!rc) rc = eng->sync();
        const double dt_s = secs(t0, Clock::now());
        if (rc) {
            std::lock_guard<std::mutex> lk(mu);
            fatal = rc;
            snprintf(fatal_msg, sizeof fatal_msg, "%s", mhd_last_error());
            for (Request *r : active) {
                r->rc = rc;
                r->done = Clock::now();
                resident -= r->bytes;
                active_cells -= r->cells();
                in_flight--;
                finished.push_back(r);
            }
            active.clear();
            stop = true;
            service_done = true;
            cv.notify_all();
            return;
        }
        est_step_ms = 0.5 * est_step_ms + 0.5 * (1e3 * dt_s / k);
        std::vector<Request *> still, done;
        for (Request *r : active) {
            r->remaining -= k;
            (r->remaining > 0 ? still : done).push_back(r);
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            busy_s += dt_s;
            t_step += dt_s;
            steps_run += k;
            for (Request *r : active) zone_updates += (double)r->nx * r->ny * r->nz * k;
            if (!done.empty()) {
                for (Request *r : done) outgoing.push_back(r);
                active.swap(still);
            }
        }
        if (!done.empty()) {
            cv.notify_all();
            std::vector<int> slots;
            for (Request *r : active) slots.push_back(r->slot);
            eng->set_active(slots);
            rebuilds++;
        }
    }
}

void mhd_server::loop_fifo()
{
    (void)hipSetDevice(device);
    for (;;) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return stop || !queue.empty(); });
            if (queue.empty()) return;
            r = queue.front();
            queue.pop_front();
        }
        r->admitted = Clock::now();
        const auto t0 = Clock::now();
        mhd_ctx *c = nullptr;
        int rc = mhd_ctx_create(&c, r->nx, r->ny, r->nz, 4, 0);
        if (!rc) rc = mhd_ctx_set_state(c, r->in_u, r->in_b);
        if (!rc) rc = mhd_ctx_step(c, r->nsteps);
        if (!rc) rc = mhd_ctx_get_state(c, r->out_u, r->out_b);
        if (c) mhd_ctx_destroy(c);
        r->rc = rc;
        r->done = Clock::now();
        std::lock_guard<std::mutex> lk(mu);
        busy_s += secs(t0, r->done);
        if (!rc) zone_updates += (double)r->nx * r->ny * r->nz * r->nsteps;
        steps_run += r->nsteps;
        last_done = r->done;
        finished.push_back(r);
    }
}

extern "C" {

int mhd_server_create(mhd_server **out, int mode, size_t hbm_budget, int max_active, long long max_active_cells)
{
    if (!out || (mode != 0 && mode != 1)) return fail(MHD_ERR_ARG, "mhd_server_create: bad argument");
    *out = nullptr;
    size_t fr = 0, tot = 0;
    hipError_t e = hipMemGetInfo(&fr, &tot);
    if (e != hipSuccess) return hip_fail(e, "hipMemGetInfo");
    mhd_server *sv = new mhd_server();
    sv->mode = mode;
    e = hipGetDevice(&sv->device);
    if (e != hipSuccess) {
        delete sv;
        return hip_fail(e, "hipGetDevice");
    }
    // default budget: 90% of the free device memory
    sv->budget = hbm_budget ? hbm_budget : (size_t)(0.9 * (double)fr);
    sv->max_active = max_active > 0 ? max_active : 256;
    // ~16 M cells keep all 256 CUs busy through every sweep (a 256^3 box)
    sv->max_cells = max_active_cells > 0 ? max_active_cells : (16LL << 20);
    if (mode == 0) {
        sv->eng = new Engine(sv->max_active, 1 << 12, 0);
        int rc = sv->eng->init();
        if (rc) {
            delete sv->eng;
            delete sv;
            return rc;
        }
        sv->th = std::thread([sv] { sv->loop_batched(); });
        sv->th_stage = std::thread([sv] { sv->loop_stage(); });
    } else {
        sv->th = std::thread([sv] { sv->loop_fifo(); });
    }
    *out = sv;
    return MHD_OK;
}

int mhd_server_submit(mhd_server *sv, int nx, int ny, int nz, int nsteps, const float *h_u, const float *h_b,
                      float *out_u, float *out_b, long long *req_id)
{
    if (!sv || !h_u || !h_b || !out_u || !out_b || nsteps < 1) return fail(MHD_ERR_ARG, "mhd_server_submit: bad argument");
    int rc = check_shape_fast3(nx, ny, nz);
    if (rc) return rc;
    Request *r = new Request();
    r->nx = nx;
    r->ny = ny;
    r->nz = nz;
    r->nsteps = nsteps;
    r->in_u = h_u;
    r->in_b = h_b;
    r->out_u = out_u;
    r->out_b = out_b;
    r->bytes = device_bytes(*r);
    r->arrived = Clock::now();
    long long id = -1;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->stop) {
            delete r;
            return fail(MHD_ERR_STATE, "mhd_server_submit: server stopped%s%s", sv->fatal ? ": " : "", sv->fatal_msg);
        }
        if (sv->mode == 0 && r->bytes > sv->budget) {
            const size_t need = r->bytes;
            delete r;
            return fail(MHD_ERR_FULL, "mhd_server_submit: request needs %zu bytes, budget %zu", need, sv->budget);
        }
        r->id = id = sv->next_id++;
        if (!sv->any) {
            sv->first_arrival = r->arrived;
            sv->any = true;
        }
        sv->queue.push_back(r);
    }
    // r may be finished and freed by another thread's poll from here on
    sv->cv.notify_all();
    if (req_id) *req_id = id;
    return MHD_OK;
}

int mhd_server_poll(mhd_server *sv, long long *ids, double *latency_s, double *queue_s, int *rcs, int max, int *n)
{
    if (!sv || !n) return fail(MHD_ERR_ARG, "mhd_server_poll: bad argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    int k = 0;
    while (k < max && !sv->finished.empty()) {
        Request *r = sv->finished.front();
        sv->finished.erase(sv->finished.begin());
        if (ids) ids[k] = r->id;
        if (latency_s) latency_s[k] = secs(r->arrived, r->done);
        if (queue_s) queue_s[k] = secs(r->arrived, r->admitted);
        if (rcs) rcs[k] = r->rc;
        delete r;
        k++;
    }
    *n = k;
    return MHD_OK;
}

int mhd_server_pending(mhd_server *sv, int *queued, int *active)
{
    if (!sv) return fail(MHD_ERR_ARG, "mhd_server_pending: null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    if (queued) *queued = (int)sv->queue.size() + sv->staging + (int)sv->ready.size();
    if (active) *active = (int)sv->active.size() + (int)sv->outgoing.size();
    return sv->fatal ? fail(sv->fatal, "server stopped: %s", sv->fatal_msg) : MHD_OK;
}

// out[0] zone-updates, [1] seconds from first arrival to last completion,
// [2] GPU busy seconds, [3] batched steps run, [4] work-list rebuilds,
// [5] bytes resident now, [6] budget
int mhd_server_stats(mhd_server *sv, double *out7)
{
    if (!sv || !out7) return fail(MHD_ERR_ARG, "mhd_server_stats: bad argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    out7[0] = sv->zone_updates;
    out7[1] = sv->any ? secs(sv->first_arrival, sv->last_done) : 0;
    out7[2] = sv->busy_s;
    out7[3] = (double)sv->steps_run;
    out7[4] = (double)sv->rebuilds;
    out7[5] = (double)sv->resident;
    out7[6] = (double)sv->budget;
    return MHD_OK;
}

// out[4]: service-thread seconds spent uploading admitted requests, rebuilding
// the work lists and step graph, stepping (launch and wait), and reading
// finished requests back (mode 0; zeros in mode 1)
int mhd_server_timing(mhd_server *sv, double *out4)
{
    if (!sv || !out4) return fail(MHD_ERR_ARG, "mhd_server_timing: bad argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    out4[0] = sv->t_upload;
    out4[1] = sv->t_rebuild;
    out4[2] = sv->t_step;
    out4[3] = sv->t_download;
    return MHD_OK;
}

int mhd_server_destroy(mhd_server *sv)
{
    if (!sv) return MHD_OK;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv.notify_all();
    if (sv->th.joinable()) sv->th.join();
    if (sv->th_stage.joinable()) sv->th_stage.join();
    for (Request *r : sv->finished) delete r;
    for (Request *r : sv->queue) delete r;
    for (Request *r : sv->ready) delete r;
    for (Request *r : sv->outgoing) delete r;
    for (Request *r : sv->active) delete r;
    delete sv->eng;
    delete sv;
    return MHD_OK;
}

}  // extern "C"
