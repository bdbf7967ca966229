// Batched engine: many independent boxes of uneven sizes advanced together,
// each with its own dt.  One step is a CFL launch, a dt launch and 12 sweep
// launches over (box, tile) work items, captured in a hipGraph and replayed.
// Schedules: the default groups the boxes by cells-per-lane K so that each
// sweep kernel is compiled for one K (and holds only its registers), and
// runs the groups' chains as concurrent branches of the graph; the others
// run one kernel per sweep over all boxes, either one block per item or
// persistent blocks pulling items from per-XCD queues (own queue first,
// then stealing).  Boxes live in slots; the active set can change between
// steps (the server admits and retires requests that way).  FP32 only, the
// reference's precision.
#pragma once
#include <mutex>
#include <vector>

#include "mhd_soa.h"

namespace mhd {

struct BoxDesc {
    float *u, *b, *b2;   // SoA state and the second field buffer
    float *fa;           // scratch plane of the streamed y/z advectbyzx
    SoaGrid g;
    SoaBox fr[3];        // the three sweep frames
    int K[3];            // cells per lane of each frame
    int KF[3], PF[3];    // the fluidx sweeps' cells per lane and pencils per wave (short pencils share a wave)
    int ntj[3];          // tiles along j of each frame
    StreamBox sb[3];     // streamed advectbyzx of frames 1 and 2
    int nwa[3], nwb[3];  // its waves (pass A, pass B)
    int streamed;        // the streamed y/z advectbyzx pays for this box alone
};

struct WorkItem {
    int box, tile;
};
struct CflItem {
    int box, e0, ne;  // units e0 .. e0+ne-1 of the box (see cfl_range)
};

// boxes with the same cells-per-lane K in every frame, and the same K and
// pencils per wave of their fluidx sweeps (schedule 3): their
// items, CFL items and active slots are contiguous ranges of the engine's
// lists, and they run exact-K kernels
struct BoxGroup {
    int K[3];
    int KF[3], PF[3];
    int item_off[3], item_n[3];
    int cfl_off, cfl_n;
    int act_off, act_n;
    int lds_f[3], lds_a[3];  // dynamic LDS bytes of fluidx / advect per frame
    int sa_off[3], sa_n[3], sb_off[3], sb_n[3];  // block items of the streamed y/z advectbyzx
    int sab_nw[3];  // waves per block of the fused stream (its items in the sa list); 0: the two passes
    int f8_off[3], f8_n[3], lds_f8[3];  // y/z fluidx in tiles of two vectors (8 pencils); f8_n 0: one-vector tiles
    int x2_off, x2_n, lds_x2;           // x-frame advectbyzx in tiles of two planes (in the f8 item list)
    int tile_adv_n[3];  // leading items of item_off[fr] whose boxes run the tile advectbyzx
    long cells;
};

class Engine {
public:
    // max_boxes slots, dt history of cap steps per slot
    Engine(int max_boxes, int cap, int avg_b);
    ~Engine();
    int init();  // allocations; MHD_OK or an error code
    hipStream_t stream() const { return st_; }

    // slots: device state of one box (48 B per cell)
    int add_box(int nx, int ny, int nz, int *slot);
    int remove_box(int slot);
    int set_state(int slot, const float *h_u, const float *h_b);
    int get_state(int slot, float *h_u, float *h_b);
    // the serving path: a slot filled (stage_in) or read back and released
    // (stage_out) on the engine's transfer stream, without waiting for the
    // compute stream; the slot must not be in the active set meanwhile.
    // Safe to call from another thread than the one that steps.
    int stage_in(int nx, int ny, int nz, const float *h_u, const float *h_b, int *slot);
    int stage_out(int slot, float *h_u, float *h_b);
    int reset_clock(int slot, double tf);
    int dts(int slot, double *h, int max, int *count);
    int time(int slot, double *t, double *t64);
    long cells(int slot) const;

    // the set of slots a step advances (rebuilds work lists; re-captures)
    int set_active(const std::vector<int> &slots);
    const std::vector<int> &active() const { return active_; }
    // rebuild the work lists and re-capture the step if the active set changed
    int prepare();
    int step(int nsteps);
    int step_timed(int nsteps, float *ms);
    int sync();

    // schedule: persistent blocks per CU (0 = occupancy limit) and policy
    // (0: persistent, one queue per XCD + stealing; 1: persistent, one
    // global queue; 2: one block per work item; 3: boxes grouped by K, each
    // group a chain of exact-K kernels, one block per item, the groups'
    // chains concurrent branches of the step graph)
    int set_schedule(int blocks_per_cu, int policy);
    int autotune(int trial_steps, int *bpc, int *policy, float *best_ms);

private:
    struct Chunk {
        float *base;  // 12 planes: u (5), b (3), b2 (3), fa (1)
        long N;
    };
    int alloc_box(long N, float **base);
    void release_box(float *base, long N);
    int claim_slot(int nx, int ny, int nz, hipStream_t st, int *slot);
    int reset_clock_on(int slot, double tf, hipStream_t st);
    int transfer(int slot, void *h_u, void *h_b, int to_device, hipStream_t st);
    int build();
    int capture();
    void drop_graph();
    int enqueue_step();
    int save_all(std::vector<float *> &saved);
    int restore_all(std::vector<float *> &saved);

    int max_boxes_, cap_, avg_;
    hipStream_t st_ = nullptr;   // compute (steps)
    hipStream_t xs_ = nullptr;   // transfers of the serving path
    std::mutex mu_;              // slots and the memory pool
    std::vector<Chunk> pool_;    // released box memory, reused by size
    float *xbuf_[2] = {nullptr, nullptr};  // AoS transfer buffers of st_ and xs_
    size_t xcap_[2] = {0, 0};
    std::vector<BoxDesc> boxes_;   // host copy, per slot
    std::vector<char> used_;  // bytes, not bits: two threads touch neighbouring slots
    std::vector<int> active_;
    std::vector<double> tf_;
    // device arrays
    BoxDesc *d_boxes_ = nullptr;
    float *d_dt_ = nullptr, *d_t_ = nullptr, *d_tf_ = nullptr, *d_dts_ = nullptr;
    double *d_t64_ = nullptr;
    unsigned long long *d_cmax_ = nullptr;
    int *d_iter_ = nullptr, *d_active_ = nullptr;
    unsigned *d_ctr_ = nullptr;        // 12 sweeps x 8 queues
    int *d_part_ = nullptr;            // 3 frames x 9 partition bounds
    WorkItem *d_items_[3] = {nullptr, nullptr, nullptr};
    WorkItem *d_sitems_[2] = {nullptr, nullptr};  // streamed advectbyzx blocks: pass A, pass B (both frames)
    int cap_sitems_[2] = {0, 0};
    WorkItem *d_f8items_ = nullptr;  // y/z fluidx tiles of 8 pencils (both frames)
    int cap_f8items_ = 0;
    CflItem *d_cfl_ = nullptr;
    int nitems_[3] = {0, 0, 0}, ncfl_ = 0, cap_items_[3] = {0, 0, 0}, cap_cfl_ = 0;
    std::vector<BoxGroup> groups_;
    std::vector<char> stream_now_;  // per slot: the y/z advectbyzx streams in the current build
    // per slot, schedule 3: the CFL accumulator holds the CFL of the slot's
    // current state (the last sweep of every step refreshes it; a set state,
    // a clock reset or a schedule change make it stale, and step() then runs
    // the standalone CFL over the slot's items cfl_slot_off_ .. + cfl_slot_n_)
    std::vector<char> cfl_fresh_;
    std::vector<int> cfl_slot_off_, cfl_slot_n_;
    int prime_cfl();
    static constexpr int kAux = 4;  // side streams of the grouped schedule
    hipStream_t aux_[kAux] = {};
    hipEvent_t evfork_ = nullptr, evjoin_[kAux] = {};
    int kmax_ = 1, lds_fluidx_ = 0, lds_advect_ = 0;
    int bpc_ = 0, policy_ = 3, grid_ = 0, ncu_ = 0;  // default: K-grouped concurrent chains (fastest measured)
    bool dirty_ = true;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
};

}  // namespace mhd
