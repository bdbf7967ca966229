// Batched form: the engine of mhd_batch.h and the mhd_batch_* entry points.
#include <math.h>
#include <string.h>

#include <algorithm>

#include "../../include/mhd_api.h"
#include "mhd_batch.h"
#include "mhd_internal.h"
#include "mhd_soa_stream.h"
#include "mhd_soa_tiles.h"

namespace mhd {

constexpr int kBW = 4;       // pencils per tile (one float4 row)
constexpr int kQueues = 8;   // one work queue per XCD
static const int kSweepFrameB[12] = {0, 0, 1, 1, 2, 2, 2, 2, 1, 1, 0, 0};
static const int kSweepOpB[12] = {0, 1, 0, 1, 0, 1, 1, 0, 1, 0, 1, 0};  // 0 fluidx, 1 advectbyzx

// the field is in b before sweep 0; each advectbyzx moves it to the other buffer
static int flip_before(int s)
{
    int n = 0;
    for (int i = 0; i < s; i++) n += kSweepOpB[i];
    return n & 1;
}

// ---------------------------------------------------------------------------
// persistent sweep kernel: blocks pull (box, tile) items from the queue of
// their XCD (blocks are dealt round-robin over the 8 XCDs), then steal from
// the other queues; the item ranges were cut so that each XCD's range is
// about equal in work and contiguous in memory.
// ---------------------------------------------------------------------------
// One work item, as a real call: inlined into the persistent loop, the tile
// body with its descriptor loaded per iteration hits a register-class error
// in the gfx950 backend (ROCm 7.2); a call costs a few hundred cycles of a
// tile's tens of thousands.
template <int FR, int KMAX, bool ADV>
__device__ __noinline__ void batch_item(const BoxDesc *__restrict__ boxes, WorkItem wi, const float *__restrict__ dt,
                                        int flip, float *smem)
{
    const BoxDesc &bd = boxes[wi.box];
    const int tile = wi.tile;
    const float *dtp = dt + wi.box;
    float *const u = bd.u;
    const float *bcur = flip ? bd.b2 : bd.b;
    float *bnext = flip ? bd.b : bd.b2;
    const SoaBox bx = bd.fr[FR];  // by value: the tile reads it many times
    const int ntj = bd.ntj[FR];
#define MHD_BITEM(KK)                                                                                                \
    if constexpr (ADV)                                                                                               \
        advect_soa_tile<float, KK, kBW, FR>(u, bcur, bnext, bx, dtp, tile, ntj, smem);                              \
    else                                                                                                             \
        fluidx_soa_tile<float, KK, kBW, FR>(u, bcur, bx, dtp, tile, ntj, smem);
    switch (bd.K[FR]) {
    case 1: { MHD_BITEM(1) } break;
    case 2: if constexpr (KMAX >= 2) { MHD_BITEM(2) } break;
    case 3: if constexpr (KMAX >= 4) { MHD_BITEM(3) } break;
    case 4: if constexpr (KMAX >= 4) { MHD_BITEM(4) } break;
    case 5: if constexpr (KMAX >= 8) { MHD_BITEM(5) } break;
    case 6: if constexpr (KMAX >= 8) { MHD_BITEM(6) } break;
    case 8: if constexpr (KMAX >= 8) { MHD_BITEM(8) } break;
    default: break;
    }
#undef MHD_BITEM
}

template <int FR, int KMAX, bool ADV>
__global__ void __launch_bounds__(64 * kBW)
    batch_sweep_kernel(const BoxDesc *__restrict__ boxes, const WorkItem *__restrict__ items,
                       const int *__restrict__ part, unsigned *__restrict__ ctr, const float *__restrict__ dt, int flip,
                       int nq)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *smem = reinterpret_cast<float *>(smem_raw);
    __shared__ int s_it;
    const int q0 = nq > 1 ? (int)(blockIdx.x % nq) : 0;
    int tried = 0;
    for (;;) {
        if (threadIdx.x == 0) {
            int it = -1;
            while (tried < nq) {
                const int q = (q0 + tried) % nq;
                const int cand = part[q] + (int)atomicAdd(&ctr[q], 1u);
                if (cand < part[q + 1]) {
                    it = cand;
                    break;
                }
                tried++;
            }
            s_it = it;
        }
        __syncthreads();
        const int it = s_it;
        __syncthreads();
        if (it < 0) break;
        batch_item<FR, KMAX, ADV>(boxes, items[it], dt, flip, smem);
        __syncthreads();  // the LDS tile is reused by the next item
    }
}

// The same sweep with one block per work item (no loop, tile body inlined):
// the hardware dispatcher balances; blocks map to items so that each XCD
// receives a contiguous range of the largest-first item list.
template <int FR, int KMAX, bool ADV>
__global__ void __launch_bounds__(64 * kBW)
    batch_grid_kernel(const BoxDesc *__restrict__ boxes, const WorkItem *__restrict__ items,
                      const float *__restrict__ dt, int flip)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *smem = reinterpret_cast<float *>(smem_raw);
    const WorkItem wi = items[xcd_tile(blockIdx.x, gridDim.x)];
    const BoxDesc &bd = boxes[wi.box];
    float *const u = bd.u;
    const float *bcur = flip ? bd.b2 : bd.b;
    float *bnext = flip ? bd.b : bd.b2;
    const float *dtp = dt + wi.box;
#define MHD_GITEM(KK)                                                                                                \
    if constexpr (ADV)                                                                                               \
        advect_soa_tile<float, KK, kBW, FR>(u, bcur, bnext, bd.fr[FR], dtp, wi.tile, bd.ntj[FR], smem);             \
    else                                                                                                             \
        fluidx_soa_tile<float, KK, kBW, FR>(u, bcur, bd.fr[FR], dtp, wi.tile, bd.ntj[FR], smem);
    switch (bd.K[FR]) {
    case 1: { MHD_GITEM(1) } break;
    case 2: if constexpr (KMAX >= 2) { MHD_GITEM(2) } break;
    case 3: if constexpr (KMAX >= 4) { MHD_GITEM(3) } break;
    case 4: if constexpr (KMAX >= 4) { MHD_GITEM(4) } break;
    case 5: if constexpr (KMAX >= 8) { MHD_GITEM(5) } break;
    case 6: if constexpr (KMAX >= 8) { MHD_GITEM(6) } break;
    case 8: if constexpr (KMAX >= 8) { MHD_GITEM(8) } break;
    default: break;
    }
#undef MHD_GITEM
}

// One K for every item (schedule 3): the kernel holds only the registers of
// that K, so small-K boxes keep their occupancy beside large-K ones.
template <int FR, int K, bool ADV>
__global__ void __launch_bounds__(64 * kBW)
    batch_group_kernel(const BoxDesc *__restrict__ boxes, const WorkItem *__restrict__ items,
                       const float *__restrict__ dt, int flip)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *smem = reinterpret_cast<float *>(smem_raw);
    const WorkItem wi = items[xcd_tile(blockIdx.x, gridDim.x)];
    const BoxDesc &bd = boxes[wi.box];
    float *const u = bd.u;
    const float *bcur = flip ? bd.b2 : bd.b;
    float *bnext = flip ? bd.b : bd.b2;
    if constexpr (ADV && FR == 0) {
        // x frame: direct loads, LDS for the fluxes only
        if (K % 4 == 0 && bd.fr[0].vec_ok)
            advect_x_tile<float, K, kBW, K % 4 == 0>(u, bcur, bnext, bd.fr[0], dt + wi.box, wi.tile, bd.ntj[0], smem);
        else
            advect_x_tile<float, K, kBW, false>(u, bcur, bnext, bd.fr[0], dt + wi.box, wi.tile, bd.ntj[0], smem);
    } else if constexpr (ADV) {
        advect_soa_tile<float, K, kBW, FR>(u, bcur, bnext, bd.fr[FR], dt + wi.box, wi.tile, bd.ntj[FR], smem);
    } else if constexpr (FR == 0) {
        // x frame: direct loads, no LDS
        if (K % 4 == 0 && bd.fr[0].vec_ok)
            fluidx_x_tile<float, K, kBW, K % 4 == 0, 0>(u, bcur, bd.fr[0], dt + wi.box, wi.tile, bd.ntj[0], smem,
                                                         nullptr);
        else
            fluidx_x_tile<float, K, kBW, false, 0>(u, bcur, bd.fr[0], dt + wi.box, wi.tile, bd.ntj[0], smem, nullptr);
    } else {
        fluidx_soa_tile<float, K, kBW, FR>(u, bcur, bd.fr[FR], dt + wi.box, wi.tile, bd.ntj[FR], smem);
    }
}

// Streamed advectbyzx of the y/z frames (schedule 3): one block per item
// (box, first wave), four waves; the field of the sweep is in b (flip 0) or b2.
template <int FR>
__global__ void __launch_bounds__(256) batch_stream_a_kernel(const BoxDesc *__restrict__ boxes,
                                                             const WorkItem *__restrict__ items,
                                                             const float *__restrict__ dt, int flip)
{
    const WorkItem wi = items[xcd_tile(blockIdx.x, gridDim.x)];
    const BoxDesc &bd = boxes[wi.box];
    const int wave = wi.tile + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (wave < bd.nwa[FR])
        advect_stream_a<float, FR>(bd.u, flip ? bd.b2 : bd.b, flip ? bd.b : bd.b2, bd.fa, bd.sb[FR], dt + wi.box,
                                   wave);
}

template <int FR>
__global__ void __launch_bounds__(256) batch_stream_b_kernel(const BoxDesc *__restrict__ boxes,
                                                             const WorkItem *__restrict__ items,
                                                             const float *__restrict__ dt, int flip)
{
    const WorkItem wi = items[xcd_tile(blockIdx.x, gridDim.x)];
    const BoxDesc &bd = boxes[wi.box];
    const int wave = wi.tile + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (wave < bd.nwb[FR])
        advect_stream_b<float, FR>(bd.u, flip ? bd.b2 : bd.b, flip ? bd.b : bd.b2, bd.fa, bd.sb[FR], dt + wi.box,
                                   wave);
}

// The step's last sweep (fluidx, x frame) with the next step's CFL
// (schedule 3): the x-frame group kernel's body, each block adding its max
// CFL speed to the box's accumulator, which the next step's dt kernel reads.
// CFLM 1 / 2: face / cell-averaged B.
// P pencils per wave (short pencils, kBW/P waves per block)
template <int K, int CFLM, int P = 1>
__global__ void __launch_bounds__(64 * kBW / P)
    batch_fluidx_x_cfl_kernel(const BoxDesc *__restrict__ boxes, const WorkItem *__restrict__ items,
                              const float *__restrict__ dt, int flip, unsigned long long *__restrict__ cmax)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *smem = reinterpret_cast<float *>(smem_raw);
    const WorkItem wi = items[xcd_tile(blockIdx.x, gridDim.x)];
    const BoxDesc &bd = boxes[wi.box];
    const float *bcur = flip ? bd.b2 : bd.b;
    if (K % 4 == 0 && bd.fr[0].vec_ok)
        fluidx_x_tile<float, K, kBW, K % 4 == 0, CFLM, P>(bd.u, bcur, bd.fr[0], dt + wi.box, wi.tile, bd.ntj[0],
                                                           smem, nullptr, cmax + wi.box);
    else
        fluidx_x_tile<float, K, kBW, false, CFLM, P>(bd.u, bcur, bd.fr[0], dt + wi.box, wi.tile, bd.ntj[0], smem,
                                                      nullptr, cmax + wi.box);
}

static const void *fluidx_x_cfl_kernel(int K, int P, int cflm)
{
#define MHD_FC(KK, PP)                                                                                               \
    return cflm == 2 ? (const void *)batch_fluidx_x_cfl_kernel<KK, 2, PP> : (const void *)batch_fluidx_x_cfl_kernel<KK, 1, PP>
    if (P == 2) {
        if (K == 1) MHD_FC(1, 2);
        if (K == 3) MHD_FC(3, 2);
        return nullptr;
    }
    if (P == 4) {
        if (K == 1) MHD_FC(1, 4);
        if (K == 3) MHD_FC(3, 4);
        return nullptr;
    }
    switch (K) {
    case 1: MHD_FC(1, 1);
    case 2: MHD_FC(2, 1);
    case 3: MHD_FC(3, 1);
    case 4: MHD_FC(4, 1);
    case 5: MHD_FC(5, 1);
    case 6: MHD_FC(6, 1);
    case 8: MHD_FC(8, 1);
    default: return nullptr;
    }
#undef MHD_FC
}

// the step's first sweep (fluidx, x frame) with P > 1 pencils per wave
template <int K, int P>
__global__ void __launch_bounds__(64 * kBW / P)
    batch_fluidx_x_p_kernel(const BoxDesc *__restrict__ boxes, const WorkItem *__restrict__ items,
                            const float *__restrict__ dt, int flip)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *smem = reinterpret_cast<float *>(smem_raw);
    const WorkItem wi = items[xcd_tile(blockIdx.x, gridDim.x)];
    const BoxDesc &bd = boxes[wi.box];
    fluidx_x_tile<float, K, kBW, false, 0, P>(bd.u, flip ? bd.b2 : bd.b, bd.fr[0], dt + wi.box, wi.tile, bd.ntj[0],
                                              smem, nullptr);
}

static const void *fluidx_x_p_kernel(int K, int P)
{
    if (P == 2) return K == 1 ? (const void *)batch_fluidx_x_p_kernel<1, 2>
                              : K == 3 ? (const void *)batch_fluidx_x_p_kernel<3, 2> : nullptr;
    if (P == 4) return K == 1 ? (const void *)batch_fluidx_x_p_kernel<1, 4>
                              : K == 3 ? (const void *)batch_fluidx_x_p_kernel<3, 4> : nullptr;
    return nullptr;
}

// x-frame advectbyzx in tiles of two planes (schedule 3): one block per item
// (box, tile of kBW pencils in a plane pair), exact K
template <int K>
__global__ void __launch_bounds__(64 * kBW)
    batch_advect_x2_kernel(const BoxDesc *__restrict__ boxes, const WorkItem *__restrict__ items,
                           const float *__restrict__ dt, int flip)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *smem = reinterpret_cast<float *>(smem_raw);
    const WorkItem wi = items[xcd_tile(blockIdx.x, gridDim.x)];
    const BoxDesc &bd = boxes[wi.box];
    const float *bcur = flip ? bd.b2 : bd.b;
    float *bnext = flip ? bd.b : bd.b2;
    if (K % 4 == 0 && bd.fr[0].vec_ok)
        advect_x2_tile<float, K, kBW, K % 4 == 0>(bd.u, bcur, bnext, bd.fr[0], dt + wi.box, wi.tile, bd.ntj[0], smem);
    else
        advect_x2_tile<float, K, kBW, false>(bd.u, bcur, bnext, bd.fr[0], dt + wi.box, wi.tile, bd.ntj[0], smem);
}

static const void *advect_x2_kernel(int K)
{
    switch (K) {
    case 1: return (const void *)batch_advect_x2_kernel<1>;
    case 2: return (const void *)batch_advect_x2_kernel<2>;
    case 3: return (const void *)batch_advect_x2_kernel<3>;
    case 4: return (const void *)batch_advect_x2_kernel<4>;
    case 5: return (const void *)batch_advect_x2_kernel<5>;
    case 6: return (const void *)batch_advect_x2_kernel<6>;
    case 8: return (const void *)batch_advect_x2_kernel<8>;
    default: return nullptr;
    }
}

// y/z fluidx in tiles of two vectors (schedule 3): one block of 8 waves per
// item (box, tile of 8 pencils), exact K
template <int FR, int K, int P = 1>
__global__ void __launch_bounds__(64 * 2 * kBW / P)
    batch_fluidx_yz_kernel(const BoxDesc *__restrict__ boxes, const WorkItem *__restrict__ items,
                           const float *__restrict__ dt, int flip)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const WorkItem wi = items[xcd_tile(blockIdx.x, gridDim.x)];
    const BoxDesc &bd = boxes[wi.box];
    const int ntj = (bd.fr[FR].f.nj + 2 * kBW - 1) / (2 * kBW);
    float *smem = reinterpret_cast<float *>(smem_raw);
    if constexpr (P > 1) {
        // several short pencils per wave
        fluidx_soa_tile<float, K, 2 * kBW, FR, 0, 0, P>(bd.u, flip ? bd.b2 : bd.b, bd.fr[FR], dt + wi.box, wi.tile,
                                                        ntj, smem);
    } else if (bd.fr[FR].f.ni == 64 * K) {
        // a pencil that fills the wave: its length as a constant (uniform branch)
        fluidx_soa_tile<float, K, 2 * kBW, FR, 0, 64 * K>(bd.u, flip ? bd.b2 : bd.b, bd.fr[FR], dt + wi.box, wi.tile,
                                                          ntj, smem);
    } else {
        fluidx_soa_tile<float, K, 2 * kBW, FR>(bd.u, flip ? bd.b2 : bd.b, bd.fr[FR], dt + wi.box, wi.tile, ntj, smem);
    }
}

static const void *fluidx_yz_kernel(int fr, int K, int P)
{
    if (P == 2 || P == 4) {
#define MHD_FYP(F, PP)                                                                                               \
    if (fr == F && P == PP) return K == 1 ? (const void *)batch_fluidx_yz_kernel<F, 1, PP>                          \
                                          : K == 3 ? (const void *)batch_fluidx_yz_kernel<F, 3, PP> : nullptr;
        MHD_FYP(1, 2)
        MHD_FYP(2, 2)
        MHD_FYP(1, 4)
        MHD_FYP(2, 4)
#undef MHD_FYP
        return nullptr;
    }
#define MHD_FY(F)                                                                                                    \
    if (fr == F) switch (K) {                                                                                        \
        case 1: return (const void *)batch_fluidx_yz_kernel<F, 1>;                                                   \
        case 2: return (const void *)batch_fluidx_yz_kernel<F, 2>;                                                   \
        case 3: return (const void *)batch_fluidx_yz_kernel<F, 3>;                                                   \
        case 4: return (const void *)batch_fluidx_yz_kernel<F, 4>;                                                   \
        case 5: return (const void *)batch_fluidx_yz_kernel<F, 5>;                                                   \
        case 6: return (const void *)batch_fluidx_yz_kernel<F, 6>;                                                   \
        case 8: return (const void *)batch_fluidx_yz_kernel<F, 8>;                                                   \
        default: return nullptr;                                                                                     \
        }
    MHD_FY(1)
    MHD_FY(2)
#undef MHD_FY
    return nullptr;
}

// Fused y/z stream (schedule 3): one block per item (box, plane pair and
// segment), sized for the group's widest box; the waves beyond a narrower
// box's pencil tiles end at once (a barrier waits only for running waves)
template <int FR, int NW>
__global__ void __launch_bounds__(64 * NW) batch_stream_ab_kernel(const BoxDesc *__restrict__ boxes,
                                                                  const WorkItem *__restrict__ items,
                                                                  const float *__restrict__ dt, int flip)
{
    __shared__ float xch[64];
    const WorkItem wi = items[xcd_tile(blockIdx.x, gridDim.x)];
    const BoxDesc &bd = boxes[wi.box];
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (w >= bd.sb[FR].ntj) return;
    const int nk2 = (bd.sb[FR].nk + 1) / 2;
    advect_stream_ab<float, FR>(bd.u, flip ? bd.b2 : bd.b, flip ? bd.b : bd.b2, bd.sb[FR], dt + wi.box,
                                wi.tile % nk2, wi.tile / nk2, w, xch);
}

// CFL over the active boxes: one block per (box, unit range); per-box max
template <bool AVG>
__global__ void __launch_bounds__(256) batch_cfl_kernel(const BoxDesc *__restrict__ boxes,
                                                        const CflItem *__restrict__ items,
                                                        unsigned long long *__restrict__ cmax)
{
    const CflItem it = items[blockIdx.x];
    const BoxDesc &bd = boxes[it.box];
    float c = cfl_range<float, AVG>(bd.u, bd.b, bd.g, (long)it.e0 + threadIdx.x, (long)it.e0 + it.ne, blockDim.x);
    c = wave_cfl_max(c);
    __shared__ float red[4];
    if (lane_id() == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = red[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = cfl_max(m, red[i]);
        atomicMax(cmax + it.box, (unsigned long long)__float_as_uint(m));
    }
}

// per-box dt (Heterogeneous_MHD/C/mhd.c:63-65) for the active slots; clears
// the CFL maxima and the sweep queues for this step.  A box whose CFL max is
// NaN (NaN-sticky, cfl_max) gets the reference's sequential max over its
// state instead (cfl_after_last_nan), by the whole block.
__global__ void __launch_bounds__(256) batch_dt_kernel(const BoxDesc *__restrict__ boxes, int avg,
                                                       const int *__restrict__ active, int nactive,
                                                       unsigned long long *cmax, float *dt, float *t, double *t64,
                                                       const float *tf, float *dts, int cap, int *iter, unsigned *ctr,
                                                       int nctr)
{
    __shared__ int nan_box[256];
    __shared__ int n_nan;
    for (int base = 0; base < nactive; base += blockDim.x) {
        const int a = base + (int)threadIdx.x;
        const float ca = a < nactive ? __uint_as_float((unsigned)cmax[active[a]]) : 0.0f;
        const bool bad = ca != ca;
        if (!__syncthreads_or(bad)) continue;
        if (threadIdx.x == 0) n_nan = 0;
        __syncthreads();
        if (bad) nan_box[atomicAdd(&n_nan, 1)] = active[a];
        __syncthreads();
        for (int q = 0; q < n_nan; q++) {
            const BoxDesc &bd = boxes[nan_box[q]];
            const long ncell = (long)bd.g.nx * bd.g.ny * bd.g.nz;
            const float c = avg ? cfl_after_last_nan<float>(ncell, SoaSpeed<float, true>{bd.u, bd.b, bd.g})
                                : cfl_after_last_nan<float>(ncell, SoaSpeed<float, false>{bd.u, bd.b, bd.g});
            if (threadIdx.x == 0) cmax[nan_box[q]] = (unsigned long long)__float_as_uint(c);
        }
        __syncthreads();
    }
    for (int a = threadIdx.x; a < nactive; a += blockDim.x) {
        const int bx = active[a];
        const float c = __uint_as_float((unsigned)cmax[bx]);
        cmax[bx] = 0ull;
        const float cfl = (float)(1.0 / (double)c);
        float d = (float)(0.9 * (double)cfl);
        if (d > (tf[bx] - t[bx]) / 2.f) d = (tf[bx] - t[bx]) / 2.f;
        t[bx] = t[bx] + 2.f * d;
        t64[bx] = t64[bx] + 2.0 * (double)d;
        dt[bx] = d;
        const int it = iter[bx];
        if (it < cap) dts[(long)bx * cap + it] = d;
        iter[bx] = it + 1;
    }
    for (int i = threadIdx.x; i < nctr; i += blockDim.x) ctr[i] = 0u;
}

// ---------------------------------------------------------------------------
// host: launch helpers
// ---------------------------------------------------------------------------
template <int FR, int KMAX, bool ADV> static const void *kern_ptr()
{
    return (const void *)batch_sweep_kernel<FR, KMAX, ADV>;
}

template <int FR, int KMAX, bool ADV> static const void *grid_ptr()
{
    return (const void *)batch_grid_kernel<FR, KMAX, ADV>;
}

static const void *grid_kernel(int fr, int kmax, bool adv)
{
#define MHD_GP(F)                                                                                                    \
    if (fr == F) {                                                                                                   \
        if (kmax <= 2) return adv ? grid_ptr<F, 2, true>() : grid_ptr<F, 2, false>();                              \
        if (kmax <= 4) return adv ? grid_ptr<F, 4, true>() : grid_ptr<F, 4, false>();                              \
        return adv ? grid_ptr<F, 8, true>() : grid_ptr<F, 8, false>();                                              \
    }
    MHD_GP(0)
    MHD_GP(1)
    MHD_GP(2)
#undef MHD_GP
    return nullptr;
}

static hipError_t launch_grid(int fr, int kmax, bool adv, int nitems, size_t lds, hipStream_t st,
                              const BoxDesc *boxes, const WorkItem *items, const float *dt, int flip)
{
#define MHD_LG(F, KM, A)                                                                                             \
    hipLaunchKernelGGL((batch_grid_kernel<F, KM, A>), dim3(nitems), dim3(64 * kBW), lds, st, boxes, items, dt, flip)
#define MHD_LGF(F)                                                                                                   \
    if (fr == F) {                                                                                                   \
        if (kmax <= 2) {                                                                                             \
            if (adv) MHD_LG(F, 2, true); else MHD_LG(F, 2, false);                                                   \
        } else if (kmax <= 4) {                                                                                      \
            if (adv) MHD_LG(F, 4, true); else MHD_LG(F, 4, false);                                                   \
        } else {                                                                                                     \
            if (adv) MHD_LG(F, 8, true); else MHD_LG(F, 8, false);                                                   \
        }                                                                                                            \
    }
    MHD_LGF(0)
    MHD_LGF(1)
    MHD_LGF(2)
#undef MHD_LGF
#undef MHD_LG
    return hipGetLastError();
}

template <int FR, int K> static const void *group_ptr(bool adv)
{
    return adv ? (const void *)batch_group_kernel<FR, K, true> : (const void *)batch_group_kernel<FR, K, false>;
}

static const void *group_kernel(int fr, int K, bool adv)
{
#define MHD_GK(F)                                                                                                    \
    if (fr == F) switch (K) {                                                                                        \
        case 1: return group_ptr<F, 1>(adv);                                                                         \
        case 2: return group_ptr<F, 2>(adv);                                                                         \
        case 3: return group_ptr<F, 3>(adv);                                                                         \
        case 4: return group_ptr<F, 4>(adv);                                                                         \
        case 5: return group_ptr<F, 5>(adv);                                                                         \
        case 6: return group_ptr<F, 6>(adv);                                                                         \
        case 8: return group_ptr<F, 8>(adv);                                                                         \
        default: return nullptr;                                                                                     \
        }
    MHD_GK(0)
    MHD_GK(1)
    MHD_GK(2)
#undef MHD_GK
    return nullptr;
}

static hipError_t launch_group(int fr, int K, bool adv, int nitems, size_t lds, hipStream_t st,
                               const BoxDesc *boxes, const WorkItem *items, const float *dt, int flip)
{
    const void *k = group_kernel(fr, K, adv);
    if (!k) return hipErrorInvalidValue;
    void *args[] = {(void *)&boxes, (void *)&items, (void *)&dt, (void *)&flip};
    return hipLaunchKernel(k, dim3(nitems), dim3(64 * kBW), args, lds, st);
}

static const void *sweep_kernel(int fr, int kmax, bool adv)
{
#define MHD_KP(F)                                                                                                    \
    if (fr == F) {                                                                                                   \
        if (kmax <= 2) return adv ? kern_ptr<F, 2, true>() : kern_ptr<F, 2, false>();                              \
        if (kmax <= 4) return adv ? kern_ptr<F, 4, true>() : kern_ptr<F, 4, false>();                              \
        return adv ? kern_ptr<F, 8, true>() : kern_ptr<F, 8, false>();                                              \
    }
    MHD_KP(0)
    MHD_KP(1)
    MHD_KP(2)
#undef MHD_KP
    return nullptr;
}

static hipError_t launch_sweep(int fr, int kmax, bool adv, int grid, size_t lds, hipStream_t st, const BoxDesc *boxes,
                               const WorkItem *items, const int *part, unsigned *ctr, const float *dt, int flip, int nq)
{
#define MHD_LS(F, KM, A)                                                                                             \
    hipLaunchKernelGGL((batch_sweep_kernel<F, KM, A>), dim3(grid), dim3(64 * kBW), lds, st, boxes, items, part, ctr,  \
                       dt, flip, nq)
#define MHD_LF(F)                                                                                                    \
    if (fr == F) {                                                                                                   \
        if (kmax <= 2) {                                                                                             \
            if (adv) MHD_LS(F, 2, true); else MHD_LS(F, 2, false);                                                   \
        } else if (kmax <= 4) {                                                                                      \
            if (adv) MHD_LS(F, 4, true); else MHD_LS(F, 4, false);                                                   \
        } else {                                                                                                     \
            if (adv) MHD_LS(F, 8, true); else MHD_LS(F, 8, false);                                                   \
        }                                                                                                            \
    }
    MHD_LF(0)
    MHD_LF(1)
    MHD_LF(2)
#undef MHD_LF
#undef MHD_LS
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
Engine::Engine(int max_boxes, int cap, int avg_b) : max_boxes_(max_boxes), cap_(cap), avg_(avg_b ? 1 : 0) {}

Engine::~Engine()
{
    if (st_) (void)hipStreamSynchronize(st_);
    drop_graph();
    if (xs_) (void)hipStreamSynchronize(xs_);
    for (size_t s = 0; s < boxes_.size(); s++)
        if (used_[s]) (void)hipFree(boxes_[s].u);
    for (const Chunk &c : pool_) (void)hipFree(c.base);
    for (float *x : xbuf_)
        if (x) (void)hipFree(x);
    void *ps[] = {d_boxes_, d_dt_, d_t_, d_tf_, d_dts_, d_t64_, d_cmax_, d_iter_, d_active_, d_ctr_, d_part_,
                  d_items_[0], d_items_[1], d_items_[2], d_cfl_, d_sitems_[0], d_sitems_[1], d_f8items_};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (evfork_) (void)hipEventDestroy(evfork_);
    for (int i = 0; i < kAux; i++) {
        if (evjoin_[i]) (void)hipEventDestroy(evjoin_[i]);
        if (aux_[i]) {
            (void)hipStreamSynchronize(aux_[i]);
            (void)hipStreamDestroy(aux_[i]);
        }
    }
    if (st_) (void)hipStreamDestroy(st_);
    if (xs_) (void)hipStreamDestroy(xs_);
}

int Engine::init()
{
    if (max_boxes_ < 1 || cap_ < 1) return fail(MHD_ERR_ARG, "engine: bad capacity");
    hipError_t e = hipStreamCreate(&st_);
    if (e == hipSuccess) e = hipStreamCreate(&xs_);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    const size_t B = max_boxes_;
    struct A {
        void **p;
        size_t n;
    } al[] = {{(void **)&d_boxes_, B * sizeof(BoxDesc)}, {(void **)&d_dt_, B * 4},   {(void **)&d_t_, B * 4},
              {(void **)&d_tf_, B * 4},                   {(void **)&d_dts_, B * cap_ * 4},
              {(void **)&d_t64_, B * 8},                  {(void **)&d_cmax_, B * 8}, {(void **)&d_iter_, B * 4},
              {(void **)&d_active_, B * 4},               {(void **)&d_ctr_, 12 * kQueues * 4},
              {(void **)&d_part_, 3 * (kQueues + 1) * 4}};
    for (auto &a : al) {
        e = hipMalloc(a.p, a.n);
        if (e != hipSuccess) return fail(MHD_ERR_NOMEM, "engine: hipMalloc(%zu): %s", a.n, hipGetErrorString(e));
    }
    e = hipMemsetAsync(d_cmax_, 0, B * 8, st_);
    if (e == hipSuccess) e = hipMemsetAsync(d_ctr_, 0, 12 * kQueues * 4, st_);
    if (e == hipSuccess) e = hipEventCreate(&ev0_);
    if (e == hipSuccess) e = hipEventCreate(&ev1_);
    if (e != hipSuccess) return hip_fail(e, "engine init");
    boxes_.assign(B, BoxDesc());
    used_.assign(B, false);
    tf_.assign(B, INFINITY);
    cfl_fresh_.assign(B, 0);
    cfl_slot_off_.assign(B, 0);
    cfl_slot_n_.assign(B, 0);
    hipDeviceProp_t prop;
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    ncu_ = prop.multiProcessorCount;
    for (int fr = 0; fr < 3; fr++)
        for (int km : {2, 4, 8})
            for (int adv = 0; adv < 2; adv++) {
                // the largest tile (n = 512) needs 76 KB; the persistent
                // kernels also hold a little static LDS, so stay below 160 KB
                e = hipFuncSetAttribute(sweep_kernel(fr, km, adv), hipFuncAttributeMaxDynamicSharedMemorySize,
                                        128 * 1024);
                if (e == hipSuccess)
                    e = hipFuncSetAttribute(grid_kernel(fr, km, adv), hipFuncAttributeMaxDynamicSharedMemorySize,
                                            128 * 1024);
                if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute");
            }
    for (int fr = 0; fr < 3 && e == hipSuccess; fr++)
        for (int K : {1, 2, 3, 4, 5, 6, 8}) {
            for (int adv = 0; adv < 2 && e == hipSuccess; adv++)
                e = hipFuncSetAttribute(group_kernel(fr, K, adv), hipFuncAttributeMaxDynamicSharedMemorySize,
                                        128 * 1024);
            if (fr > 0 && e == hipSuccess)
                e = hipFuncSetAttribute(fluidx_yz_kernel(fr, K, 1), hipFuncAttributeMaxDynamicSharedMemorySize,
                                        160 * 1024);
            if (fr == 0 && e == hipSuccess)
                e = hipFuncSetAttribute(advect_x2_kernel(K), hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024);
            for (int cflm = 1; fr == 0 && cflm <= 2 && e == hipSuccess; cflm++)
                e = hipFuncSetAttribute(fluidx_x_cfl_kernel(K, 1, cflm), hipFuncAttributeMaxDynamicSharedMemorySize,
                                        128 * 1024);
        }
    // the short-pencil fluidx kernels (n <= 96: a few tens of KB of LDS)
    for (int P : {2, 4})
        for (int K : {1, 3}) {
            for (int fr = 1; fr < 3 && e == hipSuccess; fr++)
                e = hipFuncSetAttribute(fluidx_yz_kernel(fr, K, P), hipFuncAttributeMaxDynamicSharedMemorySize,
                                        64 * 1024);
            for (int cflm = 1; cflm <= 2 && e == hipSuccess; cflm++)
                e = hipFuncSetAttribute(fluidx_x_cfl_kernel(K, P, cflm), hipFuncAttributeMaxDynamicSharedMemorySize,
                                        64 * 1024);
        }
    if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute");
    // side streams (default priority) and the fork/join events of the
    // grouped schedule's step graph
    for (int i = 0; i < kAux && e == hipSuccess; i++) {
        e = hipStreamCreate(&aux_[i]);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&evjoin_[i], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&evfork_, hipEventDisableTiming);
    if (e != hipSuccess) return hip_fail(e, "engine side streams");
    e = hipStreamSynchronize(st_);
    return e == hipSuccess ? MHD_OK : hip_fail(e, "engine init");
}

long Engine::cells(int slot) const
{
    const SoaGrid &g = boxes_[slot].g;
    return (long)g.nx * g.ny * g.nz;
}

// Box memory: one allocation of 12 planes (u, b, b2, the EMF scratch fa).  Released boxes go to
// a pool and are reused for boxes of N <= chunk <= 2N cells per plane, so a
// serving loop turning requests over calls no hipFree (which waits for the
// whole device).  On an allocation failure the pool is emptied and the
// allocation retried.
int Engine::alloc_box(long N, float **base)
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        int best = -1;
        for (size_t i = 0; i < pool_.size(); i++)
            if (pool_[i].N >= N && pool_[i].N <= 2 * N && (best < 0 || pool_[i].N < pool_[best].N)) best = (int)i;
        if (best >= 0) {
            *base = pool_[best].base;
            pool_.erase(pool_.begin() + best);
            return MHD_OK;
        }
    }
    const size_t bytes = 12 * (size_t)N * 4;
    hipError_t e = hipMalloc((void **)base, bytes);
    if (e != hipSuccess) {
        std::vector<Chunk> drop;
        {
            std::lock_guard<std::mutex> lk(mu_);
            drop.swap(pool_);
        }
        for (const Chunk &c : drop) (void)hipFree(c.base);
        e = hipMalloc((void **)base, bytes);
    }
    return e == hipSuccess ? MHD_OK : fail(MHD_ERR_NOMEM, "engine: box of %ld cells: %s", N, hipGetErrorString(e));
}

void Engine::release_box(float *base, long N)
{
    std::lock_guard<std::mutex> lk(mu_);
    pool_.push_back({base, N});
    if (pool_.size() > 32) {  // keep the pool small: the oldest chunk goes
        (void)hipFree(pool_.front().base);
        pool_.erase(pool_.begin());
    }
}

int Engine::claim_slot(int nx, int ny, int nz, hipStream_t st, int *slot)
{
    int rc = check_shape_fast3(nx, ny, nz);
    if (rc) return rc;
    BoxDesc d;
    memset(&d, 0, sizeof d);
    d.g = make_grid(nx, ny, nz, soa_pad_y(4), soa_pad_z(4));
    const long N = d.g.N;
    float *base = nullptr;
    rc = alloc_box(N, &base);
    if (rc) return rc;
    d.u = base;
    d.b = base + 5 * N;
    d.b2 = base + 8 * N;
    d.fa = base + 11 * N;
    for (int fr = 0; fr < 3; fr++) {
        d.fr[fr] = make_soa_box<float>(fr, d.g);
        d.K[fr] = pick_k(d.fr[fr].f.ni);
        d.PF[fr] = pick_p(d.fr[fr].f.ni);
        d.KF[fr] = d.PF[fr] > 1 ? d.fr[fr].f.ni * d.PF[fr] / 64 : d.K[fr];
        d.ntj[fr] = (d.fr[fr].f.nj + kBW - 1) / kBW;
        if (fr > 0) {
            d.sb[fr] = make_stream_box(fr, d.g);
            d.nwa[fr] = stream_waves_a(d.sb[fr]);
            d.nwb[fr] = stream_waves_b(d.sb[fr]);
        }
    }
    d.streamed = stream_pays(d.sb[1]) && stream_pays(d.sb[2]);
    int s = -1;
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (int i = 0; i < max_boxes_; i++)
            if (!used_[i]) {
                s = i;
                break;
            }
        if (s >= 0) {
            boxes_[s] = d;
            used_[s] = true;
        }
    }
    if (s < 0) {
        release_box(base, N);
        return fail(MHD_ERR_FULL, "engine: all %d slots in use", max_boxes_);
    }
    hipError_t e = hipMemcpyAsync(d_boxes_ + s, &boxes_[s], sizeof(BoxDesc), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) rc = reset_clock_on(s, INFINITY, st);
    else rc = hip_fail(e, "engine add_box");
    if (rc) {
        remove_box(s);
        return rc;
    }
    *slot = s;
    return MHD_OK;
}

int Engine::add_box(int nx, int ny, int nz, int *slot)
{
    hipError_t e = hipStreamSynchronize(st_);
    if (e != hipSuccess) return hip_fail(e, "engine add_box");
    return claim_slot(nx, ny, nz, st_, slot);
}

int Engine::remove_box(int slot)
{
    if (slot < 0 || slot >= max_boxes_ || !used_[slot]) return fail(MHD_ERR_ARG, "engine: bad slot %d", slot);
    hipError_t e = hipStreamSynchronize(st_);
    if (e == hipSuccess) e = hipStreamSynchronize(xs_);
    if (e != hipSuccess) return hip_fail(e, "engine remove_box");
    auto it = std::find(active_.begin(), active_.end(), slot);
    if (it != active_.end()) {
        active_.erase(it);
        dirty_ = true;
    }
    BoxDesc d;
    {
        std::lock_guard<std::mutex> lk(mu_);
        d = boxes_[slot];
        memset(&boxes_[slot], 0, sizeof(BoxDesc));
        used_[slot] = false;
    }
    release_box(d.u, d.g.N);
    return MHD_OK;
}

// AoS host arrays <-> the slot's SoA planes through the stream's AoS buffer
int Engine::transfer(int slot, void *h_u, void *h_b, int to_device, hipStream_t st)
{
    BoxDesc &d = boxes_[slot];
    const long nc = (long)d.g.nx * d.g.ny * d.g.nz;
    const int xi = st == xs_ ? 1 : 0;
    const size_t need = 5 * (size_t)nc * 4;
    hipError_t e = hipSuccess;
    if (xcap_[xi] < need) {
        if (xbuf_[xi]) (void)hipFree(xbuf_[xi]);
        xbuf_[xi] = nullptr;
        xcap_[xi] = 0;
        e = hipMalloc((void **)&xbuf_[xi], need);
        if (e != hipSuccess) return fail(MHD_ERR_NOMEM, "state transfer buffer: %s", hipGetErrorString(e));
        xcap_[xi] = need;
    }
    float *tmp = xbuf_[xi];
    struct P {
        void *h;
        float *d;
        int nc;
    } ps[2] = {{h_u, d.u, 5}, {h_b, d.b, 3}};
    if (to_device) cfl_fresh_[slot] = 0;
    for (auto &p : ps) {
        if (e != hipSuccess) break;
        if (to_device) {
            e = hipMemcpyAsync(tmp, p.h, (size_t)nc * p.nc * 4, hipMemcpyHostToDevice, st);
            if (e == hipSuccess) e = launch_aos_soa<float>(p.d, tmp, p.nc, d.g, 1, st);
        } else {
            e = launch_aos_soa<float>(p.d, tmp, p.nc, d.g, 0, st);
            if (e == hipSuccess) e = hipMemcpyAsync(p.h, tmp, (size_t)nc * p.nc * 4, hipMemcpyDeviceToHost, st);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    return e == hipSuccess ? MHD_OK : hip_fail(e, "state transfer");
}

int Engine::set_state(int slot, const float *h_u, const float *h_b)
{
    if (slot < 0 || slot >= max_boxes_ || !used_[slot] || !h_u || !h_b) return fail(MHD_ERR_ARG, "set_state");
    hipError_t e = hipStreamSynchronize(st_);
    if (e != hipSuccess) return hip_fail(e, "set_state");
    return transfer(slot, const_cast<float *>(h_u), const_cast<float *>(h_b), 1, st_);
}

int Engine::get_state(int slot, float *h_u, float *h_b)
{
    if (slot < 0 || slot >= max_boxes_ || !used_[slot] || !h_u || !h_b) return fail(MHD_ERR_ARG, "get_state");
    return transfer(slot, h_u, h_b, 0, st_);
}

int Engine::stage_in(int nx, int ny, int nz, const float *h_u, const float *h_b, int *slot)
{
    if (!h_u || !h_b || !slot) return fail(MHD_ERR_ARG, "stage_in: bad argument");
    int s = -1;
    int rc = claim_slot(nx, ny, nz, xs_, &s);
    if (rc) return rc;
    rc = transfer(s, const_cast<float *>(h_u), const_cast<float *>(h_b), 1, xs_);
    if (rc) {
        BoxDesc d;
        {
            std::lock_guard<std::mutex> lk(mu_);
            d = boxes_[s];
            memset(&boxes_[s], 0, sizeof(BoxDesc));
            used_[s] = false;
        }
        release_box(d.u, d.g.N);
        return rc;
    }
    *slot = s;
    return MHD_OK;
}

int Engine::stage_out(int slot, float *h_u, float *h_b)
{
    if (slot < 0 || slot >= max_boxes_ || !used_[slot]) return fail(MHD_ERR_ARG, "stage_out: bad slot %d", slot);
    const int rc = (h_u && h_b) ? transfer(slot, h_u, h_b, 0, xs_) : MHD_OK;
    BoxDesc d;
    {
        std::lock_guard<std::mutex> lk(mu_);
        d = boxes_[slot];
        memset(&boxes_[slot], 0, sizeof(BoxDesc));
        used_[slot] = false;
    }
    release_box(d.u, d.g.N);
    return rc;
}

int Engine::reset_clock(int slot, double tf)
{
    if (slot < 0 || slot >= max_boxes_ || !used_[slot]) return fail(MHD_ERR_ARG, "reset_clock: bad slot");
    return reset_clock_on(slot, tf, st_);
}

// the slot's clock entries: tf, t = dt = t64 = 0, no steps, CFL max 0 (small
// fills on the given stream, outside any captured graph)
int Engine::reset_clock_on(int slot, double tf, hipStream_t st)
{
    tf_[slot] = tf;
    cfl_fresh_[slot] = 0;
    const float tff = (float)tf;
    hipError_t e = hipMemcpyAsync(d_tf_ + slot, &tff, 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(d_t_ + slot, 0, 4, st);
    if (e == hipSuccess) e = hipMemsetAsync(d_dt_ + slot, 0, 4, st);
    if (e == hipSuccess) e = hipMemsetAsync(d_t64_ + slot, 0, 8, st);
    if (e == hipSuccess) e = hipMemsetAsync(d_iter_ + slot, 0, 4, st);
    if (e == hipSuccess) e = hipMemsetAsync(d_cmax_ + slot, 0, 8, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? MHD_OK : hip_fail(e, "reset_clock");
}

int Engine::dts(int slot, double *h, int max, int *count)
{
    if (slot < 0 || slot >= max_boxes_ || !used_[slot] || !count) return fail(MHD_ERR_ARG, "dts: bad argument");
    int it = 0;
    hipError_t e = hipStreamSynchronize(st_);
    if (e == hipSuccess) e = hipMemcpy(&it, d_iter_ + slot, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "dts");
    *count = it;
    int n = std::min(std::min(it, max), cap_);
    if (n > 0 && h) {
        std::vector<float> f(n);
        e = hipMemcpy(f.data(), d_dts_ + (long)slot * cap_, n * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "dts copy");
        for (int i = 0; i < n; i++) h[i] = f[i];
    }
    return MHD_OK;
}

int Engine::time(int slot, double *t, double *t64)
{
    if (slot < 0 || slot >= max_boxes_ || !used_[slot]) return fail(MHD_ERR_ARG, "time: bad slot");
    float tf = 0;
    double td = 0;
    hipError_t e = hipStreamSynchronize(st_);
    if (e == hipSuccess) e = hipMemcpy(&tf, d_t_ + slot, 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&td, d_t64_ + slot, 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "time");
    if (t) *t = tf;
    if (t64) *t64 = td;
    return MHD_OK;
}

int Engine::set_active(const std::vector<int> &slots)
{
    for (int s : slots)
        if (s < 0 || s >= max_boxes_ || !used_[s]) return fail(MHD_ERR_ARG, "set_active: bad slot %d", s);
    active_ = slots;
    dirty_ = true;
    return MHD_OK;
}

int Engine::set_schedule(int bpc, int policy)
{
    if (bpc < 0 || policy < 0 || policy > 3) return fail(MHD_ERR_ARG, "set_schedule: bad argument");
    bpc_ = bpc;
    policy_ = policy;
    dirty_ = true;
    // the other schedules do not keep the CFL accumulators current
    for (auto &f : cfl_fresh_) f = 0;
    return MHD_OK;
}

template <typename X> static hipError_t ensure_cap(X **p, int *cap, int need)
{
    if (need <= *cap) return hipSuccess;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    int c = std::max(need, 2 * *cap);
    hipError_t e = hipMalloc((void **)p, sizeof(X) * (size_t)c);
    *cap = e == hipSuccess ? c : 0;
    return e;
}

int Engine::build()
{
    hipError_t e = hipStreamSynchronize(st_);
    if (e != hipSuccess) return hip_fail(e, "engine build");
    drop_graph();
    // largest boxes first: the queues then hand out the longest items first
    std::vector<int> order = active_;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cells(a) > cells(b); });
    // schedule 3: boxes grouped by their K in every frame, the groups with
    // the most cells first; the lists below then hold each group contiguously
    groups_.clear();
    if (policy_ == 3) {
        const BoxGroup G0{};
        std::vector<std::vector<int>> members;
        for (int s : order) {
            size_t g = 0;
            for (; g < groups_.size(); g++)
                if (memcmp(groups_[g].K, boxes_[s].K, sizeof G0.K) == 0 &&
                    memcmp(groups_[g].KF, boxes_[s].KF, sizeof G0.KF) == 0 &&
                    memcmp(groups_[g].PF, boxes_[s].PF, sizeof G0.PF) == 0)
                    break;
            if (g == groups_.size()) {
                BoxGroup G;
                memset(&G, 0, sizeof G);
                for (int fr = 0; fr < 3; fr++) {
                    G.K[fr] = boxes_[s].K[fr];
                    G.KF[fr] = boxes_[s].KF[fr];
                    G.PF[fr] = boxes_[s].PF[fr];
                }
                groups_.push_back(G);
                members.emplace_back();
            }
            members[g].push_back(s);
            groups_[g].cells += cells(s);
        }
        // The streamed y/z advectbyzx needs many waves in flight: a box streams
        // if it has 2048 waves of its own or if its group has them together
        // (the group's boxes run in one launch).  Within a group the boxes that
        // run tiles come first, so that their advect items lead each frame's
        // item range.
        stream_now_.assign(max_boxes_, 0);
        for (auto &m : members) {
            long w1 = 0, w2 = 0;
            for (int s : m) {
                w1 += boxes_[s].nwa[1];
                w2 += boxes_[s].nwa[2];
            }
            const bool together = w1 >= 2048 && w2 >= 2048;
            for (int s : m) stream_now_[s] = together || boxes_[s].streamed;
            std::stable_partition(m.begin(), m.end(), [&](int s) { return !stream_now_[s]; });
        }
        // short pencils keep several per wave only where that pays for the
        // group's launch as a whole (p_pays over its boxes' pencils); every
        // box of a group with PF > 1 has the same pencil length
        for (size_t g = 0; g < groups_.size(); g++)
            for (int fr = 0; fr < 3; fr++) {
                BoxGroup &G = groups_[g];
                if (G.PF[fr] == 1) continue;
                long pencils = 0;
                for (int s : members[g]) pencils += (long)boxes_[s].fr[fr].f.nj * boxes_[s].fr[fr].f.nk;
                if (p_pays(boxes_[members[g][0]].fr[fr].f.ni, pencils) == 1) {
                    G.PF[fr] = 1;
                    G.KF[fr] = G.K[fr];
                }
            }
        std::vector<int> gi(groups_.size());
        for (size_t g = 0; g < gi.size(); g++) gi[g] = (int)g;
        std::stable_sort(gi.begin(), gi.end(), [&](int a, int b) { return groups_[a].cells > groups_[b].cells; });
        std::vector<BoxGroup> sorted;
        order.clear();
        for (int g : gi) {
            BoxGroup G = groups_[g];
            G.act_off = (int)order.size();
            G.act_n = (int)members[g].size();
            for (int s : members[g]) order.push_back(s);
            sorted.push_back(G);
        }
        groups_.swap(sorted);
    }
    kmax_ = 1;
    lds_fluidx_ = lds_advect_ = 0;
    std::vector<WorkItem> items[3];
    std::vector<int> part(3 * (kQueues + 1), 0);
    for (int fr = 0; fr < 3; fr++) {
        std::vector<double> cost;
        size_t gnext = 0;
        for (size_t oi = 0; oi < order.size(); oi++) {
            const int s = order[oi];
            if (gnext < groups_.size() && (int)oi == groups_[gnext].act_off) groups_[gnext++].item_off[fr] = (int)items[fr].size();
            const BoxDesc &d = boxes_[s];
            const int nt = d.ntj[fr] * d.fr[fr].f.nk;
            kmax_ = std::max(kmax_, d.K[fr]);
            // (the x-frame fluidx with the fused CFL needs fluidx_x_words)
            const int lf = (int)(4 * std::max(fluidx_soa_words<kBW>(d.fr[fr].f.ni),
                                              fr == 0 ? fluidx_x_words<kBW>(d.fr[fr].f.ni, 1) : 0)),
                      la = (int)(4 * (7 * kBW + 2) * d.fr[fr].f.ni);
            // schedule 3 runs the x-frame advectbyzx with direct loads (smaller LDS)
            const int lag = fr == 0 ? (int)(4 * advect_x_words<kBW>(d.fr[fr].f.ni)) : la;
            lds_fluidx_ = std::max(lds_fluidx_, lf);
            lds_advect_ = std::max(lds_advect_, la);
            if (gnext > 0) {
                BoxGroup &G = groups_[gnext - 1];
                G.lds_f[fr] = std::max(G.lds_f[fr], lf);
                G.lds_a[fr] = std::max(G.lds_a[fr], lag);
            }
            for (int t = 0; t < nt; t++) {
                items[fr].push_back({s, t});
                cost.push_back((double)d.K[fr]);  // lane-slots per tile ~ K (W pencils x 64 lanes x K)
            }
            if (gnext > 0 && !stream_now_[s])
                groups_[gnext - 1].tile_adv_n[fr] = (int)items[fr].size() - groups_[gnext - 1].item_off[fr];
        }
        for (size_t g = 0; g < groups_.size(); g++)
            groups_[g].item_n[fr] =
                (g + 1 < groups_.size() ? groups_[g + 1].item_off[fr] : (int)items[fr].size()) - groups_[g].item_off[fr];
        const int n = (int)items[fr].size();
        int *pt = part.data() + fr * (kQueues + 1);
        if (policy_ == 0) {
            double tot = 0;
            for (double c : cost) tot += c;
            double acc = 0;
            int q = 1;
            pt[0] = 0;
            for (int i = 0; i < n && q < kQueues; i++) {
                acc += cost[i];
                while (q < kQueues && acc >= tot * q / kQueues) pt[q++] = i + 1;
            }
            for (; q <= kQueues; q++) pt[q] = n;
        } else {
            pt[0] = 0;
            for (int q = 1; q <= kQueues; q++) pt[q] = n;  // queue 0 holds everything
        }
        nitems_[fr] = n;
        e = ensure_cap(&d_items_[fr], &cap_items_[fr], std::max(n, 1));
        if (e == hipSuccess && n) e = hipMemcpyAsync(d_items_[fr], items[fr].data(), n * sizeof(WorkItem),
                                                      hipMemcpyHostToDevice, st_);
        if (e != hipSuccess) return hip_fail(e, "engine build items");
    }
    // streamed y/z advectbyzx (schedule 3), per group and frame: the fused
    // stream where every streamed box of the group allows it, one item per
    // (box, plane pair and segment) block; else the two passes, one item per
    // block of four waves of each box, pass A and pass B in their own lists
    std::vector<WorkItem> sitems[2];
    for (size_t g = 0; g < groups_.size(); g++) {
        BoxGroup &G = groups_[g];
        for (int fr = 1; fr < 3; fr++) {
            G.sa_off[fr] = (int)sitems[0].size();
            G.sb_off[fr] = (int)sitems[1].size();
            int nw = 0;
            bool fuse = true;
            for (int oi = G.act_off; oi < G.act_off + G.act_n; oi++) {
                if (!stream_now_[order[oi]]) continue;
                const StreamBox &sb = boxes_[order[oi]].sb[fr];
                nw = std::max(nw, sb.ntj);
                fuse = fuse && stream_fused(sb);
            }
            G.sab_nw[fr] = fuse ? nw : 0;
            for (int oi = G.act_off; oi < G.act_off + G.act_n; oi++) {
                const BoxDesc &d = boxes_[order[oi]];
                if (!stream_now_[order[oi]]) continue;
                if (G.sab_nw[fr]) {
                    const int nb = (d.sb[fr].nk + 1) / 2 * d.sb[fr].nseg;
                    for (int b = 0; b < nb; b++) sitems[0].push_back({order[oi], b});
                } else {
                    for (int w = 0; w < d.nwa[fr]; w += 4) sitems[0].push_back({order[oi], w});
                    for (int w = 0; w < d.nwb[fr]; w += 4) sitems[1].push_back({order[oi], w});
                }
            }
            G.sa_n[fr] = (int)sitems[0].size() - G.sa_off[fr];
            G.sb_n[fr] = (int)sitems[1].size() - G.sb_off[fr];
        }
    }
    // y/z fluidx of a group in tiles of 8 pencils where two such tiles of its
    // longest pencil fit a CU's LDS (as in the single-box path)
    std::vector<WorkItem> f8items;
    for (size_t g = 0; g < groups_.size(); g++) {
        BoxGroup &G = groups_[g];
        for (int fr = 1; fr < 3; fr++) {
            G.f8_off[fr] = (int)f8items.size();
            int lds = 0;
            for (int oi = G.act_off; oi < G.act_off + G.act_n; oi++)
                lds = std::max(lds, (int)(4 * fluidx_soa_words<2 * kBW>(boxes_[order[oi]].fr[fr].f.ni)));
            G.lds_f8[fr] = lds;
            if (lds > 80 * 1024) continue;
            for (int oi = G.act_off; oi < G.act_off + G.act_n; oi++) {
                const Frame &f = boxes_[order[oi]].fr[fr].f;
                const int nt = (f.nj + 2 * kBW - 1) / (2 * kBW) * f.nk;
                for (int t = 0; t < nt; t++) f8items.push_back({order[oi], t});
            }
            G.f8_n[fr] = (int)f8items.size() - G.f8_off[fr];
        }
        // x-frame advectbyzx tiles of two planes
        G.x2_off = (int)f8items.size();
        G.lds_x2 = 0;
        for (int oi = G.act_off; oi < G.act_off + G.act_n; oi++) {
            const BoxDesc &d = boxes_[order[oi]];
            G.lds_x2 = std::max(G.lds_x2, (int)(4 * advect_x2_words<kBW>(d.fr[0].f.ni)));
            const int nt = d.ntj[0] * ((d.fr[0].f.nk + 1) / 2);
            for (int t = 0; t < nt; t++) f8items.push_back({order[oi], t});
        }
        G.x2_n = (int)f8items.size() - G.x2_off;
    }
    {
        const int n = (int)f8items.size();
        e = ensure_cap(&d_f8items_, &cap_f8items_, std::max(n, 1));
        if (e == hipSuccess && n)
            e = hipMemcpyAsync(d_f8items_, f8items.data(), n * sizeof(WorkItem), hipMemcpyHostToDevice, st_);
        if (e != hipSuccess) return hip_fail(e, "engine build items");
    }
    for (int p = 0; p < 2; p++) {
        const int n = (int)sitems[p].size();
        e = ensure_cap(&d_sitems_[p], &cap_sitems_[p], std::max(n, 1));
        if (e == hipSuccess && n)
            e = hipMemcpyAsync(d_sitems_[p], sitems[p].data(), n * sizeof(WorkItem), hipMemcpyHostToDevice, st_);
        if (e != hipSuccess) return hip_fail(e, "engine build stream items");
    }
    // CFL items: ranges of 1024 units (4 per thread; a unit is a 16-byte
    // vector of a row, or a cell where rows are not vector-aligned)
    std::vector<CflItem> cfl;
    {
        size_t gnext = 0;
        for (size_t oi = 0; oi < order.size(); oi++) {
            const int s = order[oi];
            if (gnext < groups_.size() && (int)oi == groups_[gnext].act_off) groups_[gnext++].cfl_off = (int)cfl.size();
            const long units = cfl_units<float>(boxes_[s].g);
            cfl_slot_off_[s] = (int)cfl.size();
            for (long e = 0; e < units; e += 1024) cfl.push_back({s, (int)e, (int)std::min(1024L, units - e)});
            cfl_slot_n_[s] = (int)cfl.size() - cfl_slot_off_[s];
        }
        for (size_t g = 0; g < groups_.size(); g++)
            groups_[g].cfl_n = (g + 1 < groups_.size() ? groups_[g + 1].cfl_off : (int)cfl.size()) - groups_[g].cfl_off;
    }
    ncfl_ = (int)cfl.size();
    e = ensure_cap(&d_cfl_, &cap_cfl_, std::max(ncfl_, 1));
    if (e == hipSuccess && ncfl_) e = hipMemcpyAsync(d_cfl_, cfl.data(), ncfl_ * sizeof(CflItem), hipMemcpyHostToDevice, st_);
    if (e == hipSuccess) e = hipMemcpyAsync(d_part_, part.data(), part.size() * 4, hipMemcpyHostToDevice, st_);
    // the active slots in list order (a permutation of active_; grouped under schedule 3)
    if (e == hipSuccess && !order.empty())
        e = hipMemcpyAsync(d_active_, order.data(), order.size() * 4, hipMemcpyHostToDevice, st_);
    if (e != hipSuccess) return hip_fail(e, "engine build upload");
    // persistent grid: blocks per CU from the occupancy of the heaviest variant
    int occ = 1 << 30;
    for (int fr = 0; fr < 3; fr++)
        for (int adv = 0; adv < 2; adv++) {
            int nb = 0;
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, sweep_kernel(fr, kmax_, adv), 64 * kBW,
                                                             adv ? lds_advect_ : lds_fluidx_);
            if (e != hipSuccess) return hip_fail(e, "occupancy");
            occ = std::min(occ, std::max(nb, 1));
        }
    const int bpc = bpc_ > 0 ? std::min(bpc_, occ) : occ;
    grid_ = std::max(1, bpc * ncu_);
    e = hipStreamSynchronize(st_);
    if (e != hipSuccess) return hip_fail(e, "engine build");
    dirty_ = false;
    return MHD_OK;
}

int Engine::enqueue_step()
{
    if (policy_ == 3) {
        // fork: each group's chain (CFL, dt, 12 sweeps) on a side stream;
        // groups beyond the side streams queue behind others; join at the end
        hipError_t e = hipEventRecord(evfork_, st_);
        const int ng = (int)groups_.size(), ns = std::min(ng, kAux);
        for (int i = 0; i < ns && e == hipSuccess; i++) e = hipStreamWaitEvent(aux_[i], evfork_, 0);
        for (int g = 0; g < ng && e == hipSuccess; g++) {
            const BoxGroup &G = groups_[g];
            hipStream_t S = aux_[g % kAux];
            // the CFL of the state the step starts from is already in the
            // accumulator: the previous step's last sweep put it there (or the
            // standalone CFL run by step() after the state was set)
            if (e == hipSuccess) {
                hipLaunchKernelGGL(batch_dt_kernel, dim3(1), dim3(256), 0, S, d_boxes_, avg_, d_active_ + G.act_off, G.act_n, d_cmax_,
                                   d_dt_, d_t_, d_t64_, d_tf_, d_dts_, cap_, d_iter_, d_ctr_, 0);
                e = hipGetLastError();
            }
            for (int sw = 0; sw < 12 && e == hipSuccess; sw++) {
                const int fr = kSweepFrameB[sw];
                const bool adv = kSweepOpB[sw] == 1;
                const int flip = flip_before(sw);
                if (adv && fr > 0) {
                    // y/z frames: tiles for the small boxes, the streamed passes for the others
                    if (G.tile_adv_n[fr] > 0)
                        e = launch_group(fr, G.K[fr], true, G.tile_adv_n[fr], G.lds_a[fr], S, d_boxes_,
                                         d_items_[fr] + G.item_off[fr], d_dt_, flip);
                    if (e != hipSuccess || G.sa_n[fr] == 0) continue;
                    if (G.sab_nw[fr]) {
                        const dim3 grid(G.sa_n[fr]), block(64 * G.sab_nw[fr]);
                        const WorkItem *it = d_sitems_[0] + G.sa_off[fr];
#define MHD_SAB(FR, NW) hipLaunchKernelGGL((batch_stream_ab_kernel<FR, NW>), grid, block, 0, S, d_boxes_, it, d_dt_, flip)
                        if (G.sab_nw[fr] <= 4) {
                            if (fr == 1) MHD_SAB(1, 4); else MHD_SAB(2, 4);
                        } else {
                            if (fr == 1) MHD_SAB(1, 16); else MHD_SAB(2, 16);
                        }
#undef MHD_SAB
                    } else if (fr == 1) {
                        hipLaunchKernelGGL(batch_stream_a_kernel<1>, dim3(G.sa_n[1]), dim3(256), 0, S, d_boxes_,
                                           d_sitems_[0] + G.sa_off[1], d_dt_, flip);
                        hipLaunchKernelGGL(batch_stream_b_kernel<1>, dim3(G.sb_n[1]), dim3(256), 0, S, d_boxes_,
                                           d_sitems_[1] + G.sb_off[1], d_dt_, flip);
                    } else {
                        hipLaunchKernelGGL(batch_stream_a_kernel<2>, dim3(G.sa_n[2]), dim3(256), 0, S, d_boxes_,
                                           d_sitems_[0] + G.sa_off[2], d_dt_, flip);
                        hipLaunchKernelGGL(batch_stream_b_kernel<2>, dim3(G.sb_n[2]), dim3(256), 0, S, d_boxes_,
                                           d_sitems_[1] + G.sb_off[2], d_dt_, flip);
                    }
                    e = hipGetLastError();
                } else if (adv && fr == 0 && G.x2_n > 0) {
                    const void *k = advect_x2_kernel(G.K[0]);
                    const WorkItem *it = d_f8items_ + G.x2_off;
                    void *args[] = {(void *)&d_boxes_, (void *)&it, (void *)&d_dt_, (void *)&flip};
                    e = k ? hipLaunchKernel(k, dim3(G.x2_n), dim3(64 * kBW), args, G.lds_x2, S) : hipErrorInvalidValue;
                } else if (sw == 11) {
                    // the last sweep (fluidx, x frame) also forms the next step's CFL
                    const void *k = fluidx_x_cfl_kernel(G.KF[0], G.PF[0], avg_ ? 2 : 1);
                    const WorkItem *it = d_items_[0] + G.item_off[0];
                    void *args[] = {(void *)&d_boxes_, (void *)&it, (void *)&d_dt_, (void *)&flip, (void *)&d_cmax_};
                    e = k ? hipLaunchKernel(k, dim3(G.item_n[0]), dim3(64 * kBW / G.PF[0]), args, G.lds_f[0], S)
                          : hipErrorInvalidValue;
                } else if (!adv && fr > 0 && G.f8_n[fr] > 0) {
                    const void *k = fluidx_yz_kernel(fr, G.KF[fr], G.PF[fr]);
                    const WorkItem *it = d_f8items_ + G.f8_off[fr];
                    void *args[] = {(void *)&d_boxes_, (void *)&it, (void *)&d_dt_, (void *)&flip};
                    e = k ? hipLaunchKernel(k, dim3(G.f8_n[fr]), dim3(64 * 2 * kBW / G.PF[fr]), args, G.lds_f8[fr], S)
                          : hipErrorInvalidValue;
                } else if (!adv && fr == 0 && G.PF[0] > 1) {
                    const void *k = fluidx_x_p_kernel(G.KF[0], G.PF[0]);
                    const WorkItem *it = d_items_[0] + G.item_off[0];
                    void *args[] = {(void *)&d_boxes_, (void *)&it, (void *)&d_dt_, (void *)&flip};
                    e = k ? hipLaunchKernel(k, dim3(G.item_n[0]), dim3(64 * kBW / G.PF[0]), args, 0, S)
                          : hipErrorInvalidValue;
                } else {
                    e = launch_group(fr, G.K[fr], adv, G.item_n[fr], adv ? G.lds_a[fr] : G.lds_f[fr], S, d_boxes_,
                                     d_items_[fr] + G.item_off[fr], d_dt_, flip);
                }
            }
        }
        for (int i = 0; i < ns && e == hipSuccess; i++) {
            e = hipEventRecord(evjoin_[i], aux_[i]);
            if (e == hipSuccess) e = hipStreamWaitEvent(st_, evjoin_[i], 0);
        }
        return e == hipSuccess ? MHD_OK : hip_fail(e, "engine step (grouped)");
    }
    if (avg_)
        hipLaunchKernelGGL((batch_cfl_kernel<true>), dim3(ncfl_), dim3(256), 0, st_, d_boxes_, d_cfl_, d_cmax_);
    else
        hipLaunchKernelGGL((batch_cfl_kernel<false>), dim3(ncfl_), dim3(256), 0, st_, d_boxes_, d_cfl_, d_cmax_);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(batch_dt_kernel, dim3(1), dim3(256), 0, st_, d_boxes_, avg_, d_active_, (int)active_.size(), d_cmax_, d_dt_,
                           d_t_, d_t64_, d_tf_, d_dts_, cap_, d_iter_, d_ctr_, 12 * kQueues);
        e = hipGetLastError();
    }
    const int nq = policy_ == 0 ? kQueues : 1;
    for (int s = 0; s < 12 && e == hipSuccess; s++) {
        const int fr = kSweepFrameB[s];
        const bool adv = kSweepOpB[s] == 1;
        const size_t lds = adv ? lds_advect_ : lds_fluidx_;
        if (policy_ == 2)
            e = launch_grid(fr, kmax_, adv, nitems_[fr], lds, st_, d_boxes_, d_items_[fr], d_dt_, flip_before(s));
        else
            e = launch_sweep(fr, kmax_, adv, grid_, lds, st_, d_boxes_, d_items_[fr], d_part_ + fr * (kQueues + 1),
                             d_ctr_ + s * kQueues, d_dt_, flip_before(s), nq);
    }
    return e == hipSuccess ? MHD_OK : hip_fail(e, "engine step");
}

void Engine::drop_graph()
{
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
}

int Engine::capture()
{
    hipError_t e = hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    const int rc = enqueue_step();
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(st_, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(e, "hipGraphInstantiate");
    }
    graph_ = g;
    exec_ = x;
    return MHD_OK;
}

int Engine::prepare()
{
    if (active_.empty()) return MHD_OK;
    if (dirty_) {
        int rc = build();
        if (rc) return rc;
    }
    return exec_ ? MHD_OK : capture();
}

// Schedule 3: the standalone CFL, on the compute stream before the step's
// replay, for the active slots whose accumulator is stale (the accumulator is
// cleared first; outside any captured graph)
int Engine::prime_cfl()
{
    if (policy_ != 3) return MHD_OK;
    hipError_t e = hipSuccess;
    for (int s : active_) {
        if (cfl_fresh_[s] || e != hipSuccess) continue;
        e = hipMemsetAsync(d_cmax_ + s, 0, 8, st_);
        if (e == hipSuccess && cfl_slot_n_[s] > 0) {
            if (avg_)
                hipLaunchKernelGGL((batch_cfl_kernel<true>), dim3(cfl_slot_n_[s]), dim3(256), 0, st_, d_boxes_,
                                   d_cfl_ + cfl_slot_off_[s], d_cmax_);
            else
                hipLaunchKernelGGL((batch_cfl_kernel<false>), dim3(cfl_slot_n_[s]), dim3(256), 0, st_, d_boxes_,
                                   d_cfl_ + cfl_slot_off_[s], d_cmax_);
            e = hipGetLastError();
        }
        if (e == hipSuccess) cfl_fresh_[s] = 1;
    }
    return e == hipSuccess ? MHD_OK : hip_fail(e, "engine CFL");
}

int Engine::step(int nsteps)
{
    if (nsteps < 0) return fail(MHD_ERR_ARG, "step: nsteps < 0");
    if (active_.empty() || nsteps == 0) return MHD_OK;
    int rc = prepare();
    if (rc) return rc;
    rc = prime_cfl();
    if (rc) return rc;
    for (int i = 0; i < nsteps; i++) {
        hipError_t e = hipGraphLaunch(exec_, st_);
        if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    }
    return MHD_OK;
}

int Engine::step_timed(int nsteps, float *ms)
{
    if (!ms) return fail(MHD_ERR_ARG, "step_timed: null ms");
    if (dirty_) {
        int rc = build();
        if (rc) return rc;
    }
    if (!exec_ && !active_.empty()) {
        int rc = capture();
        if (rc) return rc;
    }
    hipError_t e = hipEventRecord(ev0_, st_);
    int rc = e == hipSuccess ? step(nsteps) : hip_fail(e, "hipEventRecord");
    if (rc) return rc;
    e = hipEventRecord(ev1_, st_);
    if (e == hipSuccess) e = hipEventSynchronize(ev1_);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, ev0_, ev1_);
    return e == hipSuccess ? MHD_OK : hip_fail(e, "step_timed");
}

int Engine::sync()
{
    hipError_t e = hipStreamSynchronize(st_);
    return e == hipSuccess ? MHD_OK : hip_fail(e, "engine sync");
}

// device copies of every active box's state and of the per-slot clocks
int Engine::save_all(std::vector<float *> &saved)
{
    hipError_t e = hipStreamSynchronize(st_);
    const size_t B = max_boxes_;
    for (int s : active_) {
        const size_t n = 8 * (size_t)boxes_[s].g.N;
        float *p = nullptr;
        if (e == hipSuccess) e = hipMalloc((void **)&p, n * 4);
        if (e == hipSuccess) {
            saved.push_back(p);
            e = hipMemcpyAsync(p, boxes_[s].u, 5 * (size_t)boxes_[s].g.N * 4, hipMemcpyDeviceToDevice, st_);
            if (e == hipSuccess)
                e = hipMemcpyAsync(p + 5 * (size_t)boxes_[s].g.N, boxes_[s].b, 3 * (size_t)boxes_[s].g.N * 4,
                                   hipMemcpyDeviceToDevice, st_);
        }
    }
    float *clk = nullptr;
    if (e == hipSuccess) e = hipMalloc((void **)&clk, B * (4 + 4 + 8 + 4));
    if (e == hipSuccess) {
        saved.push_back(clk);
        char *c = (char *)clk;
        e = hipMemcpyAsync(c, d_t_, B * 4, hipMemcpyDeviceToDevice, st_);
        if (e == hipSuccess) e = hipMemcpyAsync(c + B * 4, d_dt_, B * 4, hipMemcpyDeviceToDevice, st_);
        if (e == hipSuccess) e = hipMemcpyAsync(c + B * 8, d_t64_, B * 8, hipMemcpyDeviceToDevice, st_);
        if (e == hipSuccess) e = hipMemcpyAsync(c + B * 16, d_iter_, B * 4, hipMemcpyDeviceToDevice, st_);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st_);
    return e == hipSuccess ? MHD_OK : hip_fail(e, "save state");
}

int Engine::restore_all(std::vector<float *> &saved)
{
    hipError_t e = hipStreamSynchronize(st_);
    const size_t B = max_boxes_;
    size_t i = 0;
    for (int s : active_) {
        if (i >= saved.size()) break;
        float *p = saved[i++];
        if (e == hipSuccess)
            e = hipMemcpyAsync(boxes_[s].u, p, 5 * (size_t)boxes_[s].g.N * 4, hipMemcpyDeviceToDevice, st_);
        if (e == hipSuccess)
            e = hipMemcpyAsync(boxes_[s].b, p + 5 * (size_t)boxes_[s].g.N, 3 * (size_t)boxes_[s].g.N * 4,
                               hipMemcpyDeviceToDevice, st_);
    }
    if (i < saved.size() && e == hipSuccess) {
        char *c = (char *)saved[i];
        e = hipMemcpyAsync(d_t_, c, B * 4, hipMemcpyDeviceToDevice, st_);
        if (e == hipSuccess) e = hipMemcpyAsync(d_dt_, c + B * 4, B * 4, hipMemcpyDeviceToDevice, st_);
        if (e == hipSuccess) e = hipMemcpyAsync(d_t64_, c + B * 8, B * 8, hipMemcpyDeviceToDevice, st_);
        if (e == hipSuccess) e = hipMemcpyAsync(d_iter_, c + B * 16, B * 4, hipMemcpyDeviceToDevice, st_);
    }
    if (e == hipSuccess) e = hipMemsetAsync(d_cmax_, 0, B * 8, st_);
    if (e == hipSuccess) e = hipStreamSynchronize(st_);
    for (auto &f : cfl_fresh_) f = 0;
    for (float *p : saved) (void)hipFree(p);
    saved.clear();
    return e == hipSuccess ? MHD_OK : hip_fail(e, "restore state");
}

// Times candidate schedules on the batch itself (trial_steps each, after one
// warm-up step), then restores every box's state and clock and keeps the
// fastest.  Candidates: the persistent sweeps with {1, 2, 3, occupancy}
// blocks per CU under both queue policies, one block per work item, and the
// K-grouped concurrent chains.
int Engine::autotune(int trial_steps, int *bpc_out, int *policy_out, float *best_ms)
{
    if (active_.empty()) return fail(MHD_ERR_STATE, "autotune: no active boxes");
    if (trial_steps < 1) trial_steps = 1;
    std::vector<float *> saved;
    int rc = save_all(saved);
    if (rc) return rc;
    float best = INFINITY;
    int bb = 0, bp = 0;
    for (int policy = 0; policy < 4 && !rc; policy++)
        for (int bpc : {1, 2, 3, 0}) {
            if (policy >= 2 && bpc != 0) continue;
            rc = set_schedule(bpc, policy);
            if (!rc) rc = step(1);
            float ms = 0;
            if (!rc) rc = step_timed(trial_steps, &ms);
            if (rc) break;
            if (ms < best) {
                best = ms;
                bb = bpc;
                bp = policy;
            }
        }
    const int rc2 = restore_all(saved);
    if (rc) return rc;
    if (rc2) return rc2;
    rc = set_schedule(bb, bp);
    if (bpc_out) *bpc_out = bb;
    if (policy_out) *policy_out = bp;
    if (best_ms) *best_ms = best / trial_steps;
    return rc;
}

}  // namespace mhd

// ===========================================================================
// C ABI: mhd_batch_*
// ===========================================================================
using namespace mhd;

struct mhd_batch {
    Engine *eng;
    std::vector<int> slot;
};

extern "C" {

int mhd_batch_create(mhd_batch **out, int nbox, const int *dims, int avg_b)
{
    if (!out || !dims || nbox < 1) return fail(MHD_ERR_ARG, "mhd_batch_create: bad argument");
    *out = nullptr;
    for (int i = 0; i < nbox; i++) {
        int rc = check_shape_fast3(dims[3 * i], dims[3 * i + 1], dims[3 * i + 2]);
        if (rc) return rc;
    }
    mhd_batch *bt = new mhd_batch();
    bt->eng = new Engine(nbox, 1 << 14, avg_b);
    int rc = bt->eng->init();
    for (int i = 0; i < nbox && !rc; i++) {
        int s = -1;
        rc = bt->eng->add_box(dims[3 * i], dims[3 * i + 1], dims[3 * i + 2], &s);
        bt->slot.push_back(s);
    }
    if (!rc) rc = bt->eng->set_active(bt->slot);
    if (rc) {
        delete bt->eng;
        delete bt;
        return rc;
    }
    *out = bt;
    return MHD_OK;
}

int mhd_batch_destroy(mhd_batch *bt)
{
    if (!bt) return MHD_OK;
    delete bt->eng;
    delete bt;
    return MHD_OK;
}

static int bslot(mhd_batch *bt, int box)
{
    if (!bt || box < 0 || box >= (int)bt->slot.size()) return -1;
    return bt->slot[box];
}

int mhd_batch_set_state(mhd_batch *bt, int box, const float *h_u, const float *h_b)
{
    const int s = bslot(bt, box);
    if (s < 0) return fail(MHD_ERR_ARG, "mhd_batch_set_state: bad box");
    return bt->eng->set_state(s, h_u, h_b);
}

int mhd_batch_get_state(mhd_batch *bt, int box, float *h_u, float *h_b)
{
    const int s = bslot(bt, box);
    if (s < 0) return fail(MHD_ERR_ARG, "mhd_batch_get_state: bad box");
    return bt->eng->get_state(s, h_u, h_b);
}

int mhd_batch_step(mhd_batch *bt, int nsteps)
{
    if (!bt) return fail(MHD_ERR_ARG, "mhd_batch_step: null batch");
    return bt->eng->step(nsteps);
}

int mhd_batch_step_timed(mhd_batch *bt, int nsteps, float *ms)
{
    if (!bt) return fail(MHD_ERR_ARG, "mhd_batch_step_timed: null batch");
    return bt->eng->step_timed(nsteps, ms);
}

int mhd_batch_sync(mhd_batch *bt)
{
    if (!bt) return fail(MHD_ERR_ARG, "mhd_batch_sync: null batch");
    return bt->eng->sync();
}

int mhd_batch_dts(mhd_batch *bt, int box, double *h_dts, int max, int *count)
{
    const int s = bslot(bt, box);
    if (s < 0) return fail(MHD_ERR_ARG, "mhd_batch_dts: bad box");
    return bt->eng->dts(s, h_dts, max, count);
}

int mhd_batch_set_schedule(mhd_batch *bt, int blocks_per_cu, int policy)
{
    if (!bt) return fail(MHD_ERR_ARG, "mhd_batch_set_schedule: null batch");
    return bt->eng->set_schedule(blocks_per_cu, policy);
}

int mhd_batch_autotune(mhd_batch *bt, int trial_steps, int *blocks_per_cu, int *policy, float *ms_per_step)
{
    if (!bt) return fail(MHD_ERR_ARG, "mhd_batch_autotune: null batch");
    return bt->eng->autotune(trial_steps, blocks_per_cu, policy, ms_per_step);
}

}  // extern "C"
