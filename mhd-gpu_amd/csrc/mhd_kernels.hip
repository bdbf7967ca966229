// gfx950 kernels of the relaxing-TVD MHD timestep.
//
// Work decomposition (all sweeps): one workgroup owns a tile of W pencils
// that are neighbours along j (the next pencil along j is the fluid sweep's
// "jp" and the CT sweep's "j+1"), one wavefront per pencil, each lane holding
// K consecutive cells of the pencil in registers (n = L*K, L <= 64 lanes).
// The tile is staged through LDS with coalesced loads: for the x frame a
// pencil is a contiguous run of n cells, for the y and z frames (sj == 1) each
// row i of the tile is a contiguous run of W cells.  LDS groups of K cells are
// padded to an odd word count so that lane-strided reads are conflict free.
// Neighbours along the pencil come from lane shuffles; the pencil-global max
// of the freezing speed is a wave reduction.
#include <stdint.h>

#include <initializer_list>

#include "mhd_kernels.h"
#include "mhd_aos_x.h"
#include "mhd_core.h"
#include "mhd_generic.h"

namespace mhd {

// padded LDS offset of cell i, component c, in a pencil of C-component cells
template <int K, int C> __device__ __forceinline__ int pidx(int i, int c)
{
    return (i / K) * (C * K + 1) + (i % K) * C + c;
}
template <int K, int C> __host__ __device__ constexpr int plen(int n) { return (n / K) * (C * K + 1); }

// j >= 0 reduced mod nj (j < nj + W + 2 in every caller, so the loop runs at
// most a few times even for boxes narrower than a tile)
__device__ __forceinline__ int wrapj(int j, int nj)
{
    while (j >= nj) j -= nj;
    return j;
}

// Load all C components of nslot pencils (j = j0+s mod nj, plane k) into LDS
// slot s (pitch P), physical component p stored as frame component inv[p].
template <typename T, int K, int C, int S, bool XMAJ, class Inv>
__device__ __forceinline__ void load_full(T *lds, int P, const T *__restrict__ g, const Frame &f, int j0,
                                          int nslot, int k)
{
    const int n = f.ni;
    const int tid = threadIdx.x, nt = blockDim.x;
    if constexpr (XMAJ) {
        for (int s = 0; s < nslot; s++) {
            const long base = ((long)wrapj(j0 + s, f.nj) * f.sj + (long)k * f.sk) * C;
            for (int e = tid; e < n * C; e += nt) {
                const int i = e / C, p = e - i * C;
                lds[s * P + pidx<K, C>(i, Inv::at(p))] = g[base + e];
            }
        }
    } else {
        const long kb = (long)k * f.sk;
        for (int e = tid; e < n * S * C; e += nt) {
            const int i = e / (S * C), r = e - i * (S * C);
            const int s = r / C, p = r - s * C;
            if (s < nslot) {
                const long cell = kb + (long)i * f.si + wrapj(j0 + s, f.nj);
                lds[s * P + pidx<K, C>(i, Inv::at(p))] = g[cell * C + p];
            }
        }
    }
}

// Store frame components back (inverse of load_full), nslot pencils.
template <typename T, int K, int C, int S, bool XMAJ, class Inv>
__device__ __forceinline__ void store_full(const T *lds, int P, T *__restrict__ g, const Frame &f, int j0,
                                           int nslot, int k)
{
    const int n = f.ni;
    const int tid = threadIdx.x, nt = blockDim.x;
    if constexpr (XMAJ) {
        for (int s = 0; s < nslot; s++) {
            const long base = ((long)(j0 + s) * f.sj + (long)k * f.sk) * C;
            for (int e = tid; e < n * C; e += nt) {
                const int i = e / C, p = e - i * C;
                g[base + e] = lds[s * P + pidx<K, C>(i, Inv::at(p))];
            }
        }
    } else {
        const long kb = (long)k * f.sk;
        for (int e = tid; e < n * S * C; e += nt) {
            const int i = e / (S * C), r = e - i * (S * C);
            const int s = r / C, p = r - s * C;
            if (s < nslot) {
                const long cell = kb + (long)i * f.si + (j0 + s);
                g[cell * C + p] = lds[s * P + pidx<K, C>(i, Inv::at(p))];
            }
        }
    }
}

// Load one physical component p (of C) of nslot pencils starting at pencil
// j0 (mod nj) in plane k, single-word cells, pitch P.
template <typename T, int K, int C, int S, bool XMAJ>
__device__ __forceinline__ void load_comp(T *lds, int P, const T *__restrict__ g, int p, const Frame &f, int j0,
                                          int nslot, int k)
{
    const int n = f.ni;
    const int tid = threadIdx.x, nt = blockDim.x;
    const long kb = (long)k * f.sk;
    if constexpr (XMAJ) {
        for (int s = 0; s < nslot; s++) {
            const long base = (long)wrapj(j0 + s, f.nj) * f.sj + kb;
            for (int i = tid; i < n; i += nt) lds[s * P + pidx<K, 1>(i, 0)] = g[(base + i) * C + p];
        }
    } else {
        for (int e = tid; e < n * S; e += nt) {
            const int i = e / S, s = e - i * S;
            if (s < nslot) {
                const long cell = kb + (long)i * f.si + wrapj(j0 + s, f.nj);
                lds[s * P + pidx<K, 1>(i, 0)] = g[cell * C + p];
            }
        }
    }
}

// physical component p -> frame component, as selects (p varies per lane)
template <int FR> struct InvU {
    static __device__ __forceinline__ int at(int p)
    {
        return p == 0 ? CompInv<FR>::u[0]
                      : p == 1 ? CompInv<FR>::u[1]
                               : p == 2 ? CompInv<FR>::u[2] : p == 3 ? CompInv<FR>::u[3] : CompInv<FR>::u[4];
    }
};
template <int FR> struct InvB {
    static __device__ __forceinline__ int at(int p)
    {
        return p == 0 ? CompInv<FR>::b[0] : p == 1 ? CompInv<FR>::b[1] : CompInv<FR>::b[2];
    }
};

// ---------------------------------------------------------------------------
// fluidx: Heterogeneous_MHD/C/fluidx.c:19-102 + tvd1.c:8-118 + mhdflux.c.
// ---------------------------------------------------------------------------
template <typename T, int K, int W> __host__ __device__ constexpr int fluidx_lds_words(int n)
{
    return W * plen<K, 5>(n) + (W + 1) * plen<K, 3>(n) + W * plen<K, 1>(n);
}

template <typename T, int K, int W, int FR>
__global__ void __launch_bounds__(64 * W) fluidx_kernel(T *__restrict__ u, const T *__restrict__ b, Frame f,
                                                        const T *__restrict__ dtp, int ntj)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *smem = reinterpret_cast<T *>(smem_raw);
    constexpr bool XM = (FR == 0);
    const int n = f.ni, L = n / K;
    const int PU = plen<K, 5>(n), PB = plen<K, 3>(n), PK = plen<K, 1>(n);
    T *su = smem, *sb = su + W * PU, *sk = sb + (W + 1) * PB;

    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int k = tile / ntj, j0 = (tile - k * ntj) * W;
    const int nvalid = min(W, f.nj - j0);
    const int kp = (k + 1 == f.nk) ? 0 : k + 1;

    load_full<T, K, 5, W, XM, InvU<FR>>(su, PU, u, f, j0, nvalid, k);
    load_full<T, K, 3, W + 1, XM, InvB<FR>>(sb, PB, b, f, j0, nvalid + 1, k);
    load_comp<T, K, 3, W, XM>(sk, PK, b, Comp<FR>::b[2], f, j0, nvalid, kp);
    __syncthreads();

    const int w = threadIdx.x >> 6;
    if (w < nvalid) {
        const int l = lane_id();
        const int lc = l % L;  // lanes >= L shadow lane l % L and store nothing
        const int src_next = lc + 1 == L ? 0 : lc + 1, src_prev = lc == 0 ? L - 1 : lc - 1;
        const T dt = *dtp;
        T uu[K][5], bx[K][3];
        const T *pu = su + w * PU + lc * (5 * K + 1);
#pragma unroll
        for (int q = 0; q < K; q++)
#pragma unroll
            for (int m = 0; m < 5; m++) uu[q][m] = pu[q * 5 + m];
        // face B to cell centre (fluidx.c:71-81)
#pragma unroll
        for (int q = 0; q < K; q++) {
            const int i = lc * K + q, ip = i + 1 == n ? 0 : i + 1;
            const T b1 = sb[w * PB + pidx<K, 3>(i, 0)], b1p = sb[w * PB + pidx<K, 3>(ip, 0)];
            const T b2 = sb[w * PB + pidx<K, 3>(i, 1)], b2p = sb[(w + 1) * PB + pidx<K, 3>(i, 1)];
            const T b3 = sb[w * PB + pidx<K, 3>(i, 2)], b3p = sk[w * PK + pidx<K, 1>(i, 0)];
            bx[q][0] = (T)0.5 * b1 + (T)0.5 * b1p;
            bx[q][1] = (T)0.5 * b2 + b2p / (T)2;
            bx[q][2] = (T)0.5 * b3 + b3p / (T)2;
        }
        tvd1_pencil<T, K>(uu, bx, dt, src_next, src_prev);
        if (l < L) {
            T *po = su + w * PU + lc * (5 * K + 1);
#pragma unroll
            for (int q = 0; q < K; q++)
#pragma unroll
                for (int m = 0; m < 5; m++) po[q * 5 + m] = uu[q][m];
        }
    }
    __syncthreads();
    store_full<T, K, 5, W, XM, InvU<FR>>(su, PU, u, f, j0, nvalid, k);
}

// ---------------------------------------------------------------------------
// advectbyzx: Heterogeneous_MHD/C/advectbyzx.c:14-23 with A1/A2/B1/B2 fused.
// Per tile the fluxes FA of pencils j0..j0+nvalid (neighbour j+1 needed by
// b1), FB at k and at k+1 are computed (tvdb each), then
// b1 = (((b1 - FA_j) + FA_{j+1}) - FB_k) + FB_{k+1}   (A1.c:55, A2.c:23, B1.c:54, B2.c:23).
// ---------------------------------------------------------------------------
template <typename T, int K, int W> __host__ __device__ constexpr int advect_lds_words(int n)
{
    // rho & momentum: (W+2) slots at k, W at k-1, W at k+1; b: W+1 slots (3 comps);
    // b3 at k+1: W slots; fluxes FA (W+1), FB0 (W), FB1 (W)
    return 2 * (3 * W + 2) * plen<K, 1>(n) + (W + 1) * plen<K, 3>(n) + W * plen<K, 1>(n) +
           (3 * W + 1) * plen<K, 1>(n);
}

// b is read from bin and written to bout (distinct buffers): a tile reads the
// old b2 of the next pencil along j and the old b3 of the next plane, which
// other tiles update, exactly as the reference reads its update_b copy
// (advectbyzxA1.c:43, B1.c:43) while writing b.
template <typename T, int K, int W, int FR>
__global__ void __launch_bounds__(64 * W) advect_kernel(const T *__restrict__ u, const T *__restrict__ bin,
                                                        T *__restrict__ bout, Frame f, const T *__restrict__ dtp,
                                                        int ntj)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *smem = reinterpret_cast<T *>(smem_raw);
    constexpr bool XM = (FR == 0);
    const int n = f.ni, L = n / K;
    const int PK = plen<K, 1>(n), PB = plen<K, 3>(n);
    // slot layout of the rho/momentum rows: 0..W+1 at k (pencil j0-1+s),
    // W+2..2W+1 at k-1 (pencil j0+s), 2W+2..3W+1 at k+1 (pencil j0+s)
    T *sr = smem, *sm = sr + (3 * W + 2) * PK;
    T *sb = sm + (3 * W + 2) * PK, *sb3p = sb + (W + 1) * PB;
    T *sF = sb3p + W * PK;  // FA: slots 0..W, FB0: W+1..2W, FB1: 2W+1..3W

    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int k = tile / ntj, j0 = (tile - k * ntj) * W;
    const int nvalid = min(W, f.nj - j0);
    const int kp = (k + 1 == f.nk) ? 0 : k + 1, km = k == 0 ? f.nk - 1 : k - 1;
    constexpr int pm = Comp<FR>::u[1];

    load_comp<T, K, 5, W + 2, XM>(sr, PK, u, 0, f, j0 - 1 + f.nj, nvalid + 2, k);
    load_comp<T, K, 5, W + 2, XM>(sm, PK, u, pm, f, j0 - 1 + f.nj, nvalid + 2, k);
    load_comp<T, K, 5, W, XM>(sr + (W + 2) * PK, PK, u, 0, f, j0, nvalid, km);
    load_comp<T, K, 5, W, XM>(sm + (W + 2) * PK, PK, u, pm, f, j0, nvalid, km);
    load_comp<T, K, 5, W, XM>(sr + (2 * W + 2) * PK, PK, u, 0, f, j0, nvalid, kp);
    load_comp<T, K, 5, W, XM>(sm + (2 * W + 2) * PK, PK, u, pm, f, j0, nvalid, kp);
    load_full<T, K, 3, W + 1, XM, InvB<FR>>(sb, PB, bin, f, j0, nvalid + 1, k);
    load_comp<T, K, 3, W, XM>(sb3p, PK, bin, Comp<FR>::b[2], f, j0, nvalid, kp);
    __syncthreads();

    const int w = threadIdx.x >> 6;
    const int l = lane_id();
    const int lc = l % L;
    const int src_next = lc + 1 == L ? 0 : lc + 1, src_prev = lc == 0 ? L - 1 : lc - 1;
    const T dt = *dtp;
    const int nitem = 3 * nvalid + 1;
    for (int it = w; it < nitem; it += W) {
        // item -> (rows of the two velocity pencils, field source, outputs)
        int ra, rb, fslot, fkind;  // fkind 0: b2 of sb slot, 1: b3 of sb slot, 2: b3 at k+1
        int Fslot;
        if (it <= nvalid) {  // FA(s): pencils j-1 (row s) and j (row s+1) at k, field b2
            ra = it;
            rb = it + 1;
            fslot = it;
            fkind = 0;
            Fslot = it;
        } else if (it <= 2 * nvalid) {  // FB0(s): pencil (s,k-1) and (s,k), field b3
            const int s = it - nvalid - 1;
            ra = W + 2 + s;
            rb = s + 1;
            fslot = s;
            fkind = 1;
            Fslot = W + 1 + s;
        } else {  // FB1(s): pencil (s,k) and (s,k+1), field b3 at k+1
            const int s = it - 2 * nvalid - 1;
            ra = s + 1;
            rb = 2 * W + 2 + s;
            fslot = s;
            fkind = 2;
            Fslot = 2 * W + 1 + s;
        }
        T vx[K], vxm[K], vxn[K], vs[K], bin[K], bnew[K], F[K];
#pragma unroll
        for (int q = 0; q < K; q++) {
            const int o = pidx<K, 1>(lc * K + q, 0);
            // A1.c:34 (u2(jm)+u2(j))/(u1(jm)+u1(j)); the minus-side pencil first
            vx[q] = (sm[ra * PK + o] + sm[rb * PK + o]) / (sr[ra * PK + o] + sr[rb * PK + o]);
        }
        shift_prev<T, K>(vx, vxm, src_prev);
        shift_next<T, K>(vx, vxn, src_next);
#pragma unroll
        for (int q = 0; q < K; q++) {
            // A1.c:41 (vx[i-1] + vx[i+1] + 2.0*vx)/4.0
            vs[q] = avg121<T>(vxm[q] + vxn[q], vx[q]);
            const int i = lc * K + q;
            if (fkind == 0)
                bin[q] = sb[fslot * PB + pidx<K, 3>(i, 1)];
            else if (fkind == 1)
                bin[q] = sb[fslot * PB + pidx<K, 3>(i, 2)];
            else
                bin[q] = sb3p[fslot * PK + pidx<K, 1>(i, 0)];
        }
        tvdb_pencil<T, K>(bin, vs, dt, bnew, F, src_next, src_prev);
        if (l < L) {
#pragma unroll
            for (int q = 0; q < K; q++) {
                const int i = lc * K + q;
                sF[Fslot * PK + pidx<K, 1>(i, 0)] = F[q];
                // new b2 (FA) / b3 (FB0) of the tile's own pencils; the item only
                // reads its own field row, so it is overwritten in place
                if (fkind == 0 && fslot < nvalid) sb[fslot * PB + pidx<K, 3>(i, 1)] = bnew[q];
                if (fkind == 1) sb[fslot * PB + pidx<K, 3>(i, 2)] = bnew[q];
            }
        }
    }
    __syncthreads();
    if (w < nvalid && l < L) {
#pragma unroll
        for (int q = 0; q < K; q++) {
            const int i = lc * K + q;
            const int o = pidx<K, 1>(i, 0);
            T b1 = sb[w * PB + pidx<K, 3>(i, 0)];
            b1 = b1 - sF[w * PK + o];
            b1 = b1 + sF[(w + 1) * PK + o];
            b1 = b1 - sF[(W + 1 + w) * PK + o];
            b1 = b1 + sF[(2 * W + 1 + w) * PK + o];
            sb[w * PB + pidx<K, 3>(i, 0)] = b1;
        }
    }
    __syncthreads();
    store_full<T, K, 3, W, XM, InvB<FR>>(sb, PB, bout, f, j0, nvalid, k);
}

// ---------------------------------------------------------------------------
// cfl: Heterogeneous_MHD/C/calcfl.c:13-65 -> max over cells, as the
// NaN-sticky unsigned bits of speed_bits in *cmax (finite speeds are >= 0, so
// their unsigned order is the float order; any NaN outranks them).  AVG
// selects B^2 of the cell-averaged field that calcfl.c:38-43 forms, with the
// C path's arithmetic.
// ---------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ unsigned long long tobits(T x)
{
    if constexpr (sizeof(T) == 4)
        return (unsigned long long)__float_as_uint(x);
    else
        return (unsigned long long)__double_as_longlong(x);
}
template <typename T> __device__ __forceinline__ T frombits(unsigned long long x)
{
    if constexpr (sizeof(T) == 4)
        return __uint_as_float((unsigned)x);
    else
        return __longlong_as_double((long long)x);
}

template <typename T, bool AVG>
__global__ void __launch_bounds__(256) cfl_kernel(const T *__restrict__ u, const T *__restrict__ b, int nx, int ny,
                                                  int nz, unsigned long long *cmax)
{
    const long ncell = (long)nx * ny * nz;
    const AosSpeed<T, AVG> speed{u, b, nx, ny, nz};
    T c = (T)0;
    for (long s = (long)blockIdx.x * blockDim.x + threadIdx.x; s < ncell; s += (long)gridDim.x * blockDim.x)
        c = cfl_max(c, speed(s));  // max_2num(c, t), NaN-sticky
    c = wave_cfl_max(c);
    __shared__ T red[4];
    if (lane_id() == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        T m = red[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = cfl_max(m, red[i]);
        atomicMax(cmax, tobits(m));
    }
}

// The step's dt on the device (Heterogeneous_MHD/C/mhd.c:63-65) from the CFL
// max: cfl = 1/c in double rounded to T, dt = 0.9*cfl in double rounded to
// T, dt clamped to (tf-t)/2, t += 2*dt (in T; an FP64 copy of t is
// accumulated beside it), dt recorded in the history.
template <typename T> __device__ __forceinline__ void dt_update(T c, T *dt, T *t, double *t64, T tf, T *dts, int *iter,
                                                                int cap)
{
    const T cfl = (T)(1.0 / (double)c);
    T d = (T)(0.9 * (double)cfl);
    if (d > (tf - *t) / (T)2) d = (tf - *t) / (T)2;
    *t = *t + (T)2 * d;
    *t64 = *t64 + 2.0 * (double)d;
    *dt = d;
    const int it = *iter;
    if (it < cap) dts[it] = d;
    *iter = it + 1;
}

// The dt of a device-resident box.  The CFL max comes as block maxima
// (part[0..np)) of the previous step's fused last sweep, beside the
// accumulator, which a standalone CFL fills after the state was set from
// outside; both are NaN-sticky (cfl_max), and a NaN max is replaced by the
// reference's sequential one over the state (cfl_after_last_nan).  Clears
// the accumulator.  Deterministic: the max is order-independent.
template <typename T>
__global__ void __launch_bounds__(1024) dt_part_kernel(unsigned long long *cmax, const T *__restrict__ part, int np,
                                                       T *dt, T *t, double *t64, T tf, T *dts, int *iter, int cap,
                                                       const T *__restrict__ u, const T *__restrict__ b, SoaGrid g,
                                                       int avg)
{
    // four independent loads in flight per thread
    T c0 = (T)0, c1 = (T)0, c2 = (T)0, c3 = (T)0;
    const int nt = blockDim.x;
    int i = threadIdx.x;
    for (; i + 3 * nt < np; i += 4 * nt) {
        c0 = cfl_max(c0, part[i]);
        c1 = cfl_max(c1, part[i + nt]);
        c2 = cfl_max(c2, part[i + 2 * nt]);
        c3 = cfl_max(c3, part[i + 3 * nt]);
    }
    for (; i < np; i += nt) c0 = cfl_max(c0, part[i]);
    T c = wave_cfl_max(cfl_max(cfl_max(c0, c1), cfl_max(c2, c3)));
    __shared__ T red[16];
    if (lane_id() == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < nt / 64; w++) c = cfl_max(c, red[w]);
        red[0] = cfl_max(c, speed_from_bits<T>(*cmax));
    }
    __syncthreads();
    c = red[0];
    if (c != c) {
        const long ncell = (long)g.nx * g.ny * g.nz;
        c = avg ? cfl_after_last_nan<T>(ncell, SoaSpeed<T, true>{u, b, g})
                : cfl_after_last_nan<T>(ncell, SoaSpeed<T, false>{u, b, g});
    }
    if (threadIdx.x != 0) return;
    *cmax = 0ull;
    dt_update<T>(c, dt, t, t64, tf, dts, iter, cap);
}

template <typename T> __global__ void __launch_bounds__(256) fill_zero_kernel(T *p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = (T)0;
}

// cfl value 1/c (calcfl.c:61) for the step-level entry point; a NaN max is
// replaced by the reference's sequential one over the AoS state
template <typename T>
__global__ void __launch_bounds__(1024) cfl_value_kernel(unsigned long long *cmax, T *out, const T *__restrict__ u,
                                                         const T *__restrict__ b, int nx, int ny, int nz, int avg)
{
    T c = frombits<T>(*cmax);
    if (c != c) {
        const long ncell = (long)nx * ny * nz;
        c = avg ? cfl_after_last_nan<T>(ncell, AosSpeed<T, true>{u, b, nx, ny, nz})
                : cfl_after_last_nan<T>(ncell, AosSpeed<T, false>{u, b, nx, ny, nz});
    }
    if (threadIdx.x != 0) return;
    *cmax = 0ull;
    *out = (T)(1.0 / (double)c);
}

// plain streaming copy (16 B per lane per iteration), the calibration
// yardstick of the HBM byte counters
__global__ void __launch_bounds__(256) copy_kernel(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t n16)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st)
{
    const size_t n16 = bytes / 16;
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, st, (uint4 *)dst, (const uint4 *)src, n16);
    return hipGetLastError();
}

// n floats, any count and 4-byte alignment
__global__ void __launch_bounds__(256) copy_f32_kernel(float *__restrict__ dst, const float *__restrict__ src, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_copy_f32(float *dst, const float *src, long n, hipStream_t st)
{
    const long nb = (n + 255) / 256;
    hipLaunchKernelGGL(copy_f32_kernel, dim3(nb < 4096 ? (nb > 0 ? nb : 1) : 4096), dim3(256), 0, st, dst, src, n);
    return hipGetLastError();
}

// max |1 - D*rcp(D)| of v_rcp_f64 over every finite nonzero float-valued
// D, as FP64 bits (non-negative doubles order like their bit patterns); the
// bound the FP32 limiter quotient (limiter_quotient) rests on
__global__ void __launch_bounds__(256) rcp_scan_kernel(unsigned long long *maxbits)
{
    const unsigned long long nt = (unsigned long long)gridDim.x * blockDim.x;
    double m = 0.0;
    for (unsigned long long b = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; b < (1ull << 32); b += nt) {
        const float s = __uint_as_float((unsigned)b);
        if (!__builtin_isfinite(s) || s == 0.0f) continue;
        const double D = (double)s;
        const double e = fabs(fma(-D, __builtin_amdgcn_rcp(D), 1.0));
        m = e > m ? e : m;
    }
    atomicMax(maxbits, (unsigned long long)__double_as_longlong(m));
}

hipError_t launch_rcp_scan(unsigned long long *maxbits, hipStream_t st)
{
    hipLaunchKernelGGL(rcp_scan_kernel, dim3(4096), dim3(256), 0, st, maxbits);
    return hipGetLastError();
}

// The shared-divisor quotient quot_by(a, recip_of(d)) against the IEEE
// division a/d for d = 1 + (d0 + blockIdx.x*stride)*2^-23 and every a in
// [1, 2): block b takes one divisor, its threads all 2^23 dividends.
__global__ void __launch_bounds__(256) div_pair_scan_kernel(unsigned d0, unsigned stride,
                                                            unsigned long long *mismatches)
{
    const unsigned dm = d0 + blockIdx.x * stride;
    const float d = __uint_as_float(0x3f800000u | dm);
    const Recip r = recip_of(d);
    unsigned long long bad = 0;
    for (unsigned am = threadIdx.x; am < (1u << 23); am += 256) {
        const float a = __uint_as_float(0x3f800000u | am);
        if (__float_as_uint(quot_by(a, r)) != __float_as_uint(a / d)) bad++;
    }
    if (bad) atomicAdd(mismatches, bad);
}

// v_rcp_f32 of +-m*2^e, m in [1, 2), e in [-40, 40]: +-2^-e times its value
// at m, bit for bit (the scaling and the sign pass through exactly)
__global__ void __launch_bounds__(256) rcp_equivariance_kernel(unsigned long long *mismatches)
{
    const unsigned long long nt = (unsigned long long)gridDim.x * blockDim.x;
    unsigned long long bad = 0;
    for (unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; t < (81ull << 23);
         t += nt) {
        const int e = (int)(t >> 23) - 40;
        const unsigned m = (unsigned)(t & 0x7fffffu);
        const float base = __uint_as_float(0x3f800000u | m);
        const float d = __uint_as_float((unsigned)(127 + e) << 23 | m);
        const float yd = __builtin_amdgcn_rcpf(d);
        const float ys = ldexpf(__builtin_amdgcn_rcpf(base), -e);
        const float yn = __builtin_amdgcn_rcpf(-d);
        if (__float_as_uint(yd) != __float_as_uint(ys) || __float_as_uint(yn) != (__float_as_uint(yd) ^ 0x80000000u))
            bad++;
    }
    if (bad) atomicAdd(mismatches, bad);
}

hipError_t launch_div_scan(unsigned stride, unsigned long long *mismatches, hipStream_t st)
{
    // divisor mantissas 0, stride, 2*stride, ... < 2^23, at most 2^19 per
    // launch (about 2 s of work each), each launch finished before the next
    const unsigned nd = ((1u << 23) + stride - 1) / stride;
    const unsigned per = 1u << 19;
    for (unsigned b0 = 0; b0 < nd; b0 += per) {
        const unsigned nb = min(per, nd - b0);
        hipLaunchKernelGGL(div_pair_scan_kernel, dim3(nb), dim3(256), 0, st, b0 * stride, stride, mismatches);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(rcp_equivariance_kernel, dim3(4096), dim3(256), 0, st, mismatches + 1);
    return hipGetLastError();
}

// sqrt_float_valued against the IEEE double square root on every positive
// finite float x (bit patterns 1 .. 0x7f7fffff), grid-stride
__global__ void __launch_bounds__(256) sqrt_scan_kernel(unsigned long long *mismatches)
{
    const unsigned last = 0x7f7fffffu, step = gridDim.x * 256u;
    unsigned long long bad = 0;
    for (unsigned i = blockIdx.x * 256u + threadIdx.x + 1u; i <= last; i += step) {
        const double x = (double)__uint_as_float(i);
        bad += __double_as_longlong(sqrt_float_valued(x)) != __double_as_longlong(__builtin_sqrt(x));
        if (i > last - step) break;
    }
    if (bad) atomicAdd(mismatches, bad);
}

hipError_t launch_sqrt_scan(unsigned long long *mismatches, hipStream_t st)
{
    hipLaunchKernelGGL(sqrt_scan_kernel, dim3(8192), dim3(256), 0, st, mismatches);
    return hipGetLastError();
}

// Both FP32 forms of the van Leer limiter (tvd1.c:71-81, tvdb.c:72-76) on
// given operand pairs: the branching one of the tiles and the branch-free
// one of the streamed tvdb
__global__ void __launch_bounds__(256) limiter_eval_kernel(const float *__restrict__ p, const float *__restrict__ m,
                                                           float *__restrict__ out, float *__restrict__ out_nb, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        out[i] = vanleer<float>(p[i], m[i]);
        out_nb[i] = vanleer_nb<float>(p[i], m[i]);
    }
}

hipError_t launch_limiter_eval(const float *p, const float *m, float *out, float *out_nb, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(limiter_eval_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, m, out, out_nb, n);
    return hipGetLastError();
}

__global__ void zero_u64_kernel(unsigned long long *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ---------------------------------------------------------------------------
// Transposes of the step-level API (transpose12.c:7-47, transpose13.c:9-62):
// out cell (b + a*sao + c3*sco) = perm(in cell (a + b*sbi + c3*sci)), a
// contiguous in the input, b contiguous in the output.  32x32 cell tiles
// through LDS, both sides coalesced.
// ---------------------------------------------------------------------------
template <typename T, int C>
__global__ void __launch_bounds__(256) transpose_kernel(T *__restrict__ out, const T *__restrict__ in, int na, int nb,
                                                        long sbi, long sci, long sao, long sco, int p0, int p1,
                                                        int p2, int p3, int p4)
{
    __shared__ T tile[32][32 * C + 1];
    const int perm[5] = {p0, p1, p2, p3, p4};
    const int a0 = blockIdx.x * 32, b0 = blockIdx.y * 32, c3 = blockIdx.z;
    const int tid = threadIdx.x;
    for (int e = tid; e < 32 * 32 * C; e += 256) {
        const int bb = e / (32 * C), r = e - bb * (32 * C);
        const int aa = r / C, p = r - aa * C;
        if (a0 + aa < na && b0 + bb < nb)
            tile[bb][aa * C + p] = in[((long)(a0 + aa) + (long)(b0 + bb) * sbi + (long)c3 * sci) * C + p];
    }
    __syncthreads();
    for (int e = tid; e < 32 * 32 * C; e += 256) {
        const int aa = e / (32 * C), r = e - aa * (32 * C);
        const int bb = r / C, q = r - bb * C;
        if (a0 + aa < na && b0 + bb < nb)
            out[((long)(b0 + bb) + (long)(a0 + aa) * sao + (long)c3 * sco) * C + q] = tile[bb][aa * C + perm[q]];
    }
}

// ---------------------------------------------------------------------------
// The same permutation into two buffers at once, for square planes (na == nb,
// so sao == sbi and sco == sci): the transposed tile (A,B) lands exactly on
// the input positions of tile (B,A).  One block reads both tiles of a pair
// (A <= B) of plane c3 into LDS, then writes each transposed onto the other's
// positions, in io (the input, updated in place: the reference's copy back,
// HIP/cu/cuda_transpose12.cu:33-34) and in out.  No other block touches those
// positions.  VEC: whole 32-cell rows as 16-byte vectors (n % 32 == 0, every
// row 16-byte aligned); otherwise one value at a time with bounds checks.
// ---------------------------------------------------------------------------
template <typename T, int C, bool VEC>
__global__ void __launch_bounds__(256) transpose_pair_kernel(T *__restrict__ out, T *io, int n, int nt, long sb,
                                                             long sc, int p0, int p1, int p2, int p3, int p4)
{
    constexpr int PITCH = 32 * C + 4;  // row pitch in values (16-byte rows for the vector path)
    __shared__ __align__(16) T tile[2][32][PITCH];
    const int perm[5] = {p0, p1, p2, p3, p4};
    // pair index -> (A, B), A <= B, in the upper triangle of nt x nt tiles
    const int pi = blockIdx.x, c3 = blockIdx.y;
    int B = (int)((sqrtf(8.0f * pi + 1.0f) - 1.0f) * 0.5f);
    while ((B + 1) * (B + 2) / 2 <= pi) B++;
    while (B * (B + 1) / 2 > pi) B--;
    const int A = pi - B * (B + 1) / 2;
    const int ntile = A == B ? 1 : 2;
    const long pl = (long)c3 * sc;
    const int tid = threadIdx.x;
    for (int tt = 0; tt < ntile; tt++) {
        const int a0 = (tt == 0 ? A : B) * 32, b0 = (tt == 0 ? B : A) * 32;
        if constexpr (VEC && sizeof(T) == 4) {
            constexpr int RV = 8 * C;  // 16-byte vectors per 32-cell row
            for (int e = tid; e < 32 * RV; e += 256) {
                const int bb = e / RV, v = e - bb * RV;
                const float4 x =
                    reinterpret_cast<const float4 *>(io + ((long)a0 + (long)(b0 + bb) * sb + pl) * C)[v];
                *reinterpret_cast<float4 *>(&tile[tt][bb][4 * v]) = x;
            }
        } else {
            for (int e = tid; e < 32 * 32 * C; e += 256) {
                const int bb = e / (32 * C), r = e - bb * (32 * C);
                const int aa = r / C;
                if (a0 + aa < n && b0 + bb < n)
                    tile[tt][bb][r] = io[((long)(a0 + aa) + (long)(b0 + bb) * sb + pl) * C + (r - aa * C)];
            }
        }
    }
    __syncthreads();
    for (int tt = 0; tt < ntile; tt++) {
        // tile tt holds a in [a0, a0+32), b in [b0, b0+32); its transpose
        // goes to rows a (output cell b + a*sb), b contiguous
        const int a0 = (tt == 0 ? A : B) * 32, b0 = (tt == 0 ? B : A) * 32;
        if constexpr (VEC && sizeof(T) == 4) {
            constexpr int RV = 8 * C;
            for (int e = tid; e < 32 * RV; e += 256) {
                const int aa = e / RV, v = e - aa * RV;
                float y[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int f = 4 * v + i, bb = f / C, q = f - bb * C;
                    y[i] = tile[tt][bb][aa * C + perm[q]];
                }
                const long o = ((long)b0 + (long)(a0 + aa) * sb + pl) * C;
                const float4 x = make_float4(y[0], y[1], y[2], y[3]);
                reinterpret_cast<float4 *>(out + o)[v] = x;
                reinterpret_cast<float4 *>(io + o)[v] = x;
            }
        } else {
            for (int e = tid; e < 32 * 32 * C; e += 256) {
                const int aa = e / (32 * C), r = e - aa * (32 * C);
                const int bb = r / C, q = r - bb * C;
                if (a0 + aa < n && b0 + bb < n) {
                    const long o = ((long)(b0 + bb) + (long)(a0 + aa) * sb + pl) * C + q;
                    const T x = tile[tt][bb][aa * C + perm[q]];
                    out[o] = x;
                    io[o] = x;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 monitors: the deterministic reduction of np partial records (mass,
// energy, momentum x, y, z summed; max |div B| max-ed; 6 doubles each).
// Thread t takes records t, t+nt, ... in order, then a fixed tree over the
// threads, so the result does not depend on timing.  out[6] when given;
// with hist, also hist[6*(*iter - 1) ..], the record of the step just taken
// (its dt kernel advanced *iter), when that index is below cap.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) monitor_reduce_kernel(const double *__restrict__ part, int np, double *out,
                                                              double *hist, const int *iter, int cap)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    const double2 *p2 = reinterpret_cast<const double2 *>(part);  // a record is three double2
    double a[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    auto add = [&](const double2 (&v)[3]) {
        a[0] = a[0] + v[0].x;
        a[1] = a[1] + v[0].y;
        a[2] = a[2] + v[1].x;
        a[3] = a[3] + v[1].y;
        a[4] = a[4] + v[2].x;
        a[5] = fmax(a[5], v[2].y);
    };
    // four records per round: their loads are issued before any addition,
    // the additions keep the record order
    int r = tid;
    for (; r + 3 * nt < np; r += 4 * nt) {
        double2 v[4][3];
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int h = 0; h < 3; h++) v[k][h] = p2[3 * (long)(r + k * nt) + h];
#pragma unroll
        for (int k = 0; k < 4; k++) add(v[k]);
    }
    for (; r < np; r += nt) {
        double2 v[3];
#pragma unroll
        for (int h = 0; h < 3; h++) v[h] = p2[3 * (long)r + h];
        add(v);
    }
    // the threads in a fixed order: a butterfly in each wave, then the waves
    // in index order (monitor_block_store)
    __shared__ double res[6];
    monitor_block_store(a, res);
    __syncthreads();
    if (tid < 6) {
        if (out) out[tid] = res[tid];
        if (hist) {
            const int idx = *iter - 1;
            if (idx >= 0 && idx < cap) hist[6 * (long)idx + tid] = res[tid];
        }
    }
}

// ===========================================================================
// host launchers
// ===========================================================================
int pick_k(int n)
{
    static const int ks[] = {1, 2, 3, 4, 5, 6, 8};
    if (n < 2) return 0;
    for (int K : ks)
        if (K * 64 >= n && n % K == 0) return K;
    return 0;
}

int pick_p(int n)
{
    for (int P : {2, 4}) {
        const int G = 64 / P;
        if (n % G == 0 && (n / G == 1 || n / G == 3)) return P;
    }
    return 1;
}

int p_pays(int n, long pencils)
{
    const int P = pick_p(n);
    if (P == 1) return 1;
    return n * P / 64 == pick_k(n) || pencils / P >= 4096 ? P : 1;
}

Frame make_frame(int fr, int nx, int ny, int nz)
{
    Frame f;
    const long sx = 1, sy = nx, sz = (long)nx * ny;
    if (fr == 0) {
        f.ni = nx; f.nj = ny; f.nk = nz; f.si = sx; f.sj = sy; f.sk = sz;
    } else if (fr == 1) {
        f.ni = ny; f.nj = nx; f.nk = nz; f.si = sy; f.sj = sx; f.sk = sz;
    } else {
        f.ni = nz; f.nj = nx; f.nk = ny; f.si = sz; f.sj = sx; f.sk = sy;
    }
    return f;
}

template <typename T> constexpr int tile_w() { return sizeof(T) == 4 ? 4 : 2; }

#define MHD_KSWITCH(K, CALL)                                                                                         \
    switch (K) {                                                                                                    \
    case 1: return CALL(1);                                                                                         \
    case 2: return CALL(2);                                                                                         \
    case 3: return CALL(3);                                                                                         \
    case 4: return CALL(4);                                                                                         \
    case 5: return CALL(5);                                                                                         \
    case 6: return CALL(6);                                                                                         \
    case 8: return CALL(8);                                                                                         \
    default: return hipErrorInvalidValue;                                                                           \
    }

// raise the dynamic-LDS limit of a sweep kernel once per process (160 KB on gfx950)
template <typename T, int K, int FR, int WHICH> static hipError_t sweep_attr()
{
    constexpr int W = tile_w<T>();
    static const hipError_t e = WHICH == 0
        ? hipFuncSetAttribute((const void *)fluidx_kernel<T, K, W, FR>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024)
        : hipFuncSetAttribute((const void *)advect_kernel<T, K, W, FR>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    return e;
}

template <typename T, int K, int FR> static hipError_t sweep_attr_both()
{
    hipError_t e = sweep_attr<T, K, FR, 0>();
    return e != hipSuccess ? e : sweep_attr<T, K, FR, 1>();
}

template <typename T> hipError_t prepare_frame(int fr, int n)
{
    const int K = pick_k(n);
    if (!K) return prepare_gen<T>(n, 0);
#define PA0(KK) sweep_attr_both<T, KK, 0>()
#define PA1(KK) sweep_attr_both<T, KK, 1>()
#define PA2(KK) sweep_attr_both<T, KK, 2>()
    if (fr == 0) { MHD_KSWITCH(K, PA0) }
    if (fr == 1) { MHD_KSWITCH(K, PA1) }
    MHD_KSWITCH(K, PA2)
#undef PA0
#undef PA1
#undef PA2
}

template <typename T> hipError_t prepare_sweeps(int nx, int ny, int nz)
{
    hipError_t e = prepare_frame<T>(0, nx);
    if (e == hipSuccess) e = prepare_frame<T>(1, ny);
    if (e == hipSuccess) e = prepare_frame<T>(2, nz);
    return e;
}

// the widest record loads the x-frame AoS kernels may use: 16-byte loads
// when every lane's first record (cell nx*(j + ny*k) + l*K) and the arrays
// are 16-byte aligned (nx and K multiples of 4), 8-byte loads when they are
// 8-byte aligned, else one value at a time
static int aos_vec(int nx, int K, std::initializer_list<const void *> ptrs)
{
    bool a16 = true, a8 = true;
    for (const void *p : ptrs) {
        a16 = a16 && ((uintptr_t)p % 16 == 0);
        a8 = a8 && ((uintptr_t)p % 8 == 0);
    }
    if (sizeof(float) == 4 && a16 && nx % 4 == 0 && K % 4 == 0) return 4;
    if (a8 && nx % 2 == 0 && K % 2 == 0) return 2;
    return 1;
}

template <typename T, int K, int FR>
static hipError_t fluidx_launch_k(T *u, const T *b, const Frame &f, const T *dt, hipStream_t st)
{
    constexpr int W = tile_w<T>();
    if constexpr (FR == 0) {
        // along x: whole records straight from global memory (mhd_aos_x.h)
        const int ntj = (f.nj + W - 1) / W;
        const dim3 grid(ntj * f.nk), block(64 * W);
        const int v = aos_vec(f.ni, K, {u, b});
        if (v == 4 && sizeof(T) == 4)
            hipLaunchKernelGGL((fluidx_aos_x_kernel<T, K, W, 4>), grid, block, 0, st, u, b, f, dt, ntj);
        else if (v >= 2)
            hipLaunchKernelGGL((fluidx_aos_x_kernel<T, K, W, 2>), grid, block, 0, st, u, b, f, dt, ntj);
        else
            hipLaunchKernelGGL((fluidx_aos_x_kernel<T, K, W, 1>), grid, block, 0, st, u, b, f, dt, ntj);
        return hipGetLastError();
    } else {
        const size_t lds = sizeof(T) * fluidx_lds_words<T, K, W>(f.ni);
        auto kern = fluidx_kernel<T, K, W, FR>;
        const hipError_t ea = sweep_attr<T, K, FR, 0>();
        if (ea != hipSuccess) return ea;
        const int ntj = (f.nj + W - 1) / W;
        const int nblk = ntj * f.nk;
        hipLaunchKernelGGL(kern, dim3(nblk), dim3(64 * W), lds, st, u, b, f, dt, ntj);
        return hipGetLastError();
    }
}

template <typename T, int K, int FR>
static hipError_t advect_launch_k(const T *u, const T *bin, T *bout, const Frame &f, const T *dt, hipStream_t st)
{
    constexpr int W = tile_w<T>();
    if constexpr (FR == 0) {
        // along x: tiles of two planes with direct record loads (mhd_aos_x.h),
        // eight pencils wide (measured 4% faster than four at 256^3: fewer
        // halo pencils per pencil)
        constexpr int W = 2 * tile_w<T>();
        const int ntj = (f.nj + W - 1) / W;
        const dim3 grid(ntj * ((f.nk + 1) / 2)), block(64 * W);
        const size_t lds = sizeof(T) * advect_aos_x2_words<W>(f.ni);
        const int v = aos_vec(f.ni, K, {u, bin, bout});
        hipError_t e = hipSuccess;
#define MHD_AX(V)                                                                                                    \
    do {                                                                                                             \
        static const hipError_t a = hipFuncSetAttribute((const void *)advect_aos_x2_kernel<T, K, W, V>,              \
                                                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);    \
        e = a;                                                                                                       \
        if (e == hipSuccess)                                                                                         \
            hipLaunchKernelGGL((advect_aos_x2_kernel<T, K, W, V>), grid, block, lds, st, u, bin, bout, f, dt, ntj); \
    } while (0)
        if (v == 4 && sizeof(T) == 4) MHD_AX(4);
        else if (v >= 2) MHD_AX(2);
        else MHD_AX(1);
#undef MHD_AX
        return e != hipSuccess ? e : hipGetLastError();
    } else {
        const size_t lds = sizeof(T) * advect_lds_words<T, K, W>(f.ni);
        auto kern = advect_kernel<T, K, W, FR>;
        const hipError_t ea = sweep_attr<T, K, FR, 1>();
        if (ea != hipSuccess) return ea;
        const int ntj = (f.nj + W - 1) / W;
        const int nblk = ntj * f.nk;
        hipLaunchKernelGGL(kern, dim3(nblk), dim3(64 * W), lds, st, u, bin, bout, f, dt, ntj);
        return hipGetLastError();
    }
}


template <typename T>
hipError_t launch_fluidx(T *u, const T *b, int fr, int nx, int ny, int nz, const T *dt, hipStream_t st)
{
    const Frame f = make_frame(fr, nx, ny, nz);
    const int K = pick_k(f.ni);
    if (!K) return launch_fluidx_gen<T>(u, b, f, 0, 0, fr, dt, st);
#define FX0(KK) fluidx_launch_k<T, KK, 0>(u, b, f, dt, st)
#define FX1(KK) fluidx_launch_k<T, KK, 1>(u, b, f, dt, st)
#define FX2(KK) fluidx_launch_k<T, KK, 2>(u, b, f, dt, st)
    if (fr == 0) { MHD_KSWITCH(K, FX0) }
    if (fr == 1) { MHD_KSWITCH(K, FX1) }
    MHD_KSWITCH(K, FX2)
#undef FX0
#undef FX1
#undef FX2
}

template <typename T>
hipError_t launch_advect(const T *u, const T *bin, T *bout, int fr, int nx, int ny, int nz, const T *dt,
                         hipStream_t st)
{
    const Frame f = make_frame(fr, nx, ny, nz);
    const int K = pick_k(f.ni);
    if (!K) return launch_advect_gen<T>(u, bin, bout, f, 0, 0, fr, dt, st);
#define AX0(KK) advect_launch_k<T, KK, 0>(u, bin, bout, f, dt, st)
#define AX1(KK) advect_launch_k<T, KK, 1>(u, bin, bout, f, dt, st)
#define AX2(KK) advect_launch_k<T, KK, 2>(u, bin, bout, f, dt, st)
    if (fr == 0) { MHD_KSWITCH(K, AX0) }
    if (fr == 1) { MHD_KSWITCH(K, AX1) }
    MHD_KSWITCH(K, AX2)
#undef AX0
#undef AX1
#undef AX2
}

static int cfl_blocks(long ncell)
{
    long nb = (ncell + 255) / 256;
    return (int)(nb < 4096 ? nb : 4096);
}

template <typename T>
hipError_t launch_cfl(const T *u, const T *b, int nx, int ny, int nz, int avg, unsigned long long *cmax,
                      hipStream_t st)
{
    const int nb = cfl_blocks((long)nx * ny * nz);
    if (avg)
        hipLaunchKernelGGL((cfl_kernel<T, true>), dim3(nb), dim3(256), 0, st, u, b, nx, ny, nz, cmax);
    else
        hipLaunchKernelGGL((cfl_kernel<T, false>), dim3(nb), dim3(256), 0, st, u, b, nx, ny, nz, cmax);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_dt_part(unsigned long long *cmax, const T *part, int np, T *dt, T *t, double *t64, T tf, T *dts,
                          int *iter, int cap, const T *u, const T *b, const SoaGrid &g, int avg, hipStream_t st)
{
    hipLaunchKernelGGL(dt_part_kernel<T>, dim3(1), dim3(1024), 0, st, cmax, part, np, dt, t, t64, tf, dts, iter, cap,
                       u, b, g, avg);
    return hipGetLastError();
}

template <typename T> hipError_t launch_fill_zero(T *p, int n, hipStream_t st)
{
    const int nb = (n + 255) / 256;
    hipLaunchKernelGGL(fill_zero_kernel<T>, dim3(nb < 1024 ? (nb > 0 ? nb : 1) : 1024), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_cfl_value(unsigned long long *cmax, T *out, const T *u, const T *b, int nx, int ny, int nz, int avg,
                            hipStream_t st)
{
    hipLaunchKernelGGL(cfl_value_kernel<T>, dim3(1), dim3(1024), 0, st, cmax, out, u, b, nx, ny, nz, avg);
    return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long *p, int n, hipStream_t st)
{
    hipLaunchKernelGGL(zero_u64_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_transpose(T *uo, T *bo, const T *u, const T *b, int which, int nx, int ny, int nz, hipStream_t st,
                            bool both)
{
    // which 12: (i,j,k)->(j,i,k), u (1,3,2,4,5), b (2,1,3)
    // which 13: (i,j,k)->(k,j,i), u (1,4,3,2,5), b (3,2,1)
    int na = nx, nb, nc;
    long sbi, sci, sao, sco;
    int pu[5], pb[3];
    if (which == 12) {
        nb = ny; nc = nz; sbi = nx; sci = (long)nx * ny; sao = ny; sco = (long)nx * ny;
        const int a[5] = {0, 2, 1, 3, 4}, c[3] = {1, 0, 2};
        for (int i = 0; i < 5; i++) pu[i] = a[i];
        for (int i = 0; i < 3; i++) pb[i] = c[i];
    } else {
        nb = nz; nc = ny; sbi = (long)nx * ny; sci = nx; sao = (long)ny * nz; sco = nz;
        const int a[5] = {0, 3, 2, 1, 4}, c[3] = {2, 1, 0};
        for (int i = 0; i < 5; i++) pu[i] = a[i];
        for (int i = 0; i < 3; i++) pb[i] = c[i];
    }
    if (both && na == nb) {
        // square planes: transpose in place into u, b and into uo, bo at once
        const int nt = (na + 31) / 32;
        const dim3 grid(nt * (nt + 1) / 2, nc);
        auto aligned = [](const void *p) { return (uintptr_t)p % 16 == 0; };
        const bool vec = sizeof(T) == 4 && na % 32 == 0 && aligned(u) && aligned(b) && aligned(uo) && aligned(bo);
        T *iu = const_cast<T *>(u), *ib = const_cast<T *>(b);
        if (vec) {
            hipLaunchKernelGGL((transpose_pair_kernel<T, 5, true>), grid, dim3(256), 0, st, uo, iu, na, nt, sbi, sci,
                               pu[0], pu[1], pu[2], pu[3], pu[4]);
            hipLaunchKernelGGL((transpose_pair_kernel<T, 3, true>), grid, dim3(256), 0, st, bo, ib, na, nt, sbi, sci,
                               pb[0], pb[1], pb[2], 0, 0);
        } else {
            hipLaunchKernelGGL((transpose_pair_kernel<T, 5, false>), grid, dim3(256), 0, st, uo, iu, na, nt, sbi, sci,
                               pu[0], pu[1], pu[2], pu[3], pu[4]);
            hipLaunchKernelGGL((transpose_pair_kernel<T, 3, false>), grid, dim3(256), 0, st, bo, ib, na, nt, sbi, sci,
                               pb[0], pb[1], pb[2], 0, 0);
        }
        return hipGetLastError();
    }
    dim3 grid((na + 31) / 32, (nb + 31) / 32, nc);
    hipLaunchKernelGGL((transpose_kernel<T, 5>), grid, dim3(256), 0, st, uo, u, na, nb, sbi, sci, sao, sco, pu[0],
                       pu[1], pu[2], pu[3], pu[4]);
    hipLaunchKernelGGL((transpose_kernel<T, 3>), grid, dim3(256), 0, st, bo, b, na, nb, sbi, sci, sao, sco, pb[0],
                       pb[1], pb[2], 0, 0);
    hipError_t e = hipGetLastError();
    if (both && e == hipSuccess) {
        // the copy back of the reference (HIP/cu/cuda_transpose12.cu:33-34)
        const size_t ncell = (size_t)nx * ny * nz;
        e = hipMemcpyAsync(const_cast<T *>(u), uo, ncell * 5 * sizeof(T), hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(const_cast<T *>(b), bo, ncell * 3 * sizeof(T), hipMemcpyDeviceToDevice, st);
    }
    return e;
}

// first stage of a large reduction: block g reduces records [256g, 256g+256)
// (one per thread, then monitor_block_store's fixed order) into stage[g]
__global__ void __launch_bounds__(256) monitor_stage_kernel(const double *__restrict__ part, int np,
                                                            double *__restrict__ stage)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    double a[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (r < np) {
        const double2 *p2 = reinterpret_cast<const double2 *>(part) + 3 * (long)r;
        const double2 v0 = p2[0], v1 = p2[1], v2 = p2[2];
        a[0] = v0.x;
        a[1] = v0.y;
        a[2] = v1.x;
        a[3] = v1.y;
        a[4] = v2.x;
        a[5] = v2.y;
    }
    monitor_block_store(a, stage + 6 * (long)blockIdx.x);
}

hipError_t launch_monitor_reduce(const double *part, int np, double *out, double *hist, const int *iter, int cap,
                                 hipStream_t st)
{
    if (np > 2048) {
        // one workgroup alone reads 16 K records at about 40 GB/s: a first
        // stage over 256-record groups, on as many workgroups, writes its
        // records after the np inputs (the caller's buffer holds them)
        const int ng = (np + 255) / 256;
        double *stage = const_cast<double *>(part) + 6 * (long)np;
        hipLaunchKernelGGL(monitor_stage_kernel, dim3(ng), dim3(256), 0, st, part, np, stage);
        hipLaunchKernelGGL(monitor_reduce_kernel, dim3(1), dim3(1024), 0, st, stage, ng, out, hist, iter, cap);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(monitor_reduce_kernel, dim3(1), dim3(1024), 0, st, part, np, out, hist, iter, cap);
    return hipGetLastError();
}

#define MHD_INST(T)                                                                                                  \
    template hipError_t launch_fluidx<T>(T *, const T *, int, int, int, int, const T *, hipStream_t);              \
    template hipError_t launch_advect<T>(const T *, const T *, T *, int, int, int, int, const T *, hipStream_t);              \
    template hipError_t launch_cfl<T>(const T *, const T *, int, int, int, int, unsigned long long *, hipStream_t); \
    template hipError_t launch_dt_part<T>(unsigned long long *, const T *, int, T *, T *, double *, T, T *, int *,   \
                                          int, const T *, const T *, const SoaGrid &, int, hipStream_t);            \
    template hipError_t launch_fill_zero<T>(T *, int, hipStream_t);                                                 \
    template hipError_t launch_cfl_value<T>(unsigned long long *, T *, const T *, const T *, int, int, int, int,     \
                                            hipStream_t);                                                           \
    template hipError_t launch_transpose<T>(T *, T *, const T *, const T *, int, int, int, int, hipStream_t, bool); \
    template hipError_t prepare_sweeps<T>(int, int, int);
MHD_INST(float)
MHD_INST(double)

}  // namespace mhd
