// Sweeps for every pencil length the reference accepts (2 <= n <= 1024,
// HIP/cu/cuda_fluidx.cu:15-17 runs a pencil as one block of n threads), in
// both layouts: the step-level AoS arrays and the device-resident SoA
// planes.  The fast kernels need n = L*K with K <= 8; these take any n:
// K = 2, 4, 8 or 16 cells per lane, L = ceil(n/K) lanes, the last lane
// ragged (LanesRagged: its shifts wrap the pencil at its last cell, its
// unused cells stay out of the pencil max).  A workgroup owns W pencils
// neighbouring along j, staged through LDS as in the AoS step-level kernels;
// the arithmetic is the shared core (mhd_core.h), so results equal the fast
// kernels' and the reference's.
#include "mhd_generic.h"
#include "mhd_core.h"

namespace mhd {

namespace {

// padded LDS offset of cell i, component c, in a pencil of C-component cells
// (groups of K cells padded to an odd word count)
template <int K, int C> __device__ __forceinline__ int gidx(int i, int c) { return (i / K) * (C * K + 1) + (i % K) * C + c; }
// words of one pencil slot, ragged pencils rounded up to whole lanes
template <int K, int C> __host__ __device__ constexpr int glen(int n) { return (n + K - 1) / K * (C * K + 1); }

__device__ __forceinline__ int wrapn(int j, int nj)
{
    while (j >= nj) j -= nj;
    return j;
}

// element c of a runtime component map (selects: no dynamic indexing)
template <int C> __device__ __forceinline__ int cmap(const int (&m)[C], int c)
{
    int r = m[0];
#pragma unroll
    for (int q = 1; q < C; q++) r = c == q ? m[q] : r;
    return r;
}

template <bool SOA, int C> __device__ __forceinline__ long gaddr(long cell, int p, long N)
{
    return SOA ? (long)p * N + cell : cell * C + p;
}

// Load nslot pencils (j = j0+s mod nj, plane k) into LDS slot s (pitch P):
// all C components (physical component p stored as frame component inv[p])
// when comp < 0, else the one physical component comp as a one-word cell.
// The element order follows the memory layout so that loads coalesce.
template <typename T, int K, int C, int CL, bool SOA>
__device__ void gload(T *lds, int P, const T *__restrict__ g, const Frame &f, long N, int j0, int nslot, int k,
                      const int (&inv)[C], int comp)
{
    const int n = f.ni, tid = threadIdx.x, nt = blockDim.x;
    const int nc = comp < 0 ? C : 1;
    const long kb = (long)k * f.sk;
    if (f.si == 1) {  // x frame: a pencil is contiguous
        for (int s = 0; s < nslot; s++) {
            const long base = kb + (long)wrapn(j0 + s, f.nj) * f.sj;
            for (int e = tid; e < n * nc; e += nt) {
                int i, q;
                if (SOA) {
                    q = e / n;
                    i = e - q * n;
                } else {
                    i = e / nc;
                    q = e - i * nc;
                }
                const int p = comp < 0 ? q : comp, c = comp < 0 ? cmap<C>(inv, p) : 0;
                lds[s * P + gidx<K, CL>(i, c)] = g[gaddr<SOA, C>(base + i, p, N)];
            }
        }
    } else {  // y/z frames: neighbouring pencils are neighbouring cells
        for (int e = tid; e < n * nslot * nc; e += nt) {
            int i, s, q;
            if (SOA) {
                i = e / (nslot * nc);
                const int r = e - i * nslot * nc;
                q = r / nslot;
                s = r - q * nslot;
            } else {
                i = e / (nslot * nc);
                const int r = e - i * nslot * nc;
                s = r / nc;
                q = r - s * nc;
            }
            const int p = comp < 0 ? q : comp, c = comp < 0 ? cmap<C>(inv, p) : 0;
            const long cell = kb + (long)i * f.si + (long)wrapn(j0 + s, f.nj) * f.sj;
            lds[s * P + gidx<K, CL>(i, c)] = g[gaddr<SOA, C>(cell, p, N)];
        }
    }
}

// Store all C components of nslot pencils (inverse of gload)
template <typename T, int K, int C, bool SOA>
__device__ void gstore(const T *lds, int P, T *__restrict__ g, const Frame &f, long N, int j0, int nslot, int k,
                       const int (&inv)[C])
{
    const int n = f.ni, tid = threadIdx.x, nt = blockDim.x;
    const long kb = (long)k * f.sk;
    if (f.si == 1) {
        for (int s = 0; s < nslot; s++) {
            const long base = kb + (long)(j0 + s) * f.sj;
            for (int e = tid; e < n * C; e += nt) {
                int i, p;
                if (SOA) {
                    p = e / n;
                    i = e - p * n;
                } else {
                    i = e / C;
                    p = e - i * C;
                }
                g[gaddr<SOA, C>(base + i, p, N)] = lds[s * P + gidx<K, C>(i, cmap<C>(inv, p))];
            }
        }
    } else {
        for (int e = tid; e < n * nslot * C; e += nt) {
            int i, s, p;
            i = e / (nslot * C);
            const int r = e - i * nslot * C;
            if (SOA) {
                p = r / nslot;
                s = r - p * nslot;
            } else {
                s = r / C;
                p = r - s * C;
            }
            const long cell = kb + (long)i * f.si + (long)(j0 + s) * f.sj;
            g[gaddr<SOA, C>(cell, p, N)] = lds[s * P + gidx<K, C>(i, cmap<C>(inv, p))];
        }
    }
}

template <int K, int W> __host__ __device__ constexpr int gen_fluidx_words(int n)
{
    return W * glen<K, 5>(n) + (W + 1) * glen<K, 3>(n) + W * glen<K, 1>(n);
}

template <int K, int W> __host__ __device__ constexpr int gen_advect_words(int n)
{
    // rho and momentum: W+2 slots at k, W at k-1, W at k+1; b: W+1 slots;
    // b3 at k+1: W slots; fluxes FA (W+1), FB0 (W), FB1 (W)
    return 2 * (3 * W + 2) * glen<K, 1>(n) + (W + 1) * glen<K, 3>(n) + W * glen<K, 1>(n) +
           (3 * W + 1) * glen<K, 1>(n);
}

// fluidx (Heterogeneous_MHD/C/fluidx.c:19-102, tvd1.c, mhdflux.c)
template <typename T, int K, int W, bool SOA>
__global__ void __launch_bounds__(64 * W) fluidx_gen_kernel(T *__restrict__ u, const T *__restrict__ b, Frame f,
                                                            long N, CompMap cm, const T *__restrict__ dtp, int ntj)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *smem = reinterpret_cast<T *>(smem_raw);
    const int n = f.ni, L = (n + K - 1) / K;
    const int PU = glen<K, 5>(n), PB = glen<K, 3>(n), PK = glen<K, 1>(n);
    T *su = smem, *sb = su + W * PU, *sk = sb + (W + 1) * PB;

    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int k = tile / ntj, j0 = (tile - k * ntj) * W;
    const int nvalid = min(W, f.nj - j0);
    const int kp = k + 1 == f.nk ? 0 : k + 1;

    gload<T, K, 5, 5, SOA>(su, PU, u, f, N, j0, nvalid, k, cm.iu, -1);
    gload<T, K, 3, 3, SOA>(sb, PB, b, f, N, j0, nvalid + 1, k, cm.ib, -1);
    gload<T, K, 3, 1, SOA>(sk, PK, b, f, N, j0, nvalid, kp, cm.ib, cm.b[2]);
    __syncthreads();

    const int w = threadIdx.x >> 6;
    if (w < nvalid) {
        const int l = lane_id(), lc = l % L;  // lanes >= L shadow lane l % L and store nothing
        const LanesRagged ln = ragged_lanes(n, K, lc);
        const T dt = *dtp;
        T uu[K][5], bx[K][3];
        const T *pu = su + w * PU + lc * (5 * K + 1);
#pragma unroll
        for (int q = 0; q < K; q++)
#pragma unroll
            for (int m = 0; m < 5; m++) uu[q][m] = pu[q * 5 + m];
        // face B to cell centre (fluidx.c:71-81); the cells past the pencil
        // end (i >= n) read cell 0 as their neighbour
#pragma unroll
        for (int q = 0; q < K; q++) {
            const int i = lc * K + q, ip = i + 1 >= n ? 0 : i + 1;
            const T b1 = sb[w * PB + gidx<K, 3>(i, 0)], b1p = sb[w * PB + gidx<K, 3>(ip, 0)];
            const T b2 = sb[w * PB + gidx<K, 3>(i, 1)], b2p = sb[(w + 1) * PB + gidx<K, 3>(i, 1)];
            const T b3 = sb[w * PB + gidx<K, 3>(i, 2)], b3p = sk[w * PK + gidx<K, 1>(i, 0)];
            bx[q][0] = (T)0.5 * b1 + (T)0.5 * b1p;
            bx[q][1] = (T)0.5 * b2 + b2p / (T)2;
            bx[q][2] = (T)0.5 * b3 + b3p / (T)2;
        }
        tvd1_pencil<T, K>(uu, bx, dt, ln);
        if (l < L) {
            T *po = su + w * PU + lc * (5 * K + 1);
#pragma unroll
            for (int q = 0; q < K; q++)
#pragma unroll
                for (int m = 0; m < 5; m++) po[q * 5 + m] = uu[q][m];
        }
    }
    __syncthreads();
    gstore<T, K, 5, SOA>(su, PU, u, f, N, j0, nvalid, k, cm.iu);
}

// advectbyzx (advectbyzx.c:14-23 with A1/A2/B1/B2 fused): per tile the
// fluxes FA of pencils j0..j0+nvalid (b1 needs pencil j+1), FB at k and
// k+1 (tvdb each), then b1 = (((b1 - FA_j) + FA_{j+1}) - FB_k) + FB_{k+1}
// (A1.c:55, A2.c:23, B1.c:54, B2.c:23).  Reads bin, writes bout.
template <typename T, int K, int W, bool SOA>
__global__ void __launch_bounds__(64 * W)
    advect_gen_kernel(const T *__restrict__ u, const T *__restrict__ bin, T *__restrict__ bout, Frame f, long N,
                      CompMap cm, const T *__restrict__ dtp, int ntj)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *smem = reinterpret_cast<T *>(smem_raw);
    const int n = f.ni, L = (n + K - 1) / K;
    const int PK = glen<K, 1>(n), PB = glen<K, 3>(n);
    // rho/momentum rows: 0..W+1 at k (pencil j0-1+s), W+2..2W+1 at k-1
    // (pencil j0+s), 2W+2..3W+1 at k+1 (pencil j0+s)
    T *sr = smem, *sm = sr + (3 * W + 2) * PK;
    T *sb = sm + (3 * W + 2) * PK, *sb3p = sb + (W + 1) * PB;
    T *sF = sb3p + W * PK;  // FA: slots 0..W, FB0: W+1..2W, FB1: 2W+1..3W

    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int k = tile / ntj, j0 = (tile - k * ntj) * W;
    const int nvalid = min(W, f.nj - j0);
    const int kp = k + 1 == f.nk ? 0 : k + 1, km = k == 0 ? f.nk - 1 : k - 1;
    const int pm = cm.u[1];

    gload<T, K, 5, 1, SOA>(sr, PK, u, f, N, j0 - 1 + f.nj, nvalid + 2, k, cm.iu, 0);
    gload<T, K, 5, 1, SOA>(sm, PK, u, f, N, j0 - 1 + f.nj, nvalid + 2, k, cm.iu, pm);
    gload<T, K, 5, 1, SOA>(sr + (W + 2) * PK, PK, u, f, N, j0, nvalid, km, cm.iu, 0);
    gload<T, K, 5, 1, SOA>(sm + (W + 2) * PK, PK, u, f, N, j0, nvalid, km, cm.iu, pm);
    gload<T, K, 5, 1, SOA>(sr + (2 * W + 2) * PK, PK, u, f, N, j0, nvalid, kp, cm.iu, 0);
    gload<T, K, 5, 1, SOA>(sm + (2 * W + 2) * PK, PK, u, f, N, j0, nvalid, kp, cm.iu, pm);
    gload<T, K, 3, 3, SOA>(sb, PB, bin, f, N, j0, nvalid + 1, k, cm.ib, -1);
    gload<T, K, 3, 1, SOA>(sb3p, PK, bin, f, N, j0, nvalid, kp, cm.ib, cm.b[2]);
    __syncthreads();

    const int w = threadIdx.x >> 6;
    const int l = lane_id(), lc = l % L;
    const LanesRagged ln = ragged_lanes(n, K, lc);
    const T dt = *dtp;
    const int nitem = 3 * nvalid + 1;
    for (int it = w; it < nitem; it += W) {
        int ra, rb, fslot, fkind, Fslot;  // fkind 0: b2 of sb slot, 1: b3 of sb slot, 2: b3 at k+1
        if (it <= nvalid) {  // FA(s): pencils j-1 (row s) and j (row s+1) at k, field b2
            ra = it;
            rb = it + 1;
            fslot = it;
            fkind = 0;
            Fslot = it;
        } else if (it <= 2 * nvalid) {  // FB0(s): pencils (s,k-1) and (s,k), field b3
            const int s = it - nvalid - 1;
            ra = W + 2 + s;
            rb = s + 1;
            fslot = s;
            fkind = 1;
            Fslot = W + 1 + s;
        } else {  // FB1(s): pencils (s,k) and (s,k+1), field b3 at k+1
            const int s = it - 2 * nvalid - 1;
            ra = s + 1;
            rb = 2 * W + 2 + s;
            fslot = s;
            fkind = 2;
            Fslot = 2 * W + 1 + s;
        }
        T vx[K], vxm[K], vxn[K], vs[K], b0[K], bnew[K], F[K];
#pragma unroll
        for (int q = 0; q < K; q++) {
            const int o = gidx<K, 1>(lc * K + q, 0);
            // A1.c:34 (u2(jm)+u2(j))/(u1(jm)+u1(j)); the minus-side pencil first
            vx[q] = (sm[ra * PK + o] + sm[rb * PK + o]) / (sr[ra * PK + o] + sr[rb * PK + o]);
        }
        shift_prev<T, K>(vx, vxm, ln);
        shift_next<T, K>(vx, vxn, ln);
#pragma unroll
        for (int q = 0; q < K; q++) {
            // A1.c:41 (vx[i-1] + vx[i+1] + 2.0*vx)/4.0
            vs[q] = avg121<T>(vxm[q] + vxn[q], vx[q]);
            const int i = lc * K + q;
            if (fkind == 0)
                b0[q] = sb[fslot * PB + gidx<K, 3>(i, 1)];
            else if (fkind == 1)
                b0[q] = sb[fslot * PB + gidx<K, 3>(i, 2)];
            else
                b0[q] = sb3p[fslot * PK + gidx<K, 1>(i, 0)];
        }
        tvdb_pencil<T, K>(b0, vs, dt, bnew, F, ln);
        if (l < L) {
#pragma unroll
            for (int q = 0; q < K; q++) {
                const int i = lc * K + q;
                sF[Fslot * PK + gidx<K, 1>(i, 0)] = F[q];
                // new b2 (FA) / b3 (FB0) of the tile's own pencils; the item
                // reads only its own field row, so it is overwritten in place
                if (fkind == 0 && fslot < nvalid) sb[fslot * PB + gidx<K, 3>(i, 1)] = bnew[q];
                if (fkind == 1) sb[fslot * PB + gidx<K, 3>(i, 2)] = bnew[q];
            }
        }
    }
    __syncthreads();
    if (w < nvalid && l < L) {
#pragma unroll
        for (int q = 0; q < K; q++) {
            const int i = lc * K + q;
            const int o = gidx<K, 1>(i, 0);
            T b1 = sb[w * PB + gidx<K, 3>(i, 0)];
            b1 = b1 - sF[w * PK + o];
            b1 = b1 + sF[(w + 1) * PK + o];
            b1 = b1 - sF[(W + 1 + w) * PK + o];
            b1 = b1 + sF[(2 * W + 1 + w) * PK + o];
            sb[w * PB + gidx<K, 3>(i, 0)] = b1;
        }
    }
    __syncthreads();
    gstore<T, K, 3, SOA>(sb, PB, bout, f, N, j0, nvalid, k, cm.ib);
}

// tile width (pencils per workgroup) of the generic kernels: the LDS of a
// 1024-cell FP32 pencil allows two, an FP64 pencil of up to 512 cells two
template <typename T, int K> constexpr int gen_w() { return sizeof(T) == 4 ? (K <= 8 ? 4 : 2) : (K <= 4 ? 2 : 1); }

constexpr int kGenLds = 160 * 1024;

template <typename T, int K, bool SOA> hipError_t gen_attr()
{
    static const hipError_t e = [] {
        constexpr int W = gen_w<T, K>();
        hipError_t a = hipFuncSetAttribute((const void *)fluidx_gen_kernel<T, K, W, SOA>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, kGenLds);
        if (a == hipSuccess)
            a = hipFuncSetAttribute((const void *)advect_gen_kernel<T, K, W, SOA>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kGenLds);
        return a;
    }();
    return e;
}

template <typename T, int K, bool SOA>
hipError_t fluidx_gen_k(T *u, const T *b, const Frame &f, long N, const CompMap &cm, const T *dt, hipStream_t st)
{
    constexpr int W = gen_w<T, K>();
    const size_t lds = sizeof(T) * gen_fluidx_words<K, W>(f.ni);
    if (lds > (size_t)kGenLds) return hipErrorInvalidValue;
    const hipError_t ea = gen_attr<T, K, SOA>();
    if (ea != hipSuccess) return ea;
    const int ntj = (f.nj + W - 1) / W;
    hipLaunchKernelGGL((fluidx_gen_kernel<T, K, W, SOA>), dim3(ntj * f.nk), dim3(64 * W), lds, st, u, b, f, N, cm, dt,
                       ntj);
    return hipGetLastError();
}

template <typename T, int K, bool SOA>
hipError_t advect_gen_k(const T *u, const T *bin, T *bout, const Frame &f, long N, const CompMap &cm, const T *dt,
                        hipStream_t st)
{
    constexpr int W = gen_w<T, K>();
    const size_t lds = sizeof(T) * gen_advect_words<K, W>(f.ni);
    if (lds > (size_t)kGenLds) return hipErrorInvalidValue;
    const hipError_t ea = gen_attr<T, K, SOA>();
    if (ea != hipSuccess) return ea;
    const int ntj = (f.nj + W - 1) / W;
    hipLaunchKernelGGL((advect_gen_kernel<T, K, W, SOA>), dim3(ntj * f.nk), dim3(64 * W), lds, st, u, bin, bout, f, N,
                       cm, dt, ntj);
    return hipGetLastError();
}

}  // namespace

int gen_k(int n, int prec)
{
    if (n < 2 || n > 1024) return 0;
    const int K = n <= 128 ? 2 : n <= 256 ? 4 : n <= 512 ? 8 : 16;
    return (prec == 8 && K > 8) ? 0 : K;
}

CompMap comp_map(int fr)
{
    CompMap m;
    for (int c = 0; c < 5; c++) m.u[c] = fr == 0 ? Comp<0>::u[c] : fr == 1 ? Comp<1>::u[c] : Comp<2>::u[c];
    for (int c = 0; c < 3; c++) m.b[c] = fr == 0 ? Comp<0>::b[c] : fr == 1 ? Comp<1>::b[c] : Comp<2>::b[c];
    for (int c = 0; c < 5; c++) m.iu[m.u[c]] = c;
    for (int c = 0; c < 3; c++) m.ib[m.b[c]] = c;
    return m;
}

#define GEN_KSWITCH(K, CALL)                                                                                         \
    switch (K) {                                                                                                    \
    case 2: return CALL(2);                                                                                         \
    case 4: return CALL(4);                                                                                         \
    case 8: return CALL(8);                                                                                         \
    default: break;                                                                                                 \
    }

template <typename T>
hipError_t launch_fluidx_gen(T *u, const T *b, const Frame &f, long N, int soa, int fr, const T *dt, hipStream_t st)
{
    const int K = gen_k(f.ni, sizeof(T));
    const CompMap cm = comp_map(fr);
#define GF0(KK) fluidx_gen_k<T, KK, false>(u, b, f, N, cm, dt, st)
#define GF1(KK) fluidx_gen_k<T, KK, true>(u, b, f, N, cm, dt, st)
    if (soa) {
        GEN_KSWITCH(K, GF1)
    } else {
        GEN_KSWITCH(K, GF0)
    }
    if constexpr (sizeof(T) == 4) {
        if (K == 16) return soa ? GF1(16) : GF0(16);
    }
#undef GF0
#undef GF1
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_advect_gen(const T *u, const T *bin, T *bout, const Frame &f, long N, int soa, int fr, const T *dt,
                             hipStream_t st)
{
    const int K = gen_k(f.ni, sizeof(T));
    const CompMap cm = comp_map(fr);
#define GA0(KK) advect_gen_k<T, KK, false>(u, bin, bout, f, N, cm, dt, st)
#define GA1(KK) advect_gen_k<T, KK, true>(u, bin, bout, f, N, cm, dt, st)
    if (soa) {
        GEN_KSWITCH(K, GA1)
    } else {
        GEN_KSWITCH(K, GA0)
    }
    if constexpr (sizeof(T) == 4) {
        if (K == 16) return soa ? GA1(16) : GA0(16);
    }
#undef GA0
#undef GA1
    return hipErrorInvalidValue;
}

template <typename T> hipError_t prepare_gen(int n, int soa)
{
    const int K = gen_k(n, sizeof(T));
#define GP0(KK) gen_attr<T, KK, false>()
#define GP1(KK) gen_attr<T, KK, true>()
    if (soa) {
        GEN_KSWITCH(K, GP1)
    } else {
        GEN_KSWITCH(K, GP0)
    }
    if constexpr (sizeof(T) == 4) {
        if (K == 16) return soa ? GP1(16) : GP0(16);
    }
#undef GP0
#undef GP1
    return hipErrorInvalidValue;
}

template hipError_t launch_fluidx_gen<float>(float *, const float *, const Frame &, long, int, int, const float *,
                                             hipStream_t);
template hipError_t launch_fluidx_gen<double>(double *, const double *, const Frame &, long, int, int, const double *,
                                              hipStream_t);
template hipError_t launch_advect_gen<float>(const float *, const float *, float *, const Frame &, long, int, int,
                                             const float *, hipStream_t);
template hipError_t launch_advect_gen<double>(const double *, const double *, double *, const Frame &, long, int, int,
                                              const double *, hipStream_t);
template hipError_t prepare_gen<float>(int, int);
template hipError_t prepare_gen<double>(int, int);

}  // namespace mhd
