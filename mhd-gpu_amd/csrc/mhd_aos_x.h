// Step-level sweeps along x of the reference's AoS arrays with direct loads:
// u[((k*ny+j)*nx+i)*5+c], b[...*3+c] (HIP/array_definition.h:5-7).  A pencil
// along x is n contiguous records, so a wave reads its K cells per lane as
// whole records (K*5 and K*3 consecutive values) straight from global memory,
// selects the components in registers and writes the records back the same
// way; LDS holds only what crosses waves (the EMF fluxes of advectbyzx).
// Same arithmetic, items and update order as the structure-of-arrays x-frame
// tiles (fluidx_x_tile, advect_x2_tile), which the parity tests pin.
#pragma once
#include "mhd_core.h"
#include "mhd_soa_tiles.h"

namespace mhd {

// K cells of C-component records from p (= the first record of the lane) into
// o[q][c]; V is the widest load in values (4: 16-byte loads, p 16-byte
// aligned; 2: 8-byte loads; 1: one value each)
template <typename T, int K, int C, int V>
__device__ __forceinline__ void load_recs(const T *__restrict__ p, T (&o)[K][C])
{
    constexpr int NV = K * C;
    T f[NV];
    if constexpr (V == 4 && sizeof(T) == 4 && NV % 4 == 0) {
#pragma unroll
        for (int v = 0; v < NV / 4; v++) {
            const float4 x = reinterpret_cast<const float4 *>(p)[v];
            f[4 * v] = x.x;
            f[4 * v + 1] = x.y;
            f[4 * v + 2] = x.z;
            f[4 * v + 3] = x.w;
        }
    } else if constexpr (V >= 2 && NV % 2 == 0) {
        using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
#pragma unroll
        for (int v = 0; v < NV / 2; v++) {
            const V2 x = reinterpret_cast<const V2 *>(p)[v];
            f[2 * v] = x.x;
            f[2 * v + 1] = x.y;
        }
    } else {
#pragma unroll
        for (int v = 0; v < NV; v++) f[v] = p[v];
    }
#pragma unroll
    for (int q = 0; q < K; q++)
#pragma unroll
        for (int c = 0; c < C; c++) o[q][c] = f[q * C + c];
}

template <typename T, int K, int C, int V>
__device__ __forceinline__ void store_recs(T *__restrict__ p, const T (&o)[K][C])
{
    constexpr int NV = K * C;
    T f[NV];
#pragma unroll
    for (int q = 0; q < K; q++)
#pragma unroll
        for (int c = 0; c < C; c++) f[q * C + c] = o[q][c];
    if constexpr (V == 4 && sizeof(T) == 4 && NV % 4 == 0) {
#pragma unroll
        for (int v = 0; v < NV / 4; v++)
            reinterpret_cast<float4 *>(p)[v] = make_float4(f[4 * v], f[4 * v + 1], f[4 * v + 2], f[4 * v + 3]);
    } else if constexpr (V >= 2 && NV % 2 == 0) {
        using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
#pragma unroll
        for (int v = 0; v < NV / 2; v++) {
            V2 x;
            x.x = f[2 * v];
            x.y = f[2 * v + 1];
            reinterpret_cast<V2 *>(p)[v] = x;
        }
    } else {
#pragma unroll
        for (int v = 0; v < NV; v++) p[v] = f[v];
    }
}

// fluidx along x (Heterogeneous_MHD/C/fluidx.c:19-102 + tvd1.c + mhdflux.c):
// a tile of W pencils neighbouring along j in plane k, one wave per pencil
template <typename T, int K, int W, int V>
__global__ void __launch_bounds__(64 * W, K <= 4 ? 4 : 1)
    fluidx_aos_x_kernel(T *__restrict__ u, const T *__restrict__ b, Frame f, const T *__restrict__ dtp, int ntj)
{
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int k = tile / ntj, j0 = (tile - k * ntj) * W;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int j = j0 + w;
    if (j >= f.nj) return;
    const int n = f.ni, L = n / K;
    const int l = lane_id(), lc = l % L;
    const int src_next = lc + 1 == L ? 0 : lc + 1, src_prev = lc == 0 ? L - 1 : lc - 1;
    const int jp = j + 1 == f.nj ? 0 : j + 1, kp = k + 1 == f.nk ? 0 : k + 1;
    const long cell = (long)j * f.sj + (long)k * f.sk + (long)lc * K;
    const long cj = (long)jp * f.sj + (long)k * f.sk + (long)lc * K;
    const long ck = (long)j * f.sj + (long)kp * f.sk + (long)lc * K;
    const T dt = *dtp;
    T uu[K][5], bxc[K][3];
    load_recs<T, K, 5, V>(u + cell * 5, uu);
    {
        T br[K][3], bj[K][3], bk[K][3];
        load_recs<T, K, 3, V>(b + cell * 3, br);
        load_recs<T, K, 3, V>(b + cj * 3, bj);  // b2 of pencil j+1
        load_recs<T, K, 3, V>(b + ck * 3, bk);  // b3 of plane k+1
        T b1[K], b1p[K];
#pragma unroll
        for (int q = 0; q < K; q++) b1[q] = br[q][0];
        shift_next<T, K>(b1, b1p, src_next);  // b1 at i+1, periodic along the pencil
#pragma unroll
        for (int q = 0; q < K; q++) {
            // face B to cell centre (fluidx.c:71-81)
            bxc[q][0] = (T)0.5 * b1[q] + (T)0.5 * b1p[q];
            bxc[q][1] = (T)0.5 * br[q][1] + bj[q][1] / (T)2;
            bxc[q][2] = (T)0.5 * br[q][2] + bk[q][2] / (T)2;
        }
    }
    tvd1_pencil<T, K>(uu, bxc, dt, src_next, src_prev);
    if (l < L) store_recs<T, K, 5, V>(u + cell * 5, uu);
}

// LDS words of advect_aos_x2_kernel: the shifted fluxes of a tile's items
// (5W+2 rows) and the new b2, b3 of its pencils (4W rows), which leave as
// whole records with the new b1
template <int W> __host__ __device__ constexpr int advect_aos_x2_words(int n) { return (9 * W + 2) * n; }

// advectbyzx along x over a tile of W pencils in two planes k, k+1 (one where
// k is the last plane of an odd count), as advect_x2_tile: FA of the W+1
// faces along j in each plane, FB at the faces k, k+1, k+2 of the W pencils,
// then b1 = (((b1 - FA_j) + FA_{j+1}) - FB_k) + FB_{k+1} (A1.c:55, A2.c:23,
// B1.c:54, B2.c:23).  bin and bout are distinct: a tile reads the old b of
// neighbouring pencils that other tiles update, as the reference reads its
// copy of b (advectbyzxA1.c:43, B1.c:43).
template <typename T, int K, int W, int V>
__global__ void __launch_bounds__(64 * W)
    advect_aos_x2_kernel(const T *__restrict__ u, const T *__restrict__ bin, T *__restrict__ bout, Frame f,
                         const T *__restrict__ dtp, int ntj)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *X = reinterpret_cast<T *>(smem_raw);
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int n = f.ni, L = n / K;
    // the tile order of advect_x2_tile: groups of 2^G plane pairs by all
    // pencil tiles, plane pairs fastest
    constexpr int G = MHD_X2_GROUP_LOG;
    const int nk2 = (f.nk + 1) / 2, nfull = nk2 >> G, ing = ntj << G;
    int jt, kq;
    if (tile < nfull * ing) {
        const int kg = tile / ing, r = tile - kg * ing;
        jt = r >> G;
        kq = (kg << G) + (r & ((1 << G) - 1));
    } else {
        const int rem = nk2 - (nfull << G), r = tile - nfull * ing;
        jt = r / rem;
        kq = (nfull << G) + (r - jt * rem);
    }
    const int j0 = jt * W, ns = min(W, f.nj - j0);
    const int k = 2 * kq, P = k + 1 < f.nk ? 2 : 1;
    long kpl[4];
#pragma unroll
    for (int p = 0; p < 4; p++) kpl[p] = (long)wrapn(k - 1 + p + f.nk, f.nk) * f.sk;
    const int jn = wrapn(j0 + ns, f.nj), jm = wrapn(j0 - 1 + f.nj, f.nj);
    const int nfa = P * (ns + 1), nrow = nfa + (P + 1) * ns;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), nw = blockDim.x >> 6;
    const int l = lane_id(), lc = l % L;
    const int src_next = lc + 1 == L ? 0 : lc + 1, src_prev = lc == 0 ? L - 1 : lc - 1;
    const long cell0 = (long)lc * K;
    const T dt = *dtp;
    for (int it = w; it < nrow; it += nw) {
        // the item's two pencils (minus side first), its field component and
        // whether its new field value is this tile's
        long pa, pb;
        int fc, own_row;  // own_row: LDS row of the new field value, if it is this tile's
        bool own;
        if (it < nfa) {  // FA: b2 of pencil j0+a in plane k+p, face velocity from j-1, j
            const int p = it / (ns + 1), a = it - p * (ns + 1);
            pa = (long)(a == 0 ? jm : j0 + a - 1) * f.sj + kpl[1 + p];
            pb = (long)(a == ns ? jn : j0 + a) * f.sj + kpl[1 + p];
            fc = 1;
            own = a < ns;
            own_row = p * ns + a;
        } else {  // FB at face k+g: b3 of pencil j0+s in plane k+g, velocity from planes k+g-1, k+g
            const int r = it - nfa, g = r / ns, sp = r - g * ns;
            const long js = (long)(j0 + sp) * f.sj;
            pa = js + kpl[g];
            pb = js + kpl[1 + g];
            fc = 2;
            own = g < P;
            own_row = P * ns + g * ns + sp;
        }
        T ua[K][5], ub[K][5];
        load_recs<T, K, 5, V>(u + (pa + cell0) * 5, ua);
        load_recs<T, K, 5, V>(u + (pb + cell0) * 5, ub);
        T vx[K], vxm[K], vxn[K], vs[K], bi[K], bnew[K], F[K];
#pragma unroll
        for (int q = 0; q < K; q++) {
            // A1.c:34 (u2(jm)+u2(j))/(u1(jm)+u1(j)); the minus-side pencil first
            vx[q] = face_v(ua[q][1], ub[q][1], ua[q][0], ub[q][0]);
            // the item's field component (uniform), one value per record
            bi[q] = bin[(pb + cell0 + q) * 3 + fc];
        }
        shift_prev<T, K>(vx, vxm, src_prev);
        shift_next<T, K>(vx, vxn, src_next);
#pragma unroll
        for (int q = 0; q < K; q++) vs[q] = avg121<T>(vxm[q] + vxn[q], vx[q]);
        tvdb_pencil<T, K>(bi, vs, dt, bnew, F, src_next, src_prev);
        if (l < L) {
            store_cells<T, K>(X + it * n + cell0, F);
            if (own) store_cells<T, K>(X + (nrow + own_row) * n + cell0, bnew);
        }
    }
    __syncthreads();
    // b1 of the tile's pencils, then the whole new record (b1, b2, b3)
    for (int e = w; e < P * ns; e += nw) {
        if (l >= L) continue;
        const int p = e / ns, sp = e - p * ns;
        const long pos = (long)(j0 + sp) * f.sj + kpl[1 + p] + cell0;
        T br[K][3], fa0[K], fa1[K], fb0[K], fb1[K], b2n[K], b3n[K];
        load_recs<T, K, 3, V>(bin + pos * 3, br);
        load_cells<T, K>(X + (p * (ns + 1) + sp) * n + cell0, fa0);
        load_cells<T, K>(X + (p * (ns + 1) + sp + 1) * n + cell0, fa1);
        load_cells<T, K>(X + (nfa + p * ns + sp) * n + cell0, fb0);
        load_cells<T, K>(X + (nfa + (p + 1) * ns + sp) * n + cell0, fb1);
        load_cells<T, K>(X + (nrow + p * ns + sp) * n + cell0, b2n);
        load_cells<T, K>(X + (nrow + P * ns + p * ns + sp) * n + cell0, b3n);
#pragma unroll
        for (int q = 0; q < K; q++) {
            T v = br[q][0];
            v = v - fa0[q];
            v = v + fa1[q];
            v = v - fb0[q];
            v = v + fb1[q];
            br[q][0] = v;
            br[q][1] = b2n[q];
            br[q][2] = b3n[q];
        }
        store_recs<T, K, 3, V>(bout + pos * 3, br);
    }
}

}  // namespace mhd
