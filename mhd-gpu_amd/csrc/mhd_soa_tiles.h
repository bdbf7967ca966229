// Tile bodies of the structure-of-arrays sweeps: one workgroup processes one
// tile (W pencils neighbouring along j) of one box.  Used by the one-box
// kernels (mhd_soa.hip) and by the persistent batched kernels (mhd_batch.hip).
#pragma once
#include "mhd_core.h"
#include "mhd_soa.h"

namespace mhd {

template <typename T> struct Vec16;
template <> struct Vec16<float> {
    using type = float4;
    static constexpr int n = 4;
};
template <> struct Vec16<double> {
    using type = double2;
    static constexpr int n = 2;
};
template <typename T> using V16 = typename Vec16<T>::type;

template <typename T> __device__ __forceinline__ T vget(const V16<T> &v, int e)
{
    if constexpr (sizeof(T) == 4)
        return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w;
    else
        return e == 0 ? v.x : v.y;
}

__device__ __forceinline__ int fdiv(int f, FastDiv d) { return d.one ? f : (int)__umulhi((unsigned)f, d.m); }

__device__ __forceinline__ int wrapn(int j, int nj)
{
    while (j >= nj) j -= nj;
    return j;
}

// ---------------------------------------------------------------------------
// CFL (calcfl.c:13-65) over a range of units of one box.  A unit is one
// 16-byte vector of a row (Vec16<T>::n cells) where the layout allows it,
// else one cell; units are numbered row by row (rows j + ny*k), so every
// thread of a block is busy whatever nx is.  Returns the thread's max of the
// per-cell speed (max_2num, calcfl.c:55); the speeds are >= 0.
// ---------------------------------------------------------------------------
template <typename T> __host__ __device__ __forceinline__ bool cfl_vec(const SoaGrid &g)
{
    constexpr int VE = 16 / sizeof(T);
    return g.nx % VE == 0 && g.sy % VE == 0 && g.sz % VE == 0 && g.N % VE == 0;
}
template <typename T> __host__ __device__ __forceinline__ long cfl_units(const SoaGrid &g)
{
    return (long)(cfl_vec<T>(g) ? g.nx / (16 / (int)sizeof(T)) : g.nx) * g.ny * g.nz;
}

template <typename T, bool AVG>
__device__ __forceinline__ T cfl_speed_at(const T *__restrict__ u, const T *__restrict__ b, const SoaGrid &g, long s,
                                          T b1, T b2, T b3, long sip, long sjp, long skp)
{
    const long N = g.N;
    T B;
    if constexpr (AVG) {
        // cell-averaged field (calcfl.c:38-43), C-path arithmetic
        const T bxa = half_of<T>(b1 + b[sip]);
        const T bya = half_of<T>(b2 + b[N + sjp]);
        const T bza = half_of<T>(b3 + b[2 * N + skp]);
        B = bxa * bxa + bya * bya + bza * bza;
    } else {
        B = b1 * b1 + b2 * b2 + b3 * b3;  // face values, calcfl.c:39-50
    }
    return cfl_speed<T>(u[s], u[N + s], u[2 * N + s], u[3 * N + s], u[4 * N + s], B);
}

template <typename T, bool AVG>
__device__ __forceinline__ T cfl_range(const T *__restrict__ u, const T *__restrict__ b, const SoaGrid &g, long e,
                                       long end, long stride)
{
    constexpr int VE = Vec16<T>::n;
    const bool vec = cfl_vec<T>(g);
    const unsigned nv = vec ? g.nx / VE : g.nx;
    const long N = g.N;
    T c = (T)0;
    for (; e < end; e += stride) {
        const unsigned r = (unsigned)e / nv, iu = (unsigned)e - r * nv;
        const unsigned k = r / (unsigned)g.ny, j = r - k * g.ny;
        const long row = (long)j * g.sy + (long)k * g.sz;
        const long rowjp = (long)(j + 1 == (unsigned)g.ny ? 0 : j + 1) * g.sy + (long)k * g.sz;
        const long rowkp = (long)j * g.sy + (long)(k + 1 == (unsigned)g.nz ? 0 : k + 1) * g.sz;
        if (vec) {
            const int i0 = iu * VE;
            const long s = row + i0;
            V16<T> U[5], Bv[3];
#pragma unroll
            for (int q = 0; q < 5; q++) U[q] = *reinterpret_cast<const V16<T> *>(u + q * N + s);
#pragma unroll
            for (int q = 0; q < 3; q++) Bv[q] = *reinterpret_cast<const V16<T> *>(b + q * N + s);
#pragma unroll
            for (int q = 0; q < VE; q++) {
                const T b1 = vget<T>(Bv[0], q), b2 = vget<T>(Bv[1], q), b3 = vget<T>(Bv[2], q);
                T B;
                if constexpr (AVG) {
                    const int ip = i0 + q + 1 == g.nx ? 0 : i0 + q + 1;
                    const T b1p = q + 1 < VE ? vget<T>(Bv[0], q + 1) : b[row + ip];
                    const T bxa = half_of<T>(b1 + b1p);
                    const T bya = half_of<T>(b2 + b[N + rowjp + i0 + q]);
                    const T bza = half_of<T>(b3 + b[2 * N + rowkp + i0 + q]);
                    B = bxa * bxa + bya * bya + bza * bza;
                } else {
                    B = b1 * b1 + b2 * b2 + b3 * b3;
                }
                const T sp = cfl_speed<T>(vget<T>(U[0], q), vget<T>(U[1], q), vget<T>(U[2], q), vget<T>(U[3], q),
                                          vget<T>(U[4], q), B);
                c = cfl_max(c, sp);
            }
        } else {
            const int i = iu;
            const long s = row + i;
            const T sp = cfl_speed_at<T, AVG>(u, b, g, s, b[s], b[N + s], b[2 * N + s],
                                              row + (i + 1 == g.nx ? 0 : i + 1), rowjp + i, rowkp + i);
            c = cfl_max(c, sp);
        }
    }
    return c;
}

// per-wave maxima of the fused CFL (one block-wide array per kernel)
template <typename T> __device__ __forceinline__ T *cfl_red_slot()
{
    __shared__ T red[16];
    return red;
}

// ---------------------------------------------------------------------------
// tile geometry
// ---------------------------------------------------------------------------
struct TileCtx {
    int n, ns, j0, k, kp, km;  // pencil length, valid slots, first pencil, planes
    long kb, kpb, kmb;         // cell offsets of planes k, k+1, k-1
    bool vec;                  // the 16-byte vector path applies
};

// VEW: the vector width in cells (a y/z tile row is whole vectors)
template <int W, int VEW = W> __device__ __forceinline__ TileCtx tile_ctx(const SoaBox &bx, int tile, int ntj)
{
    TileCtx t;
    const Frame &f = bx.f;
    t.n = f.ni;
    t.k = tile / ntj;
    t.j0 = (tile - t.k * ntj) * W;
    t.ns = min(W, f.nj - t.j0);
    t.kp = t.k + 1 == f.nk ? 0 : t.k + 1;
    t.km = t.k == 0 ? f.nk - 1 : t.k - 1;
    t.kb = (long)t.k * f.sk;
    t.kpb = (long)t.kp * f.sk;
    t.kmb = (long)t.km * f.sk;
    // x frame: pencils of n values, n a multiple of the vector width; y/z
    // frames: tile rows of whole vectors starting at a vector boundary
    t.vec = bx.vec_ok && (f.si == 1 || t.ns % VEW == 0);
    return t;
}

// ---------------------------------------------------------------------------
// fluidx (Heterogeneous_MHD/C/fluidx.c:19-102 + tvd1.c + mhdflux.c)
// LDS: su[W][5][n], sb[W][3][n], sb2n[n] (b2 of the pencil after the tile),
//      sbk[W][n] (b3 of the tile's pencils at k+1)
// ---------------------------------------------------------------------------
// Slot strides of the tile's LDS arrays: each slot (pencil) is padded by 16
// bytes, so that the two halves of a tile row, four slots apart, fall in
// different LDS banks when a wave writes them (a row's vector pair) or
// reads them back (the store phase).
template <typename T> __host__ __device__ constexpr int slot_pad() { return 16 / (int)sizeof(T); }
template <typename T> __host__ __device__ constexpr int su_slot(int n) { return 5 * n + slot_pad<T>(); }
template <typename T> __host__ __device__ constexpr int sb_slot(int n) { return 3 * n + slot_pad<T>(); }
template <typename T> __host__ __device__ constexpr int sk_slot(int n) { return n + slot_pad<T>(); }
template <int W, typename T = float> __host__ __device__ constexpr int fluidx_soa_words(int n)
{
    return W * (su_slot<T>(n) + sb_slot<T>(n) + sk_slot<T>(n)) + n;
}

// K consecutive cells from / to p, as 16-byte vectors when K allows (p
// then 16-byte aligned)
template <typename T, int K>
__device__ __forceinline__ void load_cells(const T *__restrict__ p, T (&o)[K])
{
    constexpr int VE = Vec16<T>::n;
    if constexpr (K % VE == 0) {
#pragma unroll
        for (int v = 0; v < K / VE; v++) {
            const V16<T> x = *reinterpret_cast<const V16<T> *>(p + v * VE);
#pragma unroll
            for (int e = 0; e < VE; e++) o[v * VE + e] = vget<T>(x, e);
        }
    } else {
#pragma unroll
        for (int q = 0; q < K; q++) o[q] = p[q];
    }
}

template <typename T, int K>
__device__ __forceinline__ void store_cells(T *__restrict__ p, const T (&o)[K])
{
    constexpr int VE = Vec16<T>::n;
    if constexpr (K % VE == 0 && sizeof(T) == 4) {
#pragma unroll
        for (int v = 0; v < K / VE; v++)
            *reinterpret_cast<float4 *>(p + v * VE) = make_float4(o[v * VE], o[v * VE + 1], o[v * VE + 2], o[v * VE + 3]);
    } else if constexpr (K % VE == 0) {
#pragma unroll
        for (int v = 0; v < K / VE; v++)
            *reinterpret_cast<double2 *>(p + v * VE) = make_double2(o[v * VE], o[v * VE + 1]);
    } else {
#pragma unroll
        for (int q = 0; q < K; q++) p[q] = o[q];
    }
}

// One 16-byte vector of row i of a y/z tile (h-th of the row's H vectors):
// the 5 u and 3 b components at plane k, b3 at k+1 and, with the row's last
// vector, b2 of the pencil after the tile
template <typename T> struct YzRow {
    V16<T> u[5], b[4];
    T b2n;
};

template <typename T, int FR>
__device__ __forceinline__ void yz_row_load(YzRow<T> &r, const T *__restrict__ u, const T *__restrict__ b,
                                            const SoaBox &bx, const TileCtx &t, int jn, int i, int h, bool last)
{
    constexpr int VE = Vec16<T>::n;
    const long N = bx.N, ro = (long)i * bx.f.si + t.j0 + h * VE;
#pragma unroll
    for (int c = 0; c < 5; c++) r.u[c] = *reinterpret_cast<const V16<T> *>(u + cu<FR>(c) * N + t.kb + ro);
#pragma unroll
    for (int c = 0; c < 3; c++) r.b[c] = *reinterpret_cast<const V16<T> *>(b + cb<FR>(c) * N + t.kb + ro);
    r.b[3] = *reinterpret_cast<const V16<T> *>(b + Comp<FR>::b[2] * N + t.kpb + ro);
    r.b2n = (T)0;
    if (last) r.b2n = b[Comp<FR>::b[1] * N + t.kb + (long)i * bx.f.si + jn];
}

// the row vector's VE values into their slots [slot][component][cell]
template <typename T>
__device__ __forceinline__ void yz_row_scatter(const YzRow<T> &r, T *su, T *sb, T *sb2n, T *sbk, int n, int i, int h,
                                               bool last)
{
    constexpr int VE = Vec16<T>::n;
#pragma unroll
    for (int v = 0; v < VE; v++) {
        const int s = h * VE + v;
#pragma unroll
        for (int c = 0; c < 5; c++) su[s * su_slot<T>(n) + c * n + i] = vget<T>(r.u[c], v);
#pragma unroll
        for (int c = 0; c < 3; c++) sb[s * sb_slot<T>(n) + c * n + i] = vget<T>(r.b[c], v);
        sbk[s * sk_slot<T>(n) + i] = vget<T>(r.b[3], v);
    }
    if (last) sb2n[i] = r.b2n;
}

// The tile's pencils from the LDS slots, one pencil per wave, or P per wave
// on lane groups (fluidx.c:61-95): face B to cell centre, tvd1, the new u
// back into su.
template <typename T, int K, int P = 1>
__device__ __forceinline__ void fluidx_lds_pencils(T *su, const T *sb, const T *sb2n, const T *sbk, int n, int ns,
                                                   const T *__restrict__ dtp)
{
    constexpr int G = 64 / P;
    const int L = P == 1 ? n / K : G;
    const int l = lane_id();
    const int w = (threadIdx.x >> 6) * P + (P == 1 ? 0 : l / G);  // the lane's slot
    if (w < ns) {
        const int lc = P == 1 ? l % L : l & (G - 1);  // with P == 1, lanes >= L shadow lane l % L and store nothing
        const int gb = P == 1 ? 0 : l & ~(G - 1);
        const int src_next = gb + (lc + 1 == L ? 0 : lc + 1), src_prev = gb + (lc == 0 ? L - 1 : lc - 1);
        const T dt = *dtp;
        // the lane's K cells as 16-byte LDS reads (lane stride K values:
        // scalar reads of that stride collide in the banks)
        T uu[K][5], bxc[K][3];
#pragma unroll
        for (int m = 0; m < 5; m++) {
            T x[K];
            load_cells<T, K>(su + w * su_slot<T>(n) + m * n + lc * K, x);
#pragma unroll
            for (int q = 0; q < K; q++) uu[q][m] = x[q];
        }
        {
            T b1[K], b1p[K], b2[K], b2p[K], b3[K], b3p[K];
            load_cells<T, K>(sb + w * sb_slot<T>(n) + lc * K, b1);
            load_cells<T, K>(sb + w * sb_slot<T>(n) + n + lc * K, b2);
            load_cells<T, K>((w + 1 < ns ? sb + (w + 1) * sb_slot<T>(n) + n : sb2n) + lc * K, b2p);
            load_cells<T, K>(sb + w * sb_slot<T>(n) + 2 * n + lc * K, b3);
            load_cells<T, K>(sbk + w * sk_slot<T>(n) + lc * K, b3p);
            shift_next<T, K>(b1, b1p, src_next);  // b1 at i+1, periodic along the pencil
#pragma unroll
            for (int q = 0; q < K; q++) {
                // face B to cell centre (fluidx.c:71-81)
                bxc[q][0] = (T)0.5 * b1[q] + (T)0.5 * b1p[q];
                bxc[q][1] = (T)0.5 * b2[q] + b2p[q] / (T)2;
                bxc[q][2] = (T)0.5 * b3[q] + b3p[q] / (T)2;
            }
        }
        if constexpr (P == 1)
            tvd1_pencil<T, K>(uu, bxc, dt, LanesFull{src_next, src_prev});
        else
            tvd1_pencil<T, K>(uu, bxc, dt, LanesGroup<G>{src_next, src_prev});
        if (l < L || P > 1) {
#pragma unroll
            for (int m = 0; m < 5; m++) {
                T x[K];
#pragma unroll
                for (int q = 0; q < K; q++) x[q] = uu[q][m];
                store_cells<T, K>(su + w * su_slot<T>(n) + m * n + lc * K, x);
            }
        }
    }
}

// The new u of a y/z tile back to HBM, component by component (the rows of
// one component leave together, which lets L2 merge the row stores of
// neighbouring tiles): a thread gathers the VE slots of one vector of row i,
// as in the loader.
template <typename T, int FR>
__device__ __forceinline__ void fluidx_yz_store(T *__restrict__ u, const T *su, const SoaBox &bx, const TileCtx &t,
                                                int n)
{
    constexpr int VE = Vec16<T>::n;
    const long N = bx.N;
    const int H = t.ns / VE, tid = threadIdx.x, nt = blockDim.x;
#pragma unroll
    for (int c = 0; c < 5; c++) {
        T *uc = u + cu<FR>(c) * N + t.j0 + t.kb;
        for (int e = tid; e < n * H; e += nt) {
            const int i = H == 2 ? e >> 1 : e, h = e - i * H;
            const int us = su_slot<T>(n);
            const T *sv = su + h * VE * us + c * n + i;
            V16<T> v;
            if constexpr (sizeof(T) == 4)
                v = make_float4(sv[0], sv[us], sv[2 * us], sv[3 * us]);
            else
                v = make_double2(sv[0], sv[us]);
            *reinterpret_cast<V16<T> *>(uc + (long)i * bx.f.si + h * VE) = v;
        }
    }
}

// CFLM (x frame only): 0 = plain sweep; 1 / 2 = also the CFL speed of the
// new state (calcfl.c:45-56, face B / cell-averaged B of OpenMP/mhd.f90:
// 112-120), the block's max written to cfl_part[tile] - the sweep is the
// step's last, so that is the next step's CFL
// NF: the pencil length when the caller fixes it at compile time (n = 64K,
// a full wave: the LDS offsets of the slots become immediates), else 0
template <typename T, int K, int W, int FR, int CFLM = 0, int NF = 0, int P = 1>
__device__ __forceinline__ void fluidx_soa_tile(T *__restrict__ u, const T *__restrict__ b, const SoaBox &bx,
                                                const T *__restrict__ dtp, int tile, int ntj, T *smem,
                                                T *__restrict__ cfl_part = nullptr)
{
    static_assert(CFLM == 0 || FR == 0, "the CFL is fused into the x sweep only");
    static_assert(NF == 0 || NF == 64 * K, "a fixed pencil length fills the wave");
    static_assert(P == 1 || (CFLM == 0 && NF == 0 && W % P == 0), "P pencils per wave: W/P waves, no fused CFL");
    constexpr int VE = Vec16<T>::n;
    static_assert(W == VE || (FR != 0 && W == 2 * VE), "tile width is one 16-byte vector (two in the y/z frames)");
    const Frame &f = bx.f;
    const TileCtx t = tile_ctx<W, VE>(bx, tile, ntj);
    const int n = NF ? NF : t.n, ns = t.ns, L = n / K;
    const long N = bx.N;
    const int us = su_slot<T>(n), bs = sb_slot<T>(n), ks = sk_slot<T>(n);
    T *su = smem, *sb = su + W * us, *sb2n = sb + W * bs, *sbk = sb2n + n;
    const int jn = wrapn(t.j0 + ns, f.nj);  // pencil after the tile (jp of the last one)
    const int tid = threadIdx.x, nt = blockDim.x;

    if (FR == 0 && t.vec) {
        // x frame: items (segment, slot, component, vector index); four items
        // per thread are loaded before any is stored
        const int n4 = n / VE;
        const int eU = ns * 5 * n4, eB = eU + ns * 3 * n4, eN = eB + n4, eK = eN + ns * n4;
        auto item = [&](int e, int &doff) -> const T * {
            if (e < eU) {
                const int sc = fdiv(e, bx.n4_div), i4 = e - sc * n4, s = sc / 5, c = sc - 5 * s;
                doff = s * us + c * n + i4 * VE;
                return u + cu<FR>(c) * N + (long)(t.j0 + s) * f.sj + t.kb + i4 * VE;
            } else if (e < eB) {
                const int r = e - eU, sc = fdiv(r, bx.n4_div), i4 = r - sc * n4, s = sc / 3, c = sc - 3 * s;
                doff = W * us + s * bs + c * n + i4 * VE;
                return b + cb<FR>(c) * N + (long)(t.j0 + s) * f.sj + t.kb + i4 * VE;
            } else if (e < eN) {
                const int i4 = e - eB;
                doff = W * (us + bs) + i4 * VE;
                return b + Comp<FR>::b[1] * N + (long)jn * f.sj + t.kb + i4 * VE;
            } else {
                const int r = e - eN, s = fdiv(r, bx.n4_div), i4 = r - s * n4;
                doff = W * (us + bs) + n + s * ks + i4 * VE;
                return b + Comp<FR>::b[2] * N + (long)(t.j0 + s) * f.sj + t.kpb + i4 * VE;
            }
        };
        for (int f0 = tid; f0 < eK; f0 += 4 * nt) {
            int d0 = -1, d1 = -1, d2 = -1, d3 = -1;
            V16<T> v0, v1, v2, v3;
            if (f0 < eK) v0 = *reinterpret_cast<const V16<T> *>(item(f0, d0));
            if (f0 + nt < eK) v1 = *reinterpret_cast<const V16<T> *>(item(f0 + nt, d1));
            if (f0 + 2 * nt < eK) v2 = *reinterpret_cast<const V16<T> *>(item(f0 + 2 * nt, d2));
            if (f0 + 3 * nt < eK) v3 = *reinterpret_cast<const V16<T> *>(item(f0 + 3 * nt, d3));
            if (d0 >= 0) *reinterpret_cast<V16<T> *>(smem + d0) = v0;
            if (d1 >= 0) *reinterpret_cast<V16<T> *>(smem + d1) = v1;
            if (d2 >= 0) *reinterpret_cast<V16<T> *>(smem + d2) = v2;
            if (d3 >= 0) *reinterpret_cast<V16<T> *>(smem + d3) = v3;
        }
    } else if (FR != 0 && t.vec) {
        // y/z frames: a row i of the tile is H = ns / VE vectors of VE slots
        // (H = 1 or 2).  A thread takes one vector of a row, adjacent threads
        // the adjacent vectors (so that the loads of a row merge into one
        // request), issues its loads (u 5, b 3, b3 at k+1; b2 of the pencil
        // after the tile with the last vector) together and scatters them
        // into the slots
        const int H = ns / VE;
        for (int e = tid; e < n * H; e += nt) {
            const int i = H == 2 ? e >> 1 : e, h = e - i * H;
            const bool last = h == H - 1;
            YzRow<T> r;
            yz_row_load<T, FR>(r, u, b, bx, t, jn, i, h, last);
            yz_row_scatter<T>(r, su, sb, sb2n, sbk, n, i, h, last);
        }
    } else {
        // general tile (partial, or vectors not aligned): one value per item
        const int eU = ns * 5 * n, eB = eU + ns * 3 * n, eN = eB + n, eK = eN + ns * n;
        for (int e = tid; e < eK; e += nt) {
            int s, c, i;
            if (e < eU) {
                const int r = fdiv(e, bx.n_div);
                i = e - r * n;
                s = r / 5;
                c = r - 5 * s;
                su[s * us + c * n + i] =
                    u[cu<FR>(c) * N + t.kb + (long)i * f.si + (long)(t.j0 + s) * f.sj];
            } else if (e < eB) {
                const int q = e - eU, r = fdiv(q, bx.n_div);
                i = q - r * n;
                s = r / 3;
                c = r - 3 * s;
                sb[s * bs + c * n + i] =
                    b[cb<FR>(c) * N + t.kb + (long)i * f.si + (long)(t.j0 + s) * f.sj];
            } else if (e < eN) {
                i = e - eB;
                sb2n[i] = b[Comp<FR>::b[1] * N + t.kb + (long)i * f.si + (long)jn * f.sj];
            } else {
                const int q = e - eN;
                s = fdiv(q, bx.n_div);
                i = q - s * n;
                sbk[s * ks + i] = b[Comp<FR>::b[2] * N + t.kpb + (long)i * f.si + (long)(t.j0 + s) * f.sj];
            }
        }
    }
    __syncthreads();
    fluidx_lds_pencils<T, K, P>(su, sb, sb2n, sbk, n, ns, dtp);
    __syncthreads();
    if constexpr (CFLM != 0) {
        // CFL speed of the new state: after the barrier (a scheduling
        // boundary, so the sweep's registers are free) from the new u in LDS;
        // shadow lanes repeat lane l % L and leave the max unchanged
        T c = (T)0;
        const int w = threadIdx.x >> 6;
        if (w < ns) {
            const int lc = lane_id() % L;
#pragma unroll
            for (int q = 0; q < K; q++) {
                const int i = lc * K + q;
                const T *ui = su + w * us + i;
                const T b1 = sb[w * bs + i], b2 = sb[w * bs + n + i], b3 = sb[w * bs + 2 * n + i];
                T B;
                if constexpr (CFLM == 2) {
                    const int ip = i + 1 == n ? 0 : i + 1;
                    const T b1p = sb[w * bs + ip];
                    const T b2p = w + 1 < ns ? sb[(w + 1) * bs + n + i] : sb2n[i];
                    const T b3p = sbk[w * ks + i];
                    const T bxa = half_of<T>(b1 + b1p);
                    const T bya = half_of<T>(b2 + b2p);
                    const T bza = half_of<T>(b3 + b3p);
                    B = bxa * bxa + bya * bya + bza * bza;
                } else {
                    B = b1 * b1 + b2 * b2 + b3 * b3;
                }
                const T sp = cfl_speed<T>(ui[0], ui[n], ui[2 * n], ui[3 * n], ui[4 * n], B);
                c = cfl_max(c, sp);  // max_2num (calcfl.c:55)
            }
        }
        c = wave_cfl_max(c);
        if (lane_id() == 0) cfl_red_slot<T>()[w] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            T m = cfl_red_slot<T>()[0];
            for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = cfl_max(m, cfl_red_slot<T>()[i]);
            cfl_part[tile] = m;
        }
    }

    // store the new u of the tile's pencils
    if (FR == 0 && t.vec) {
        const int n4 = n / VE, eU = ns * 5 * n4;
        for (int e = tid; e < eU; e += nt) {
            const int sc = fdiv(e, bx.n4_div), i4 = e - sc * n4, s = sc / 5, c = sc - 5 * s;
            *reinterpret_cast<V16<T> *>(u + cu<FR>(c) * N + (long)(t.j0 + s) * f.sj + t.kb + i4 * VE) =
                *reinterpret_cast<const V16<T> *>(su + s * us + c * n + i4 * VE);
        }
    } else if (FR != 0 && t.vec) {
        fluidx_yz_store<T, FR>(u, su, bx, t, n);
    } else {
        for (int e = tid; e < ns * 5 * n; e += nt) {
            const int r = fdiv(e, bx.n_div), i = e - r * n, s = r / 5, c = r - 5 * s;
            u[cu<FR>(c) * N + t.kb + (long)i * f.si + (long)(t.j0 + s) * f.sj] = su[s * us + c * n + i];
        }
    }
}

// ---------------------------------------------------------------------------
// advectbyzx (Heterogeneous_MHD/C/advectbyzx.c:14-23, A1/A2/B1/B2 fused).
// Rows of the flux work area X (3 ns + 1 of them, n values each):
//   row s          FA(s), s = 0..ns   face velocity between pencils j0+s-1, j0+s (plane k)
//   row ns+1+s     FB0(s)             between planes k-1 and k of pencil j0+s
//   row 2ns+1+s    FB1(s)             between planes k and k+1 of pencil j0+s
// Each row holds the smoothed-before face velocity (vx), then the shifted flux F.
// LDS: X[3W+1][n], sb[W][3][n], sb2n[n] (b2 of the pencil after the tile),
//      sb3p[W][n] (b3 of the tile's pencils at k+1)
// ---------------------------------------------------------------------------
template <int W> __host__ __device__ constexpr int advect_soa_words(int n) { return (7 * W + 2) * n; }

template <typename T> __device__ __forceinline__ T face_v(T ma, T mb, T ra, T rb) { return (ma + mb) / (ra + rb); }

template <typename T, int K, int W, int FR>
__device__ __forceinline__ void advect_soa_tile(const T *__restrict__ u, const T *__restrict__ bin,
                                                T *__restrict__ bout, const SoaBox &bx, const T *__restrict__ dtp,
                                                int tile, int ntj, T *smem)
{
    constexpr int VE = Vec16<T>::n;
    static_assert(W == VE, "tile width is one 16-byte vector");
    const Frame &f = bx.f;
    const TileCtx t = tile_ctx<W>(bx, tile, ntj);
    const int n = t.n, ns = t.ns, L = n / K;
    const long N = bx.N;
    T *X = smem, *sb = X + (3 * W + 1) * n, *sb2n = sb + 3 * W * n, *sb3p = sb2n + n;
    const int jn = wrapn(t.j0 + ns, f.nj), jm = wrapn(t.j0 - 1 + f.nj, f.nj);
    const int tid = threadIdx.x, nt = blockDim.x;
    const T *rho = u, *mom = u + Comp<FR>::u[1] * N;
    const int nrow = 3 * ns + 1;

    if (FR == 0 && t.vec) {
        const int n4 = n / VE;
        // face velocities: rows x vectors
        for (int e = tid; e < nrow * n4; e += nt) {
            const int r = fdiv(e, bx.n4_div), i4 = e - r * n4;
            long ca, cb;  // cell offsets of the two pencils (minus side first)
            if (r <= ns) {
                ca = (long)(r == 0 ? jm : t.j0 + r - 1) * f.sj + t.kb;
                cb = (long)(r == ns ? jn : t.j0 + r) * f.sj + t.kb;
            } else if (r <= 2 * ns) {
                const long js = (long)(t.j0 + r - ns - 1) * f.sj;
                ca = js + t.kmb;
                cb = js + t.kb;
            } else {
                const long js = (long)(t.j0 + r - 2 * ns - 1) * f.sj;
                ca = js + t.kb;
                cb = js + t.kpb;
            }
            const long o = i4 * VE;
            const V16<T> ma = *reinterpret_cast<const V16<T> *>(mom + ca + o);
            const V16<T> mb = *reinterpret_cast<const V16<T> *>(mom + cb + o);
            const V16<T> ra = *reinterpret_cast<const V16<T> *>(rho + ca + o);
            const V16<T> rb = *reinterpret_cast<const V16<T> *>(rho + cb + o);
            V16<T> vx;
            if constexpr (sizeof(T) == 4)
                vx = make_float4(face_v(ma.x, mb.x, ra.x, rb.x), face_v(ma.y, mb.y, ra.y, rb.y),
                                 face_v(ma.z, mb.z, ra.z, rb.z), face_v(ma.w, mb.w, ra.w, rb.w));
            else
                vx = make_double2(face_v(ma.x, mb.x, ra.x, rb.x), face_v(ma.y, mb.y, ra.y, rb.y));
            *reinterpret_cast<V16<T> *>(X + r * n + o) = vx;
        }
        // b of the tile (3 comps), b2 after the tile, b3 at k+1
        const int eB = ns * 3 * n4, eN = eB + n4, eK = eN + ns * n4;
        auto item = [&](int e, int &doff) -> const T * {
            if (e < eB) {
                const int sc = fdiv(e, bx.n4_div), i4 = e - sc * n4, s = sc / 3, c = sc - 3 * s;
                doff = (3 * W + 1) * n + (s * 3 + c) * n + i4 * VE;
                return bin + cb<FR>(c) * N + (long)(t.j0 + s) * f.sj + t.kb + i4 * VE;
            } else if (e < eN) {
                const int i4 = e - eB;
                doff = (6 * W + 1) * n + i4 * VE;
                return bin + Comp<FR>::b[1] * N + (long)jn * f.sj + t.kb + i4 * VE;
            } else {
                const int r = e - eN, s = fdiv(r, bx.n4_div), i4 = r - s * n4;
                doff = (6 * W + 2) * n + s * n + i4 * VE;
                return bin + Comp<FR>::b[2] * N + (long)(t.j0 + s) * f.sj + t.kpb + i4 * VE;
            }
        };
        for (int f0 = tid; f0 < eK; f0 += 4 * nt) {
            int d0 = -1, d1 = -1, d2 = -1, d3 = -1;
            V16<T> v0, v1, v2, v3;
            if (f0 < eK) v0 = *reinterpret_cast<const V16<T> *>(item(f0, d0));
            if (f0 + nt < eK) v1 = *reinterpret_cast<const V16<T> *>(item(f0 + nt, d1));
            if (f0 + 2 * nt < eK) v2 = *reinterpret_cast<const V16<T> *>(item(f0 + 2 * nt, d2));
            if (f0 + 3 * nt < eK) v3 = *reinterpret_cast<const V16<T> *>(item(f0 + 3 * nt, d3));
            if (d0 >= 0) *reinterpret_cast<V16<T> *>(smem + d0) = v0;
            if (d1 >= 0) *reinterpret_cast<V16<T> *>(smem + d1) = v1;
            if (d2 >= 0) *reinterpret_cast<V16<T> *>(smem + d2) = v2;
            if (d3 >= 0) *reinterpret_cast<V16<T> *>(smem + d3) = v3;
        }
    } else if (FR != 0 && t.vec) {
        // y/z frames, full tile: one item per row i does every face velocity
        for (int i = tid; i < n; i += nt) {
            const long ro = (long)i * f.si;
            const long ck = t.kb + ro + t.j0, ckm = t.kmb + ro + t.j0, ckp = t.kpb + ro + t.j0;
            const V16<T> m0 = *reinterpret_cast<const V16<T> *>(mom + ck);
            const V16<T> r0 = *reinterpret_cast<const V16<T> *>(rho + ck);
            const V16<T> mm = *reinterpret_cast<const V16<T> *>(mom + ckm);
            const V16<T> rm = *reinterpret_cast<const V16<T> *>(rho + ckm);
            const V16<T> mp = *reinterpret_cast<const V16<T> *>(mom + ckp);
            const V16<T> rp = *reinterpret_cast<const V16<T> *>(rho + ckp);
            const T mlo = mom[t.kb + ro + jm], rlo = rho[t.kb + ro + jm];
            const T mhi = mom[t.kb + ro + jn], rhi = rho[t.kb + ro + jn];
            T mk[W + 2], rk[W + 2];
            mk[0] = mlo;
            rk[0] = rlo;
#pragma unroll
            for (int s = 0; s < W; s++) {
                mk[s + 1] = vget<T>(m0, s);
                rk[s + 1] = vget<T>(r0, s);
            }
            mk[W + 1] = mhi;
            rk[W + 1] = rhi;
#pragma unroll
            for (int s = 0; s <= W; s++) X[s * n + i] = face_v(mk[s], mk[s + 1], rk[s], rk[s + 1]);
#pragma unroll
            for (int s = 0; s < W; s++) {
                X[(W + 1 + s) * n + i] = face_v(vget<T>(mm, s), mk[s + 1], vget<T>(rm, s), rk[s + 1]);
                X[(2 * W + 1 + s) * n + i] = face_v(mk[s + 1], vget<T>(mp, s), rk[s + 1], vget<T>(rp, s));
            }
        }
        const int eB = 3 * n, eK = eB + n;
        for (int f0 = tid; f0 < eK; f0 += 4 * nt) {
            V16<T> v[4];
            T *dbase[4];
            int dstride[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int e = f0 + q * nt;
                dbase[q] = nullptr;
                if (e < eK) {
                    const T *src;
                    if (e < eB) {
                        const int c = fdiv(e, bx.n_div), i = e - c * n;
                        src = bin + cb<FR>(c) * N + t.kb + (long)i * f.si + t.j0;
                        dbase[q] = sb + c * n + i;
                        dstride[q] = 3 * n;
                    } else {
                        const int i = e - eB;
                        src = bin + Comp<FR>::b[2] * N + t.kpb + (long)i * f.si + t.j0;
                        dbase[q] = sb3p + i;
                        dstride[q] = n;
                    }
                    v[q] = *reinterpret_cast<const V16<T> *>(src);
                }
            }
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (dbase[q])
#pragma unroll
                    for (int s = 0; s < W; s++) dbase[q][s * dstride[q]] = vget<T>(v[q], s);
        }
        for (int i = tid; i < n; i += nt) sb2n[i] = bin[Comp<FR>::b[1] * N + t.kb + (long)i * f.si + jn];
    } else {
        // general tile: one value per item
        for (int e = tid; e < nrow * n; e += nt) {
            const int r = fdiv(e, bx.n_div), i = e - r * n;
            const long ro = (long)i * f.si;
            long ca, cb;
            if (r <= ns) {
                ca = (long)(r == 0 ? jm : t.j0 + r - 1) * f.sj + t.kb + ro;
                cb = (long)(r == ns ? jn : t.j0 + r) * f.sj + t.kb + ro;
            } else if (r <= 2 * ns) {
                const long js = (long)(t.j0 + r - ns - 1) * f.sj + ro;
                ca = js + t.kmb;
                cb = js + t.kb;
            } else {
                const long js = (long)(t.j0 + r - 2 * ns - 1) * f.sj + ro;
                ca = js + t.kb;
                cb = js + t.kpb;
            }
            X[r * n + i] = face_v(mom[ca], mom[cb], rho[ca], rho[cb]);
        }
        const int eB = ns * 3 * n, eN = eB + n, eK = eN + ns * n;
        for (int e = tid; e < eK; e += nt) {
            if (e < eB) {
                const int r = fdiv(e, bx.n_div), i = e - r * n, s = r / 3, c = r - 3 * s;
                sb[(s * 3 + c) * n + i] = bin[cb<FR>(c) * N + t.kb + (long)i * f.si + (long)(t.j0 + s) * f.sj];
            } else if (e < eN) {
                const int i = e - eB;
                sb2n[i] = bin[Comp<FR>::b[1] * N + t.kb + (long)i * f.si + (long)jn * f.sj];
            } else {
                const int q = e - eN, s = fdiv(q, bx.n_div), i = q - s * n;
                sb3p[s * n + i] = bin[Comp<FR>::b[2] * N + t.kpb + (long)i * f.si + (long)(t.j0 + s) * f.sj];
            }
        }
    }
    __syncthreads();

    const int w = threadIdx.x >> 6;
    const int l = lane_id();
    const int lc = l % L;
    const int src_next = lc + 1 == L ? 0 : lc + 1, src_prev = lc == 0 ? L - 1 : lc - 1;
    const T dt = *dtp;
    // (the strided vector path numbers rows with W, equal to ns there)
    for (int it = w; it < nrow; it += W) {
        T vx[K], vxm[K], vxn[K], vs[K], bi[K], bnew[K], F[K];
        T *row = X + it * n;
#pragma unroll
        for (int q = 0; q < K; q++) vx[q] = row[lc * K + q];
        shift_prev<T, K>(vx, vxm, src_prev);
        shift_next<T, K>(vx, vxn, src_next);
#pragma unroll
        for (int q = 0; q < K; q++) {
            // A1.c:41 (vx[i-1] + vx[i+1] + 2.0*vx)/4.0
            vs[q] = avg121<T>(vxm[q] + vxn[q], vx[q]);
            const int i = lc * K + q;
            if (it <= ns)
                bi[q] = it < ns ? sb[(it * 3 + 1) * n + i] : sb2n[i];
            else if (it <= 2 * ns)
                bi[q] = sb[((it - ns - 1) * 3 + 2) * n + i];
            else
                bi[q] = sb3p[(it - 2 * ns - 1) * n + i];
        }
        tvdb_pencil<T, K>(bi, vs, dt, bnew, F, src_next, src_prev);
        if (l < L) {
#pragma unroll
            for (int q = 0; q < K; q++) {
                const int i = lc * K + q;
                row[i] = F[q];
                // new b2 (FA) / b3 (FB0) of the tile's own pencils; only this
                // item reads that field row, so it is overwritten in place
                if (it < ns) sb[(it * 3 + 1) * n + i] = bnew[q];
                if (it > ns && it <= 2 * ns) sb[((it - ns - 1) * 3 + 2) * n + i] = bnew[q];
            }
        }
    }
    __syncthreads();
    // b1 = (((b1 - FA_j) + FA_{j+1}) - FB_k) + FB_{k+1}  (A1.c:55, A2.c:23, B1.c:54, B2.c:23)
    if (w < ns && l < L) {
#pragma unroll
        for (int q = 0; q < K; q++) {
            const int i = lc * K + q;
            T b1 = sb[(w * 3) * n + i];
            b1 = b1 - X[w * n + i];
            b1 = b1 + X[(w + 1) * n + i];
            b1 = b1 - X[(ns + 1 + w) * n + i];
            b1 = b1 + X[(2 * ns + 1 + w) * n + i];
            sb[(w * 3) * n + i] = b1;
        }
    }
    __syncthreads();

    if (FR == 0 && t.vec) {
        const int n4 = n / VE, eB = ns * 3 * n4;
        for (int e = tid; e < eB; e += nt) {
            const int sc = fdiv(e, bx.n4_div), i4 = e - sc * n4, s = sc / 3, c = sc - 3 * s;
            *reinterpret_cast<V16<T> *>(bout + cb<FR>(c) * N + (long)(t.j0 + s) * f.sj + t.kb + i4 * VE) =
                *reinterpret_cast<const V16<T> *>(sb + (s * 3 + c) * n + i4 * VE);
        }
    } else if (FR != 0 && t.vec) {
        for (int e = tid; e < 3 * n; e += nt) {
            const int c = fdiv(e, bx.n_div), i = e - c * n;
            V16<T> v;
            if constexpr (sizeof(T) == 4)
                v = make_float4(sb[c * n + i], sb[(3 + c) * n + i], sb[(6 + c) * n + i], sb[(9 + c) * n + i]);
            else
                v = make_double2(sb[c * n + i], sb[(3 + c) * n + i]);
            *reinterpret_cast<V16<T> *>(bout + cb<FR>(c) * N + t.kb + (long)i * f.si + t.j0) = v;
        }
    } else {
        for (int e = tid; e < ns * 3 * n; e += nt) {
            const int r = fdiv(e, bx.n_div), i = e - r * n, s = r / 3, c = r - 3 * s;
            bout[cb<FR>(c) * N + t.kb + (long)i * f.si + (long)(t.j0 + s) * f.sj] = sb[(s * 3 + c) * n + i];
        }
    }
}

// ---------------------------------------------------------------------------
// advectbyzx in the x frame with direct loads: the pencils are contiguous,
// so each wave reads the rows of its tvdb items straight from global memory
// (lane l holds cells l*K .. l*K+K-1, coalesced) and only the EMF fluxes go
// through LDS, for the b1 update that combines the items of four waves.
// Same items, arithmetic and update order as advect_soa_tile.
// LDS: X[3W+1][n] (the shifted fluxes of the items)
// ---------------------------------------------------------------------------
template <int W> __host__ __device__ constexpr int advect_x_words(int n) { return (3 * W + 1) * n; }

// vec: 16-byte loads apply (rows and planes aligned, K a multiple of the
// vector width); otherwise the cells are read one by one
template <typename T, int K, int W, bool VEC>
__device__ __forceinline__ void advect_x_tile(const T *__restrict__ u, const T *__restrict__ bin,
                                              T *__restrict__ bout, const SoaBox &bx, const T *__restrict__ dtp,
                                              int tile, int ntj, T *X)
{
    constexpr int FR = 0;
    const Frame &f = bx.f;
    const TileCtx t = tile_ctx<W>(bx, tile, ntj);
    const int n = t.n, ns = t.ns, L = n / K;
    const long N = bx.N;
    const int jn = wrapn(t.j0 + ns, f.nj), jm = wrapn(t.j0 - 1 + f.nj, f.nj);
    const T *rho = u, *mom = u + Comp<FR>::u[1] * N;
    const T *b1i = bin, *b2i = bin + N, *b3i = bin + 2 * N;
    T *b1o = bout, *b2o = bout + N, *b3o = bout + 2 * N;
    const int nrow = 3 * ns + 1;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), nw = blockDim.x >> 6;
    const int l = lane_id(), lc = l % L;
    const int src_next = lc + 1 == L ? 0 : lc + 1, src_prev = lc == 0 ? L - 1 : lc - 1;
    const long cell0 = (long)lc * K;
    const T dt = *dtp;
    for (int it = w; it < nrow; it += nw) {
        // the item's two pencils (minus side first) and its field row
        long pa, pb, pf;
        const T *fld;
        if (it <= ns) {  // FA: b2 of pencil j0+it with the face velocity from j-1, j
            pa = (long)(it == 0 ? jm : t.j0 + it - 1) * f.sj + t.kb;
            pb = (long)(it == ns ? jn : t.j0 + it) * f.sj + t.kb;
            pf = pb;
            fld = b2i;
        } else if (it <= 2 * ns) {  // FB at face k: b3 of pencil j0+s
            const long js = (long)(t.j0 + it - ns - 1) * f.sj;
            pa = js + t.kmb;
            pb = js + t.kb;
            pf = pb;
            fld = b3i;
        } else {  // FB at face k+1: b3 of pencil j0+s at k+1 (flux only)
            const long js = (long)(t.j0 + it - 2 * ns - 1) * f.sj;
            pa = js + t.kb;
            pb = js + t.kpb;
            pf = pb;
            fld = b3i;
        }
        T ma[K], mb[K], ra[K], rb[K], bi[K];
        if constexpr (VEC) {
            load_cells<T, K>(mom + pa + cell0, ma);
            load_cells<T, K>(mom + pb + cell0, mb);
            load_cells<T, K>(rho + pa + cell0, ra);
            load_cells<T, K>(rho + pb + cell0, rb);
            load_cells<T, K>(fld + pf + cell0, bi);
        } else {
#pragma unroll
            for (int q = 0; q < K; q++) {
                ma[q] = mom[pa + cell0 + q];
                mb[q] = mom[pb + cell0 + q];
                ra[q] = rho[pa + cell0 + q];
                rb[q] = rho[pb + cell0 + q];
                bi[q] = fld[pf + cell0 + q];
            }
        }
        T vx[K], vxm[K], vxn[K], vs[K], bnew[K], F[K];
#pragma unroll
        for (int q = 0; q < K; q++) vx[q] = face_v(ma[q], mb[q], ra[q], rb[q]);
        shift_prev<T, K>(vx, vxm, src_prev);
        shift_next<T, K>(vx, vxn, src_next);
#pragma unroll
        for (int q = 0; q < K; q++) vs[q] = avg121<T>(vxm[q] + vxn[q], vx[q]);
        tvdb_pencil<T, K>(bi, vs, dt, bnew, F, src_next, src_prev);
        if (l < L) {
            store_cells<T, K>(X + it * n + cell0, F);  // 16-byte LDS writes when K allows
            // new b2 (FA) / b3 (FB at k) of the tile's own pencils
            if (it < ns) {
                if constexpr (VEC) store_cells<T, K>(b2o + pf + cell0, bnew);
                else
#pragma unroll
                    for (int q = 0; q < K; q++) b2o[pf + cell0 + q] = bnew[q];
            } else if (it > ns && it <= 2 * ns) {
                if constexpr (VEC) store_cells<T, K>(b3o + pf + cell0, bnew);
                else
#pragma unroll
                    for (int q = 0; q < K; q++) b3o[pf + cell0 + q] = bnew[q];
            }
        }
    }
    __syncthreads();
    // b1 = (((b1 - FA_j) + FA_{j+1}) - FB_k) + FB_{k+1}  (A1.c:55, A2.c:23, B1.c:54, B2.c:23)
    for (int s = w; s < ns; s += nw) {
        if (l >= L) continue;
        const long p = (long)(t.j0 + s) * f.sj + t.kb + cell0;
        T b1[K];
        if constexpr (VEC) load_cells<T, K>(b1i + p, b1);
        else
#pragma unroll
            for (int q = 0; q < K; q++) b1[q] = b1i[p + q];
        T fa0[K], fa1[K], fb0[K], fb1[K];
        load_cells<T, K>(X + s * n + cell0, fa0);
        load_cells<T, K>(X + (s + 1) * n + cell0, fa1);
        load_cells<T, K>(X + (ns + 1 + s) * n + cell0, fb0);
        load_cells<T, K>(X + (2 * ns + 1 + s) * n + cell0, fb1);
#pragma unroll
        for (int q = 0; q < K; q++) {
            T v = b1[q];
            v = v - fa0[q];
            v = v + fa1[q];
            v = v - fb0[q];
            v = v + fb1[q];
            b1[q] = v;
        }
        if constexpr (VEC) store_cells<T, K>(b1o + p, b1);
        else
#pragma unroll
            for (int q = 0; q < K; q++) b1o[p + q] = b1[q];
    }
}

// ---------------------------------------------------------------------------
// fluidx in the x frame with direct loads: a wave's pencil is contiguous, so
// it reads u, its face fields and the neighbours the face average needs
// (b2 of pencil j+1, b3 at k+1) straight from global memory, b1 at i+1 by a
// lane shuffle, and stores u back; no LDS and no barrier.  Same arithmetic
// as fluidx_soa_tile.  CFLM as there: the u of the tile's pencils is kept
// in LDS (5 n words per pencil) for the CFL speed after a barrier.
// P: pencils per wave (n = 64K/P, LanesGroup); the tile's W pencils then
// take W/P waves, wave w holding pencils w*P .. w*P+P-1.
// ---------------------------------------------------------------------------
// LDS of the fused CFL: the new u [W][5][n], the CFL's B^2 per cell [W][n]
// and each lane's max |div B| (one double per thread); afterwards the same
// words hold the monitor partials (6 doubles per thread, at most 64 W
// threads)
template <int W> __host__ __device__ constexpr int fluidx_x_words(int n, int cflm)
{
    return cflm ? (6 * W * n + 128 * W > 768 * W ? 6 * W * n + 128 * W : 768 * W) : 0;
}

template <typename T, int K, int W, bool VEC, int CFLM, int P = 1>
__device__ __forceinline__ void fluidx_x_tile(T *__restrict__ u, const T *__restrict__ b, const SoaBox &bx,
                                              const T *__restrict__ dtp, int tile, int ntj, T *su,
                                              T *__restrict__ cfl_part, unsigned long long *cmax = nullptr,
                                              double *__restrict__ mon_part = nullptr)
{
    static_assert(P == 1 || ((P == 2 || P == 4) && W % P == 0), "pencils per wave: 1, 2 or 4");
    constexpr int G = 64 / P;
    const Frame &f = bx.f;
    const TileCtx t = tile_ctx<W>(bx, tile, ntj);
    const int n = t.n, ns = t.ns, L = P == 1 ? n / K : G;
    const long N = bx.N;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int l = lane_id(), lc = P == 1 ? l % L : l & (G - 1);
    const int ws = w * P + (P == 1 ? 0 : l / G);  // the lane's pencil in the tile
    const int gb = P == 1 ? 0 : l & ~(G - 1);    // first lane of its group
    const int src_next = gb + (lc + 1 == L ? 0 : lc + 1), src_prev = gb + (lc == 0 ? L - 1 : lc - 1);
    const long cell0 = (long)lc * K;
    const long p = (long)(t.j0 + ws) * f.sj + t.kb + cell0;                  // this pencil
    const long pj = (long)wrapn(t.j0 + ws + 1, f.nj) * f.sj + t.kb + cell0;  // pencil j+1
    const long pk = (long)(t.j0 + ws) * f.sj + t.kpb + cell0;                 // plane k+1
    auto load = [&](const T *src, T (&o)[K]) {
        if constexpr (VEC) {
            load_cells<T, K>(src, o);
        } else {
#pragma unroll
            for (int q = 0; q < K; q++) o[q] = src[q];
        }
    };
    T *sB = su + 5 * W * n;
    double *sD = reinterpret_cast<double *>(su + 6 * W * n);
    const bool mon = mon_part != nullptr;
    if (ws < ns) {
        const T dt = *dtp;
        T uu[K][5], bxc[K][3];
        {
            T tmp[K];
#pragma unroll
            for (int m = 0; m < 5; m++) {
                load(u + m * N + p, tmp);
#pragma unroll
                for (int q = 0; q < K; q++) uu[q][m] = tmp[q];
            }
            T b1[K], b1p[K], b2[K], b2p[K], b3[K], b3p[K];
            load(b + p, b1);
            load(b + N + p, b2);
            load(b + 2 * N + p, b3);
            load(b + N + pj, b2p);
            load(b + 2 * N + pk, b3p);
            shift_next<T, K>(b1, b1p, src_next);
#pragma unroll
            for (int q = 0; q < K; q++) {
                // face B to cell centre (fluidx.c:71-81)
                bxc[q][0] = (T)0.5 * b1[q] + (T)0.5 * b1p[q];
                bxc[q][1] = (T)0.5 * b2[q] + b2p[q] / (T)2;
                bxc[q][2] = (T)0.5 * b3[q] + b3p[q] / (T)2;
            }
            if constexpr (CFLM != 0) {
                // fluidx does not write b: the CFL's B^2 of each cell and
                // the lane's max |div B| (left faces) of the monitors are
                // formed now and kept in LDS for after the sweep
                if (l < L || P > 1) {
                    T Bc[K];
#pragma unroll
                    for (int q = 0; q < K; q++) {
                        if constexpr (CFLM == 2) {
                            const T bxa = half_of<T>(b1[q] + b1p[q]);
                            const T bya = half_of<T>(b2[q] + b2p[q]);
                            const T bza = half_of<T>(b3[q] + b3p[q]);
                            Bc[q] = bxa * bxa + bya * bya + bza * bza;
                        } else {
                            Bc[q] = b1[q] * b1[q] + b2[q] * b2[q] + b3[q] * b3[q];
                        }
                    }
                    store_cells<T, K>(sB + ws * n + cell0, Bc);
                    if (mon) {
                        double dm = 0.0;
#pragma unroll
                        for (int q = 0; q < K; q++) {
                            const double d = ((double)b1p[q] - (double)b1[q]) + ((double)b2p[q] - (double)b2[q]) +
                                             ((double)b3p[q] - (double)b3[q]);
                            dm = fmax(dm, fabs(d));
                        }
                        sD[threadIdx.x] = dm;
                    }
                }
            }
        }
        if constexpr (P == 1)
            tvd1_pencil<T, K>(uu, bxc, dt, LanesFull{src_next, src_prev});
        else
            tvd1_pencil<T, K>(uu, bxc, dt, LanesGroup<G>{src_next, src_prev});
        if (l < L || P > 1) {
            // the lane's offsets again, from a copy of the lane index the
            // compiler cannot see through: it then forms the store addresses
            // here rather than keeping the load addresses live (and spilled)
            // across the sweep
            int lo = l;
            asm volatile("" : "+v"(lo));
            const int lcs = P == 1 ? lo % L : lo & (G - 1);
            const long cs = (long)lcs * K;
            const long ps = (long)(t.j0 + ws) * f.sj + t.kb + cs;
#pragma unroll
            for (int m = 0; m < 5; m++) {
                T o[K];
#pragma unroll
                for (int q = 0; q < K; q++) o[q] = uu[q][m];
                if constexpr (VEC) {
                    store_cells<T, K>(u + m * N + ps, o);
                } else {
#pragma unroll
                    for (int q = 0; q < K; q++) u[m * N + ps + q] = o[q];
                }
                if constexpr (CFLM != 0) store_cells<T, K>(su + (ws * 5 + m) * n + cs, o);
            }
        }
    }
    if constexpr (CFLM != 0) {
        // CFL speed of the new state (calcfl.c:45-56), after a barrier so the
        // sweep's registers are free: u and the cells' B^2 from LDS; shadow
        // lanes repeat lane l % L and leave the max unchanged
        __syncthreads();
        T c = (T)0;
        // with mon_part: the FP64 monitors of the new state (mass, energy,
        // momentum x, y, z, max |div B| on the left faces) of the tile's cells
        double ms[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (ws < ns) {
            T um[5][K], Bc[K];
#pragma unroll
            for (int m = 0; m < 5; m++) load_cells<T, K>(su + (ws * 5 + m) * n + cell0, um[m]);
            load_cells<T, K>(sB + ws * n + cell0, Bc);
#pragma unroll
            for (int q = 0; q < K; q++) {
                const T sp = cfl_speed<T>(um[0][q], um[1][q], um[2][q], um[3][q], um[4][q], Bc[q]);
                c = cfl_max(c, sp);  // max_2num (calcfl.c:55)
            }
            if (mon && (l < L || P > 1)) {
#pragma unroll
                for (int q = 0; q < K; q++) {
                    ms[0] += (double)um[0][q];
                    ms[1] += (double)um[4][q];
                    ms[2] += (double)um[1][q];
                    ms[3] += (double)um[2][q];
                    ms[4] += (double)um[3][q];
                }
                ms[5] = sD[threadIdx.x];
            }
        }
        c = wave_cfl_max(c);
        if (lane_id() == 0) cfl_red_slot<T>()[w] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            T m = cfl_red_slot<T>()[0];
            for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = cfl_max(m, cfl_red_slot<T>()[i]);
            // the block max: into the tile's slot, or (FP32, cmax given) into
            // the box's accumulator as float bits (all speeds >= 0: bit
            // order is value order)
            if (sizeof(T) == 4 && cmax)
                atomicMax(cmax, (unsigned long long)__float_as_uint((float)m));
            else
                cfl_part[tile] = m;
        }
        // the monitor record through the dynamic LDS, which every wave has
        // read for the last time before the barrier above (fluidx_x_words
        // leaves room for 6 doubles per thread)
        if (mon) monitor_block_store_lds(ms, mon_part + (long)tile * 6, reinterpret_cast<double *>(su));
    }
}

// ---------------------------------------------------------------------------
// advectbyzx in the x frame over a tile of W pencils in two planes k, k+1
// (one where k is the last plane of an odd count).  The items are FA of the
// W+1 faces along j in each plane, and FB at the faces k, k+1, k+2 of the W
// pencils: the face k+1 serves both planes, so a pencil costs 2.75 tvdb
// calls instead of 3.25 (advect_x_tile).  Same loads, arithmetic and update
// order as advect_x_tile.
// LDS: X[P(W+1) + (P+1)W][n] with P = 2 (the shifted fluxes of the items)
// ---------------------------------------------------------------------------
template <int W> __host__ __device__ constexpr int advect_x2_words(int n) { return (5 * W + 2) * n; }

// log2 of the plane pairs per group in advect_x2_tile's tile order.  At
// 256^3 a group of 16 plane pairs is one XCD's share of the launch (1024
// tiles); measured on MI355X: 0.130-0.133 ms per sweep with 4 plane pairs,
// 0.127-0.129 with 8, 0.122-0.125 with 16, 0.121-0.125 with 32
#ifndef MHD_X2_GROUP_LOG
#define MHD_X2_GROUP_LOG 4
#endif

template <typename T, int K, int W, bool VEC>
__device__ __forceinline__ void advect_x2_tile(const T *__restrict__ u, const T *__restrict__ bin,
                                               T *__restrict__ bout, const SoaBox &bx, const T *__restrict__ dtp,
                                               int tile, int ntj, T *X)
{
    constexpr int FR = 0;
    const Frame &f = bx.f;
    const int n = f.ni, L = n / K;
    // tiles in groups of 2^G plane pairs by all pencil tiles, the plane
    // pairs fastest: a tile's neighbours along j (pencil j0-1, j0+W) and
    // along k (planes k-1, k+2) are mostly blocks close in launch order, so
    // their rows come from L2
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
    // planes k-1 .. k+2, periodic
    long kpl[4];
#pragma unroll
    for (int p = 0; p < 4; p++) kpl[p] = (long)wrapn(k - 1 + p + f.nk, f.nk) * f.sk;
    const long N = bx.N;
    const int jn = wrapn(j0 + ns, f.nj), jm = wrapn(j0 - 1 + f.nj, f.nj);
    const T *rho = u, *mom = u + Comp<FR>::u[1] * N;
    const T *b1i = bin, *b2i = bin + N, *b3i = bin + 2 * N;
    T *b1o = bout, *b2o = bout + N, *b3o = bout + 2 * N;
    const int nfa = P * (ns + 1), nrow = nfa + (P + 1) * ns;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), nw = blockDim.x >> 6;
    const int l = lane_id(), lc = l % L;
    const int src_next = lc + 1 == L ? 0 : lc + 1, src_prev = lc == 0 ? L - 1 : lc - 1;
    const long cell0 = (long)lc * K;
    const T dt = *dtp;
    // An item: its two pencils (minus side first), its field row and its
    // output (the new field value when it is this tile's).  A wave takes the
    // items w, w + nw, ... and loads the operands of its next item while it
    // computes the current one, in two register sets that alternate.
    struct Item {
        long pa, pb;
        const T *fld;
        T *out;
        T ma[K], mb[K], ra[K], rb[K], bi[K];
    };
    auto fetch = [&](int it, Item &m) {
        m.out = nullptr;
        if (it < nfa) {  // FA: b2 of pencil j0+a in plane k+p, face velocity from j-1, j
            const int p = it / (ns + 1), a = it - p * (ns + 1);
            m.pa = (long)(a == 0 ? jm : j0 + a - 1) * f.sj + kpl[1 + p];
            m.pb = (long)(a == ns ? jn : j0 + a) * f.sj + kpl[1 + p];
            m.fld = b2i;
            if (a < ns) m.out = b2o;
        } else {  // FB at face k+g: b3 of pencil j0+s in plane k+g, velocity from planes k+g-1, k+g
            const int r = it - nfa, g = r / ns, sp = r - g * ns;
            const long js = (long)(j0 + sp) * f.sj;
            m.pa = js + kpl[g];
            m.pb = js + kpl[1 + g];
            m.fld = b3i;
            if (g < P) m.out = b3o;
        }
        if constexpr (VEC) {
            load_cells<T, K>(mom + m.pa + cell0, m.ma);
            load_cells<T, K>(mom + m.pb + cell0, m.mb);
            load_cells<T, K>(rho + m.pa + cell0, m.ra);
            load_cells<T, K>(rho + m.pb + cell0, m.rb);
            load_cells<T, K>(m.fld + m.pb + cell0, m.bi);
        } else {
#pragma unroll
            for (int q = 0; q < K; q++) {
                m.ma[q] = mom[m.pa + cell0 + q];
                m.mb[q] = mom[m.pb + cell0 + q];
                m.ra[q] = rho[m.pa + cell0 + q];
                m.rb[q] = rho[m.pb + cell0 + q];
                m.bi[q] = m.fld[m.pb + cell0 + q];
            }
        }
    };
    auto compute = [&](int it, const Item &m) {
        T vx[K], vxm[K], vxn[K], vs[K], bnew[K], F[K];
#pragma unroll
        for (int q = 0; q < K; q++) vx[q] = face_v(m.ma[q], m.mb[q], m.ra[q], m.rb[q]);
        shift_prev<T, K>(vx, vxm, src_prev);
        shift_next<T, K>(vx, vxn, src_next);
#pragma unroll
        for (int q = 0; q < K; q++) vs[q] = avg121<T>(vxm[q] + vxn[q], vx[q]);
        tvdb_pencil<T, K>(m.bi, vs, dt, bnew, F, src_next, src_prev);
        if (l < L) {
            store_cells<T, K>(X + it * n + cell0, F);
            if (m.out) {
                if constexpr (VEC) store_cells<T, K>(m.out + m.pb + cell0, bnew);
                else
#pragma unroll
                    for (int q = 0; q < K; q++) m.out[m.pb + cell0 + q] = bnew[q];
            }
        }
    };
    Item A, B;
    if (w < nrow) fetch(w, A);
    for (int it = w; it < nrow; it += 2 * nw) {
        if (it + nw < nrow) fetch(it + nw, B);
        compute(it, A);
        if (it + nw < nrow) {
            if (it + 2 * nw < nrow) fetch(it + 2 * nw, A);
            compute(it + nw, B);
        }
    }
    __syncthreads();
    // b1 = (((b1 - FA_j) + FA_{j+1}) - FB_k) + FB_{k+1}  (A1.c:55, A2.c:23, B1.c:54, B2.c:23)
    for (int e = w; e < P * ns; e += nw) {
        if (l >= L) continue;
        const int p = e / ns, sp = e - p * ns;
        const long pos = (long)(j0 + sp) * f.sj + kpl[1 + p] + cell0;
        T b1[K];
        if constexpr (VEC) load_cells<T, K>(b1i + pos, b1);
        else
#pragma unroll
            for (int q = 0; q < K; q++) b1[q] = b1i[pos + q];
        T fa0[K], fa1[K], fb0[K], fb1[K];
        load_cells<T, K>(X + (p * (ns + 1) + sp) * n + cell0, fa0);
        load_cells<T, K>(X + (p * (ns + 1) + sp + 1) * n + cell0, fa1);
        load_cells<T, K>(X + (nfa + p * ns + sp) * n + cell0, fb0);
        load_cells<T, K>(X + (nfa + (p + 1) * ns + sp) * n + cell0, fb1);
#pragma unroll
        for (int q = 0; q < K; q++) {
            T v = b1[q];
            v = v - fa0[q];
            v = v + fa1[q];
            v = v - fb0[q];
            v = v + fb1[q];
            b1[q] = v;
        }
        if constexpr (VEC) store_cells<T, K>(b1o + pos, b1);
        else
#pragma unroll
            for (int q = 0; q < K; q++) b1o[pos + q] = b1[q];
    }
}

}  // namespace mhd
