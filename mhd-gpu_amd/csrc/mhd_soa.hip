// Sweeps of the device-resident path over a structure-of-arrays state.
//
// The state is 5 planes of u and 3 planes of b, each nx*ny*nz cells, x
// fastest (plane stride N).  A workgroup owns a tile of W pencils (W*sizeof(T)
// = 16 bytes) neighbouring along j; its pencils and their neighbours are
// staged through LDS as [slot][component][cell] with 16-byte vector loads:
// in the x frame a pencil of one component is n contiguous values, in the y
// and z frames (sj == 1) row i of the tile is exactly one 16-byte vector.
// Every thread issues its loads for a batch of items before it writes any of
// them to LDS, so the loads of a tile are in flight together.  The per-pencil
// arithmetic is the shared core (mhd_core.h), unchanged.
#include <stdlib.h>

#include <algorithm>

#include "mhd_generic.h"
#include "mhd_soa_stream.h"
#include "mhd_soa_tiles.h"

namespace mhd {

// y/z frames (LDS tiles); NF and P (pencils per wave, W/P waves) as in
// fluidx_soa_tile
template <typename T, int K, int W, int FR, int NF = 0, int P = 1>
__global__ void __launch_bounds__(64 * W / P)
    fluidx_soa_kernel(T *__restrict__ u, const T *__restrict__ b, SoaBox bx, const T *__restrict__ dtp, int ntj)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    fluidx_soa_tile<T, K, W, FR, 0, NF, P>(u, b, bx, dtp, xcd_tile(blockIdx.x, gridDim.x), ntj,
                                           reinterpret_cast<T *>(smem_raw));
}

template <typename T, int K, int W, int FR>
__global__ void __launch_bounds__(64 * W)
    advect_soa_kernel(const T *__restrict__ u, const T *__restrict__ bin, T *__restrict__ bout, SoaBox bx,
                      const T *__restrict__ dtp, int ntj)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    advect_soa_tile<T, K, W, FR>(u, bin, bout, bx, dtp, xcd_tile(blockIdx.x, gridDim.x), ntj,
                                 reinterpret_cast<T *>(smem_raw));
}

// at most 128 VGPRs for K <= 4 (four waves per SIMD); P pencils per wave
template <typename T, int K, int W, bool VEC, int CFLM, int P = 1>
__global__ void __launch_bounds__(64 * W / P, K <= 4 ? 4 : 1)
    fluidx_x_kernel(T *__restrict__ u, const T *__restrict__ b, SoaBox bx, const T *__restrict__ dtp, int ntj,
                    T *__restrict__ cfl_part, double *__restrict__ mon_part)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    fluidx_x_tile<T, K, W, VEC, CFLM, P>(u, b, bx, dtp, xcd_tile(blockIdx.x, gridDim.x), ntj,
                                         reinterpret_cast<T *>(smem_raw), cfl_part, nullptr, mon_part);
}

template <typename T, int K, int W, bool VEC>
__global__ void __launch_bounds__(64 * W)
    advect_x_kernel(const T *__restrict__ u, const T *__restrict__ bin, T *__restrict__ bout, SoaBox bx,
                    const T *__restrict__ dtp, int ntj)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    advect_x2_tile<T, K, W, VEC>(u, bin, bout, bx, dtp, xcd_tile(blockIdx.x, gridDim.x), ntj,
                                 reinterpret_cast<T *>(smem_raw));
}

// column-streaming advectbyzx of the y and z frames (mhd_soa_stream.h):
// four independent waves per block, blocks dealt XCD-contiguously
template <typename T, int FR>
__global__ void __launch_bounds__(256) advect_stream_a_kernel(const T *__restrict__ u, const T *__restrict__ bin,
                                                              T *__restrict__ bout, T *__restrict__ fa, StreamBox sb,
                                                              const T *__restrict__ dtp, int nwaves)
{
    // the wave index is uniform: readfirstlane keeps every pencil index and
    // row counter derived from it in scalar registers
    const int wave = xcd_tile(blockIdx.x, gridDim.x) * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (wave < nwaves) advect_stream_a<T, FR>(u, bin, bout, fa, sb, dtp, wave);
}

template <typename T, int FR>
__global__ void __launch_bounds__(256) advect_stream_b_kernel(const T *__restrict__ u, const T *__restrict__ bin,
                                                              T *__restrict__ bout, const T *__restrict__ fa,
                                                              StreamBox sb, const T *__restrict__ dtp, int nwaves)
{
    const int wave = xcd_tile(blockIdx.x, gridDim.x) * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (wave < nwaves) advect_stream_b<T, FR>(u, bin, bout, fa, sb, dtp, wave);
}

// both passes fused: one workgroup of ntj <= NW waves per plane pair and
// segment (pencil tiles along j inside the workgroup)
template <typename T, int FR, int NW>
__global__ void __launch_bounds__(64 * NW) advect_stream_ab_kernel(const T *__restrict__ u,
                                                                   const T *__restrict__ bin, T *__restrict__ bout,
                                                                   StreamBox sb, const T *__restrict__ dtp)
{
    __shared__ T xch[64];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int nk2 = (sb.nk + 1) / 2;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    advect_stream_ab<T, FR>(u, bin, bout, sb, dtp, tile % nk2, tile / nk2, w, xch);
}

// ---------------------------------------------------------------------------
// CFL over the SoA state (calcfl.c:13-65): per-cell speed, max -> *cmax bits
// ---------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ unsigned long long soa_bits(T x)
{
    if constexpr (sizeof(T) == 4)
        return (unsigned long long)__float_as_uint(x);
    else
        return (unsigned long long)__double_as_longlong(x);
}

template <typename T, bool AVG>
__global__ void __launch_bounds__(256) cfl_soa_kernel(const T *__restrict__ u, const T *__restrict__ b, SoaGrid g,
                                                      unsigned long long *cmax)
{
    T c = cfl_range<T, AVG>(u, b, g, (long)blockIdx.x * blockDim.x + threadIdx.x, cfl_units<T>(g),
                            (long)gridDim.x * blockDim.x);
    c = wave_cfl_max(c);
    __shared__ T red[4];
    if (lane_id() == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        T m = red[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = cfl_max(m, red[i]);
        atomicMax(cmax, soa_bits(m));
    }
}

// FP64 monitors over the SoA state (mass, energy, momentum x/y/z, max |div B|)
template <typename T>
__global__ void __launch_bounds__(256) monitor_soa_kernel(const T *__restrict__ u, const T *__restrict__ b, SoaGrid g,
                                                          double *part)
{
    const long N = g.N;
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int r = blockIdx.x; r < g.ny * g.nz; r += gridDim.x) {
        const int j = r % g.ny, k = r / g.ny;
        const long row = (long)j * g.sy + (long)k * g.sz;
        for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
            const long s = row + i;
            const long sip = row + (i + 1 == g.nx ? 0 : i + 1);
            const long sjp = (long)(j + 1 == g.ny ? 0 : j + 1) * g.sy + (long)k * g.sz + i;
            const long skp = (long)j * g.sy + (long)(k + 1 == g.nz ? 0 : k + 1) * g.sz + i;
            a[0] += (double)u[s];
            a[1] += (double)u[4 * N + s];
            a[2] += (double)u[N + s];
            a[3] += (double)u[2 * N + s];
            a[4] += (double)u[3 * N + s];
            const double d = ((double)b[sip] - (double)b[s]) + ((double)b[N + sjp] - (double)b[N + s]) +
                             ((double)b[2 * N + skp] - (double)b[2 * N + s]);
            a[5] = fmax(a[5], fabs(d));
        }
    }
    monitor_block_store(a, part + (long)blockIdx.x * 6);
}

// AoS (logical cell order, C per cell) <-> SoA planes on the padded grid
template <typename T, int C>
__global__ void __launch_bounds__(256) aos_soa_kernel(T *__restrict__ soa, T *__restrict__ aos, SoaGrid g, int to_soa)
{
    for (int r = blockIdx.x; r < g.ny * g.nz; r += gridDim.x) {
        const int j = r % g.ny, k = r / g.ny;
        const long row = (long)j * g.sy + (long)k * g.sz;
        const long lrow = (long)r * g.nx;
        for (int e = threadIdx.x; e < g.nx * C; e += blockDim.x) {
            const int i = e / C, c = e - i * C;
            if (to_soa)
                soa[c * g.N + row + i] = aos[(lrow + i) * C + c];
            else
                aos[(lrow + i) * C + c] = soa[c * g.N + row + i];
        }
    }
}

// ===========================================================================
// host side
// ===========================================================================
FastDiv make_fastdiv(int d)
{
    FastDiv r;
    r.one = d <= 1;
    r.m = d <= 1 ? 0u : (unsigned)(((1ull << 32) + (unsigned long long)d - 1) / (unsigned long long)d);
    return r;
}

SoaGrid make_grid(int nx, int ny, int nz, int pad_y, int pad_z)
{
    SoaGrid g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.sy = nx + pad_y;
    g.sz = g.sy * ny + pad_z;
    g.N = g.sz * nz;
    return g;
}

template <typename T> SoaBox make_soa_box(int fr, const SoaGrid &g)
{
    SoaBox bx;
    Frame &f = bx.f;
    if (fr == 0) {
        f.ni = g.nx; f.nj = g.ny; f.nk = g.nz; f.si = 1; f.sj = g.sy; f.sk = g.sz;
    } else if (fr == 1) {
        f.ni = g.ny; f.nj = g.nx; f.nk = g.nz; f.si = g.sy; f.sj = 1; f.sk = g.sz;
    } else {
        f.ni = g.nz; f.nj = g.nx; f.nk = g.ny; f.si = g.sz; f.sj = 1; f.sk = g.sy;
    }
    bx.N = g.N;
    constexpr int VE = 16 / sizeof(T);
    bx.n_div = make_fastdiv(f.ni);
    bx.n4_div = make_fastdiv(f.ni / VE > 0 ? f.ni / VE : 1);
    // 16-byte vectors need every row, plane and tile start on a vector
    // boundary (nx, both strides and the plane size multiples of the vector
    // width) and, in the x frame, whole vectors along the pencil
    bx.vec_ok = (g.nx % VE == 0) && (g.sy % VE == 0) && (g.sz % VE == 0) && (g.N % VE == 0) && (f.ni % VE == 0);
    return bx;
}

template <typename T> constexpr int soa_w() { return 16 / sizeof(T); }

#define SOA_KSWITCH(K, CALL)                                                                                         \
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

template <typename T, int K, int FR> static hipError_t soa_attr()
{
    constexpr int W = soa_w<T>();
    static const hipError_t e = [] {
        if constexpr (FR != 0) {
            for (const void *k : {(const void *)fluidx_soa_kernel<T, K, W, FR>,
                                  (const void *)fluidx_soa_kernel<T, K, 2 * W, FR>,
                                  (const void *)fluidx_soa_kernel<T, K, 2 * W, FR, 64 * K>}) {
                const hipError_t a = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
                if (a != hipSuccess) return a;
            }
        }
        return hipFuncSetAttribute((const void *)advect_soa_kernel<T, K, W, FR>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    }();
    return e;
}

// several short pencils per wave (P > 1): the pencil length is (64/P)*K
template <typename T, int K, int FR, int P> static hipError_t fluidx_p(T *u, const T *b, const SoaBox &bx, const T *dt,
                                                                       hipStream_t st, T *cfl_part, int cflm,
                                                                       double *mon_part)
{
    constexpr int W = soa_w<T>();
    if constexpr (FR == 0) {
        const int ntj = (bx.f.nj + W - 1) / W;
        const size_t xl = sizeof(T) * fluidx_x_words<W>(bx.f.ni, cflm);
        const dim3 grid(ntj * bx.f.nk), block(64 * W / P);
#define MHD_FXP(C)                                                                                                   \
    hipLaunchKernelGGL((fluidx_x_kernel<T, K, W, false, C, P>), grid, block, xl, st, u, b, bx, dt, ntj, cfl_part,    \
                       mon_part)
        if (cflm == 1) MHD_FXP(1);
        else if (cflm == 2) MHD_FXP(2);
        else MHD_FXP(0);
#undef MHD_FXP
    } else {
        constexpr int W2 = 2 * W;
        const int ntj2 = (bx.f.nj + W2 - 1) / W2;
        const size_t lds2 = sizeof(T) * fluidx_soa_words<W2, T>(bx.f.ni);
        hipLaunchKernelGGL((fluidx_soa_kernel<T, K, W2, FR, 0, P>), dim3(ntj2 * bx.f.nk), dim3(64 * W2 / P), lds2,
                           st, u, b, bx, dt, ntj2);
    }
    return hipGetLastError();
}

template <typename T, int K, int FR>
static hipError_t fluidx_soa_k(T *u, const T *b, const SoaBox &bx, const T *dt, hipStream_t st, T *cfl_part,
                               int cflm, double *mon_part)
{
    constexpr int W = soa_w<T>();
    hipError_t e = soa_attr<T, K, FR>();
    if (e != hipSuccess) return e;
    const int ntj = (bx.f.nj + W - 1) / W;
    const size_t lds = sizeof(T) * fluidx_soa_words<W, T>(bx.f.ni);
    const dim3 grid(ntj * bx.f.nk), block(64 * W);
    if constexpr (FR == 0) {
        // x frame: direct loads, no LDS staging (LDS only for the fused CFL)
        constexpr bool kvec = K % (16 / sizeof(T)) == 0;
        const size_t xl = sizeof(T) * fluidx_x_words<W>(bx.f.ni, cflm);
#define MHD_FX(V, C)                                                                                                 \
    hipLaunchKernelGGL((fluidx_x_kernel<T, K, W, V, C>), grid, block, xl, st, u, b, bx, dt, ntj, cfl_part, mon_part)
        if (kvec && bx.vec_ok) {
            if (cflm == 1) MHD_FX(kvec, 1);
            else if (cflm == 2) MHD_FX(kvec, 2);
            else MHD_FX(kvec, 0);
        } else {
            if (cflm == 1) MHD_FX(false, 1);
            else if (cflm == 2) MHD_FX(false, 2);
            else MHD_FX(false, 0);
        }
#undef MHD_FX
    } else {
        // y/z frames: tiles two vectors wide (32-byte rows: half the L2
        // requests of 16-byte rows) where two such tiles fit a CU's LDS
        // (measured on MI355X at 256^3: 0.356-0.363 ms against 0.373 with one
        // vector; four vectors, 148 KB of LDS and one tile per CU, 0.50 ms).
        // MHD_YZ_W=1 selects the one-vector tiles.
        static const int yzw = [] { const char *v = getenv("MHD_YZ_W"); return v ? atoi(v) : 2; }();
        constexpr int W2 = 2 * W;
        const size_t lds2 = sizeof(T) * fluidx_soa_words<W2, T>(bx.f.ni);
        if (yzw >= 2 && lds2 <= 80 * 1024) {
            const int ntj2 = (bx.f.nj + W2 - 1) / W2;
            // a pencil that fills the wave: its length as a constant
            if (bx.f.ni == 64 * K)
                hipLaunchKernelGGL((fluidx_soa_kernel<T, K, W2, FR, 64 * K>), dim3(ntj2 * bx.f.nk), dim3(64 * W2),
                                   lds2, st, u, b, bx, dt, ntj2);
            else
                hipLaunchKernelGGL((fluidx_soa_kernel<T, K, W2, FR>), dim3(ntj2 * bx.f.nk), dim3(64 * W2), lds2, st,
                                   u, b, bx, dt, ntj2);
        } else {
            hipLaunchKernelGGL((fluidx_soa_kernel<T, K, W, FR>), grid, block, lds, st, u, b, bx, dt, ntj);
        }
    }
    return hipGetLastError();
}

template <typename T, int K, int FR>
static hipError_t advect_soa_k(const T *u, const T *bin, T *bout, const SoaBox &bx, const T *dt, hipStream_t st)
{
    constexpr int W = soa_w<T>();
    if constexpr (FR == 0) {
        // x frame: direct loads, LDS only for the fluxes
        // tiles of W pencils in two planes (advect_x2_tile)
        const int ntj = (bx.f.nj + W - 1) / W;
        const size_t lds = sizeof(T) * advect_x2_words<W>(bx.f.ni);
        const dim3 grid(ntj * ((bx.f.nk + 1) / 2));
        constexpr bool kvec = K % (16 / sizeof(T)) == 0;
        if (kvec && bx.vec_ok)
            hipLaunchKernelGGL((advect_x_kernel<T, K, W, kvec>), grid, dim3(64 * W), lds, st, u, bin, bout, bx, dt,
                               ntj);
        else
            hipLaunchKernelGGL((advect_x_kernel<T, K, W, false>), grid, dim3(64 * W), lds, st, u, bin, bout, bx, dt,
                               ntj);
        return hipGetLastError();
    }
    hipError_t e = soa_attr<T, K, FR>();
    if (e != hipSuccess) return e;
    const int ntj = (bx.f.nj + W - 1) / W;
    const size_t lds = sizeof(T) * advect_soa_words<W>(bx.f.ni);
    hipLaunchKernelGGL((advect_soa_kernel<T, K, W, FR>), dim3(ntj * bx.f.nk), dim3(64 * W), lds, st, u, bin, bout, bx,
                       dt, ntj);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fluidx_soa(T *u, const T *b, int fr, const SoaGrid &g, const T *dt, hipStream_t st, T *cfl_part,
                             int cflm, double *mon_part)
{
    if (cflm && fr != 0) return hipErrorInvalidValue;
    const SoaBox bx = make_soa_box<T>(fr, g);
    if constexpr (sizeof(T) == 4) {
        // short pencils that fill a wave P at a time (n = (64/P)*K), where
        // that keeps K or the launch still has a wave per SIMD slot
        const int P = p_pays(bx.f.ni, (long)bx.f.nj * bx.f.nk), KP = P > 1 ? bx.f.ni * P / 64 : 0;
#define SFP(FR, KK, PP) return fluidx_p<T, KK, FR, PP>(u, b, bx, dt, st, cfl_part, cflm, mon_part)
#define SFPF(FR)                                                                                                     \
    if (P == 2 && KP == 1) SFP(FR, 1, 2);                                                                            \
    if (P == 2 && KP == 3) SFP(FR, 3, 2);                                                                            \
    if (P == 4 && KP == 1) SFP(FR, 1, 4);                                                                            \
    if (P == 4 && KP == 3) SFP(FR, 3, 4);
        if (fr == 0) { SFPF(0) }
        else if (fr == 1) { SFPF(1) }
        else { SFPF(2) }
#undef SFPF
#undef SFP
    }
    const int K = pick_k(bx.f.ni);
    // any other pencil length: the generic kernels (no fused CFL)
    if (!K) return cflm ? hipErrorInvalidValue : launch_fluidx_gen<T>(u, b, bx.f, bx.N, 1, fr, dt, st);
#define SF0(KK) fluidx_soa_k<T, KK, 0>(u, b, bx, dt, st, cfl_part, cflm, mon_part)
#define SF1(KK) fluidx_soa_k<T, KK, 1>(u, b, bx, dt, st, nullptr, 0, nullptr)
#define SF2(KK) fluidx_soa_k<T, KK, 2>(u, b, bx, dt, st, nullptr, 0, nullptr)
    if (fr == 0) { SOA_KSWITCH(K, SF0) }
    if (fr == 1) { SOA_KSWITCH(K, SF1) }
    SOA_KSWITCH(K, SF2)
#undef SF0
#undef SF1
#undef SF2
}

template <typename T>
hipError_t launch_advect_soa(const T *u, const T *bin, T *bout, int fr, const SoaGrid &g, const T *dt,
                             hipStream_t st)
{
    const SoaBox bx = make_soa_box<T>(fr, g);
    const int K = pick_k(bx.f.ni);
    if (!K) return launch_advect_gen<T>(u, bin, bout, bx.f, bx.N, 1, fr, dt, st);
#define SA0(KK) advect_soa_k<T, KK, 0>(u, bin, bout, bx, dt, st)
#define SA1(KK) advect_soa_k<T, KK, 1>(u, bin, bout, bx, dt, st)
#define SA2(KK) advect_soa_k<T, KK, 2>(u, bin, bout, bx, dt, st)
    if (fr == 0) { SOA_KSWITCH(K, SA0) }
    if (fr == 1) { SOA_KSWITCH(K, SA1) }
    SOA_KSWITCH(K, SA2)
#undef SA0
#undef SA1
#undef SA2
}

// Rows per segment: enough segments for about 8192 waves of pass A (8 per
// SIMD: the streams are latency-bound), but at least 32 rows each (a segment
// re-reads 8 rows); MHD_STREAM_SEG (rows) overrides.
// MHD_STREAM_AB=0 keeps the two-pass form where the fused one applies
bool stream_fused(const StreamBox &sb)
{
    static const bool on = [] {
        const char *v = getenv("MHD_STREAM_AB");
        return !(v && v[0] == '0');
    }();
    return on && stream_fusable(sb);
}

StreamBox make_stream_box(int fr, const SoaGrid &g)
{
    StreamBox sb;
    sb.N = g.N;
    sb.nj = g.nx;
    if (fr == 1) {
        sb.ni = g.ny; sb.si = g.sy; sb.nk = g.nz; sb.sk = g.sz;
    } else {
        sb.ni = g.nz; sb.si = g.sz; sb.nk = g.ny; sb.sk = g.sy;
    }
    sb.ntj = (sb.nj + 63) / 64;
    static const int env_seg = [] {
        const char *v = getenv("MHD_STREAM_SEG");
        return v ? atoi(v) : 0;
    }();
    int S;
    if (env_seg > 0) {
        S = env_seg;
    } else {
        const long pencils = (long)sb.ntj * sb.nk;
        int nseg = (int)((8192 + pencils - 1) / pencils);
        nseg = std::max(1, std::min(nseg, sb.ni / 32));
        S = (sb.ni + nseg - 1) / nseg;
    }
    S = std::max(1, std::min(S, sb.ni));
    sb.S = S;
    sb.nseg = (sb.ni + S - 1) / S;
    return sb;
}

template <typename T>
hipError_t launch_advect_stream(const T *u, const T *bin, T *bout, T *fa, int fr, const SoaGrid &g, const T *dt,
                                hipStream_t st, int form)
{
    if (fr != 1 && fr != 2) return hipErrorInvalidValue;
    const StreamBox sb = make_stream_box(fr, g);
    if (form == 2 && !stream_fusable(sb)) return hipErrorInvalidValue;
    if (form == 2 || (form == 0 && stream_fused(sb))) {
        const dim3 grid(((sb.nk + 1) / 2) * sb.nseg), block(64 * sb.ntj);
#define SAB(FR, NW) hipLaunchKernelGGL((advect_stream_ab_kernel<T, FR, NW>), grid, block, 0, st, u, bin, bout, sb, dt)
        if (sb.ntj <= 4) {
            if (fr == 1) SAB(1, 4); else SAB(2, 4);
        } else {
            if (fr == 1) SAB(1, 16); else SAB(2, 16);
        }
#undef SAB
        return hipGetLastError();
    }
    const int nwa = stream_waves_a(sb), nwb = stream_waves_b(sb);
    const dim3 ga((nwa + 3) / 4), gb((nwb + 3) / 4), block(256);
    if (fr == 1) {
        hipLaunchKernelGGL((advect_stream_a_kernel<T, 1>), ga, block, 0, st, u, bin, bout, fa, sb, dt, nwa);
        hipLaunchKernelGGL((advect_stream_b_kernel<T, 1>), gb, block, 0, st, u, bin, bout, fa, sb, dt, nwb);
    } else {
        hipLaunchKernelGGL((advect_stream_a_kernel<T, 2>), ga, block, 0, st, u, bin, bout, fa, sb, dt, nwa);
        hipLaunchKernelGGL((advect_stream_b_kernel<T, 2>), gb, block, 0, st, u, bin, bout, fa, sb, dt, nwb);
    }
    return hipGetLastError();
}

template <typename T> hipError_t prepare_soa(int nx, int ny, int nz)
{
    const int n3[3] = {nx, ny, nz};
    for (int fr = 0; fr < 3; fr++) {
        const int K = pick_k(n3[fr]);
        hipError_t e = hipErrorInvalidValue;
#define PS(KK) (fr == 0 ? soa_attr<T, KK, 0>() : fr == 1 ? soa_attr<T, KK, 1>() : soa_attr<T, KK, 2>())
        switch (K) {
        case 1: e = PS(1); break;
        case 2: e = PS(2); break;
        case 3: e = PS(3); break;
        case 4: e = PS(4); break;
        case 5: e = PS(5); break;
        case 6: e = PS(6); break;
        case 8: e = PS(8); break;
        default: e = prepare_gen<T>(n3[fr], 1); break;
        }
#undef PS
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

static int grid_for(long n)
{
    const long nb = (n + 255) / 256;
    return (int)(nb < 4096 ? nb : 4096);
}

static int row_grid(const SoaGrid &g)
{
    const int rows = g.ny * g.nz;
    return rows < 8192 ? rows : 8192;
}

template <typename T>
hipError_t launch_cfl_soa(const T *u, const T *b, const SoaGrid &g, int avg, unsigned long long *cmax, hipStream_t st)
{
    // about 4 units per thread; beyond 8192 blocks the grid strides
    const long units = cfl_units<T>(g);
    const long nb = (units + 1023) / 1024;
    const int grid = (int)(nb < 8192 ? nb : 8192);
    if (avg)
        hipLaunchKernelGGL((cfl_soa_kernel<T, true>), dim3(grid), dim3(256), 0, st, u, b, g, cmax);
    else
        hipLaunchKernelGGL((cfl_soa_kernel<T, false>), dim3(grid), dim3(256), 0, st, u, b, g, cmax);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_monitors_soa(const T *u, const T *b, const SoaGrid &g, double *part, int nblk, double *out,
                               hipStream_t st)
{
    hipLaunchKernelGGL(monitor_soa_kernel<T>, dim3(nblk), dim3(256), 0, st, u, b, g, part);
    if (!out) return hipGetLastError();  // the partials only
    return launch_monitor_reduce(part, nblk, out, nullptr, nullptr, 0, st);
}

template <typename T> hipError_t launch_aos_soa(T *soa, T *aos, int ncomp, const SoaGrid &g, int to_soa, hipStream_t st)
{
    if (ncomp == 5)
        hipLaunchKernelGGL((aos_soa_kernel<T, 5>), dim3(row_grid(g)), dim3(256), 0, st, soa, aos, g, to_soa);
    else
        hipLaunchKernelGGL((aos_soa_kernel<T, 3>), dim3(row_grid(g)), dim3(256), 0, st, soa, aos, g, to_soa);
    return hipGetLastError();
}

#define SOA_INST(T)                                                                                                  \
    template SoaBox make_soa_box<T>(int, const SoaGrid &);                                                          \
    template hipError_t launch_fluidx_soa<T>(T *, const T *, int, const SoaGrid &, const T *, hipStream_t, T *,   \
                                             int, double *);                                                        \
    template hipError_t launch_advect_soa<T>(const T *, const T *, T *, int, const SoaGrid &, const T *,           \
                                             hipStream_t);                                                          \
    template hipError_t prepare_soa<T>(int, int, int);                                                              \
    template hipError_t launch_advect_stream<T>(const T *, const T *, T *, T *, int, const SoaGrid &, const T *,    \
                                                hipStream_t, int);                                                  \
    template hipError_t launch_cfl_soa<T>(const T *, const T *, const SoaGrid &, int, unsigned long long *,        \
                                          hipStream_t);                                                             \
    template hipError_t launch_monitors_soa<T>(const T *, const T *, const SoaGrid &, double *, int, double *,     \
                                               hipStream_t);                                                        \
    template hipError_t launch_aos_soa<T>(T *, T *, int, const SoaGrid &, int, hipStream_t);
SOA_INST(float)
SOA_INST(double)

}  // namespace mhd
