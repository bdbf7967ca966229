// Launchers of the structure-of-arrays sweeps (mhd_soa.hip), used by the
// device-resident context, the batched form and the server.  State: u as 5
// planes and b as 3 planes of N = nx*ny*nz cells, x fastest.
#pragma once
#include "mhd_kernels.h"

namespace mhd {

// floor(f / d) as the high word of f * m, m = ceil(2^32 / d), exact for
// 0 <= f < 2^32 / d; d == 1 is flagged (its m does not fit 32 bits)
struct FastDiv {
    unsigned m;
    int one;
};
FastDiv make_fastdiv(int d);

SoaGrid make_grid(int nx, int ny, int nz, int pad_y, int pad_z);

// one box in one sweep frame
struct SoaBox {
    Frame f;
    long N;           // plane stride in cells
    FastDiv n_div;    // division by the pencil length
    FastDiv n4_div;   // division by the pencil length in 16-byte vectors
    bool vec_ok;      // 16-byte vector access applies (nx and n multiples of 16/sizeof(T))
};
template <typename T> SoaBox make_soa_box(int fr, const SoaGrid &g);

// Geometry of a column-streamed advectbyzx (mhd_soa_stream.h): frame 1
// (i,j,k) = (y,x,z), frame 2 (z,x,y)
struct StreamBox {
    long N, si, sk;   // plane stride, row stride along i, stride along k
    int ni, nj, nk;   // pencil length, pencils along j (= nx), along k
    int ntj, nseg, S; // 64-pencil tiles along j, segments per pencil, rows per segment
};
StreamBox make_stream_box(int fr, const SoaGrid &g);
// waves of the two passes: A one pencil per lane, B two planes per lane
inline int stream_waves_a(const StreamBox &sb) { return sb.ntj * sb.nk * sb.nseg; }
inline int stream_waves_b(const StreamBox &sb) { return sb.ntj * ((sb.nk + 1) / 2) * sb.nseg; }
// The streamed passes need many waves to hide their per-row latency: below
// 2048 waves of pass A the LDS-tile kernel is faster (measured on MI355X:
// 128^3, 1024 waves, 0.64 ms/step streamed against 0.52 with tiles; 192^3,
// 3456 waves, 1.48 against 1.55).
inline bool stream_pays(const StreamBox &sb) { return stream_waves_a(sb) >= 2048; }
// both passes in one kernel: the pencils of a row fit one workgroup
inline bool stream_fusable(const StreamBox &sb) { return sb.ntj <= 16; }
bool stream_fused(const StreamBox &sb);

// cflm (x frame only): 1 / 2 also writes the block maxima of the new state's
// CFL speed (face / cell-averaged B) to cfl_part, one per tile (the x
// frame's ntj * nk tiles), and, with mon_part, the tiles' FP64 monitor
// partials (6 per tile: mass, energy, momentum x, y, z, max |div B|)
template <typename T>
hipError_t launch_fluidx_soa(T *u, const T *b, int fr, const SoaGrid &g, const T *dt, hipStream_t st,
                             T *cfl_part = nullptr, int cflm = 0, double *mon_part = nullptr);
// tiles of the x-frame fluidx (entries of its cfl_part)
template <typename T> inline int fluidx_tiles_x(const SoaGrid &g)
{
    const int W = 16 / sizeof(T);
    return (g.ny + W - 1) / W * g.nz;
}
// reads bin, writes bout (distinct buffers)
template <typename T>
hipError_t launch_advect_soa(const T *u, const T *bin, T *bout, int fr, const SoaGrid &g, const T *dt,
                             hipStream_t st);
// advectbyzx of the y (fr 1) or z (fr 2) frame as a column stream: form 0
// fused where stream_fused allows, else two passes; 1 two passes (fa is
// their scratch plane of N values); 2 fused (hipErrorInvalidValue if the
// pencils of a row exceed one workgroup)
template <typename T>
hipError_t launch_advect_stream(const T *u, const T *bin, T *bout, T *fa, int fr, const SoaGrid &g, const T *dt,
                                hipStream_t st, int form = 0);
template <typename T> hipError_t prepare_soa(int nx, int ny, int nz);
template <typename T>
hipError_t launch_cfl_soa(const T *u, const T *b, const SoaGrid &g, int avg, unsigned long long *cmax, hipStream_t st);
template <typename T>
hipError_t launch_monitors_soa(const T *u, const T *b, const SoaGrid &g, double *part, int nblk, double *out,
                               hipStream_t st);
// ncomp 5 or 3; to_soa 1: aos -> soa planes, 0: soa planes -> aos
template <typename T> hipError_t launch_aos_soa(T *soa, T *aos, int ncomp, const SoaGrid &g, int to_soa, hipStream_t st);

}  // namespace mhd
