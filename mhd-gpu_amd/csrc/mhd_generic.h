// Launchers of the sweeps for any pencil length 2 <= n <= 1024
// (mhd_generic.hip): the fallback where the fast kernels' n = L*K, K <= 8
// does not hold, in the step-level AoS layout (soa = 0, N unused) or over
// the device-resident SoA planes (soa = 1, planes N apart).
#pragma once
#include "mhd_common.h"

namespace mhd {

// frame component -> physical component (u, b) and the inverse maps
struct CompMap {
    int u[5], b[3], iu[5], ib[3];
};
CompMap comp_map(int fr);

// cells per lane of the generic sweeps for a pencil of n cells in a
// precision of prec bytes; 0 if unsupported (n outside [2, 1024], or an FP64
// pencil above 512 cells, whose LDS tiles would not fit)
int gen_k(int n, int prec);

template <typename T>
hipError_t launch_fluidx_gen(T *u, const T *b, const Frame &f, long N, int soa, int fr, const T *dt, hipStream_t st);
// reads bin, writes bout (distinct buffers)
template <typename T>
hipError_t launch_advect_gen(const T *u, const T *bin, T *bout, const Frame &f, long N, int soa, int fr, const T *dt,
                             hipStream_t st);
// the kernels' LDS limit, set before any stream capture
template <typename T> hipError_t prepare_gen(int n, int soa);

}  // namespace mhd
