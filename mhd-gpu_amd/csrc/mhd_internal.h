// Internal helpers shared by the C-ABI translation units.
#pragma once
#include <stdarg.h>

#include "mhd_kernels.h"

namespace mhd {
// record the calling thread's error message and return code
int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *where);
// MHD_OK when the x sweep of an nx*ny*nz box in precision prec (4 or 8
// bytes) has a supported pencil length
int check_shape(int nx, int ny, int nz, int prec);
// the same for all three sweep directions of a box
int check_shape3(int nx, int ny, int nz, int prec);
// FP32 boxes whose three pencil lengths all take the fast kernels (n = L*K,
// K <= 8), as the batched engine needs
int check_shape_fast3(int nx, int ny, int nz);
// default SoA padding (cells) of rows and planes for a precision (4 or 8)
int soa_pad_y(int prec);
int soa_pad_z(int prec);
// set the dynamic-LDS limit of the sweep kernels a box will use, so that no
// attribute call happens inside a stream capture
template <typename T> hipError_t prepare_sweeps(int nx, int ny, int nz);
}  // namespace mhd
