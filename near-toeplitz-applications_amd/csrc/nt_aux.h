// Launchers of the stand-alone (unfused) kernels; see nt_aux_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nt {

struct PermuteGeom {
  int64_t d2, dq;         // output extents of the last axis and of axis q
  int64_t in_2, in_r;     // input strides of output axes 2 and r (axis q has unit stride)
  int64_t out_q, out_r;   // output strides of axes q and r
};

hipError_t launch_compact_rhs(const double* f, double* rhs, int64_t total, int64_t nx, double dx,
                              hipStream_t st);
hipError_t launch_adi_rhs(const double* u, double* d, int64_t N, double dx, double dt,
                          hipStream_t st);
hipError_t launch_permute(const double* a, double* b, const int64_t shape[3], const int perm[3],
                          hipStream_t st);

}  // namespace nt
