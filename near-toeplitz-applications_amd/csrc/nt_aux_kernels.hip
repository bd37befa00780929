// Stand-alone kernels behind the reference's separate launch sites, kept so
// that the reference call sequences run unchanged on this library:
//   compact/rhs.py:8-53       computeRHS            -> nt_compact_rhs
//   compact/permute.py:6-52   permuteKernel         -> nt_permute (any 3-D shape)
//   adi/rhs.py:9-55           applyVerticalStencil  -> nt_adi_rhs
//   adi/transpose.py:8-50     transposeCoalesced    -> nt_transpose
// The fused paths (nt_compact_derivative, nt_adi_steps) do not use them.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "nt_aux.h"

namespace nt {

__global__ void __launch_bounds__(256)
compact_rhs_kernel(const double* __restrict__ f, double* __restrict__ rhs, int64_t total, int64_t nx,
                   double c34, double c12) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += stride) {
    const int64_t x = i % nx;
    double v;
    if (x == 0) {
      v = c12 * (-5.0 * f[i] + 4.0 * f[i + 1] + f[i + 2]);
    } else if (x == nx - 1) {
      v = -c12 * (-5.0 * f[i] + 4.0 * f[i - 1] + f[i - 2]);
    } else {
      v = c34 * (f[i + 1] - f[i - 1]);
    }
    rhs[i] = v;
  }
}

__global__ void __launch_bounds__(256)
adi_rhs_kernel(const double* __restrict__ u, double* __restrict__ d, int64_t N, double bottom,
               double center, double top) {
  const int64_t total = N * N;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += stride) {
    const int64_t y = i / N, x = i - y * N;
    if (x > 0 && x < N - 1 && y > 0 && y < N - 1)
      d[i] = bottom * u[i - N] + center * u[i] + top * u[i + N];
  }
}

// out[i0][i1][i2] = in[i0*s0 + i1*s1 + i2*s2] where s_d is the input stride of
// the input axis placed at output position d.  When the input's unit-stride
// axis is not the output's last axis (q = its output position), a 32x32 LDS
// tile over output axes (q, 2) keeps both the reads and the writes coalesced.
constexpr int PT = 32;
constexpr int PR = 8;

__global__ void __launch_bounds__(PT * PR)
permute_tile_kernel(const double* __restrict__ in, double* __restrict__ out, PermuteGeom g) {
  __shared__ double tile[PT][PT + 1];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int64_t t2 = static_cast<int64_t>(blockIdx.x) * PT;   // output axis 2
  const int64_t tq = static_cast<int64_t>(blockIdx.y) * PT;   // output axis q
  const int64_t ir = blockIdx.z;                              // the remaining output axis
  // read: i_q fastest (unit stride in the input)
  for (int j = 0; j < PT; j += PR) {
    const int64_t iq = tq + tx, i2 = t2 + ty + j;
    if (iq < g.dq && i2 < g.d2) tile[ty + j][tx] = in[ir * g.in_r + iq + i2 * g.in_2];
  }
  __syncthreads();
  // write: i2 fastest (unit stride in the output)
  for (int j = 0; j < PT; j += PR) {
    const int64_t i2 = t2 + tx, iq = tq + ty + j;
    if (iq < g.dq && i2 < g.d2) out[ir * g.out_r + iq * g.out_q + i2] = tile[tx][ty + j];
  }
}

__global__ void __launch_bounds__(256)
permute_rows_kernel(const double* __restrict__ in, double* __restrict__ out, int64_t d0, int64_t d1,
                    int64_t d2, int64_t s0, int64_t s1) {
  // the input's unit-stride axis stays last: copy rows
  const int64_t total = d0 * d1 * d2;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += stride) {
    const int64_t i2 = i % d2, r = i / d2;
    const int64_t i1 = r % d1, i0 = r / d1;
    out[i] = in[i0 * s0 + i1 * s1 + i2];
  }
}

static unsigned grid_for(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

hipError_t launch_compact_rhs(const double* f, double* rhs, int64_t total, int64_t nx, double dx,
                              hipStream_t st) {
  hipLaunchKernelGGL(compact_rhs_kernel, dim3(grid_for(total)), dim3(256), 0, st, f, rhs, total,
                     nx, 3.0 / (4.0 * dx), 1.0 / (2.0 * dx));
  return hipGetLastError();
}

hipError_t launch_adi_rhs(const double* u, double* d, int64_t N, double dx, double dt,
                          hipStream_t st) {
  const double bottom = -1.0 / (dx * dx), top = -1.0 / (dx * dx);
  const double center = 2.0 * (-1.0 / dt + 1.0 / (dx * dx));
  hipLaunchKernelGGL(adi_rhs_kernel, dim3(grid_for(N * N)), dim3(256), 0, st, u, d, N, bottom,
                     center, top);
  return hipGetLastError();
}

hipError_t launch_permute(const double* a, double* b, const int64_t shape[3], const int perm[3],
                          hipStream_t st) {
  const int64_t in_stride[3] = {shape[1] * shape[2], shape[2], 1};
  const int64_t od[3] = {shape[perm[0]], shape[perm[1]], shape[perm[2]]};
  const int64_t s[3] = {in_stride[perm[0]], in_stride[perm[1]], in_stride[perm[2]]};
  const int64_t total = od[0] * od[1] * od[2];
  if (total == 0) return hipSuccess;
  if (perm[2] == 2) {
    hipLaunchKernelGGL(permute_rows_kernel, dim3(grid_for(total)), dim3(256), 0, st, a, b, od[0],
                       od[1], od[2], s[0], s[1]);
    return hipGetLastError();
  }
  const int q = (perm[0] == 2) ? 0 : 1;   // output position of the input's unit-stride axis
  const int r = 1 - q;                    // the other non-last output axis
  const int64_t out_stride[3] = {od[1] * od[2], od[2], 1};
  PermuteGeom g;
  g.d2 = od[2];
  g.dq = od[q];
  g.in_2 = s[2];
  g.in_r = s[r];
  g.out_q = out_stride[q];
  g.out_r = out_stride[r];
  if (od[r] > 65535) return hipErrorInvalidValue;
  const dim3 grid(static_cast<unsigned>((od[2] + PT - 1) / PT),
                  static_cast<unsigned>((od[q] + PT - 1) / PT), static_cast<unsigned>(od[r]));
  hipLaunchKernelGGL(permute_tile_kernel, grid, dim3(PT, PR), 0, st, a, b, g);
  return hipGetLastError();
}

}  // namespace nt
