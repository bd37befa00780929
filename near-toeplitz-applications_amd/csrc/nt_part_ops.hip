// Per-subdomain steps of the partition method, unfused, for callers that run
// the reference's sequence one step at a time (run-parallel-compact.py:72-76):
//
//   nt_part_rhs_ghosted  parallelcompact/rhs.py:8-73   computeRHSKernel
//   nt_part_params       parallelcompact/reduced.py:11-196  negateAndCopyFaces +
//                        reduced system + Thomas, for the P subdomains of one
//                        device at once (the reference gathers them over MPI)
//   nt_part_sum          parallelcompact/sum.py:7-51   sumSolutionsKernel
//
// All three are streaming kernels (HBM bound, 16 B per point for the RHS and
// 24 B per point for the in-place sum; the reduced solve touches two faces
// per line and subdomain).  The fused one-launch path is nt_part_solve_lines.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/ntoeplitz.h"
#include "nt_error.h"
#include "nt_part_kernel.h"

namespace nt {

constexpr int OPS_TPB = 256;

// One thread per output point; block b covers points xc*256 .. xc*256+255 of
// row b / xchunks, so the row (and its ghosted offset) is one scalar division.
__global__ void __launch_bounds__(OPS_TPB)
part_rhs_ghosted_kernel(const double* __restrict__ fl, double* __restrict__ x, int64_t ny, int64_t nx,
                        int64_t xchunks, double s1, double s2, int first, int last) {
  const int64_t row = blockIdx.x / xchunks;
  const int64_t xi = (blockIdx.x - row * xchunks) * OPS_TPB + threadIdx.x;
  if (xi >= nx) return;
  const int64_t z = row / ny, y = row - z * ny;
  const double* __restrict__ c = fl + ((z + 1) * (ny + 2) + (y + 1)) * (nx + 2) + xi + 1;
  double v = s1 * (c[1] - c[-1]);
  if (first && xi == 0) v = s2 * (-5.0 * c[0] + 4.0 * c[1] + c[2]);
  if (last && xi == nx - 1) v = -s2 * (-5.0 * c[0] + 4.0 * c[-1] + c[-2]);
  x[row * nx + xi] = v;
}

__global__ void __launch_bounds__(OPS_TPB)
part_sum_kernel(const double* __restrict__ xu, const double* __restrict__ xl, double* __restrict__ xr,
                const double* __restrict__ alpha, const double* __restrict__ beta, int64_t nx, int64_t xchunks) {
  const int64_t row = blockIdx.x / xchunks;
  const int64_t xi = (blockIdx.x - row * xchunks) * OPS_TPB + threadIdx.x;
  if (xi >= nx) return;
  const int64_t i = row * nx + xi;
  xr[i] = xr[i] + alpha[row] * xu[xi] + beta[row] * xl[xi];
}

// One thread per line: the 2P faces of the line's subdomains (negated; zero
// where there is no neighbour), the reduced system of reduced.py:163-178 built
// from the secondary solutions' end values, a serial Thomas sweep
// (reduced.py:11-49), alpha_p = u[2p], beta_p = u[2p+1].
template <int P>
__global__ void __launch_bounds__(OPS_TPB)
part_params_kernel(const double* __restrict__ xu, const double* __restrict__ xl, const double* __restrict__ xr,
                   double* __restrict__ alpha, double* __restrict__ beta, int64_t lines, int64_t m) {
  const int64_t line = static_cast<int64_t>(blockIdx.x) * OPS_TPB + threadIdx.x;
  if (line >= lines) return;
  constexpr int Q = 2 * P;
  double a[Q], b[Q], c[Q], d[Q];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const double u0 = xu[p * m], u1 = xu[p * m + m - 1];
    const double l0 = xl[p * m], l1 = xl[p * m + m - 1];
    a[2 * p] = -1.0;
    a[2 * p + 1] = u1;
    b[2 * p] = u0;
    b[2 * p + 1] = l1;
    c[2 * p] = l0;
    c[2 * p + 1] = -1.0;
    const double* __restrict__ s = xr + (p * lines + line) * m;
    d[2 * p] = (p == 0) ? 0.0 : -s[0];
    d[2 * p + 1] = (p == P - 1) ? 0.0 : -s[m - 1];
  }
  a[0] = 0.0;
  c[0] = 0.0;
  b[0] = 1.0;
  a[Q - 1] = 0.0;
  c[Q - 1] = 0.0;
  b[Q - 1] = 1.0;
  a[1] = 0.0;
  c[Q - 2] = 0.0;
  double c2[Q];
  c2[0] = c[0] / b[0];
  d[0] = d[0] / b[0];
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    const double bmac = b[i] - a[i] * c2[i - 1];
    c2[i] = c[i] / bmac;
    d[i] = (d[i] - a[i] * d[i - 1]) / bmac;
  }
#pragma unroll
  for (int i = Q - 2; i >= 0; --i) d[i] = d[i] - c2[i] * d[i + 1];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    alpha[p * lines + line] = d[2 * p];
    beta[p * lines + line] = d[2 * p + 1];
  }
}

template <int P>
hipError_t launch_params(const double* xu, const double* xl, const double* xr, double* al, double* be,
                         int64_t lines, int64_t m, hipStream_t st) {
  const int64_t grid = (lines + OPS_TPB - 1) / OPS_TPB;
  hipLaunchKernelGGL(part_params_kernel<P>, dim3(static_cast<unsigned>(grid)), dim3(OPS_TPB), 0, st, xu, xl, xr,
                     al, be, lines, m);
  return hipGetLastError();
}

}  // namespace nt

namespace {

bool rows_grid(int64_t rows, int64_t nx, int64_t* xchunks, int64_t* grid) {
  *xchunks = (nx + nt::OPS_TPB - 1) / nt::OPS_TPB;
  *grid = rows * *xchunks;
  return *grid > 0 && *grid <= 0x7fffffffLL;
}

}  // namespace

extern "C" {

int nt_part_rhs_ghosted(const void* f_local, void* x, int64_t nz, int64_t ny, int64_t nx, double dx, int mx,
                        int npx, void* stream) {
  if (!f_local || !x) return nt::set_error(NT_EINVAL, "%s", "nt_part_rhs_ghosted: null pointer");
  if (nz <= 0 || ny <= 0 || nx < 3)
    return nt::set_error(NT_EINVAL, "nt_part_rhs_ghosted: extents (%lld, %lld, %lld); nx must be >= 3",
                         (long long)nz, (long long)ny, (long long)nx);
  if (npx < 1 || mx < 0 || mx >= npx)
    return nt::set_error(NT_EINVAL, "nt_part_rhs_ghosted: rank %d of %d", mx, npx);
  if (!(std::isfinite(dx) && dx != 0.0)) return nt::set_error(NT_EINVAL, "%s", "nt_part_rhs_ghosted: dx");
  int64_t xchunks, grid;
  if (!rows_grid(nz * ny, nx, &xchunks, &grid))
    return nt::set_error(NT_EINVAL, "%s", "nt_part_rhs_ghosted: field too large for one launch");
  hipLaunchKernelGGL(nt::part_rhs_ghosted_kernel, dim3(static_cast<unsigned>(grid)), dim3(nt::OPS_TPB), 0,
                     static_cast<hipStream_t>(stream), static_cast<const double*>(f_local), static_cast<double*>(x),
                     ny, nx, xchunks, 3.0 / (4 * dx), 1.0 / (2 * dx), mx == 0 ? 1 : 0, mx == npx - 1 ? 1 : 0);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return nt::set_error(NT_EHIP, "nt_part_rhs_ghosted: %s", hipGetErrorString(e));
  return nt::set_error(NT_OK, "%s", "");
}

int nt_part_sum(const void* xU, const void* xL, void* xR, const void* alpha, const void* beta, int64_t nz,
                int64_t ny, int64_t nx, void* stream) {
  if (!xU || !xL || !xR || !alpha || !beta) return nt::set_error(NT_EINVAL, "%s", "nt_part_sum: null pointer");
  if (nz <= 0 || ny <= 0 || nx <= 0) return nt::set_error(NT_EINVAL, "%s", "nt_part_sum: extents");
  int64_t xchunks, grid;
  if (!rows_grid(nz * ny, nx, &xchunks, &grid))
    return nt::set_error(NT_EINVAL, "%s", "nt_part_sum: field too large for one launch");
  hipLaunchKernelGGL(nt::part_sum_kernel, dim3(static_cast<unsigned>(grid)), dim3(nt::OPS_TPB), 0,
                     static_cast<hipStream_t>(stream), static_cast<const double*>(xU),
                     static_cast<const double*>(xL), static_cast<double*>(xR), static_cast<const double*>(alpha),
                     static_cast<const double*>(beta), nx, xchunks);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return nt::set_error(NT_EHIP, "nt_part_sum: %s", hipGetErrorString(e));
  return nt::set_error(NT_OK, "%s", "");
}

int nt_part_params(const void* xU, const void* xL, const void* xR, void* alpha, void* beta, int nsub, int64_t nz,
                   int64_t ny, int64_t m, void* stream) {
  if (!xU || !xL || !xR || !alpha || !beta) return nt::set_error(NT_EINVAL, "%s", "nt_part_params: null pointer");
  if (nsub < 1 || nsub > NT_PART_MAX_P)
    return nt::set_error(NT_EINVAL, "nt_part_params: %d subdomains (1..%d)", nsub, NT_PART_MAX_P);
  if (nz <= 0 || ny <= 0 || m <= 0) return nt::set_error(NT_EINVAL, "%s", "nt_part_params: extents");
  const int64_t lines = nz * ny;
  if ((lines + nt::OPS_TPB - 1) / nt::OPS_TPB > 0x7fffffffLL)
    return nt::set_error(NT_EINVAL, "%s", "nt_part_params: too many lines for one launch");
  const auto* u = static_cast<const double*>(xU);
  const auto* l = static_cast<const double*>(xL);
  const auto* r = static_cast<const double*>(xR);
  auto* al = static_cast<double*>(alpha);
  auto* be = static_cast<double*>(beta);
  const auto st = static_cast<hipStream_t>(stream);
  hipError_t e = hipSuccess;
  switch (nsub) {
    case 1: e = nt::launch_params<1>(u, l, r, al, be, lines, m, st); break;
    case 2: e = nt::launch_params<2>(u, l, r, al, be, lines, m, st); break;
    case 3: e = nt::launch_params<3>(u, l, r, al, be, lines, m, st); break;
    case 4: e = nt::launch_params<4>(u, l, r, al, be, lines, m, st); break;
    case 5: e = nt::launch_params<5>(u, l, r, al, be, lines, m, st); break;
    case 6: e = nt::launch_params<6>(u, l, r, al, be, lines, m, st); break;
    case 7: e = nt::launch_params<7>(u, l, r, al, be, lines, m, st); break;
    case 8: e = nt::launch_params<8>(u, l, r, al, be, lines, m, st); break;
    case 9: e = nt::launch_params<9>(u, l, r, al, be, lines, m, st); break;
    case 10: e = nt::launch_params<10>(u, l, r, al, be, lines, m, st); break;
    case 11: e = nt::launch_params<11>(u, l, r, al, be, lines, m, st); break;
    case 12: e = nt::launch_params<12>(u, l, r, al, be, lines, m, st); break;
    case 13: e = nt::launch_params<13>(u, l, r, al, be, lines, m, st); break;
    case 14: e = nt::launch_params<14>(u, l, r, al, be, lines, m, st); break;
    case 15: e = nt::launch_params<15>(u, l, r, al, be, lines, m, st); break;
    default: e = nt::launch_params<16>(u, l, r, al, be, lines, m, st); break;
  }
  if (e != hipSuccess) return nt::set_error(NT_EHIP, "nt_part_params: %s", hipGetErrorString(e));
  return nt::set_error(NT_OK, "%s", "");
}

}  // extern "C"
