// One-GPU partitioned (superposition) solve: the per-subdomain solves, the
// reduced 2P-unknown system and the correction sum of the reference's
// distributed compact solver (parallelcompact/misc.py:21-46 secondary
// solutions, reduced.py:135-196 reduced system, sum.py:7-32 correction),
// fused into one launch with no host round trip.
//
// A workgroup owns one line of n = P*m points (P <= 16 subdomains of m points),
// one 64-lane wavefront per subdomain, so every wavefront reads a single
// (uniform) level table: the first subdomain keeps the physical first row,
// the last keeps the physical last row, the others use the Toeplitz principal
// block.  Steps: stage the line (and, for ADI, its two neighbour lines)
// through LDS; form the RHS in line coordinates (the "ghost" values of
// parallelcompact/rhs.py are simply the neighbouring subdomain's values);
// solve each subdomain in registers; put the negated end values of every
// subdomain in LDS; one lane runs the reduced system's Thomas sweep with
// factors precomputed at plan time; every subdomain adds
// alpha*xU + beta*xL and the line goes back to HBM.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "nt_line_kernel.h"

#define NT_PART_MAX_P 16
#define NT_PART_MAX_N 2048

namespace nt {

// Device-resident plan data of a partitioned solve (one allocation).
struct NtPartDev {
  NtPlanT<double> plan[3];          // subdomain kinds: 0 first, 1 interior, 2 last
  int32_t nlev;                     // reductions of an m-point subdomain
  int32_t m, P, n;
  double b1, c1, ai, bi, ci, an, bn;
  // reduced system (2P unknowns): a[i] sub-diagonal, ib[i] = 1/(b[i] - a[i] c2[i-1]), c2[i]
  double red_a[2 * NT_PART_MAX_P];
  double red_ib[2 * NT_PART_MAX_P];
  double red_c2[2 * NT_PART_MAX_P];
  // followed in the same allocation by xU[3][m] and xL[3][m]
};

struct PartArgs {
  const NtPartDev* dev;
  const double* xu;     // xU[kind*m + i]
  const double* xl;     // xL[kind*m + i]
  const void* in;
  void* out;
  int64_t lines;        // staged lines (ADI: including the two boundary lines)
  int64_t lstride;      // distance between lines (elements)
  double s1, s2;        // producer scalars as in LineArgs
};

template <int R, int PROD>
__global__ void __launch_bounds__(64 * NT_PART_MAX_P) part_kernel(const PartArgs args) {
  constexpr int HALO = (PROD == P_ADI) ? 1 : 0;
  constexpr int ROWS = 1 + 2 * HALO;
  constexpr int CP = (R >= 4) ? 2 : 0;
  // LDS layout of a staged line: subdomain q, element el -> q*SUBP + el + (el/R)*CP
  constexpr int LMAX = NT_PART_MAX_N + (NT_PART_MAX_N / R) * CP + NT_PART_MAX_P * (R + CP + 2);
  __shared__ double lds[ROWS * LMAX];
  __shared__ double red[2 * NT_PART_MAX_P];

  const NtPartDev& D = *args.dev;
  const int m = D.m, P = D.P, n = D.n;
  const int tid = threadIdx.x;
  const int nthreads = blockDim.x;
  const int p = __builtin_amdgcn_readfirstlane(tid >> 6);     // my subdomain (wave-uniform)
  const int lane = tid & 63;
  const int kind = (P == 1) ? 0 : (p == 0 ? 0 : (p == P - 1 ? 2 : 1));
  const NtPlanT<double>& PL = D.plan[kind];
  const int64_t line = static_cast<int64_t>(blockIdx.x) + HALO;   // staged line index of the system
  const int SUBP = ((m + R - 1) / R) * (R + CP) + 2;             // LDS stride between subdomains
  const int LSTR = P * SUBP;                                       // LDS stride between staged lines

  auto lds_at = [&](int row, int e) -> int {     // e: position along the whole line
    const int q = e / m, el = e - q * m;
    return row * LSTR + q * SUBP + el + (el / R) * CP;
  };

  // ---- stage the line (+ halo lines) -------------------------------------------
  const double* __restrict__ in = static_cast<const double*>(args.in);
  for (int idx = tid; idx < ROWS * n; idx += nthreads) {
    const int r = idx / n, e = idx - r * n;
    lds[lds_at(r, e)] = in[(line - HALO + r) * args.lstride + e];
  }
  __syncthreads();

  // ---- right-hand side in line coordinates ------------------------------------------
  const int my_el0 = lane * R;          // first row of my chunk inside the subdomain
  double d[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int el = my_el0 + k;
    const int e = p * m + el;
    double v = 0.0;
    if (el < m) {
      const double c = lds[lds_at(HALO, e)];
      if (PROD == P_COMPACT) {
        if (e == 0) {
          v = args.s2 * (-5.0 * c + 4.0 * lds[lds_at(0, 1)] + lds[lds_at(0, 2)]);
        } else if (e == n - 1) {
          v = -args.s2 * (-5.0 * c + 4.0 * lds[lds_at(0, n - 2)] + lds[lds_at(0, n - 3)]);
        } else {
          v = args.s1 * (lds[lds_at(0, e + 1)] - lds[lds_at(0, e - 1)]);
        }
      } else if (PROD == P_ADI) {
        v = (e == 0 || e == n - 1)
                ? c
                : args.s1 * (lds[lds_at(0, e)] + lds[lds_at(2, e)]) + args.s2 * c;
      } else {
        v = c;
      }
    }
    d[k] = v;
  }

  // ---- subdomain solve -----------------------------------------------------------
  cr_solve_regs<double, R, 64>(d, PL, D.nlev, lane);

  // ---- faces: negated first/last value of each subdomain (reduced.py:51-88) -------
  if (lane == 0) red[2 * p] = (p == 0) ? 0.0 : -d[0];
  {
    const int last_lane = (m - 1) / R, last_k = (m - 1) % R;
    double v = d[0];
#pragma unroll
    for (int k = 1; k < R; ++k) v = (k == last_k) ? d[k] : v;
    if (lane == last_lane) red[2 * p + 1] = (p == P - 1) ? 0.0 : -v;
  }
  __syncthreads();

  // ---- reduced system: serial Thomas sweep with plan-time factors (reduced.py:11-49)
  if (tid == 0) {
    const int N2 = 2 * P;
    double y = red[0] * D.red_ib[0];
    red[0] = y;
    for (int i = 1; i < N2; ++i) {
      y = (red[i] - D.red_a[i] * y) * D.red_ib[i];
      red[i] = y;
    }
    double x = red[N2 - 1];
    for (int i = N2 - 2; i >= 0; --i) {
      x = red[i] - D.red_c2[i] * x;
      red[i] = x;
    }
  }
  __syncthreads();

  // ---- correction x = xR + alpha xU + beta xL (sum.py:30) and store -----------------
  const double alpha = red[2 * p], beta = red[2 * p + 1];
  const double* __restrict__ xu = args.xu + kind * m;
  const double* __restrict__ xl = args.xl + kind * m;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int el = my_el0 + k;
    if (el < m) lds[lds_at(HALO, p * m + el)] = d[k] + alpha * xu[el] + beta * xl[el];
  }
  __syncthreads();
  double* __restrict__ out = static_cast<double*>(args.out);
  for (int e = tid; e < n; e += nthreads) out[line * args.lstride + e] = lds[lds_at(HALO, e)];
}

}  // namespace nt
