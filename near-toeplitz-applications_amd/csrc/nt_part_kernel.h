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
// subdomain in LDS; solve the reduced system by multiplying with its inverse
// (the same for every line, computed at plan time), one output per thread, so
// no thread walks a serial sweep; every subdomain adds alpha*xU + beta*xL and
// the line goes back to HBM.  Each wavefront stages and stores its own
// subdomain, so no LDS index needs a division by m.
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
  // inverse of the reduced system (2P unknowns), row-major with row stride 2P
  double red_inv[4 * NT_PART_MAX_P * NT_PART_MAX_P];
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
  constexpr int LR = Log2<R>::value;
  // LDS layout of a staged line: subdomain q, element el -> q*SUBP + el + (el/R)*CP
  constexpr int LMAX = NT_PART_MAX_N + (NT_PART_MAX_N / R) * CP + NT_PART_MAX_P * (R + CP + 2);
  __shared__ double lds[ROWS * LMAX];
  __shared__ double red[2 * NT_PART_MAX_P];
  __shared__ double sol[2 * NT_PART_MAX_P];

  const NtPartDev& D = *args.dev;
  const int m = D.m, P = D.P, n = D.n;
  const int tid = threadIdx.x;
  const int p = __builtin_amdgcn_readfirstlane(tid >> 6);     // my subdomain (wave-uniform)
  const int lane = tid & 63;
  const int kind = (P == 1) ? 0 : (p == 0 ? 0 : (p == P - 1 ? 2 : 1));
  const NtPlanT<double>& PL = D.plan[kind];
  const int64_t line = static_cast<int64_t>(blockIdx.x) + HALO;   // staged line index of the system
  const int SUBP = ((m + R - 1) / R) * (R + CP) + 2;             // LDS stride between subdomains
  const int LSTR = P * SUBP;                                       // LDS stride between staged lines
  auto at = [&](int row, int q, int el) -> int { return row * LSTR + q * SUBP + el + (el >> LR) * CP; };

  // ---- stage: wavefront p copies its own subdomain of each staged line -----------
  // (m <= 64 R, so R loads per lane cover it; all are issued before the LDS writes)
  const double* __restrict__ in = static_cast<const double*>(args.in);
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const double* __restrict__ src = in + (line - HALO + r) * args.lstride + p * m;
    double v[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int el = lane + 64 * k;
      v[k] = (el < m) ? __builtin_nontemporal_load(src + el) : 0.0;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int el = lane + 64 * k;
      if (el < m) lds[at(r, p, el)] = v[k];
    }
  }
  __syncthreads();

  // ---- right-hand side in line coordinates: the neighbouring subdomain's values
  // are the ghost points of parallelcompact/rhs.py -------------------------------------
  const int my_el0 = lane * R;          // first row of my chunk inside the subdomain
  double d[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int el = my_el0 + k;
    double v = 0.0;
    if (el < m) {
      const double c = lds[at(HALO, p, el)];
      if (PROD == P_COMPACT) {
        const int e = p * m + el;
        if (e == 0) {
          v = args.s2 * (-5.0 * c + 4.0 * lds[at(0, 0, 1)] + lds[at(0, 0, 2)]);
        } else if (e == n - 1) {
          v = -args.s2 * (-5.0 * c + 4.0 * lds[at(0, P - 1, m - 2)] + lds[at(0, P - 1, m - 3)]);
        } else {
          // neighbours across a subdomain edge (never past the line's ends: e == 0
          // and e == n-1 took the closures above; the clamps keep every index in range)
          const int pl = (p > 0) ? p - 1 : 0, pr = (p < P - 1) ? p + 1 : P - 1;
          const double fl = (el > 0) ? lds[at(0, p, el - 1)] : lds[at(0, pl, m - 1)];
          const double fr = (el < m - 1) ? lds[at(0, p, el + 1)] : lds[at(0, pr, 0)];
          v = args.s1 * (fr - fl);
        }
      } else if (PROD == P_ADI) {
        const int e = p * m + el;
        v = (e == 0 || e == n - 1) ? c : args.s1 * (lds[at(0, p, el)] + lds[at(2, p, el)]) + args.s2 * c;
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

  // ---- reduced system: (alpha_0, beta_0, alpha_1, ...) = inverse * faces -----------
  const int N2 = 2 * P;
  if (tid < N2) {
    const double* __restrict__ row = D.red_inv + tid * N2;
    double u = 0.0;
    for (int j = 0; j < N2; ++j) u += row[j] * red[j];
    sol[tid] = u;
  }
  __syncthreads();

  // ---- correction x = xR + alpha xU + beta xL (sum.py:30) and store -----------------
  const double alpha = sol[2 * p], beta = sol[2 * p + 1];
  const double* __restrict__ xu = args.xu + kind * m;
  const double* __restrict__ xl = args.xl + kind * m;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int el = my_el0 + k;
    if (el < m) lds[at(HALO, p, el)] = d[k] + alpha * xu[el] + beta * xl[el];
  }
  wave_lds_sync();                       // this wavefront reads back only its own subdomain
  double* __restrict__ dst = static_cast<double*>(args.out) + line * args.lstride + p * m;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int el = lane + 64 * k;
    if (el < m) __builtin_nontemporal_store(lds[at(HALO, p, el)], dst + el);
  }
}

}  // namespace nt
