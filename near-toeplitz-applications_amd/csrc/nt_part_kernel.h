// One-GPU partitioned (superposition) solve: the per-subdomain solves, the
// reduced 2P-unknown system and the correction sum of the reference's
// distributed compact solver (parallelcompact/misc.py:21-46 secondary
// solutions, reduced.py:135-196 reduced system, sum.py:7-32 correction),
// fused into one launch with no host round trip.
//
// A workgroup owns one line of n = P*m points (P <= 16 subdomains of m points).
// Each subdomain is solved by a segment of W lanes (W*R >= m), so a wavefront
// holds 64/W subdomains.  A wavefront reads one (uniform) level table, so the
// first subdomain (which keeps the physical first row) and the last (physical
// last row) get a wavefront each, and the interior subdomains (the Toeplitz
// principal block) fill the wavefronts in between.  Steps:
//  * stage the line through LDS; the ADI right-hand side (the explicit stencil
//    across lines) is formed while loading;
//  * form the compact RHS in line coordinates (the "ghost" values of
//    parallelcompact/rhs.py are simply the neighbouring subdomain's values);
//  * solve each subdomain in registers;
//  * put the negated end values of every subdomain in LDS;
//  * solve the reduced system by multiplying with its inverse (the same for
//    every line, computed at plan time), one output per thread, so no thread
//    walks a serial sweep;
//  * every subdomain adds alpha*xU + beta*xL and the line goes back to HBM.
// Each segment stages and stores its own subdomain, so no LDS index needs a
// division by m.
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

// Wavefronts per line for P subdomains, 64/W per wavefront: one for the first,
// one for the last, the interior ones packed in between.
__host__ __device__ inline int part_waves(int P, int W) {
  const int S = 64 / W;
  return (P == 1) ? 1 : 2 + (P - 2 + S - 1) / S;
}

// Threads per workgroup at most: subdomains of more than 512 points (R >= 16 on
// 64 lanes) leave room for at most 3 on a line of NT_PART_MAX_N points.
template <int R>
constexpr int part_max_threads() {
  return (R >= 16) ? 256 : 64 * NT_PART_MAX_P;
}

template <int R, int W, int PROD>
__global__ void __launch_bounds__(part_max_threads<R>()) part_kernel(const PartArgs args) {
  constexpr int HALO = (PROD == P_ADI) ? 1 : 0;
  constexpr int S = 64 / W;                      // subdomains per wavefront
  constexpr int CP = (R >= 4) ? 2 : 0;
  constexpr int LR = Log2<R>::value;
  constexpr int LW = Log2<W>::value;
  // LDS layout of the staged line: subdomain q, element el -> q*SUBP + el + (el/R)*CP
  constexpr int LMAX = NT_PART_MAX_N + (NT_PART_MAX_N / R) * CP + NT_PART_MAX_P * (R + CP + 2);
  __shared__ double lds[LMAX];
  __shared__ double red[2 * NT_PART_MAX_P];
  __shared__ double sol[2 * NT_PART_MAX_P];

  const NtPartDev& D = *args.dev;
  const int m = D.m, P = D.P, n = D.n;
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);        // my wavefront (uniform)
  const int nw = __builtin_amdgcn_readfirstlane(static_cast<int>(blockDim.x) >> 6);
  const int lane = tid & 63;
  const int seg = lane >> LW;                    // my segment in the wavefront
  const int sl = lane & (W - 1);                 // my lane in the segment
  // my subdomain q (-1: an idle segment) and the kind of the wavefront's level table
  int kind, q;
  if (P == 1 || w == 0) {
    kind = 0;
    q = (seg == 0) ? 0 : -1;
  } else if (w == nw - 1) {
    kind = 2;
    q = (seg == 0) ? P - 1 : -1;
  } else {
    kind = 1;
    q = 1 + (w - 1) * S + seg;
    if (q > P - 2) q = -1;
  }
  const bool active = q >= 0;
  const int qs = active ? q : 0;                 // a valid index for idle segments' addresses
  const NtPlanT<double>& PL = D.plan[kind];
  const int64_t line = static_cast<int64_t>(blockIdx.x) + HALO;   // staged line index of the system
  const int SUBP = ((m + R - 1) / R) * (R + CP) + 2;             // LDS stride between subdomains
  auto at = [&](int qq, int el) -> int { return qq * SUBP + el + (el >> LR) * CP; };

  // ---- stage: each segment copies its own subdomain (m <= W R, so R loads per
  // lane cover it; all are issued before the LDS writes).  ADI: the right-hand
  // side of the x-implicit half step is formed here from the line and its two
  // neighbours at the same positions.
  const double* __restrict__ src = static_cast<const double*>(args.in) + line * args.lstride + qs * m;
  {
    double v[R];
    if (PROD == P_ADI) {
      double up[R], dn[R];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int el = sl + W * k;
        const bool ok = active && el < m;
        v[k] = ok ? __builtin_nontemporal_load(src + el) : 0.0;
        up[k] = ok ? __builtin_nontemporal_load(src - args.lstride + el) : 0.0;
        dn[k] = ok ? __builtin_nontemporal_load(src + args.lstride + el) : 0.0;
      }
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int e = qs * m + sl + W * k;
        v[k] = (e == 0 || e == n - 1) ? v[k] : args.s1 * (up[k] + dn[k]) + args.s2 * v[k];
      }
    } else {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int el = sl + W * k;
        v[k] = (active && el < m) ? __builtin_nontemporal_load(src + el) : 0.0;
      }
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int el = sl + W * k;
      if (active && el < m) lds[at(q, el)] = v[k];
    }
  }
  __syncthreads();

  // ---- right-hand side in line coordinates: the neighbouring subdomain's values
  // are the ghost points of parallelcompact/rhs.py ---------------------------------
  const int my_el0 = sl * R;            // first row of my chunk inside the subdomain
  double d[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int el = my_el0 + k;
    double v = 0.0;
    if (active && el < m) {
      const double c = lds[at(q, el)];
      if (PROD == P_COMPACT) {
        const int e = q * m + el;
        if (e == 0) {
          v = args.s2 * (-5.0 * c + 4.0 * lds[at(0, 1)] + lds[at(0, 2)]);
        } else if (e == n - 1) {
          v = -args.s2 * (-5.0 * c + 4.0 * lds[at(P - 1, m - 2)] + lds[at(P - 1, m - 3)]);
        } else {
          // neighbours across a subdomain edge (never past the line's ends: e == 0
          // and e == n-1 took the closures above; the clamps keep every index in range)
          const int pl = (q > 0) ? q - 1 : 0, pr = (q < P - 1) ? q + 1 : P - 1;
          const double fl = (el > 0) ? lds[at(q, el - 1)] : lds[at(pl, m - 1)];
          const double fr = (el < m - 1) ? lds[at(q, el + 1)] : lds[at(pr, 0)];
          v = args.s1 * (fr - fl);
        }
      } else {
        v = c;                          // plain solve, or the ADI RHS formed at staging
      }
    }
    d[k] = v;
  }

  // ---- subdomain solve (idle segments solve zeros) --------------------------------
  cr_solve_regs<double, R, W>(d, PL, D.nlev, sl);

  // ---- faces: negated first/last value of each subdomain (reduced.py:51-88) -------
  if (active && sl == 0) red[2 * q] = (q == 0) ? 0.0 : -d[0];
  {
    const int last_lane = (m - 1) / R, last_k = (m - 1) % R;
    double v = d[0];
#pragma unroll
    for (int k = 1; k < R; ++k) v = (k == last_k) ? d[k] : v;
    if (active && sl == last_lane) red[2 * q + 1] = (q == P - 1) ? 0.0 : -v;
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
  const double alpha = sol[2 * qs], beta = sol[2 * qs + 1];
  const double* __restrict__ xu = args.xu + kind * m;
  const double* __restrict__ xl = args.xl + kind * m;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int el = my_el0 + k;
    if (active && el < m) lds[at(q, el)] = d[k] + alpha * xu[el] + beta * xl[el];
  }
  wave_lds_sync();                       // a segment reads back only its own subdomain
  double* __restrict__ dst = static_cast<double*>(args.out) + line * args.lstride + qs * m;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int el = sl + W * k;
    if (active && el < m) __builtin_nontemporal_store(lds[at(q, el)], dst + el);
  }
}

}  // namespace nt
