// Strided lines with one lane per line: the block-partitioned lane solve.
//
// A workgroup of NT_LANE_P wavefronts owns 64 lines that are adjacent in
// memory (y- or z-lines of a C-ordered field); wavefront w holds one block of
// rows of all 64 lines, one line per lane, in registers: M rows for w < j and
// M - 1 for the others (the plan's m and j; M = ceil(n / 16)), where the last
// register row is padding that is never loaded or stored and holds zero.
// Each row access of a wavefront is one 512-byte run inside one plane (FP32
// loads: 256 bytes), so z-lines (rows a plane apart) touch one page per load
// instruction, where a 16-line LDS tile touches eight.  Nothing is
// transposed.  line_kernel_lanes below is the FP64 solve; line_kernel_lanes_f32r
// further down the FP32 solve with one FP64-residual sweep for FP32 storage.
//
// Per workgroup (the factors are the plan's NtLaneTab, read by scalar loads
// from the kernel arguments):
//   1. load M rows; compact: first and last row to LDS, barrier, then the
//      right-hand side in registers (compact/rhs.py:21-36, with the one-sided
//      closures at the line ends);
//   2. Thomas sweep of the block -> y (RHS-independent pivots);
//   3. the reduced system's right-hand side (-y of the block's first and last
//      rows) to LDS, barrier; wavefront 0 solves the 2P - 2 tridiagonal rows
//      of every lane's line in place, barrier;
//   4. x = y + alpha uh + beta lh (alpha: the previous block's last value,
//      beta: the next block's first value), stored.
// The arithmetic is the host's lane_solve() in nt_plan.cpp.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "nt_line_kernel.h"
#include "nt_plan.h"

namespace nt {

struct LaneArgs {
  const void* in;    // TIn: FP64 or FP32 values
  double* out;
  int64_t estride;   // distance between the rows of a line (elements)
  int64_t ostride;   // distance between slabs
  int64_t phys;      // lines per slab
  int64_t groups;    // groups of 64 lines per slab
  double s1, s2;     // compact producer: 3/(4h), 1/(2h)
  NtLaneTab t;
};

// The arguments read in place from the kernel-argument segment, so that every
// table read is a scalar load.
__device__ __forceinline__ const LaneArgs& lane_args() {
  return *(const LaneArgs*)(__builtin_amdgcn_kernarg_segment_ptr());
}

// EVEN: every block has M rows (n = 16 M), so no block carries a padding row.
// TIn: the stored values (double, or float widened at load); the solve is FP64.
template <typename TIn, int M, int PROD, bool EVEN>
__global__ void __launch_bounds__(NT_LANE_P * 64) line_kernel_lanes(LaneArgs args_in) {
  (void)args_in;
  static_assert(M >= 4 && M <= NT_LANE_MAX_M, "rows per block");
  static_assert(PROD == P_IDENT || PROD == P_COMPACT, "producer");
  constexpr int NW = NT_LANE_P, NR = NT_LANE_NR;
  // The reduced system's loops unrolled by two keep its LDS reads from being
  // issued all at once, which holds the kernel to 64 VGPRs (two workgroups per
  // CU) up to 16 rows per lane (14 with padding rows): 52 instead of 98 VGPRs at
  // 16.  Above that one workgroup fills a CU either way, and the fully unrolled
  // loops hide the LDS latency of the serial reduced solve.
  constexpr int RED_UNROLL = (M <= (EVEN ? 16 : 14)) ? 2 : NR;
  const LaneArgs& A = lane_args();
  __shared__ double edge[2][NW][64];
  __shared__ double red[NR][64];
  const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  const int jl = A.t.j;
  const bool full = EVEN || w < jl;           // this block has M rows (else M - 1)
  const bool lastw = (w == NW - 1);
  const int64_t slab = static_cast<int64_t>(blockIdx.x) / A.groups;
  const int64_t g0 = (static_cast<int64_t>(blockIdx.x) - slab * A.groups) * 64;
  const int64_t left = A.phys - g0;   // lines in this group: 1..64 used
  const bool on = lane < left;
  // lanes past the last line re-read the last line and store nothing
  const uint32_t off = static_cast<uint32_t>(on ? lane : left - 1) * 8u;
  const int first_row = w * (M - 1) + (w < jl ? w : jl);
  const int64_t row0 = slab * A.ostride + g0 + static_cast<int64_t>(first_row) * A.estride;
  const int64_t rs = A.estride * 8;
  const int64_t rsi = A.estride * static_cast<int64_t>(sizeof(TIn));
  const uint32_t offi = off / (8 / sizeof(TIn));
  const char* src = reinterpret_cast<const char*>(static_cast<const TIn*>(A.in) + row0);
  double v[M];
#pragma unroll
  for (int r = 0; r < M - 1; ++r)
    v[r] = static_cast<double>(__builtin_nontemporal_load(reinterpret_cast<const TIn*>(src + r * rsi + offi)));
  if (full)
    v[M - 1] = static_cast<double>(
        __builtin_nontemporal_load(reinterpret_cast<const TIn*>(src + (M - 1) * rsi + offi)));
  else
    v[M - 1] = 0.0;

  if constexpr (PROD == P_COMPACT) {
    edge[0][w][lane] = v[0];
    edge[1][w][lane] = full ? v[M - 1] : v[M - 2];
    __syncthreads();
    const double fl = (w > 0) ? edge[1][w > 0 ? w - 1 : 0][lane] : 0.0;
    const double fr = (w < NW - 1) ? edge[0][w < NW - 1 ? w + 1 : 0][lane] : 0.0;
    const double s1 = A.s1, s2 = A.s2;
    const double b0 = s2 * (-5.0 * v[0] + 4.0 * v[1] + v[2]);
    const double bn = full ? -s2 * (-5.0 * v[M - 1] + 4.0 * v[M - 2] + v[M - 3])
                           : -s2 * (-5.0 * v[M - 2] + 4.0 * v[M - 3] + v[M - 4]);
    double prev = fl;
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const double cur = v[r];
      double nxt;
      if (r + 2 < M) nxt = v[r + 1];
      else if (r + 2 == M) nxt = full ? v[M - 1] : fr;
      else nxt = fr;
      v[r] = s1 * (nxt - prev);
      prev = cur;
    }
    if (!full) v[M - 1] = 0.0;
    if (w == 0) v[0] = b0;
    if (lastw) {
      if (full) v[M - 1] = bn;
      else v[M - 2] = bn;
    }
  }

  // block sweep: the first block has its own pivots, the last its own last row
  {
    const double* ip = A.t.ip[w == 0 ? 0 : 1];
    const double* cp = A.t.cp[w == 0 ? 0 : 1];
    const double ai = A.t.ai;
    v[0] *= ip[0];
#pragma unroll
    for (int r = 1; r < M - 2; ++r) v[r] = (v[r] - ai * v[r - 1]) * ip[r];
    {
      const bool endrow = lastw && !full;     // row M-2 is the matrix's last row
      const double a2 = endrow ? A.t.an : ai;
      const double i2 = endrow ? A.t.ip_last : ip[M - 2];
      v[M - 2] = (v[M - 2] - a2 * v[M - 3]) * i2;
    }
    if (full) {
      const double a1 = lastw ? A.t.an : ai;
      const double i1 = lastw ? A.t.ip_last : ip[M - 1];
      v[M - 1] = (v[M - 1] - a1 * v[M - 2]) * i1;
    }
#pragma unroll
    for (int r = M - 2; r >= 0; --r) v[r] -= cp[r] * v[r + 1];
  }

  // reduced system: row 2w-1 is this block's first row, row 2w its last
  if (w >= 1) red[2 * w - 1][lane] = -v[0];
  if (w <= NW - 2) red[2 * w][lane] = full ? -v[M - 1] : -v[M - 2];
  __syncthreads();
  if (w == 0) {
    double g = red[0][lane] * A.t.rip[0];
    red[0][lane] = g;
#pragma unroll RED_UNROLL
    for (int i = 1; i < NR; ++i) {
      g = (red[i][lane] - A.t.rsub[i] * g) * A.t.rip[i];
      red[i][lane] = g;
    }
#pragma unroll RED_UNROLL
    for (int i = NR - 2; i >= 0; --i) {
      g = red[i][lane] - A.t.rcp[i] * g;
      red[i][lane] = g;
    }
  }
  __syncthreads();
  const double alpha = (w >= 1) ? red[w >= 1 ? 2 * w - 1 : 0][lane] : 0.0;
  const double beta = (w <= NW - 2) ? red[w <= NW - 2 ? 2 * w : 0][lane] : 0.0;
  const double* uh = A.t.uh[lastw ? 2 : (full ? 0 : 1)];
  const double* lh = A.t.lh[w == 0 ? 0 : (full ? 1 : 2)];
  char* dst = reinterpret_cast<char*>(A.out + row0);
  if (on) {
#pragma unroll
    for (int r = 0; r < M - 1; ++r) {
      const double x = v[r] + alpha * uh[r] + beta * lh[r];
      __builtin_nontemporal_store(x, reinterpret_cast<double*>(dst + r * rs + off));
    }
    if (full) {
      const double x = v[M - 1] + alpha * uh[M - 1] + beta * lh[M - 1];
      __builtin_nontemporal_store(x, reinterpret_cast<double*>(dst + (M - 1) * rs + off));
    }
  }
}

// FP32-stored compact derivative with the FP32 solve and one FP64-residual
// refinement sweep, one lane per line, with the blocks of line_kernel_lanes
// (M or M - 1 rows; EVEN: all M).  A lane keeps its block's FP32 field values
// and FP32 solution in registers (a padding row holds zero throughout):
//   1. load the FP32 values; the block's end rows through LDS; barrier;
//   2. x = FP32(d), d the compact right-hand side formed in FP64
//      (compact/rhs.py:21-36, one-sided closures at the line ends);
//   3. FP32 block sweep, reduced system (wavefront 0, LDS), correction -> x;
//   4. x's end rows through LDS; r = FP32(d - A x) formed in FP64, each row
//      written over the field value that no later row needs;
//   5. the same FP32 solve of r -> e;
//   6. out = x + e in FP64.
// Steps 2 and 4 are the arithmetic of the general FP32 + refinement path
// (compact_solve_from_f32 in nt_line_kernel.h); the solves are the lane solve
// in FP32 with the plan's factors rounded to FP32.
struct LaneArgs32 {
  const float* in;
  double* out;
  int64_t estride, ostride, phys, groups;
  double s1, s2;
  double b1, c1, ai, bi, ci, an, bn;   // the FP64 matrix of the residual
  NtLaneTab32 t;
};

__device__ __forceinline__ const LaneArgs32& lane32_args() {
  return *(const LaneArgs32*)(__builtin_amdgcn_kernarg_segment_ptr());
}

template <int M, bool EVEN>
__global__ void __launch_bounds__(NT_LANE_P * 64) line_kernel_lanes_f32r(LaneArgs32 args_in) {
  (void)args_in;
  static_assert(M >= 4 && M <= NT_LANE_MAX_M, "rows per block");
  constexpr int NW = NT_LANE_P, NR = NT_LANE_NR;
  // unrolled by two while that keeps two workgroups per CU (at most 64 VGPRs:
  // 13 rows per lane), fully above (see line_kernel_lanes)
  constexpr int RED_UNROLL = (M <= 13) ? 2 : NR;
  const LaneArgs32& A = lane32_args();
  __shared__ float edge[2][NW][64];
  __shared__ float red[NR][64];
  const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  const bool lastw = (w == NW - 1);
  const int jl = A.t.j;
  const bool full = EVEN || w < jl;           // this block has M rows (else M - 1)
  const bool endrow2 = lastw && !full;        // row M-2 is the line's last row
  const int64_t slab = static_cast<int64_t>(blockIdx.x) / A.groups;
  const int64_t g0 = (static_cast<int64_t>(blockIdx.x) - slab * A.groups) * 64;
  const int64_t left = A.phys - g0;   // lines in this group: 1..64 used
  const bool on = lane < left;
  const int64_t lc = on ? lane : left - 1;   // lanes past the last line re-read it, store nothing
  const int first_row = w * (M - 1) + (w < jl ? w : jl);
  const int64_t row0 = slab * A.ostride + g0 + static_cast<int64_t>(first_row) * A.estride;
  const int64_t es = A.estride;

  // 1. load
  float f[M];
  {
    const float* src = A.in + row0 + lc;
#pragma unroll
    for (int r = 0; r < M - 1; ++r) f[r] = __builtin_nontemporal_load(src + r * es);
    if (full)
      f[M - 1] = __builtin_nontemporal_load(src + (M - 1) * es);
    else
      f[M - 1] = 0.0f;
  }
  edge[0][w][lane] = f[0];
  edge[1][w][lane] = full ? f[M - 1] : f[M - 2];
  __syncthreads();
  const float fleft = (w > 0) ? edge[1][w > 0 ? w - 1 : 0][lane] : 0.0f;
  const float fright = (w < NW - 1) ? edge[0][w < NW - 1 ? w + 1 : 0][lane] : 0.0f;
  const double s1 = A.s1, s2 = A.s2;
  const double b0 = (w == 0) ? s2 * (-5.0 * static_cast<double>(f[0]) + 4.0 * static_cast<double>(f[1]) +
                                     static_cast<double>(f[2]))
                             : 0.0;
  const double bn = !lastw ? 0.0
                    : full ? -s2 * (-5.0 * static_cast<double>(f[M - 1]) + 4.0 * static_cast<double>(f[M - 2]) +
                                    static_cast<double>(f[M - 3]))
                           : -s2 * (-5.0 * static_cast<double>(f[M - 2]) + 4.0 * static_cast<double>(f[M - 3]) +
                                    static_cast<double>(f[M - 4]));
  // right-hand side of row r from its neighbours' field values, in FP64
  auto rhs = [&](int r, float fm, float fp) -> double {
    double d = s1 * (static_cast<double>(fp) - static_cast<double>(fm));
    if (r == 0 && w == 0) d = b0;
    if (lastw && (full ? r == M - 1 : r == M - 2)) d = bn;
    if (!full && r == M - 1) d = 0.0;   // padding
    return d;
  };

  // 2. x = FP32(d)
  float x[M];
#pragma unroll
  for (int r = 0; r < M; ++r) {
    const float fm = (r > 0) ? f[r > 0 ? r - 1 : 0] : fleft;
    const float fp = (r + 2 < M) ? f[r + 2 < M ? r + 1 : 0] : ((r + 2 == M && full) ? f[M - 1] : fright);
    x[r] = static_cast<float>(rhs(r, fm, fp));
  }

  const float* ip = A.t.ip[w == 0 ? 0 : 1];
  const float* cp = A.t.cp[w == 0 ? 0 : 1];
  const float* uh = A.t.uh[lastw ? 2 : (full ? 0 : 1)];
  const float* lh = A.t.lh[w == 0 ? 0 : (full ? 1 : 2)];
  const float ai = A.t.ai;
  // the reduced system of the 64 lines: this block's first and last values
  // in, alpha (previous block's last) and beta (next block's first) out
  auto reduced = [&](float first, float last, float& alpha, float& beta) {
    if (w >= 1) red[2 * w - 1][lane] = -first;
    if (w <= NW - 2) red[2 * w][lane] = -last;
    __syncthreads();
    if (w == 0) {
      float g = red[0][lane] * A.t.rip[0];
      red[0][lane] = g;
#pragma unroll RED_UNROLL
      for (int i = 1; i < NR; ++i) {
        g = (red[i][lane] - A.t.rsub[i] * g) * A.t.rip[i];
        red[i][lane] = g;
      }
#pragma unroll RED_UNROLL
      for (int i = NR - 2; i >= 0; --i) {
        g = red[i][lane] - A.t.rcp[i] * g;
        red[i][lane] = g;
      }
    }
    __syncthreads();
    alpha = (w >= 1) ? red[w >= 1 ? 2 * w - 1 : 0][lane] : 0.0f;
    beta = (w <= NW - 2) ? red[w <= NW - 2 ? 2 * w : 0][lane] : 0.0f;
  };
  // FP32 block sweep of v in place (a padding row stays zero), then the
  // reduced system and the correction
  auto solve32 = [&](float (&v)[M]) {
    v[0] *= ip[0];
#pragma unroll
    for (int r = 1; r < M - 2; ++r) v[r] = (v[r] - ai * v[r - 1]) * ip[r];
    v[M - 2] = (v[M - 2] - (endrow2 ? A.t.an : ai) * v[M - 3]) * (endrow2 ? A.t.ip_last : ip[M - 2]);
    if (full) v[M - 1] = (v[M - 1] - (lastw ? A.t.an : ai) * v[M - 2]) * (lastw ? A.t.ip_last : ip[M - 1]);
#pragma unroll
    for (int r = M - 2; r >= 0; --r) v[r] -= cp[r] * v[r + 1];
    float alpha, beta;
    reduced(v[0], full ? v[M - 1] : v[M - 2], alpha, beta);
#pragma unroll
    for (int r = 0; r < M; ++r) v[r] += alpha * uh[r] + beta * lh[r];
  };

  // 3. FP32 solve of x
  solve32(x);

  // 4. residual in FP64, over the field values (the edge rows were read before
  // the reduced system's first barrier, so they can be reused)
  edge[0][w][lane] = x[0];
  edge[1][w][lane] = full ? x[M - 1] : x[M - 2];
  __syncthreads();
  {
    const float xl = (w > 0) ? edge[1][w > 0 ? w - 1 : 0][lane] : 0.0f;
    const float xr = (w < NW - 1) ? edge[0][w < NW - 1 ? w + 1 : 0][lane] : 0.0f;
    const double b1 = A.b1, c1 = A.c1, aI = A.ai, bI = A.bi, cI = A.ci, aN = A.an, bN = A.bn;
    float pending = 0.0f;
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const float fm = (r > 0) ? f[r > 0 ? r - 1 : 0] : fleft;
      const float fp = (r + 2 < M) ? f[r + 2 < M ? r + 1 : 0] : ((r + 2 == M && full) ? f[M - 1] : fright);
      const double d = rhs(r, fm, fp);
      const double xm = static_cast<double>((r > 0) ? x[r > 0 ? r - 1 : 0] : xl);
      const double xc = static_cast<double>(x[r]);
      const double xp = static_cast<double>((r + 2 < M) ? x[r + 2 < M ? r + 1 : 0]
                                                        : ((r + 2 == M && full) ? x[M - 1] : xr));
      double ax = aI * xm + bI * xc + cI * xp;
      if (r == 0 && w == 0) ax = b1 * xc + c1 * xp;
      if (lastw && (full ? r == M - 1 : r == M - 2)) ax = aN * xm + bN * xc;
      float rr = static_cast<float>(d - ax);
      if (!full && r == M - 1) rr = 0.0f;   // padding
      // row r-1's field value is used for the last time by row r
      if (r > 0) f[r > 0 ? r - 1 : 0] = pending;
      pending = rr;
    }
    f[M - 1] = pending;
  }

  // 5. FP32 solve of r -> e, in f
  solve32(f);

  // 6. out = x + e
  if (on) {
    double* dst = A.out + row0 + lc;
#pragma unroll
    for (int r = 0; r < M - 1; ++r)
      __builtin_nontemporal_store(static_cast<double>(x[r]) + static_cast<double>(f[r]), dst + r * es);
    if (full)
      __builtin_nontemporal_store(static_cast<double>(x[M - 1]) + static_cast<double>(f[M - 1]), dst + (M - 1) * es);
  }
}

}  // namespace nt
