// One contiguous line per workgroup, spread over NW wavefronts.
//
// A line of n = 128 * NW points (a power of two, 256 .. 2048) is split into NW
// blocks of 128 rows; wavefront w holds block w, lane l rows 128w + 2l and
// 128w + 2l + 1.  Each lane therefore reads and writes its two rows as one
// 16-byte access, and each wavefront moves one contiguous 1 KB piece of the
// line: no LDS transpose, and the access pattern of a copy that spreads a line
// over its wavefronts, which runs faster on MI355X than one wavefront moving
// a whole 4 KB line as four 1 KB pieces (measured copies of a 512^3 FP64
// field: 6.7 against 6.2 TB/s).
//
// The cyclic reduction is the one of nt_cr.h (reference neato.py:34-123:
// the first reduction keeps rows 1, 3, 5, ..., every later one rows
// 2^s-1 mod 2^s), split at the block boundary:
//  * Reductions 1..7 run inside each block (in-lane, then lane shuffles).  A
//    block's last row (local row 127) is the only row whose right neighbour
//    lies in the next block, and its intermediate values feed nothing but its
//    own later reductions.  It is reduced with that neighbour term left out;
//    the next block adds up the missing terms, E = sum_s k2_s * (its local row
//    2^(s-1)-1 after reduction s-1), from values it holds itself.
//  * One LDS exchange publishes every block's last row and E.  Each wavefront
//    then corrects the NW last rows, reduces them (reductions 8..nlev), solves
//    the top 2x2 system and substitutes back among them, redundantly: all of
//    them hold every block's solved last row afterwards.
//  * Back substitution inside each block needs only its own last row and the
//    previous block's, both known.
// One workgroup barrier for the solve, one for the right-hand side's
// neighbours across block edges, and one per refinement sweep (the residual's
// neighbours).  Results differ from the one-wavefront solve only by rounding
// order (the deferred terms of the last row).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "nt_line_kernel.h"

namespace nt {

typedef float nt_flt2 __attribute__((ext_vector_type(2)));

// Solve the FULL system of n = 128 * NW rows held as d[0..1] by lane `lane` of
// wavefront `w` (rows 128w + 2 lane + {0, 1}).  `xch` is LDS scratch of 2 NW
// values; every wavefront of the workgroup must call this together.
template <typename T, int NW>
__device__ __forceinline__ void cr_solve_xw(T (&d)[2], const NtPlanT<T>& P, int w, int lane, T* xch) {
  static_assert(NW >= 2 && (NW & (NW - 1)) == 0 && NW <= 16, "2..16 wavefronts, a power of two");
  constexpr int LB = Log2<NW>::value;
  const bool first_blk = (w == 0);
  const bool last_blk = (w == NW - 1);
  const bool l63 = (lane == 63);

  // ---- reduction 1 (in-lane): row 2l+1 from rows 2l and 2l+2 ----------------
  T E;                                  // wave-uniform: terms owed to the previous block
  {
    level_fence();
    const NtLevel<T>& Lr = P.lv[opaque_zero()];
    const T k1 = uni(Lr.k1), k2 = uni(Lr.k2), k1f = uni(Lr.k1f), kl = uni(Lr.kl);
    E = k2 * shidx<64>(d[0], 0);        // local row 0, unreduced
    const T nb = from_next<64>(d[0]);
    T K1 = (first_blk && lane == 0) ? k1f : k1;
    K1 = (l63 && last_blk) ? kl : K1;
    const T K2 = l63 ? T(0) : k2;       // local row 127: right term deferred (last block: none)
    d[1] = d[1] - K1 * d[0] - K2 * nb;
  }

  // ---- reductions 2..7 (lane shuffles inside the block) ----------------------
  T v = d[1];
#pragma unroll
  for (int s = 2; s <= 7; ++s) {
    const int dl = 1 << (s - 2);
    level_fence();
    const NtLevel<T>& Lr = P.lv[s - 1 + opaque_zero()];
    const T k1 = uni(Lr.k1), k2 = uni(Lr.k2), k1f = uni(Lr.k1f), kl = uni(Lr.kl);
    E += k2 * shidx<64>(v, dl - 1);     // local row 2^(s-1)-1 after reduction s-1
    const T left = shup<64>(v, dl);
    const T right = shdn<64>(v, dl);
    const bool kept = ((lane + 1) & ((dl << 1) - 1)) == 0;
    T K1 = (first_blk && lane == 2 * dl - 1) ? k1f : k1;
    K1 = (l63 && last_blk) ? kl : K1;
    const T K2 = l63 ? T(0) : k2;
    const T nv = v - K1 * left - K2 * right;
    v = kept ? nv : v;
  }

  // ---- exchange: every block's last row and its E ---------------------------
  if (l63) xch[w] = v;
  if (lane == 0) xch[NW + w] = E;
  __syncthreads();
  T y[NW];
#pragma unroll
  for (int u = 0; u < NW; ++u) y[u] = xch[u] - ((u + 1 < NW) ? xch[NW + u + 1] : T(0));

  // ---- the NW last rows: reductions 8..nlev, top 2x2, back substitution -------
#pragma unroll
  for (int s = 8; s <= 6 + LB; ++s) {
    const int hb = 1 << (s - 8);
    level_fence();
    const NtLevel<T>& Lr = P.lv[s - 1 + opaque_zero()];
    const T k1 = uni(Lr.k1), k2 = uni(Lr.k2), k1f = uni(Lr.k1f), kl = uni(Lr.kl);
#pragma unroll
    for (int u = 2 * hb - 1; u < NW; u += 2 * hb) {
      if (u == NW - 1) y[u] = y[u] - kl * y[u - hb];
      else y[u] = y[u] - ((u == 2 * hb - 1) ? k1f : k1) * y[u - hb] - k2 * y[u + hb];
    }
  }
  {
    level_fence();
    const NtTop<T>& Tp = (&P.top)[opaque_zero()];
    const T t0 = uni(Tp.t[0]), t1 = uni(Tp.t[1]), t2 = uni(Tp.t[2]), t3 = uni(Tp.t[3]);
    constexpr int u0 = NW / 2 - 1, u1 = NW - 1;
    const T x0 = y[u0] * t0 - y[u1] * t1;
    const T x1 = y[u1] * t2 - y[u0] * t3;
    y[u0] = x0;
    y[u1] = x1;
  }
#pragma unroll
  for (int s = 6 + LB; s >= 8; --s) {
    const int hb = 1 << (s - 8);
    level_fence();
    const NtLevel<T>& Lr = P.lv[s - 1 + opaque_zero()];
    const T a = uni(Lr.a), c = uni(Lr.c), ib = uni(Lr.ib), cf = uni(Lr.cf), ibf = uni(Lr.ibf);
#pragma unroll
    for (int u = hb - 1; u < NW; u += 2 * hb) {
      if (u == hb - 1) y[u] = (y[u] - cf * y[u + hb]) * ibf;
      else y[u] = (y[u] - a * y[u - hb] - c * y[u + hb]) * ib;
    }
  }
  T xb = y[0], xp = T(0);               // my block's last row, the previous block's
#pragma unroll
  for (int u = 1; u < NW; ++u) {
    xb = (w == u) ? y[u] : xb;
    xp = (w == u) ? y[u - 1] : xp;
  }

  // ---- back substitution inside the block -----------------------------------
  v = l63 ? xb : v;
#pragma unroll
  for (int s = 7; s >= 2; --s) {
    const int dl = 1 << (s - 2);
    level_fence();
    const NtLevel<T>& Lr = P.lv[s - 1 + opaque_zero()];
    const T a = uni(Lr.a), c = uni(Lr.c), ib = uni(Lr.ib), cf = uni(Lr.cf), ibf = uni(Lr.ibf);
    const bool elim = ((lane + 1) & ((dl << 1) - 1)) == dl;
    const bool is_first = (lane == dl - 1);     // left neighbour: the previous block's last row
    const T left = is_first ? xp : shup<64>(v, dl);
    const T right = shdn<64>(v, dl);
    T nv = (v - a * left - c * right) * ib;
    const T tf = (v - cf * right) * ibf;        // row 0 of the level (first block)
    nv = (is_first && first_blk) ? tf : nv;
    v = elim ? nv : v;
  }
  d[1] = v;
  {
    level_fence();
    const NtLevel<T>& Lr = P.lv[opaque_zero()];
    const T a = uni(Lr.a), c = uni(Lr.c), ib = uni(Lr.ib), cf = uni(Lr.cf), ibf = uni(Lr.ibf);
    T prev = from_prev<64>(v);
    prev = (lane == 0) ? xp : prev;
    const T nv = (d[0] - a * prev - c * v) * ib;
    const T tf = (d[0] - cf * v) * ibf;
    d[0] = (lane == 0 && first_blk) ? tf : nv;
  }
}

// The kernel: one line per workgroup (lines q + HALO of slab 0 .. outer-1),
// FP64 (or FP32) input, FP64 output, n = 128 * NW.
template <typename TIn, int NW, int PROD, int PREC>
__global__ void __launch_bounds__(NW * 64) line_kernel_xw(const LineArgs args_by_value) {
  (void)args_by_value;
  const LineArgs& args = kernel_args();
  constexpr int N = NW * 128;
  constexpr int HALO = (PROD == P_ADI) ? 1 : 0;
  __shared__ double xf[2 * NW];        // right-hand side neighbours across block edges
  __shared__ double xs64[2 * NW];      // solve exchange, FP64
  __shared__ float xs32[2 * NW];       // solve exchange, FP32
  __shared__ double xx[2 * NW];        // residual neighbours across block edges

  const NtPlanK& P = args.plan;
  const int tid = threadIdx.x;
  const int w = tid >> 6;
  const int lane = tid & 63;
  const int64_t nsys = args.phys - 2 * HALO;
  const int64_t o = static_cast<int64_t>(blockIdx.x) / nsys;
  const int64_t p = static_cast<int64_t>(blockIdx.x) - o * nsys;
  const int64_t base = o * args.ostride + (p + HALO) * args.pstride;
  const TIn* __restrict__ in = static_cast<const TIn*>(args.in) + base;
  double* __restrict__ out = static_cast<double*>(args.out) + base;

  // ---- load my two rows (one 16- or 8-byte access per lane) -----------------
  double f0, f1;
  if constexpr (sizeof(TIn) == 8) {
    const nt_dbl2 t = __builtin_nontemporal_load(reinterpret_cast<const nt_dbl2*>(in) + tid);
    f0 = t.x;
    f1 = t.y;
  } else {
    const nt_flt2 t = __builtin_nontemporal_load(reinterpret_cast<const nt_flt2*>(in) + tid);
    f0 = static_cast<double>(t.x);
    f1 = static_cast<double>(t.y);
  }
  const int e0 = 2 * tid;              // my first row

  // ---- right-hand side -------------------------------------------------------
  double d[2];
  if constexpr (PROD == P_COMPACT) {
    // rows e0-1 and e0+2: the neighbouring lanes, across a block edge through LDS
    if (lane == 0) xf[w] = f0;
    if (lane == 63) xf[NW + w] = f1;
    __syncthreads();
    double fl = from_prev<64>(f1);
    double fr = from_next<64>(f0);
    if (lane == 0) fl = (w > 0) ? xf[NW + w - 1] : 0.0;
    if (lane == 63) fr = (w < NW - 1) ? xf[w + 1] : 0.0;
    // compact/rhs.py:21-36: interior s1 (f[i+1] - f[i-1]); closures at rows 0 and n-1
    d[0] = args.s1 * (f1 - fl);
    d[1] = args.s1 * (fr - f0);
    const double b0 = args.s2 * (-5.0 * f0 + 4.0 * f1 + fr);
    const double bn = -args.s2 * (-5.0 * f1 + 4.0 * f0 + fl);
    d[0] = (e0 == 0) ? b0 : d[0];
    d[1] = (e0 + 1 == N - 1) ? bn : d[1];
  } else if constexpr (PROD == P_ADI) {
    // adi/rhs.py:21-24: the explicit stencil across the neighbouring lines,
    // identity rows at both ends of the line
    double u0, u1, v0, v1;
    if constexpr (sizeof(TIn) == 8) {
      const nt_dbl2 tu = reinterpret_cast<const nt_dbl2*>(in - args.pstride)[tid];
      const nt_dbl2 tv = reinterpret_cast<const nt_dbl2*>(in + args.pstride)[tid];
      u0 = tu.x; u1 = tu.y; v0 = tv.x; v1 = tv.y;
    } else {
      const nt_flt2 tu = reinterpret_cast<const nt_flt2*>(in - args.pstride)[tid];
      const nt_flt2 tv = reinterpret_cast<const nt_flt2*>(in + args.pstride)[tid];
      u0 = tu.x; u1 = tu.y; v0 = tv.x; v1 = tv.y;
    }
    d[0] = (e0 == 0) ? f0 : args.s1 * (u0 + v0) + args.s2 * f0;
    d[1] = (e0 + 1 >= N - 1) ? f1 : args.s1 * (u1 + v1) + args.s2 * f1;
  } else {
    d[0] = f0;
    d[1] = f1;
  }

  // ---- solve -----------------------------------------------------------------
  double x[2];
  if constexpr (PREC == PREC_F64) {
    x[0] = d[0];
    x[1] = d[1];
    cr_solve_xw<double, NW>(x, P.d, w, lane, xs64);
  } else {
    float s[2] = {static_cast<float>(d[0]), static_cast<float>(d[1])};
    cr_solve_xw<float, NW>(s, P.f, w, lane, xs32);
    x[0] = static_cast<double>(s[0]);
    x[1] = static_cast<double>(s[1]);
    for (int it = 0; it < args.sweeps; ++it) {
      // r = d - A x in FP64; x's neighbours across a block edge through LDS
      if (lane == 0) xx[w] = x[0];
      if (lane == 63) xx[NW + w] = x[1];
      __syncthreads();
      double xl = from_prev<64>(x[1]);
      double xr = from_next<64>(x[0]);
      if (lane == 0) xl = (w > 0) ? xx[NW + w - 1] : 0.0;
      if (lane == 63) xr = (w < NW - 1) ? xx[w + 1] : 0.0;
      s[0] = static_cast<float>(d[0] - apply_row(P, e0, N, xl, x[0], x[1]));
      s[1] = static_cast<float>(d[1] - apply_row(P, e0 + 1, N, x[0], x[1], xr));
      cr_solve_xw<float, NW>(s, P.f, w, lane, xs32);
      x[0] += static_cast<double>(s[0]);
      x[1] += static_cast<double>(s[1]);
    }
  }

  // ---- store my two rows -------------------------------------------------------
  nt_dbl2 t;
  t.x = x[0];
  t.y = x[1];
  __builtin_nontemporal_store(t, reinterpret_cast<nt_dbl2*>(out) + tid);
}

}  // namespace nt
