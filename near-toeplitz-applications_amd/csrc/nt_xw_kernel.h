// One contiguous line per workgroup, spread over NW wavefronts.
//
// A line of n = 64 R NW points (a power of two) is split into NW blocks of 64 R
// rows; wavefront w holds block w, lane l its rows l R .. l R + R - 1.  At
// R = 2 each lane reads and writes its two rows as one 16-byte access; longer
// blocks go through the wavefront's own LDS slice with coalesced 16-byte
// accesses, as in the one-wavefront kernel.
//
// The cyclic reduction is the one of nt_cr.h (reference neato.py:34-123:
// the first reduction keeps rows 1, 3, 5, ..., every later one rows
// 2^s-1 mod 2^s), split at the block boundary:
//  * Reductions 1..log2(64 R) run inside each block (in-lane, then lane
//    shuffles).  A block's last row is the only row whose right neighbour lies
//    in the next block, and its intermediate values feed nothing but its own
//    later reductions.  It is reduced with that neighbour term left out; the
//    next block adds up the missing terms, E = sum_s k2_s * (its local row
//    2^(s-1)-1 after reduction s-1), from values it holds itself.
//  * One LDS exchange publishes every block's last row and E.  Each wavefront
//    then corrects the NW last rows, reduces them, solves the top 2x2 system
//    and substitutes back among them, redundantly: all of them hold every
//    block's solved last row afterwards.
//  * Back substitution inside each block needs only its own last row and the
//    previous block's, both known.
// One workgroup barrier for the solve, one for the compact right-hand side's
// neighbours across block edges, and one per refinement sweep (the residual's
// neighbours).  Results differ from the one-wavefront solve only by rounding
// order (the deferred terms of the last row).
//
// Measured on MI355X (2^27-point fields): at two rows per lane (one 16-byte
// access per lane, no LDS staging) the extra shuffles and selects make it
// compute-bound, 2x slower than one wavefront per line at 512 points.  With two
// wavefronts per line it takes 2048-point lines with FP32 + refinement in 366 us
// against 428 us for the persistent one-wavefront kernel (FP64: 358 against
// 366), and loses at 1024 points (408 against 375 us) and on the ADI 2048^2 x half
// step (step 54.7 against 52.1 us; 4 and 8 wavefronts per line: 57 and 60).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "nt_line_kernel.h"

namespace nt {

typedef float nt_flt2 __attribute__((ext_vector_type(2)));

// Solve the FULL system of n = 64 R NW rows held as d[0..R-1] by lane `lane` of
// wavefront `w` (rows 64 R w + R lane + k).  `xch` is LDS scratch of 2 NW
// values; every wavefront of the workgroup must call this together.
template <typename T, int R, int NW>
__device__ __forceinline__ void cr_solve_xw(T (&d)[R], const NtPlanT<T>& P, int w, int lane, T* xch) {
  static_assert(NW >= 2 && (NW & (NW - 1)) == 0 && NW <= 16, "2..16 wavefronts, a power of two");
  static_assert(R >= 2 && (R & (R - 1)) == 0, "rows per lane: a power of two >= 2");
  constexpr int LB = Log2<NW>::value;
  constexpr int LR = Log2<R>::value;
  constexpr int BL = LR + 6;                    // reductions inside a block
  const bool first_blk = (w == 0);
  const bool last_blk = (w == NW - 1);
  const bool l0 = (lane == 0);
  const bool l63 = (lane == 63);

  // ---- in-lane reductions 1..LR ------------------------------------------------
  T E = T(0);                                   // wave-uniform: terms owed to the previous block
#pragma unroll
  for (int s = 1; s <= LR; ++s) {
    const int h = 1 << (s - 1);
    level_fence();
    const NtLevel<T>& Lr = P.lv[s - 1 + opaque_zero()];
    const T k1 = uni(Lr.k1), k2 = uni(Lr.k2), k1f = uni(Lr.k1f), kl = uni(Lr.kl);
    E += k2 * shidx<64>(d[h - 1], 0);           // local row h-1 after reduction s-1
    const T nb = from_next<64>(d[h - 1]);       // row (lane+1) R + h - 1
#pragma unroll
    for (int p = (1 << s) - 1; p < R; p += (1 << s)) {
      const T left = d[p - h];
      const T right = (p + h < R) ? d[p + h] : nb;
      T K1 = k1, K2 = k2;
      if (p == (1 << s) - 1) K1 = (first_blk && l0) ? k1f : K1;
      if (p == R - 1) {
        K1 = (last_blk && l63) ? kl : K1;
        K2 = l63 ? T(0) : K2;                   // the block's last row: right term deferred (or none)
      }
      d[p] = d[p] - K1 * left - K2 * right;
    }
  }

  // ---- reductions LR+1..LR+6 (lane shuffles inside the block) -------------------
  T v = d[R - 1];
#pragma unroll
  for (int s = LR + 1; s <= BL; ++s) {
    const int dl = 1 << (s - 1 - LR);
    level_fence();
    const NtLevel<T>& Lr = P.lv[s - 1 + opaque_zero()];
    const T k1 = uni(Lr.k1), k2 = uni(Lr.k2), k1f = uni(Lr.k1f), kl = uni(Lr.kl);
    E += k2 * shidx<64>(v, dl - 1);             // local row 2^(s-1)-1 after reduction s-1
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
  if (l0) xch[NW + w] = E;
  __syncthreads();
  T y[NW];
#pragma unroll
  for (int u = 0; u < NW; ++u) y[u] = xch[u] - ((u + 1 < NW) ? xch[NW + u + 1] : T(0));

  // ---- the NW last rows: reductions BL+1..nlev, top 2x2, back substitution -----
#pragma unroll
  for (int s = BL + 1; s <= BL + LB - 1; ++s) {
    const int hb = 1 << (s - BL - 1);
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
  for (int s = BL + LB - 1; s >= BL + 1; --s) {
    const int hb = 1 << (s - BL - 1);
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

  // ---- back substitution inside the block, lane shuffles ---------------------
  v = l63 ? xb : v;
#pragma unroll
  for (int s = BL; s >= LR + 1; --s) {
    const int dl = 1 << (s - 1 - LR);
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
  d[R - 1] = v;

  // ---- back substitution inside the block, in-lane ---------------------------
  T prev = from_prev<64>(v);                    // solved row lane R - 1
  prev = l0 ? xp : prev;
#pragma unroll
  for (int s = LR; s >= 1; --s) {
    const int h = 1 << (s - 1);
    level_fence();
    const NtLevel<T>& Lr = P.lv[s - 1 + opaque_zero()];
    const T a = uni(Lr.a), c = uni(Lr.c), ib = uni(Lr.ib), cf = uni(Lr.cf), ibf = uni(Lr.ibf);
#pragma unroll
    for (int p = h - 1; p < R; p += 2 * h) {
      const T left = (p - h >= 0) ? d[p - h] : prev;
      const T right = d[p + h];
      T nv = (d[p] - a * left - c * right) * ib;
      if (p == h - 1) {
        const T fst = (d[p] - cf * right) * ibf;   // row 0 of the system
        nv = (first_blk && l0) ? fst : nv;
      }
      d[p] = nv;
    }
  }
}

// The kernel: one line per workgroup (lines q + HALO of slab 0 .. outer-1),
// FP64 (or FP32) input, FP64 output, n = 64 R NW.
template <typename TIn, int R, int NW, int PROD, int PREC>
__global__ void __launch_bounds__(NW * 64) line_kernel_xw(const LineArgs args_by_value) {
  (void)args_by_value;
  const LineArgs& args = kernel_args();
  constexpr int N = NW * 64 * R;
  constexpr int HALO = (PROD == P_ADI) ? 1 : 0;
  constexpr bool STAGE = (R > 2);                // blocks through LDS (R = 2: one access per lane)
  constexpr int CS = R + 2;                      // LDS doubles per chunk (2 pad: bank spread)
  constexpr int LS = STAGE ? 64 * CS : 1;        // LDS doubles per block
  constexpr int VIN = 16 / static_cast<int>(sizeof(TIn));
  static_assert(!STAGE || R % VIN == 0, "whole 16-byte loads inside a chunk");
  static_assert(!(STAGE && PROD == P_ADI && sizeof(TIn) != 8), "the staged ADI stencil reads FP64 lines");
  __shared__ __attribute__((aligned(16))) double lds_all[NW * LS];
  __shared__ double xf[2 * NW];                  // right-hand side neighbours across block edges
  __shared__ double xs64[2 * NW];                // solve exchange, FP64
  __shared__ float xs32[2 * NW];                 // solve exchange, FP32
  __shared__ double xx[2 * NW];                  // residual neighbours across block edges

  const NtPlanK& P = args.plan;
  const int tid = threadIdx.x;
  const int w = tid >> 6;
  const int lane = tid & 63;
  const int64_t nsys = args.phys - 2 * HALO;
  const int64_t o = static_cast<int64_t>(blockIdx.x) / nsys;
  const int64_t p = static_cast<int64_t>(blockIdx.x) - o * nsys;
  const int64_t base = o * args.ostride + (p + HALO) * args.pstride;
  const TIn* __restrict__ in = static_cast<const TIn*>(args.in) + base + w * 64 * R;   // my block
  double* __restrict__ out = static_cast<double*>(args.out) + base + w * 64 * R;
  double* __restrict__ L = lds_all + (STAGE ? w * LS : 0);
  auto at = [](int e) -> int { return (e / R) * CS + (e % R); };
  const int e0 = (w * 64 + lane) * R;            // my first row of the line

  // ---- load my rows: f = input (compact, solve) or the ADI right-hand side ----
  double f[R];
  if constexpr (!STAGE) {
    double u[2] = {0.0, 0.0}, v[2] = {0.0, 0.0};
    if constexpr (sizeof(TIn) == 8) {
      const nt_dbl2 t = __builtin_nontemporal_load(reinterpret_cast<const nt_dbl2*>(in) + lane);
      f[0] = t.x;
      f[1] = t.y;
      if constexpr (PROD == P_ADI) {
        const nt_dbl2 tu = reinterpret_cast<const nt_dbl2*>(in - args.pstride)[lane];
        const nt_dbl2 tv = reinterpret_cast<const nt_dbl2*>(in + args.pstride)[lane];
        u[0] = tu.x; u[1] = tu.y; v[0] = tv.x; v[1] = tv.y;
      }
    } else {
      const nt_flt2 t = __builtin_nontemporal_load(reinterpret_cast<const nt_flt2*>(in) + lane);
      f[0] = static_cast<double>(t.x);
      f[1] = static_cast<double>(t.y);
      if constexpr (PROD == P_ADI) {
        const nt_flt2 tu = reinterpret_cast<const nt_flt2*>(in - args.pstride)[lane];
        const nt_flt2 tv = reinterpret_cast<const nt_flt2*>(in + args.pstride)[lane];
        u[0] = tu.x; u[1] = tu.y; v[0] = tv.x; v[1] = tv.y;
      }
    }
    if constexpr (PROD == P_ADI) {
      // adi/rhs.py:21-24: the explicit stencil across the neighbouring lines,
      // identity rows at both ends of the line
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int e = e0 + k;
        f[k] = (e == 0 || e >= N - 1) ? f[k] : args.s1 * (u[k] + v[k]) + args.s2 * f[k];
      }
    }
  } else {
    // coalesced 16-byte accesses into my LDS slice (instruction i covers
    // 1 KB .. of the block), the ADI stencil formed while streaming
    constexpr int NV = 64 * R / VIN / 64;        // 16-byte loads per lane
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int j = lane + 64 * i;
      if constexpr (sizeof(TIn) == 8) {
        const nt_dbl2 t = __builtin_nontemporal_load(reinterpret_cast<const nt_dbl2*>(in) + j);
        double a0 = t.x, a1 = t.y;
        if constexpr (PROD == P_ADI) {
          const nt_dbl2 tu = reinterpret_cast<const nt_dbl2*>(in - args.pstride)[j];
          const nt_dbl2 tv = reinterpret_cast<const nt_dbl2*>(in + args.pstride)[j];
          const int e = w * 64 * R + 2 * j;
          a0 = (e == 0) ? a0 : args.s1 * (tu.x + tv.x) + args.s2 * a0;
          a1 = (e + 1 >= N - 1) ? a1 : args.s1 * (tu.y + tv.y) + args.s2 * a1;
        }
        L[at(2 * j)] = a0;
        L[at(2 * j + 1)] = a1;
      } else {
        const nt_flt4 t = __builtin_nontemporal_load(reinterpret_cast<const nt_flt4*>(in) + j);
        L[at(4 * j)] = static_cast<double>(t.x);
        L[at(4 * j + 1)] = static_cast<double>(t.y);
        L[at(4 * j + 2)] = static_cast<double>(t.z);
        L[at(4 * j + 3)] = static_cast<double>(t.w);
      }
    }
    wave_lds_sync();
    const double* __restrict__ mine = L + lane * CS;
#pragma unroll
    for (int k = 0; k < R; k += 2) {
      const nt_dbl2 t = *reinterpret_cast<const nt_dbl2*>(mine + k);
      f[k] = t.x;
      f[k + 1] = t.y;
    }
  }

  // ---- right-hand side -------------------------------------------------------
  double d[R];
  if constexpr (PROD == P_COMPACT) {
    // rows e0-1 and e0+R: the neighbouring lanes, across a block edge through LDS
    if (lane == 0) xf[w] = f[0];
    if (lane == 63) xf[NW + w] = f[R - 1];
    __syncthreads();
    double fl = from_prev<64>(f[R - 1]);
    double fr = from_next<64>(f[0]);
    if (lane == 0) fl = (w > 0) ? xf[NW + w - 1] : 0.0;
    if (lane == 63) fr = (w < NW - 1) ? xf[w + 1] : 0.0;
    // closures from the line's first and last three values (compact/rhs.py:21-36)
    double b0 = 0.0, bn = 0.0;
    if constexpr (STAGE) {
      const double bnd[6] = {L[at(0)], L[at(1)], L[at(2)], L[at(64 * R - 1)], L[at(64 * R - 2)], L[at(64 * R - 3)]};
      compact_closures(bnd, args.s2, b0, bn);    // valid in the first / last block
    } else {
      b0 = args.s2 * (-5.0 * f[0] + 4.0 * f[1] + fr);
      bn = -args.s2 * (-5.0 * f[1] + 4.0 * f[0] + fl);
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int e = e0 + k;
      const double left = (k > 0) ? f[k - 1] : fl;
      const double right = (k + 1 < R) ? f[k + 1] : fr;
      double v = args.s1 * (right - left);
      if (k == 0) v = (e == 0) ? b0 : v;
      if (k == R - 1) v = (e == N - 1) ? bn : v;
      d[k] = v;
    }
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k) d[k] = f[k];
  }

  // ---- solve -----------------------------------------------------------------
  double x[R];
  if constexpr (PREC == PREC_F64) {
#pragma unroll
    for (int k = 0; k < R; ++k) x[k] = d[k];
    cr_solve_xw<double, R, NW>(x, P.d, w, lane, xs64);
  } else {
    float s[R];
#pragma unroll
    for (int k = 0; k < R; ++k) s[k] = static_cast<float>(d[k]);
    cr_solve_xw<float, R, NW>(s, P.f, w, lane, xs32);
#pragma unroll
    for (int k = 0; k < R; ++k) x[k] = static_cast<double>(s[k]);
    for (int it = 0; it < args.sweeps; ++it) {
      // r = d - A x in FP64; x's neighbours across a block edge through LDS
      if (lane == 0) xx[w] = x[0];
      if (lane == 63) xx[NW + w] = x[R - 1];
      __syncthreads();
      double xl = from_prev<64>(x[R - 1]);
      double xr = from_next<64>(x[0]);
      if (lane == 0) xl = (w > 0) ? xx[NW + w - 1] : 0.0;
      if (lane == 63) xr = (w < NW - 1) ? xx[w + 1] : 0.0;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const double left = (k > 0) ? x[k - 1] : xl;
        const double right = (k + 1 < R) ? x[k + 1] : xr;
        s[k] = static_cast<float>(d[k] - apply_row(P, e0 + k, N, left, x[k], right));
      }
      cr_solve_xw<float, R, NW>(s, P.f, w, lane, xs32);
#pragma unroll
      for (int k = 0; k < R; ++k) x[k] += static_cast<double>(s[k]);
    }
  }

  // ---- store my rows -----------------------------------------------------------
  if constexpr (!STAGE) {
    nt_dbl2 t;
    t.x = x[0];
    t.y = x[1];
    __builtin_nontemporal_store(t, reinterpret_cast<nt_dbl2*>(out) + lane);
  } else {
    wave_lds_sync();                             // every lane has read its chunk
    double* __restrict__ mw = L + lane * CS;
#pragma unroll
    for (int k = 0; k < R; k += 2) {
      nt_dbl2 t;
      t.x = x[k];
      t.y = x[k + 1];
      *reinterpret_cast<nt_dbl2*>(mw + k) = t;
    }
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < 64 * R / 2 / 64; ++i) {
      const int j = lane + 64 * i;
      nt_dbl2 t;
      t.x = L[at(2 * j)];
      t.y = L[at(2 * j + 1)];
      __builtin_nontemporal_store(t, reinterpret_cast<nt_dbl2*>(out) + j);
    }
  }
}

}  // namespace nt
