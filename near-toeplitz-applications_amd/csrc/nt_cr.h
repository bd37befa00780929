// Register-resident cyclic reduction for one near-Toeplitz system per segment
// of W lanes of a 64-wide wavefront.
//
// Layout: lane l of the segment holds rows l*R .. l*R+R-1 of its system in
// d[0..R-1] (rows >= n are padding and are never touched by the factors).
// The first log2(R) reductions run inside a lane (one neighbour value crosses
// from lane l+1 per level); the remaining reductions run across lanes, each
// lane holding its row R-1, with wavefront shuffles at lane distance
// 2^(s-1)/R and no barrier.  Back substitution mirrors them; the in-lane back
// substitution needs only the solved row of lane l-1, shuffled once.
//
// Which row is kept, eliminated, first or last at each level follows the
// reference kernel's phase order (neato.py:34-123): the first reduction keeps
// rows 1, 3, 5, ..., every later one keeps rows 2^s-1 mod 2^s, the top system is
// solved directly and back substitution walks the levels down to the even rows
// of level 0 (the reference's "last substitution" at store, neato.py:112-123).
#pragma once
#include <hip/hip_runtime.h>

#include "nt_plan.h"

namespace nt {

template <int X>
struct Log2 {
  static constexpr int value = 1 + Log2<X / 2>::value;
};
template <>
struct Log2<1> {
  static constexpr int value = 0;
};

// Segment-local shuffles: lanes whose source falls outside the segment read
// their own value (HIP semantics of __shfl_up/__shfl_down with a width).
template <int W, typename T>
__device__ __forceinline__ T shup(T v, int delta) {
  return __shfl_up(v, static_cast<unsigned>(delta), W);
}
template <int W, typename T>
__device__ __forceinline__ T shdn(T v, int delta) {
  return __shfl_down(v, static_cast<unsigned>(delta), W);
}
template <int W, typename T>
__device__ __forceinline__ T shidx(T v, int src) {
  return __shfl(v, src, W);
}

// Solve in place.  `lane` is the lane index inside the segment (0..W-1).
// Requires n <= W*R and R <= 2^nlev (top rows sit at position R-1 of a lane);
// the host picks (W, R) accordingly.
template <typename T, int R, int W>
__device__ __forceinline__ void cr_solve_regs(T (&d)[R], const NtPlanT<T>& P, int nlev, int lane) {
  constexpr int LR = Log2<R>::value;
  const int row0 = lane * R;

  // ---- forward reduction, in-lane levels -------------------------------------
#pragma unroll
  for (int s = 1; s <= LR; ++s) {
    const int h = 1 << (s - 1);
    const NtLevel<T>& L = P.lv[s - 1];
    const T nb = shdn<W>(d[h - 1], 1);      // row (lane+1)*R + h - 1
#pragma unroll
    for (int p = (1 << s) - 1; p < R; p += (1 << s)) {
      const int jp = ((row0 + p + 1) >> (s - 1)) - 1;   // index at level s-1 (odd)
      if (jp < L.m) {
        const T left = d[p - h];
        const T right = (p + h < R) ? d[p + h] : nb;
        const T K1 = (jp == 1) ? L.k1f : ((jp == L.m - 1) ? L.kl : L.k1);
        const T K2 = (jp == L.m - 1) ? T(0) : ((jp == L.m - 2) ? L.kl : L.k2);
        d[p] = d[p] - K1 * left - K2 * right;
      }
    }
  }

  // ---- forward reduction, cross-lane levels ----------------------------------
  constexpr int LW = Log2<W>::value;
  T v = d[R - 1];
#pragma unroll
  for (int s = LR + 1; s <= LR + LW; ++s) {
    if (s > nlev) break;
    const int sh = s - 1 - LR;
    const int dl = 1 << sh;
    const NtLevel<T>& L = P.lv[s - 1];
    const T left = shup<W>(v, dl);
    const T right = shdn<W>(v, dl);
    const int jp = ((lane + 1) >> sh) - 1;
    const bool kept = (((lane + 1) & ((dl << 1) - 1)) == 0) && (jp < L.m);
    const T K1 = (jp == 1) ? L.k1f : ((jp == L.m - 1) ? L.kl : L.k1);
    const T K2 = (jp == L.m - 1) ? T(0) : ((jp == L.m - 2) ? L.kl : L.k2);
    const T nv = v - K1 * left - K2 * right;
    v = kept ? nv : v;
  }

  // ---- top system (m = 2 or 3 equations at level nlev) ------------------------
  {
    const NtTop<T>& Tp = P.top;
    const int H = 1 << nlev;                 // row spacing at the top level
    const int l0 = H / R - 1;                // lane holding top row 0
    const int l1 = 2 * H / R - 1;
    const T d0 = shidx<W>(v, l0);
    const T d1 = shidx<W>(v, l1);
    if (Tp.m == 2) {
      const T x0 = d0 * Tp.t[0] - d1 * Tp.t[1];
      const T x1 = d1 * Tp.t[2] - d0 * Tp.t[3];
      v = (lane == l0) ? x0 : ((lane == l1) ? x1 : v);
    } else {
      const int l2 = 3 * H / R - 1;
      const T d2 = shidx<W>(v, l2 < W ? l2 : l1);
      const T x1 = (d1 - Tp.t[0] * d0 - Tp.t[1] * d2) * Tp.t[2];
      const T x0 = (d0 - Tp.t[3] * x1) * Tp.t[4];
      const T x2 = (d2 - Tp.t[5] * x1) * Tp.t[6];
      v = (lane == l0) ? x0 : ((lane == l1) ? x1 : ((lane == l2) ? x2 : v));
    }
  }

  // ---- back substitution, cross-lane levels ----------------------------------
#pragma unroll
  for (int s = LR + LW; s >= LR + 1; --s) {
    if (s > nlev) continue;
    const int sh = s - 1 - LR;
    const int dl = 1 << sh;
    const NtLevel<T>& L = P.lv[s - 1];
    const T left = shup<W>(v, dl);
    const T right = shdn<W>(v, dl);
    const int jp = ((lane + 1) >> sh) - 1;
    const bool elim = (((lane + 1) & (dl - 1)) == 0) && ((jp & 1) == 0) && (jp < L.m);
    const T A = (jp == 0) ? T(0) : ((jp == L.m - 1) ? L.al : L.a);
    const T C = (jp == L.m - 1) ? T(0) : ((jp == 0) ? L.cf : L.c);
    const T IB = (jp == 0) ? L.ibf : ((jp == L.m - 1) ? L.ibl : L.ib);
    const T nv = (v - A * left - C * right) * IB;
    v = elim ? nv : v;
  }
  d[R - 1] = v;

  // ---- back substitution, in-lane levels --------------------------------------
  const T prev = shup<W>(v, 1);              // solved row lane*R - 1 (lane 0: unused)
#pragma unroll
  for (int s = LR; s >= 1; --s) {
    const int h = 1 << (s - 1);
    const NtLevel<T>& L = P.lv[s - 1];
#pragma unroll
    for (int p = h - 1; p < R; p += 2 * h) {
      const int jp = ((row0 + p + 1) >> (s - 1)) - 1;   // even
      if (jp < L.m) {
        const T left = (p - h >= 0) ? d[p - h] : prev;
        const T right = d[(p + h < R) ? p + h : p];
        const T A = (jp == 0) ? T(0) : ((jp == L.m - 1) ? L.al : L.a);
        const T C = (jp == L.m - 1) ? T(0) : ((jp == 0) ? L.cf : L.c);
        const T IB = (jp == 0) ? L.ibf : ((jp == L.m - 1) ? L.ibl : L.ib);
        d[p] = (d[p] - A * left - C * right) * IB;
      }
    }
  }
}

}  // namespace nt
