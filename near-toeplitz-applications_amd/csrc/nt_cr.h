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
//
// The code is branch-free: every per-row decision is a select, so a wavefront
// never splits.  With FULL (n == W*R, n a power of two, every level even) the
// level structure is known at compile time and only the first and last lanes
// select special factors.  Distance-1 shuffles use DPP lane shifts (no LDS
// traffic); longer distances use ds_bpermute; broadcasts from a known lane use
// v_readlane when a segment is the whole wavefront.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

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

// ---- cross-lane moves --------------------------------------------------------------
// 32-bit halves of a double (DPP and readlane move 32-bit registers)
__device__ __forceinline__ int lo32(double v) { return __double2loint(v); }
__device__ __forceinline__ int hi32(double v) { return __double2hiint(v); }
__device__ __forceinline__ double mkdbl(int hi, int lo) { return __hiloint2double(hi, lo); }

// DPP control words (GFX9 encoding)
constexpr int DPP_ROW_SHL1 = 0x101;
constexpr int DPP_ROW_SHR1 = 0x111;
constexpr int DPP_WAVE_SHL1 = 0x130;
constexpr int DPP_WAVE_SHR1 = 0x138;

// lanes without a source lane read 0 (bound_ctrl)
template <int CTRL>
__device__ __forceinline__ int dpp32(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true);
}
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(dpp32<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  return mkdbl(dpp32<CTRL>(hi32(v)), dpp32<CTRL>(lo32(v)));
}

// value of lane+1 / lane-1.  Lanes whose neighbour is outside the segment get
// an unspecified value (0 or another segment's); callers never use it (they
// select it away).
template <int W, typename T>
__device__ __forceinline__ T from_next(T v) {
  if constexpr (W == 16) return dpp<DPP_ROW_SHL1>(v);
  else return dpp<DPP_WAVE_SHL1>(v);
}
template <int W, typename T>
__device__ __forceinline__ T from_prev(T v) {
  if constexpr (W == 16) return dpp<DPP_ROW_SHR1>(v);
  else return dpp<DPP_WAVE_SHR1>(v);
}

// ds_bpermute of both halves from wavefront lane `src` (any lane, wraps mod 64)
__device__ __forceinline__ double bperm(double v, int src) {
  const int a = (src & 63) << 2;
  return mkdbl(__builtin_amdgcn_ds_bpermute(a, hi32(v)), __builtin_amdgcn_ds_bpermute(a, lo32(v)));
}
__device__ __forceinline__ float bperm(float v, int src) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute((src & 63) << 2, __float_as_int(v)));
}

// value of lane-delta / lane+delta.  Lanes whose source is outside the
// segment read an unspecified value, which callers select away.
template <int W, typename T>
__device__ __forceinline__ T shup(T v, int delta) {
  if (delta == 1) return from_prev<W>(v);
  return bperm(v, static_cast<int>(__lane_id()) - delta);
}
template <int W, typename T>
__device__ __forceinline__ T shdn(T v, int delta) {
  if (delta == 1) return from_next<W>(v);
  return bperm(v, static_cast<int>(__lane_id()) + delta);
}

// value of lane `src` of my segment; `src` is the same for every segment
template <int W, typename T>
__device__ __forceinline__ T shidx(T v, int src) {
  if constexpr (W == 64) {
    if constexpr (sizeof(T) == 8) {
      return mkdbl(__builtin_amdgcn_readlane(hi32(v), src), __builtin_amdgcn_readlane(lo32(v), src));
    } else {
      return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
    }
  } else {
    return bperm(v, static_cast<int>(__lane_id() & ~(W - 1)) + src);
  }
}

// A wave-uniform value read from the launch arguments, pinned to scalar
// registers (keeps the compiler from turning a select between two
// coefficients into a per-lane load from a selected address).
__device__ __forceinline__ double uni(const double& x) {
  return mkdbl(__builtin_amdgcn_readfirstlane(hi32(x)), __builtin_amdgcn_readfirstlane(lo32(x)));
}
__device__ __forceinline__ float uni(const float& x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ int uni(const int& x) { return __builtin_amdgcn_readfirstlane(x); }

// A zero the compiler cannot see through.  Indexing a level's factors with it
// keeps their scalar loads next to their use: otherwise every level's loads are
// hoisted to the kernel entry and the scalar registers spill.
__device__ __forceinline__ int opaque_zero() {
  int z;
  asm volatile("s_mov_b32 %0, 0" : "=s"(z));
  return z;
}
__device__ __forceinline__ void level_fence() { __builtin_amdgcn_sched_barrier(0); }

// A per-lane zero the compiler cannot see through (keeps values derived from
// it from being hoisted out of a loop).
__device__ __forceinline__ int opaque_vzero() {
  int z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}

// Inside a branch on a wave-uniform condition: keeps the branch (an empty
// volatile asm cannot be speculated), so a patch that applies to one register
// position is not computed and selected at every position.
#ifndef NT_NO_KEEP_BRANCH
#define NT_KEEP_BRANCH() asm volatile("")
#else
#define NT_KEEP_BRANCH() ((void)0)
#endif

// ---- the solve ----------------------------------------------------------------------
// Solve in place.  `lane` is the lane index inside the segment (0..W-1).
// Requires n <= W*R and R <= 2^nlev (top rows sit at position R-1 of a lane);
// the host picks (W, R) accordingly.  FULL: n == W*R, n a power of two.
template <typename T, int R, int W, bool FULL = false>
__device__ __forceinline__ void cr_solve_regs(T (&d)[R], const NtPlanT<T>& P, int nlev, int lane) {
  constexpr int LR = Log2<R>::value;
  constexpr int LW = Log2<W>::value;
  const int row0 = lane * R;
  const bool first_lane = (lane == 0);
  const bool last_lane = (lane == W - 1);

  // Every kept (forward) or eliminated (back) row is first updated as an
  // interior row; the few rows whose coefficients differ are then patched.
  // * The first row of each level sits at lane 0, a compile-time position.
  // * The row that touches the end of the system is the last row (m even) or
  //   the one before it (m odd, where it absorbs the eliminated last row).
  //   Its level-0 index is the same for every lane, so its lane and register
  //   position are scalars.  In the general path a scalar branch per register
  //   position patches it, instead of a per-row compare and select.
  // * Rows past the end of the system (general path) are updated too.  They
  //   stay finite (the padding starts at zero), and no row inside the system
  //   reads them: the last row's right neighbour enters with factor zero.

  // ---- forward reduction, in-lane levels -------------------------------------
#pragma unroll
  for (int s = 1; s <= LR; ++s) {
    const int h = 1 << (s - 1);
    level_fence();
    const NtLevel<T>& Lr = P.lv[s - 1 + opaque_zero()];
    const T k1 = uni(Lr.k1), k2 = uni(Lr.k2), k1f = uni(Lr.k1f), kl = uni(Lr.kl);
    // general path: the special row (level-0 index r_sp) and its factors
    int r_sp = -1;
    T K1s = k1, K2s = k2;
    if constexpr (!FULL) {
      const int m = uni(Lr.m);
      const bool odd = (m & 1) != 0;
      r_sp = (((odd ? m - 2 : m - 1) + 1) << (s - 1)) - 1;
      K1s = odd ? k1 : kl;
      K2s = odd ? kl : T(0);
    }
    const T nb = from_next<W>(d[h - 1]);         // row (lane+1)*R + h - 1
#pragma unroll
    for (int p = (1 << s) - 1; p < R; p += (1 << s)) {
      const T left = d[p - h];
      const T right = (p + h < R) ? d[p + h] : nb;
      T K1 = k1;
      if (p == (1 << s) - 1) K1 = first_lane ? k1f : K1;
      T nv;
      if constexpr (FULL) {
        // the last row of the system: last lane, last position
        if (p == R - 1) {
          K1 = last_lane ? kl : K1;
          nv = last_lane ? d[p] - K1 * left : d[p] - K1 * left - k2 * right;
        } else {
          nv = d[p] - K1 * left - k2 * right;
        }
      } else {
        nv = d[p] - K1 * left - k2 * right;
        if ((r_sp & (R - 1)) == p) {             // scalar branch
          NT_KEEP_BRANCH();
          const T sp = d[p] - K1s * left - K2s * right;
          nv = (row0 + p == r_sp) ? sp : nv;
        }
      }
      d[p] = nv;
    }
  }

  // ---- forward reduction, cross-lane levels ----------------------------------
  T v = d[R - 1];
#pragma unroll
  for (int s = LR + 1; s <= LR + LW; ++s) {
    if (!FULL && s > nlev) break;
    if (FULL && s > LR + LW - 1) break;          // n = W*R: nlev = LR + LW - 1
    const int sh = s - 1 - LR;
    const int dl = 1 << sh;
    level_fence();
    const NtLevel<T>& Lr = P.lv[s - 1 + opaque_zero()];
    const T k1 = uni(Lr.k1), k2 = uni(Lr.k2), k1f = uni(Lr.k1f), kl = uni(Lr.kl);
    int l_sp = W - 1;                            // lane of the special row
    T K1s = kl, K2s = T(0);
    if constexpr (!FULL) {
      const int m = uni(Lr.m);
      const bool odd = (m & 1) != 0;
      l_sp = (((odd ? m - 2 : m - 1) + 1) << sh) - 1;
      K1s = odd ? k1 : kl;
      K2s = odd ? kl : T(0);
    }
    const T left = shup<W>(v, dl);
    const T right = shdn<W>(v, dl);
    const bool kept = ((lane + 1) & ((dl << 1) - 1)) == 0;
    const bool is_sp = (lane == l_sp);
    const T K1 = (lane == 2 * dl - 1) ? k1f : (is_sp ? K1s : k1);
    const T K2 = is_sp ? K2s : k2;
    const T nv = v - K1 * left - K2 * right;
    v = kept ? nv : v;
  }

  // ---- top system (m = 2 or 3 equations at level nlev) ------------------------
  {
    level_fence();
    const NtTop<T>& Tp = (&P.top)[opaque_zero()];
    const T t0 = uni(Tp.t[0]), t1 = uni(Tp.t[1]), t2 = uni(Tp.t[2]), t3 = uni(Tp.t[3]);
    const int tm = FULL ? 2 : uni(Tp.m);
    const int nl = FULL ? (LR + LW - 1) : nlev;
    const int H = 1 << nl;                      // row spacing at the top level
    const int l0 = H / R - 1;                   // lane holding top row 0
    const int l1 = 2 * H / R - 1;
    const T d0 = shidx<W>(v, l0);
    const T d1 = shidx<W>(v, l1);
    if (tm == 2) {
      const T x0 = d0 * t0 - d1 * t1;
      const T x1 = d1 * t2 - d0 * t3;
      v = (lane == l0) ? x0 : ((lane == l1) ? x1 : v);
    } else {
      const int l2 = 3 * H / R - 1;
      const T d2 = shidx<W>(v, l2 < W ? l2 : l1);
      const T x1 = (d1 - t0 * d0 - t1 * d2) * t2;
      const T x0 = (d0 - t3 * x1) * uni(Tp.t[4]);
      const T x2 = (d2 - uni(Tp.t[5]) * x1) * uni(Tp.t[6]);
      v = (lane == l0) ? x0 : ((lane == l1) ? x1 : ((lane == l2) ? x2 : v));
    }
  }

  // ---- back substitution, cross-lane levels ----------------------------------
  // interior x = (v - a*xl - c*xr) / b; row 0: (v - cF*xr) / bF; the last row
  // (m odd, so it is eliminated): (v - aL*xl) / bL
#pragma unroll
  for (int s = LR + LW; s >= LR + 1; --s) {
    if (!FULL && s > nlev) continue;
    if (FULL && s > LR + LW - 1) continue;
    const int sh = s - 1 - LR;
    const int dl = 1 << sh;
    level_fence();
    const NtLevel<T>& Lr = P.lv[s - 1 + opaque_zero()];
    const T a = uni(Lr.a), c = uni(Lr.c), ib = uni(Lr.ib), cf = uni(Lr.cf), ibf = uni(Lr.ibf);
    const T left = shup<W>(v, dl);
    const T right = shdn<W>(v, dl);
    const bool elim = ((lane + 1) & ((dl << 1) - 1)) == dl;   // (lane+1)/dl odd
    const bool is_first = (lane == dl - 1);
    T nv = (v - a * left - c * right) * ib;
    const T t_first = (v - cf * right) * ibf;
    nv = is_first ? t_first : nv;
    if constexpr (!FULL) {
      const int m = uni(Lr.m);
      if (m & 1) {                               // scalar branch
        const T al = uni(Lr.al), ibl = uni(Lr.ibl);
        const T t_last = (v - al * left) * ibl;
        nv = (lane == (m << sh) - 1) ? t_last : nv;
      }
    }
    v = elim ? nv : v;
  }
  d[R - 1] = v;

  // ---- back substitution, in-lane levels --------------------------------------
  const T prev = from_prev<W>(v);              // solved row lane*R - 1 (lane 0: unused)
#pragma unroll
  for (int s = LR; s >= 1; --s) {
    const int h = 1 << (s - 1);
    level_fence();
    const NtLevel<T>& Lr = P.lv[s - 1 + opaque_zero()];
    const T a = uni(Lr.a), c = uni(Lr.c), ib = uni(Lr.ib), cf = uni(Lr.cf), ibf = uni(Lr.ibf);
    // general path: level-0 index of the last row when it is eliminated (m odd)
    int r_last = -1;
    T al = T(0), ibl = T(0);
    if constexpr (!FULL) {
      const int m = uni(Lr.m);
      if (m & 1) {
        r_last = (m << (s - 1)) - 1;
        al = uni(Lr.al);
        ibl = uni(Lr.ibl);
      }
    }
#pragma unroll
    for (int p = h - 1; p < R; p += 2 * h) {
      const T left = (p - h >= 0) ? d[p - h] : prev;
      const T right = d[(p + h < R) ? p + h : p];
      T nv = (d[p] - a * left - c * right) * ib;
      if (p == h - 1) {                          // row 0 of the system: lane 0
        const T fst = (d[p] - cf * right) * ibf;
        nv = first_lane ? fst : nv;
      }
      if constexpr (!FULL) {
        if (r_last >= 0 && (r_last & (R - 1)) == p) {   // scalar branch
          NT_KEEP_BRANCH();
          const T lst = (d[p] - al * left) * ibl;
          nv = (row0 + p == r_last) ? lst : nv;
        }
      }
      d[p] = nv;
    }
  }
}

}  // namespace nt
