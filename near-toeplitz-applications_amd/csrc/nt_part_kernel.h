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

// ---- tiles of lines: strided (y-, z-lines, the ADI y-implicit half step) and
// contiguous (x-lines, the ADI x-implicit half step) ------------------------------------
// A workgroup owns a tile of TL lines (TL = 4, 8 or 16, a power of two) that
// are adjacent in memory: for strided lines their rows are runs of TL
// consecutive elements, read with coalesced accesses; contiguous lines are
// TL consecutive rows, read along each line.  The line's right-hand side is
// formed while staging (for the compact derivative from the point and its
// neighbours along the line, for ADI from the neighbouring lines at the same
// position), so the subdomain solves read only their own chunks and overwrite
// them with their solutions.  The TL x P subdomain solves are spread over the
// wavefronts, S = 64/W per wavefront, ordered subdomain-major so that the S
// segments of a wavefront share one subdomain index and with it one (uniform)
// level table, and every segment has work.  The reduced systems of all TL
// lines are solved together; each subdomain's owner then applies the
// correction x = xR + alpha xU + beta xL to its chunk, and the tile is stored.
//
// LDS (dynamic): line l, point e -> l*LP + e + (e >> log2 R) * CP, with
// LP = n + (n >> log2 R) * CP + 2; then the faces and the reduced solutions,
// TL x 2P doubles each.
struct PartTileArgs {
  const NtPartDev* dev;
  const double* xu;     // xU[kind*m + i]
  const double* xl;     // xL[kind*m + i]
  const void* in;
  void* out;
  int64_t lstride;      // between adjacent lines (strided: 1; contiguous: the line length)
  int64_t estride;      // between the points of a line (strided: the row stride; contiguous: 1)
  int64_t ostride;      // between slabs
  int64_t phys;         // lines per slab, including ADI boundary lines
  int64_t tiles;        // tiles per slab
  double s1, s2;        // producer scalars as in LineArgs
  int32_t tl_log2;      // log2 of the lines per tile
  int32_t lp;           // LDS doubles per staged line
  int32_t vec;          // strided: pairs of adjacent lines move as 16-byte vectors
  int32_t pad_;
};

template <int R>
__host__ __device__ constexpr int part_tile_cp() {
  return (R >= 2) ? 1 : 0;
}

// right-hand side of point e of a staged line from the point c and its two
// neighbours a, b (compact: along the line, e-1 and e+1, or the closure
// points; ADI: the neighbouring lines at e)
template <int PROD>
__device__ __forceinline__ double part_rhs(const PartTileArgs& args, int e, int n, double a, double c, double b) {
  if (PROD == P_ADI) return (e == 0 || e == n - 1) ? c : args.s1 * (a + b) + args.s2 * c;
  if (PROD == P_COMPACT) {
    // interior 3/(4h) (f[e+1] - f[e-1]); closures (parallelcompact/rhs.py:42-52)
    if (e == 0) return args.s2 * (-5.0 * c + 4.0 * a + b);
    if (e == n - 1) return -args.s2 * (-5.0 * c + 4.0 * a + b);
    return args.s1 * (b - a);
  }
  return c;
}

// Strided tiles with 16-byte aligned rows: 1 pairs the staging loads, 2 the stores.
#ifndef NT_PART_VEC
#define NT_PART_VEC 3
#endif

template <int R, int W, int PROD, int LAYOUT, bool FULL>
__global__ void __launch_bounds__(1024) part_tile_kernel(const PartTileArgs args) {
  extern __shared__ double lds_dyn[];
  constexpr int HALO = (PROD == P_ADI) ? 1 : 0;
  constexpr int S = 64 / W;                      // subdomains per wavefront
  constexpr int CP = part_tile_cp<R>();
  constexpr int LR = Log2<R>::value;
  constexpr int LW = Log2<W>::value;
  const NtPartDev& D = *args.dev;
  const int m = FULL ? W * R : D.m;
  const int P = D.P, n = D.n;
  const int TLS = args.tl_log2;
  const int TL = 1 << TLS;
  const int LP = args.lp;
  const int NP2 = 2 * P;
  double* __restrict__ lds = lds_dyn;
  double* __restrict__ red = lds_dyn + TL * LP;  // faces, TL x 2P
  double* __restrict__ sol = red + TL * NP2;     // reduced solutions, TL x 2P
  const int tid = threadIdx.x;
  const int nt = blockDim.x;
  const int nw = nt >> 6;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int seg = lane >> LW;
  const int sl = lane & (W - 1);
  auto at = [&](int l, int e) -> int { return l * LP + e + (e >> LR) * CP; };

  // tile geometry: staged lines lo .. hi - 1 of the tile are the slab's lines
  // c0 + lo .. c0 + hi - 1.  Paired strided ADI tiles (args.vec) start at line 0,
  // so their rows are 16-byte aligned, and leave the boundary lines 0 and phys - 1
  // out (lo = 1 in the first tile, hi = phys - 1 - c0 in the last); other tiles
  // start at the first solved line.
  const bool colal = (HALO != 0 && LAYOUT == STRIDED && (NT_PART_VEC & 1) && args.vec);
  const int64_t t = blockIdx.x;
  const int64_t o = t / args.tiles;
  const int64_t p0 = (t - o * args.tiles) << TLS;
  const int64_t c0 = colal ? p0 : p0 + HALO;
  const int lo = (colal && p0 == 0) ? 1 : 0;
  const int64_t rv = (colal ? args.phys - 1 : args.phys - HALO) - c0;
  const int hi = rv < TL ? static_cast<int>(rv) : TL;
  auto valid = [&](int l) -> bool { return l >= lo && l < hi; };
  const int64_t ls = args.lstride, es = args.estride;
  const int64_t base = o * args.ostride + c0 * ls;
  const double* __restrict__ in = static_cast<const double*>(args.in) + base;
  double* __restrict__ out = static_cast<double*>(args.out) + base;

  // ---- stage the right-hand side ------------------------------------------------------
  // neighbours of point (l, e): compact along the line (closure points at its
  // ends), ADI the neighbouring lines
  auto nb = [&](int e, int side) -> int64_t {   // element offset of the side-th neighbour
    if (PROD == P_ADI) return (side == 0) ? -ls : ls;
    const int en = (e == 0) ? (side == 0 ? 1 : 2) : ((e == n - 1) ? (side == 0 ? n - 2 : n - 3)
                                                                  : (side == 0 ? e - 1 : e + 1));
    return static_cast<int64_t>(en - e) * es;
  };
  // Loads for PU points per thread are issued before any of them is used.
  constexpr int PU = 8;
  const int total = n << TLS;
  if (LAYOUT == STRIDED && (NT_PART_VEC & 1) && args.vec) {
    // two adjacent lines l, l + 1 (l even) of a row per thread, TL / 2 consecutive
    // threads per row: their points are one 16-byte aligned vector (c0 is even).
    // ADI takes the outer neighbours, lines l - 1 and l + 2, as scalars; the pair
    // is loaded when either line is solved (both are lines of the slab).
    constexpr int PV = PU / 2;                   // pairs per thread: the same points in flight
    const int total2 = total >> 1;
    const int hm = (TL >> 1) - 1;
    for (int i0 = tid; i0 < total2; i0 += PV * nt) {
      nt_dbl2 c[PV], a[PV], b[PV];
#pragma unroll
      for (int u = 0; u < PV; ++u) {
        const int idx = i0 + u * nt;
        const int l = (idx & hm) << 1, e = idx >> (TLS - 1);
        const bool v0 = idx < total2 && valid(l), v1 = idx < total2 && valid(l + 1);
        c[u].x = c[u].y = a[u].x = a[u].y = b[u].x = b[u].y = 0.0;
        const double* p = in + static_cast<int64_t>(e) * es + l;
        if (PROD == P_ADI) {
          if (v0 || v1) c[u] = __builtin_nontemporal_load(reinterpret_cast<const nt_dbl2*>(p));
          if (v0) a[u].x = p[-1];
          if (v1) b[u].x = p[2];
        } else if (v0 && v1) {
          c[u] = __builtin_nontemporal_load(reinterpret_cast<const nt_dbl2*>(p));
          if (PROD != P_IDENT) {
            a[u] = *reinterpret_cast<const nt_dbl2*>(p + nb(e, 0));
            b[u] = *reinterpret_cast<const nt_dbl2*>(p + nb(e, 1));
          }
        } else if (v0) {
          c[u].x = p[0];
          if (PROD != P_IDENT) {
            a[u].x = p[nb(e, 0)];
            b[u].x = p[nb(e, 1)];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < PV; ++u) {
        const int idx = i0 + u * nt;
        const int l = (idx & hm) << 1, e = idx >> (TLS - 1);
        if (idx >= total2) continue;
        double r0, r1;
        if (PROD == P_ADI) {
          r0 = part_rhs<PROD>(args, e, n, a[u].x, c[u].x, c[u].y);
          r1 = part_rhs<PROD>(args, e, n, c[u].x, c[u].y, b[u].x);
        } else {
          r0 = part_rhs<PROD>(args, e, n, a[u].x, c[u].x, b[u].x);
          r1 = part_rhs<PROD>(args, e, n, a[u].y, c[u].y, b[u].y);
        }
        if (valid(l)) lds[at(l, e)] = r0;
        if (valid(l + 1)) lds[at(l + 1, e)] = r1;
      }
    }
  } else if (LAYOUT == STRIDED) {
    // TL points of a row per TL consecutive threads
    for (int i0 = tid; i0 < total; i0 += PU * nt) {
      double c[PU], a[PU], b[PU];
#pragma unroll
      for (int u = 0; u < PU; ++u) {
        const int idx = i0 + u * nt;
        const int l = idx & (TL - 1), e = idx >> TLS;
        c[u] = a[u] = b[u] = 0.0;
        if (idx < total && valid(l)) {
          const double* p = in + static_cast<int64_t>(e) * es + l;
          c[u] = __builtin_nontemporal_load(p);
          if (PROD != P_IDENT) {
            a[u] = p[nb(e, 0)];
            b[u] = p[nb(e, 1)];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < PU; ++u) {
        const int idx = i0 + u * nt;
        const int l = idx & (TL - 1), e = idx >> TLS;
        if (idx < total && valid(l)) lds[at(l, e)] = part_rhs<PROD>(args, e, n, a[u], c[u], b[u]);
      }
    }
  } else if ((NT_PART_VEC & 1) && args.vec) {
    // contiguous lines, points e, e + 1 (e even) of a line per thread as one
    // 16-byte vector; ADI's neighbours (the adjacent lines) are vectors too, a
    // compact pair's outer neighbours e - 1 and e + 2 scalars (n is even, >= 4)
    constexpr int PV = PU / 2;
    const int n2 = n >> 1;
    const int total2 = total >> 1;
    const int step_l = nt / n2, step_e = nt - step_l * n2;
    int l0 = tid / n2, e0 = tid - l0 * n2;     // e in pairs
    for (int i0 = tid; i0 < total2; i0 += PV * nt) {
      nt_dbl2 c[PV], a[PV], b[PV];
      int lu[PV], eu[PV];
      int l = l0, e = e0;
#pragma unroll
      for (int u = 0; u < PV; ++u) {
        lu[u] = l;
        eu[u] = 2 * e;
        c[u].x = c[u].y = a[u].x = a[u].y = b[u].x = b[u].y = 0.0;
        if (valid(l)) {
          const int ee = 2 * e;
          const double* p = in + l * ls + ee;
          c[u] = __builtin_nontemporal_load(reinterpret_cast<const nt_dbl2*>(p));
          if (PROD == P_ADI) {
            a[u] = *reinterpret_cast<const nt_dbl2*>(p - ls);
            b[u] = *reinterpret_cast<const nt_dbl2*>(p + ls);
          } else if (PROD == P_COMPACT) {
            if (ee > 0) a[u].x = p[-1];
            if (ee + 2 < n) b[u].x = p[2];
          }
        }
        l += step_l;
        e += step_e;
        if (e >= n2) {
          e -= n2;
          ++l;
        }
      }
      l0 = l;
      e0 = e;
#pragma unroll
      for (int u = 0; u < PV; ++u) {
        if (!valid(lu[u])) continue;
        const int ee = eu[u];
        double r0, r1;
        if (PROD == P_ADI) {
          r0 = part_rhs<PROD>(args, ee, n, a[u].x, c[u].x, b[u].x);
          r1 = part_rhs<PROD>(args, ee + 1, n, a[u].y, c[u].y, b[u].y);
        } else if (PROD == P_COMPACT) {
          // neighbours as nb(): closures take f[1], f[2] at e = 0 and f[n-2], f[n-3] at n - 1
          const double left = a[u].x, right = b[u].x;
          r0 = (ee == 0) ? part_rhs<PROD>(args, 0, n, c[u].y, c[u].x, right)
                         : part_rhs<PROD>(args, ee, n, left, c[u].x, c[u].y);
          r1 = (ee + 1 == n - 1) ? part_rhs<PROD>(args, n - 1, n, c[u].x, c[u].y, left)
                                 : part_rhs<PROD>(args, ee + 1, n, c[u].x, c[u].y, right);
        } else {
          r0 = c[u].x;
          r1 = c[u].y;
        }
        lds[at(lu[u], ee)] = r0;
        lds[at(lu[u], ee + 1)] = r1;
      }
    }
  } else {
    // consecutive threads along each line; point idx of the tile is (idx / n, idx % n),
    // advanced incrementally from one division per thread
    const int step_l = nt / n, step_e = nt - step_l * n;
    int l0 = tid / n, e0 = tid - l0 * n;
    for (int i0 = tid; i0 < total; i0 += PU * nt) {
      double c[PU], a[PU], b[PU];
      int lu[PU], eu[PU];
      int l = l0, e = e0;
#pragma unroll
      for (int u = 0; u < PU; ++u) {
        lu[u] = l;
        eu[u] = e;
        c[u] = a[u] = b[u] = 0.0;
        if (valid(l)) {
          const double* p = in + l * ls + e;
          c[u] = __builtin_nontemporal_load(p);
          if (PROD != P_IDENT) {
            a[u] = p[nb(e, 0)];
            b[u] = p[nb(e, 1)];
          }
        }
        l += step_l;
        e += step_e;
        if (e >= n) {
          e -= n;
          ++l;
        }
      }
      l0 = l;
      e0 = e;
#pragma unroll
      for (int u = 0; u < PU; ++u)
        if (valid(lu[u])) lds[at(lu[u], eu[u])] = part_rhs<PROD>(args, eu[u], n, a[u], c[u], b[u]);
    }
  }
  __syncthreads();

  // ---- subdomain solves: task tau = q * TL + l, S per wavefront ----------------------
  const int ntasks = P << TLS;
  const int last_lane = (m - 1) / R, last_k = (m - 1) % R;
  for (int wt = w; wt * S < ntasks; wt += nw) {
    const int tau0 = wt * S;                     // uniform
    const int q = tau0 >> TLS;                   // the same for the S segments
    const int l = (tau0 & (TL - 1)) + seg;
    const int kind = (P == 1) ? 0 : (q == 0 ? 0 : (q == P - 1 ? 2 : 1));
    const bool active = valid(l);
    const int e0 = q * m + sl * R;
    double d[R];
#pragma unroll
    for (int k = 0; k < R; ++k) d[k] = (active && (FULL || sl * R + k < m)) ? lds[at(l, e0 + k)] : 0.0;
    cr_solve_regs<double, R, W, FULL>(d, D.plan[kind], D.nlev, sl);
#pragma unroll
    for (int k = 0; k < R; ++k)
      if (active && (FULL || sl * R + k < m)) lds[at(l, e0 + k)] = d[k];
    // faces: negated first/last value of the subdomain (reduced.py:51-88)
    if (active && sl == 0) red[l * NP2 + 2 * q] = (q == 0) ? 0.0 : -d[0];
    double v = d[0];
#pragma unroll
    for (int k = 1; k < R; ++k) v = (k == last_k) ? d[k] : v;
    if (active && sl == last_lane) red[l * NP2 + 2 * q + 1] = (q == P - 1) ? 0.0 : -v;
  }
  __syncthreads();

  // ---- reduced systems: (alpha_0, beta_0, alpha_1, ...) = inverse * faces, per line --
  for (int i = tid; i < (NP2 << TLS); i += nt) {
    const int l = i / NP2, r = i - l * NP2;
    double u = 0.0;
    if (valid(l)) {
      const double* __restrict__ row = D.red_inv + r * NP2;
      const double* __restrict__ f = red + l * NP2;
      for (int j = 0; j < NP2; ++j) u += row[j] * f[j];
    }
    sol[i] = u;
  }
  __syncthreads();

  // ---- correction x = xR + alpha xU + beta xL (sum.py:30), by each chunk's owner -------
  for (int wt = w; wt * S < ntasks; wt += nw) {
    const int tau0 = wt * S;
    const int q = tau0 >> TLS;
    const int l = (tau0 & (TL - 1)) + seg;
    const int kind = (P == 1) ? 0 : (q == 0 ? 0 : (q == P - 1 ? 2 : 1));
    if (valid(l)) {
      const double alpha = sol[l * NP2 + 2 * q], beta = sol[l * NP2 + 2 * q + 1];
      const double* __restrict__ xu = args.xu + kind * m + sl * R;
      const double* __restrict__ xl = args.xl + kind * m + sl * R;
      const int e0 = q * m + sl * R;
#pragma unroll
      for (int k = 0; k < R; ++k)
        if (FULL || sl * R + k < m) lds[at(l, e0 + k)] += alpha * xu[k] + beta * xl[k];
    }
  }
  __syncthreads();

  // ---- store ---------------------------------------------------------------------------
  if (LAYOUT == STRIDED && (NT_PART_VEC & 2) && args.vec && (HALO == 0 || colal)) {
    // lines l, l + 1 (l even) of a row as one aligned 16-byte store (c0 is even),
    // TL / 2 threads per row; a pair with one solved line stores that line alone
    const int hm = (TL >> 1) - 1;
    for (int idx = tid; idx < (total >> 1); idx += nt) {
      const int l = (idx & hm) << 1, e = idx >> (TLS - 1);
      double* __restrict__ row = out + static_cast<int64_t>(e) * es;
      const bool v0 = valid(l), v1 = valid(l + 1);
      if (v0 && v1) {
        nt_dbl2 v;
        v.x = lds[at(l, e)];
        v.y = lds[at(l + 1, e)];
        __builtin_nontemporal_store(v, reinterpret_cast<nt_dbl2*>(row + l));
      } else if (v0) {
        __builtin_nontemporal_store(lds[at(l, e)], row + l);
      } else if (v1) {
        __builtin_nontemporal_store(lds[at(l + 1, e)], row + l + 1);
      }
    }
  } else if (LAYOUT == STRIDED) {
    for (int idx = tid; idx < total; idx += nt) {
      const int l = idx & (TL - 1), e = idx >> TLS;
      if (valid(l)) __builtin_nontemporal_store(lds[at(l, e)], out + static_cast<int64_t>(e) * es + l);
    }
  } else if ((NT_PART_VEC & 2) && args.vec) {
    for (int l = lo; l < hi; ++l) {
      double* __restrict__ dst = out + l * ls;
      for (int e = 2 * tid; e < n; e += 2 * nt) {
        nt_dbl2 v;
        v.x = lds[at(l, e)];
        v.y = lds[at(l, e + 1)];
        __builtin_nontemporal_store(v, reinterpret_cast<nt_dbl2*>(dst + e));
      }
    }
  } else {
    for (int l = lo; l < hi; ++l) {
      double* __restrict__ dst = out + l * ls;
      for (int e = tid; e < n; e += nt) __builtin_nontemporal_store(lds[at(l, e)], dst + e);
    }
  }
}

}  // namespace nt
