// The line kernel: every batched near-Toeplitz solve in the library is one
// launch of this template.
//
// A workgroup owns a tile of TS systems (lines) that are adjacent in memory:
//   CONTIG   lines are contiguous rows (x-lines): system p starts at p*pstride,
//            elements are unit-stride;
//   STRIDED  lines are strided (y- or z-lines): TS consecutive "inner" indices
//            (unit stride between systems), elements at estride.
// The line is streamed through LDS in blocks of E elements for all TS (+ halo)
// systems of the tile with fully coalesced global accesses, and each lane
// copies its own chunk of R consecutive rows into registers (so the
// permute/transpose of the reference, compact/permute.py and adi/transpose.py,
// never touches HBM).  The producer forms the right-hand side on the fly
// (reference compact/rhs.py:21-36, adi/rhs.py:21-24), the solve runs in
// registers (nt_cr.h), and the solution goes back through LDS to HBM with the
// same coalesced pattern: each field element is read once and written once.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "nt_cr.h"
#include "nt_plan.h"

namespace nt {

enum Layout : int { CONTIG = 0, STRIDED = 1 };
enum Producer : int { P_IDENT = 0, P_COMPACT = 1, P_ADI = 2 };
enum Precision : int { PREC_F64 = 0, PREC_F32R = 1 };

struct LineArgs {
  NtPlanK plan;
  const void* in;
  void* out;
  int64_t pstride;     // distance between adjacent systems (CONTIG: line length)
  int64_t estride;     // distance between adjacent elements (CONTIG: 1)
  int64_t ostride;     // distance between outer slabs
  int64_t phys;        // lines per slab including halo lines (ADI boundary lines)
  int64_t tiles;       // tiles per slab
  int64_t total;       // tiles over all slabs (persistent kernels)
  double s1, s2;       // producer scalars: compact 3/(4h), 1/(2h); ADI r/2, 1-r
  int32_t sweeps;      // refinement sweeps (PREC_F32R)
  int32_t pad_;
  const NtLaneTab* lane;   // host copy of the plan's lane-solve tables or null (never read on the device)
  const NtLaneTab32* lane32;   // their FP32 form (every block full), or null
};

// The launch arguments read in place from the kernel-argument segment (the
// struct is the kernel's only argument, at offset 0).  Accessing them through
// this pointer keeps every level-table read a scalar load from constant
// memory; using the by-value parameter with a run-time level index makes the
// compiler copy the whole table to scratch.
__device__ __forceinline__ const LineArgs& kernel_args() {
  return *(const LineArgs*)(__builtin_amdgcn_kernarg_segment_ptr());
}

template <typename T>
__device__ __forceinline__ double ld_as_double(const T* p) {
  return static_cast<double>(*p);
}

template <typename T>
__device__ __forceinline__ void st_from_double(T* p, double v) {
  *p = static_cast<T>(v);
}

// (A x)[e] of the level-0 near-Toeplitz matrix for row e of my chunk (left and
// right: rows e-1 and e+1); the same expressions as in residual().  The three
// row forms are computed from uniform (scalar) loads and the result selected:
// with the coefficient loads inside branches on e, the compiler merges them
// into one per-lane vector load from a selected address, and that load waits
// behind every store still in flight (the vector-memory counter is in order),
// so the refinement sweep could not overlap the previous stores.
__device__ __forceinline__ double apply_row(const NtPlanK& P, int e, int n, double left, double xe, double right) {
  const double first = P.b1 * xe + P.c1 * right;
  const double last = P.an * left + P.bn * xe;
  const double inner = P.ai * left + P.bi * xe + P.ci * right;
  return (e == 0) ? first : ((e == n - 1) ? last : inner);
}

// residual r = d - A x for the level-0 near-Toeplitz matrix, chunked layout,
// formed in FP64 and rounded to the type of r (the FP32 correction solve takes
// it as float, so no FP64 copy of r is kept)
template <int R, int W, typename TR>
__device__ __forceinline__ void residual(const double (&d)[R], const double (&x)[R], TR (&r)[R],
                                         const NtPlanK& P, int lane) {
  const double xl = shup<W>(x[R - 1], 1);
  const double xr = shdn<W>(x[0], 1);
  const int n = P.n;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int e = lane * R + k;
    const double left = (k > 0) ? x[k - 1] : xl;
    const double right = (k + 1 < R) ? x[k + 1] : xr;
    const double ax = apply_row(P, e, n, left, x[k], right);
    r[k] = static_cast<TR>((e < n) ? d[k] - ax : 0.0);
  }
}


// Compact RHS (compact/rhs.py:21-36) of row k of my chunk from the chunk
// values f (any floating type), the neighbour rows fl (row my_e0-1) and fr
// (row my_e0+R), and the closures b0 (first row) and bn (last row).  FULL
// (n == W*R): only the first row of lane 0 and the last row of the last lane
// are closures.
template <int R, bool FULL, typename TF>
__device__ __forceinline__ double compact_rhs_elem(int k, const TF (&f)[R], double fl, double fr, double b0,
                                                   double bn, int my_e0, int n, double s1) {
  const int e = my_e0 + k;
  const double left = (k > 0) ? static_cast<double>(f[k - 1]) : fl;
  const double right = (k + 1 < R) ? static_cast<double>(f[k + 1]) : fr;
  const double in_v = s1 * (right - left);
  if constexpr (FULL) {
    double v = in_v;
    if (k == 0) v = (e == 0) ? b0 : v;
    if (k == R - 1) v = (e == n - 1) ? bn : v;
    return v;
  } else {
    return (e == 0) ? b0 : ((e == n - 1) ? bn : ((e < n) ? in_v : 0.0));
  }
}

// closures b0 = s2 (-5 f0 + 4 f1 + f2), bn = -s2 (-5 f[n-1] + 4 f[n-2] + f[n-3])
// from bnd = f[0], f[1], f[2], f[n-1], f[n-2], f[n-3]
template <typename TB>
__device__ __forceinline__ void compact_closures(const TB (&bnd)[6], double s2, double& b0, double& bn) {
  b0 = s2 * (-5.0 * static_cast<double>(bnd[0]) + 4.0 * static_cast<double>(bnd[1]) + static_cast<double>(bnd[2]));
  bn = -s2 * (-5.0 * static_cast<double>(bnd[3]) + 4.0 * static_cast<double>(bnd[4]) + static_cast<double>(bnd[5]));
}

// Compact RHS of my whole chunk (see compact_rhs_elem).
template <int R, bool FULL = false>
__device__ __forceinline__ void compact_rhs_chunk(const double (&f)[R], double fl, double fr,
                                                  const double (&bnd)[6], int my_e0, int n,
                                                  double s1, double s2, double (&d)[R]) {
  double b0, bn;
  compact_closures(bnd, s2, b0, bn);
#pragma unroll
  for (int k = 0; k < R; ++k) d[k] = compact_rhs_elem<R, FULL>(k, f, fl, fr, b0, bn, my_e0, n, s1);
}

// Compact derivative of my chunk straight from FP32 chunk values f: the
// right-hand side is formed in FP64 where it is used, so with FP32 + refinement
// no FP64 copy of it stays live through the solves (it is recomputed from f
// for each residual).  Gives the same bits as compact_rhs_chunk + solve_chunk.
template <int R, int W, int PREC, bool FULL>
__device__ __forceinline__ void compact_solve_from_f32(const float (&f)[R], const float (&bnd)[6], int my_e0,
                                                       int n, double s1, double s2, double (&x)[R],
                                                       const NtPlanK& P, int lane, int sweeps) {
  const double fl = shup<W>(static_cast<double>(f[R - 1]), 1);
  const double fr = shdn<W>(static_cast<double>(f[0]), 1);
  double b0, bn;
  compact_closures(bnd, s2, b0, bn);
  if constexpr (PREC == PREC_F64) {
#pragma unroll
    for (int k = 0; k < R; ++k) x[k] = compact_rhs_elem<R, FULL>(k, f, fl, fr, b0, bn, my_e0, n, s1);
    cr_solve_regs<double, R, W, FULL>(x, P.d, P.nlev, lane);
  } else {
    float s[R];
#pragma unroll
    for (int k = 0; k < R; ++k)
      s[k] = static_cast<float>(compact_rhs_elem<R, FULL>(k, f, fl, fr, b0, bn, my_e0, n, s1));
    cr_solve_regs<float, R, W, FULL>(s, P.f, P.nlev, lane);
#pragma unroll
    for (int k = 0; k < R; ++k) x[k] = static_cast<double>(s[k]);
    for (int it = 0; it < sweeps; ++it) {
      const double xl = shup<W>(x[R - 1], 1);
      const double xr = shdn<W>(x[0], 1);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int e = lane * R + k;
        const double left = (k > 0) ? x[k - 1] : xl;
        const double right = (k + 1 < R) ? x[k + 1] : xr;
        const double dk = compact_rhs_elem<R, FULL>(k, f, fl, fr, b0, bn, my_e0, n, s1);
        s[k] = static_cast<float>((e < n) ? dk - apply_row(P, e, n, left, x[k], right) : 0.0);
      }
      cr_solve_regs<float, R, W, FULL>(s, P.f, P.nlev, lane);
#pragma unroll
      for (int k = 0; k < R; ++k) x[k] += static_cast<double>(s[k]);
    }
  }
}

// Solve A x = d for my chunk: FP64 CR, or FP32 CR + `sweeps` FP64-residual
// refinement sweeps (x += A32^{-1} (d - A x), residual in FP64).
template <int R, int W, int PREC, bool FULL = false>
__device__ __forceinline__ void solve_chunk(const double (&d)[R], double (&x)[R], const NtPlanK& P,
                                            int lane, int sweeps) {
  if (PREC == PREC_F64) {
#pragma unroll
    for (int k = 0; k < R; ++k) x[k] = d[k];
    cr_solve_regs<double, R, W, FULL>(x, P.d, P.nlev, lane);
  } else {
    float s[R];
#pragma unroll
    for (int k = 0; k < R; ++k) s[k] = static_cast<float>(d[k]);
    cr_solve_regs<float, R, W, FULL>(s, P.f, P.nlev, lane);
#pragma unroll
    for (int k = 0; k < R; ++k) x[k] = static_cast<double>(s[k]);
    for (int it = 0; it < sweeps; ++it) {
      residual<R, W>(d, x, s, P, lane);
      cr_solve_regs<float, R, W, FULL>(s, P.f, P.nlev, lane);
#pragma unroll
      for (int k = 0; k < R; ++k) x[k] += static_cast<double>(s[k]);
    }
  }
}

template <typename TIn, typename TOut, int R, int W, int WAVES, int E, int LAYOUT, int PROD, int PREC,
          bool FULL>
__global__ void __launch_bounds__(WAVES * 64) line_kernel(const LineArgs args_by_value) {
  (void)args_by_value;
  const LineArgs& args = kernel_args();
  constexpr int NT = WAVES * 64;
  constexpr int SPW = 64 / W;                 // systems (segments) per wave
  constexpr int TS = WAVES * SPW;             // systems per tile
  constexpr int HALO = (PROD == P_ADI) ? 1 : 0;
  constexpr int ROWS = TS + 2 * HALO;         // lines staged per block
  constexpr int CP = (R >= 4) ? 2 : 0;        // pad doubles after each R-chunk (LDS banks)
  constexpr int EP = (E / R) * (R + CP) + 2;  // LDS stride between staged lines
  static_assert(E % R == 0, "a chunk must not straddle blocks");

  __shared__ __attribute__((aligned(16))) double lds[ROWS * EP];

  const NtPlanK& P = args.plan;
  const int n = P.n;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int seg = (tid & 63) / W;
  const int lane = tid & (W - 1);
  const int ql = wave * SPW + seg;            // tile-local system

  const int64_t bid = blockIdx.x;
  const int64_t o = bid / args.tiles;
  const int64_t p0 = (bid - o * args.tiles) * TS;   // first staged line of the tile
  const int64_t nsys = args.phys - 2 * HALO;
  const bool active = (p0 + ql) < nsys;
  const int64_t base = o * args.ostride + p0 * args.pstride;
  const TIn* __restrict__ in = static_cast<const TIn*>(args.in) + base;
  TOut* __restrict__ out = static_cast<TOut*>(args.out) + base;
  const int64_t rows_valid64 = args.phys - p0;
  const int rows_valid = rows_valid64 < ROWS ? static_cast<int>(rows_valid64) : ROWS;

  const int my_e0 = lane * R;                 // first row of my chunk
  const int my_blk = my_e0 / E;
  const int nblk = (n + E - 1) / E;

  double f[R];
#pragma unroll
  for (int k = 0; k < R; ++k) f[k] = 0.0;
  // compact closures need f[0..2] and f[n-1], f[n-2], f[n-3] of the line
  double bnd[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};

  auto lds_at = [&](int row, int el) -> int { return row * EP + el + (el / R) * CP; };

  // ---------------- load: blocks of E elements of all staged lines ----------
  for (int b = 0; b < nblk; ++b) {
    const int e0 = b * E;
    if (b > 0) __syncthreads();
    if (LAYOUT == CONTIG) {
#pragma unroll 4
      for (int idx = tid; idx < ROWS * E; idx += NT) {
        const int r = idx / E, el = idx % E;
        if (r < rows_valid && e0 + el < n)
          lds[lds_at(r, el)] = ld_as_double(in + r * args.pstride + (e0 + el));
      }
    } else {
#pragma unroll 4
      for (int idx = tid; idx < ROWS * E; idx += NT) {
        const int r = idx % ROWS, el = idx / ROWS;
        if (r < rows_valid && e0 + el < n)
          lds[lds_at(r, el)] = ld_as_double(in + r + (int64_t)(e0 + el) * args.estride);
      }
    }
    __syncthreads();
    if (my_blk == b && active) {
      const int el0 = my_e0 - e0;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int e = my_e0 + k;
        if (e < n) {
          const double c = lds[lds_at(ql + HALO, el0 + k)];
          if (PROD == P_ADI) {
            const double up = lds[lds_at(ql, el0 + k)];
            const double dn = lds[lds_at(ql + 2, el0 + k)];
            f[k] = (e == 0 || e == n - 1) ? c : args.s1 * (up + dn) + args.s2 * c;
          } else {
            f[k] = c;
          }
        }
      }
    }
    if (PROD == P_COMPACT && active) {
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        const int e = (i < 3) ? i : n + 2 - i;
        if (e >= e0 && e < e0 + E) bnd[i] = lds[lds_at(ql + HALO, e - e0)];
      }
    }
  }

  // ---------------- right-hand side and solve ---------------------------------
  double d[R];
  if (PROD == P_COMPACT) {
    const double fl = shup<W>(f[R - 1], 1);      // row my_e0 - 1
    const double fr = shdn<W>(f[0], 1);          // row my_e0 + R
    compact_rhs_chunk<R, FULL>(f, fl, fr, bnd, my_e0, n, args.s1, args.s2, d);
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k) d[k] = f[k];
  }
  double x[R];
  solve_chunk<R, W, PREC, FULL>(d, x, P, lane, args.sweeps);

  // ---------------- store: blocks back through LDS ---------------------------
  for (int b = 0; b < nblk; ++b) {
    const int e0 = b * E;
    __syncthreads();
    if (my_blk == b && active) {
      const int el0 = my_e0 - e0;
#pragma unroll
      for (int k = 0; k < R; ++k)
        if (my_e0 + k < n) lds[lds_at(ql + HALO, el0 + k)] = x[k];
    }
    __syncthreads();
    if (LAYOUT == CONTIG) {
#pragma unroll 4
      for (int idx = tid; idx < TS * E; idx += NT) {
        const int r = idx / E, el = idx % E;
        if (p0 + r < nsys && e0 + el < n)
          st_from_double(out + (r + HALO) * args.pstride + (e0 + el), lds[lds_at(r + HALO, el)]);
      }
    } else {
#pragma unroll 4
      for (int idx = tid; idx < TS * E; idx += NT) {
        const int r = idx % TS, el = idx / TS;
        if (p0 + r < nsys && e0 + el < n)
          st_from_double(out + (r + HALO) + (int64_t)(e0 + el) * args.estride,
                         lds[lds_at(r + HALO, el)]);
      }
    }
  }
}


// ---- contiguous lines without LDS staging ---------------------------------------
// Chunk loads/stores: 16-byte accesses for whole, aligned chunks (the host only
// picks this kernel when lines start 16-byte aligned and n is a multiple of the
// vector width), element accesses for the partial chunk at the end of a line.
template <typename T, int R>
__device__ __forceinline__ void load_chunk(const T* __restrict__ line, int e0, int n, double (&f)[R]) {
  constexpr int V = 16 / sizeof(T);
  if (R % V == 0 && e0 + R <= n) {
    using VT = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
    const VT* p = reinterpret_cast<const VT*>(line + e0);
#pragma unroll
    for (int v = 0; v < R / V; ++v) {
      const VT t = p[v];
      if constexpr (sizeof(T) == 8) {
        f[v * 2] = t.x;
        f[v * 2 + 1] = t.y;
      } else {
        f[v * 4] = t.x;
        f[v * 4 + 1] = t.y;
        f[v * 4 + 2] = t.z;
        f[v * 4 + 3] = t.w;
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k) f[k] = (e0 + k < n) ? static_cast<double>(line[e0 + k]) : 0.0;
  }
}

template <typename T, int R>
__device__ __forceinline__ void store_chunk(T* __restrict__ line, int e0, int n, const double (&x)[R]) {
  constexpr int V = 16 / sizeof(T);
  if (R % V == 0 && e0 + R <= n) {
    using VT = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
    VT* p = reinterpret_cast<VT*>(line + e0);
#pragma unroll
    for (int v = 0; v < R / V; ++v) {
      VT t;
      if constexpr (sizeof(T) == 8) {
        t.x = x[v * 2];
        t.y = x[v * 2 + 1];
      } else {
        t.x = static_cast<float>(x[v * 4]);
        t.y = static_cast<float>(x[v * 4 + 1]);
        t.z = static_cast<float>(x[v * 4 + 2]);
        t.w = static_cast<float>(x[v * 4 + 3]);
      }
      p[v] = t;
    }
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k)
      if (e0 + k < n) line[e0 + k] = static_cast<T>(x[k]);
  }
}

// Every lane reads its R consecutive rows directly; the lanes of a segment
// together read the line's contiguous bytes, neighbour rows cross lanes by
// shuffles, and there is no LDS and no barrier.  Each segment (W lanes) owns
// one line, a workgroup TS = WAVES*64/W lines.
template <typename TIn, typename TOut, int R, int W, int WAVES, int PROD, int PREC, bool FULL>
__global__ void __launch_bounds__(WAVES * 64) line_kernel_direct(const LineArgs args_by_value) {
  (void)args_by_value;
  const LineArgs& args = kernel_args();
  constexpr int SPW = 64 / W;
  constexpr int TS = WAVES * SPW;
  constexpr int HALO = (PROD == P_ADI) ? 1 : 0;
  const NtPlanK& P = args.plan;
  const int n = P.n;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int seg = (tid & 63) / W;
  const int lane = tid & (W - 1);
  const int64_t nsys = args.phys - 2 * HALO;
  const int64_t q = static_cast<int64_t>(blockIdx.x) * TS + wave * SPW + seg;
  if (q >= nsys) return;   // the whole segment leaves together; no barrier follows
  const int64_t row = q + HALO;
  const TIn* __restrict__ line = static_cast<const TIn*>(args.in) + row * args.pstride;
  TOut* __restrict__ oline = static_cast<TOut*>(args.out) + row * args.pstride;
  const int my_e0 = lane * R;

  double f[R];
  load_chunk<TIn, R>(line, my_e0, n, f);
  double d[R];
  if (PROD == P_COMPACT) {
    double bnd[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) bnd[i] = static_cast<double>(line[(i < 3) ? i : n + 2 - i]);
    const double fl = shup<W>(f[R - 1], 1);
    const double fr = shdn<W>(f[0], 1);
    compact_rhs_chunk<R, FULL>(f, fl, fr, bnd, my_e0, n, args.s1, args.s2, d);
  } else if (PROD == P_ADI) {
    double up[R], dn[R];
    load_chunk<TIn, R>(line - args.pstride, my_e0, n, up);
    load_chunk<TIn, R>(line + args.pstride, my_e0, n, dn);
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int e = my_e0 + k;
      d[k] = (e == 0 || e >= n - 1) ? f[k] : args.s1 * (up[k] + dn[k]) + args.s2 * f[k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k) d[k] = f[k];
  }
  double x[R];
  solve_chunk<R, W, PREC, FULL>(d, x, P, lane, args.sweeps);
  store_chunk<TOut, R>(oline, my_e0, n, x);
}

// ---- contiguous lines: wave-coalesced streaming loads + wave-private LDS transpose --
// The lines of one wavefront are read and written with fully coalesced
// nontemporal 16-byte accesses (instruction k of the wave covers bytes
// 1024k .. 1024k+1023 of a line), and the wavefront's own LDS slice turns that
// order into the chunked register layout of the solve and back.  No
// workgroup barrier: each wavefront owns its lines and its LDS slice.
typedef double nt_dbl2 __attribute__((ext_vector_type(2)));
typedef float nt_flt4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename TIn, typename TOut, int R, int W, int WAVES, int PROD, int PREC, bool FULL>
__global__ void __launch_bounds__(WAVES * 64) line_kernel_wct(const LineArgs args_by_value) {
  (void)args_by_value;
  const LineArgs& args = kernel_args();
  constexpr int SPW = 64 / W;                 // lines per wavefront
  constexpr int HALO = (PROD == P_ADI) ? 1 : 0;
  constexpr int CS = R + 2;                   // LDS doubles per chunk (2 pad: bank spread)
  constexpr int LS = W * CS;                  // LDS doubles per line
  static_assert(R >= 2, "chunk pairs need R >= 2");
  __shared__ __attribute__((aligned(16))) double lds_all[WAVES * SPW * LS];

  const NtPlanK& P = args.plan;
  const int n = P.n;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int l64 = tid & 63;
  const int seg = l64 / W;
  const int lane = l64 & (W - 1);
  double* __restrict__ L = lds_all + wave * SPW * LS;
  const int64_t nsys = args.phys - 2 * HALO;
  const int64_t q0 = (static_cast<int64_t>(blockIdx.x) * WAVES + wave) * SPW;   // first line of the wave
  if (q0 >= nsys) return;                     // the whole wavefront leaves together
  const int nl = (nsys - q0 < SPW) ? static_cast<int>(nsys - q0) : SPW;
  auto at = [&](int s, int e) -> int { return s * LS + (e / R) * CS + (e % R); };

  // ---- coalesced streaming load of the wave's lines into LDS --------------------
  // (FULL: n == W*R, so each loop runs a compile-time number of times)
  const int nv2 = FULL ? W * R / 2 : n / 2;
  const int nv4 = FULL ? W * R / 4 : n / 4;
  for (int s = 0; s < nl; ++s) {
    const TIn* __restrict__ src = static_cast<const TIn*>(args.in) + (q0 + s + HALO) * args.pstride;
    if constexpr (sizeof(TIn) == 8 && PROD == P_ADI) {
      // The explicit stencil across lines is formed while streaming: the two
      // neighbouring lines are read with the same coalesced 16-byte accesses
      // (per-lane chunk reads of them touch 64 cache lines per instruction, and
      // the 2048-point ADI half step waited on that address traffic).  Plain
      // loads: each line is read again as a neighbour by two other wavefronts.
      const nt_dbl2* v = reinterpret_cast<const nt_dbl2*>(src);
      const nt_dbl2* vu = reinterpret_cast<const nt_dbl2*>(src - args.pstride);
      const nt_dbl2* vd = reinterpret_cast<const nt_dbl2*>(src + args.pstride);
#pragma unroll
      for (int j = l64; j < nv2; j += 64) {
        const nt_dbl2 c = v[j], u = vu[j], w = vd[j];
        const int e = 2 * j;
        L[at(s, e)] = (e == 0 || e >= n - 1) ? c.x : args.s1 * (u.x + w.x) + args.s2 * c.x;
        L[at(s, e + 1)] = (e + 1 >= n - 1) ? c.y : args.s1 * (u.y + w.y) + args.s2 * c.y;
      }
    } else if constexpr (sizeof(TIn) == 8) {
      const nt_dbl2* v = reinterpret_cast<const nt_dbl2*>(src);
#pragma unroll
      for (int j = l64; j < nv2; j += 64) {
        const nt_dbl2 t = __builtin_nontemporal_load(v + j);
        L[at(s, 2 * j)] = t.x;
        L[at(s, 2 * j + 1)] = t.y;
      }
    } else {
      const nt_flt4* v = reinterpret_cast<const nt_flt4*>(src);
#pragma unroll
      for (int j = l64; j < nv4; j += 64) {
        const nt_flt4 t = __builtin_nontemporal_load(v + j);
        L[at(s, 4 * j)] = t.x;
        L[at(s, 4 * j + 1)] = t.y;
        L[at(s, 4 * j + 2)] = t.z;
        L[at(s, 4 * j + 3)] = t.w;
      }
    }
  }
  wave_lds_sync();

  // ---- my chunk, right-hand side, solve ---------------------------------------------
  // A segment past the wavefront's last line (SPW > 1, the grid's last
  // wavefront) solves zeros: its LDS slice holds no line, and every value a
  // shuffle brings across a segment boundary must be finite.
  const bool seg_on = (SPW == 1) || seg < nl;
  const int my_e0 = lane * R;
  double f[R];
  double* __restrict__ mine = L + seg * LS + lane * CS;     // 16-byte aligned chunk
  if (FULL || my_e0 + R <= n) {
#pragma unroll
    for (int k = 0; k < R; k += 2) {
      const nt_dbl2 t = *reinterpret_cast<const nt_dbl2*>(mine + k);
      f[k] = seg_on ? t.x : 0.0;
      f[k + 1] = seg_on ? t.y : 0.0;
    }
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k) f[k] = (seg_on && my_e0 + k < n) ? mine[k] : 0.0;
  }
  double d[R];
  if (PROD == P_COMPACT) {
    double bnd[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) bnd[i] = seg_on ? L[at(seg, (i < 3) ? i : n + 2 - i)] : 0.0;
    const double fl = shup<W>(f[R - 1], 1);
    const double fr = shdn<W>(f[0], 1);
    compact_rhs_chunk<R, FULL>(f, fl, fr, bnd, my_e0, n, args.s1, args.s2, d);
  } else if (PROD == P_ADI && sizeof(TIn) != 8) {
    const TIn* __restrict__ line =
        static_cast<const TIn*>(args.in) + (q0 + (seg_on ? seg : 0) + HALO) * args.pstride;
    double up[R], dn[R];
    load_chunk<TIn, R>(line - args.pstride, my_e0, n, up);
    load_chunk<TIn, R>(line + args.pstride, my_e0, n, dn);
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int e = my_e0 + k;
      d[k] = (e == 0 || e >= n - 1) ? f[k] : args.s1 * (up[k] + dn[k]) + args.s2 * f[k];
    }
  } else {                                    // solve as staged (FP64 ADI: the RHS already)
#pragma unroll
    for (int k = 0; k < R; ++k) d[k] = f[k];
  }
  double x[R];
  solve_chunk<R, W, PREC, FULL>(d, x, P, lane, args.sweeps);
  wave_lds_sync();
  if (FULL || my_e0 + R <= n) {
#pragma unroll
    for (int k = 0; k < R; k += 2) {
      nt_dbl2 t;
      t.x = x[k];
      t.y = x[k + 1];
      *reinterpret_cast<nt_dbl2*>(mine + k) = t;
    }
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k)
      if (my_e0 + k < n) mine[k] = x[k];
  }
  wave_lds_sync();

  // ---- coalesced streaming store ------------------------------------------------
  for (int s = 0; s < nl; ++s) {
    TOut* __restrict__ dst = static_cast<TOut*>(args.out) + (q0 + s + HALO) * args.pstride;
    if constexpr (sizeof(TOut) == 8) {
      nt_dbl2* v = reinterpret_cast<nt_dbl2*>(dst);
#pragma unroll
      for (int j = l64; j < nv2; j += 64) {
        nt_dbl2 t;
        t.x = L[at(s, 2 * j)];
        t.y = L[at(s, 2 * j + 1)];
        __builtin_nontemporal_store(t, v + j);
      }
    } else {
      nt_flt4* v = reinterpret_cast<nt_flt4*>(dst);
#pragma unroll
      for (int j = l64; j < nv4; j += 64) {
        nt_flt4 t;
        t.x = static_cast<float>(L[at(s, 4 * j)]);
        t.y = static_cast<float>(L[at(s, 4 * j + 1)]);
        t.z = static_cast<float>(L[at(s, 4 * j + 2)]);
        t.w = static_cast<float>(L[at(s, 4 * j + 3)]);
        __builtin_nontemporal_store(t, v + j);
      }
    }
  }
}

// Persistent form of line_kernel_wct for whole power-of-two lines (n == 64 R,
// one wavefront per workgroup, no halo lines).  Each wavefront walks lines
// q, q + G, q + 2G, ... (G = the grid) and loads the next line into registers
// while it solves and stores the current one.  The deposit of the next line
// sits at the end of the loop body, after the stores, so its wait covers that
// line's loads only and the stores keep draining; the last line is peeled off
// the loop, so every load in the body is unconditional.  The grid is a few
// wavefronts per CU: fewer lines in flight stream faster, and the prefetch
// covers the solve that a low occupancy would otherwise expose.
template <typename TIn, int R, int PROD, int PREC>
__global__ void __launch_bounds__(64) line_kernel_wcp(const LineArgs args_by_value) {
  (void)args_by_value;
  static_assert(PROD == P_IDENT || PROD == P_COMPACT, "lines without halo lines");
  const LineArgs& args = kernel_args();
  constexpr int W = 64;
  constexpr int N = W * R;
  constexpr int CS = R + 2;                   // LDS doubles per chunk (2 pad: bank spread)
  constexpr int VIN = 16 / static_cast<int>(sizeof(TIn));
  constexpr int NV = N / VIN / 64;            // 16-byte loads per lane and line
  static_assert(R >= VIN && N % (VIN * 64) == 0, "whole 16-byte loads inside chunks");
  using VT = std::conditional_t<sizeof(TIn) == 8, nt_dbl2, nt_flt4>;
  __shared__ __attribute__((aligned(16))) double L[W * CS];

  const NtPlanK& P = args.plan;
  const int lane = threadIdx.x;
  const int64_t nsys = args.phys;
  const int64_t G = gridDim.x;
  int64_t q = blockIdx.x;                     // the launcher keeps the grid <= nsys
  auto at = [](int e) -> int { return (e / R) * CS + (e % R); };

  VT pf[NV];
  auto fetch = [&](int64_t line) {
    const VT* __restrict__ v = reinterpret_cast<const VT*>(static_cast<const TIn*>(args.in) + line * args.pstride);
#pragma unroll
    for (int k = 0; k < NV; ++k) pf[k] = __builtin_nontemporal_load(v + k * 64 + lane);
  };
  auto deposit = [&]() {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const int e = VIN * (k * 64 + lane);
      if constexpr (VIN == 2) {
        L[at(e)] = pf[k].x;
        L[at(e + 1)] = pf[k].y;
      } else {
        L[at(e)] = static_cast<double>(pf[k].x);
        L[at(e + 1)] = static_cast<double>(pf[k].y);
        L[at(e + 2)] = static_cast<double>(pf[k].z);
        L[at(e + 3)] = static_cast<double>(pf[k].w);
      }
    }
  };
  auto solve_store = [&](int64_t line) {
    const int my_e0 = lane * R;
    double* __restrict__ mine = L + lane * CS;
    double f[R];
#pragma unroll
    for (int k = 0; k < R; k += 2) {
      const nt_dbl2 t = *reinterpret_cast<const nt_dbl2*>(mine + k);
      f[k] = t.x;
      f[k + 1] = t.y;
    }
    double d[R];
    if constexpr (PROD == P_COMPACT) {
      double bnd[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) bnd[i] = L[at((i < 3) ? i : N + 2 - i)];
      const double fl = shup<W>(f[R - 1], 1);
      const double fr = shdn<W>(f[0], 1);
      compact_rhs_chunk<R, true>(f, fl, fr, bnd, my_e0, N, args.s1, args.s2, d);
    } else {
#pragma unroll
      for (int k = 0; k < R; ++k) d[k] = f[k];
    }
    double x[R];
    solve_chunk<R, W, PREC, true>(d, x, P, lane, args.sweeps);
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < R; k += 2) {
      nt_dbl2 t;
      t.x = x[k];
      t.y = x[k + 1];
      *reinterpret_cast<nt_dbl2*>(mine + k) = t;
    }
    wave_lds_sync();
    nt_dbl2* __restrict__ dst = reinterpret_cast<nt_dbl2*>(static_cast<double*>(args.out) + line * args.pstride);
#pragma unroll
    for (int j = lane; j < N / 2; j += 64) {
      nt_dbl2 t;
      t.x = L[at(2 * j)];
      t.y = L[at(2 * j + 1)];
      __builtin_nontemporal_store(t, dst + j);
    }
  };

  fetch(q);
  deposit();
  while (q + G < nsys) {                      // wave-uniform
    const int64_t qn = q + G;
    wave_lds_sync();
    fetch(qn);
    solve_store(q);
    wave_lds_sync();
    deposit();
    q = qn;
  }
  wave_lds_sync();
  solve_store(q);
}

// ---- strided lines (y/z): workgroup tile of TS adjacent lines ------------------------
// The tile is TS lines that are adjacent in memory (unit stride between
// lines, estride between elements).  Blocks of E elements of all staged
// lines go through LDS: a block row of the tile is TS+2*HALO contiguous values,
// read with 16-byte nontemporal loads when aligned; each lane then copies its
// chunk.  Lines are solved by segments of W lanes, so the chunk is R = n/W
// rows (small for W = 64 or 32).  One workgroup barrier separates loading and
// copying each block; blocks alternate between two LDS buffers.
// Blocks of a strided tile kept in flight (see the pipeline in the kernel):
// deeper for small per-thread blocks, where the registers are cheap.
#ifndef NT_STRIDED_DEPTH
#define NT_STRIDED_DEPTH(PF) (((PF) <= 2) ? 4 : (((PF) <= 4) ? 2 : 1))
#endif

template <typename TIn, typename TOut, int R, int W, int WAVES, int E, int PROD, int PREC, bool FULL>
__global__ void __launch_bounds__(WAVES * 64) line_kernel_strided(const LineArgs args_by_value) {
  (void)args_by_value;
  const LineArgs& args = kernel_args();
  constexpr int NT = WAVES * 64;
  constexpr int SPW = 64 / W;
  constexpr int TS = WAVES * SPW;             // lines per tile
  constexpr int HALO = (PROD == P_ADI) ? 1 : 0;
  constexpr int ROWS = TS + 2 * HALO;         // staged lines
  constexpr int CP = (R >= 4) ? 2 : 0;
  constexpr int EP = (E / R) * (R + CP) + 2;  // LDS stride between staged lines (doubles)
  constexpr int BUF = ROWS * EP;
  static_assert(E % R == 0, "a chunk must not straddle blocks");
  __shared__ __attribute__((aligned(16))) double lds[2 * BUF];

  const NtPlanK& P = args.plan;
  const int n = FULL ? W * R : P.n;
  // refreshed through an opaque move in every block of the load loop, so the
  // per-thread offsets derived from it are recomputed, not held across the loop
  int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int seg = (tid & 63) / W;
  const int lane = tid & (W - 1);
  const int ql = wave * SPW + seg;            // tile-local line

  const int64_t bid = blockIdx.x;
  const int64_t o = bid / args.tiles;
  const int64_t p0 = (bid - o * args.tiles) * TS;   // first line of the tile
  const int64_t nsys = args.phys - 2 * HALO;
  const bool active = (p0 + ql) < nsys;
  const int64_t base = o * args.ostride + p0;        // pstride == 1 for strided lines
  const TIn* __restrict__ in = static_cast<const TIn*>(args.in) + base;
  TOut* __restrict__ out = static_cast<TOut*>(args.out) + base;
  const int64_t rows_valid64 = args.phys - p0;
  const int rows_valid = rows_valid64 < ROWS ? static_cast<int>(rows_valid64) : ROWS;
  // 16-byte row loads: whole tile in range, aligned, no halo
  constexpr int VIN = 16 / sizeof(TIn);
  const bool vec_in = (HALO == 0) && (rows_valid == ROWS) && (ROWS % VIN == 0) &&
                      ((reinterpret_cast<uintptr_t>(in) % 16) == 0) &&
                      ((args.estride * sizeof(TIn)) % 16 == 0);
  constexpr int VOUT = 16 / sizeof(TOut);
  const bool vec_out = (HALO == 0) && (p0 + TS <= nsys) && (TS % VOUT == 0) &&
                       ((reinterpret_cast<uintptr_t>(out) % 16) == 0) &&
                       ((args.estride * sizeof(TOut)) % 16 == 0);

  const int my_e0 = lane * R;
  const int my_blk = my_e0 / E;
  const int nblk = (n + E - 1) / E;
  auto at = [&](int buf, int row, int el) -> int { return buf * BUF + row * EP + el + (el / R) * CP; };

  double f[R];
#pragma unroll
  for (int k = 0; k < R; ++k) f[k] = 0.0;
  double bnd[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};

  // Block loads are software-pipelined DEPTH blocks deep: the values of block
  // b+DEPTH are fetched into registers as soon as block b has been moved to
  // LDS, so up to DEPTH blocks are in flight while earlier ones are copied.
  // PF: values per thread per block; the register ring holds DEPTH blocks.
  constexpr int VPR = ROWS / (16 / static_cast<int>(sizeof(TIn)));   // vectors per block row
  constexpr int PFV = (VPR * E + NT - 1) / NT;                        // vector loads per thread
  constexpr int PFS = (ROWS * E + NT - 1) / NT;                       // scalar loads per thread
  constexpr int VIN_ = 16 / static_cast<int>(sizeof(TIn));
  constexpr int PF = (PFV * VIN_ > PFS) ? PFV * VIN_ : PFS;
  constexpr int NBMAX = (W * R + E - 1) / E;
  constexpr int DEPTH_ = NT_STRIDED_DEPTH(PF);
  constexpr int DEPTH = (DEPTH_ < NBMAX) ? DEPTH_ : NBMAX;
  double ring[DEPTH][PF];
  auto fetch = [&](int b, double (&pf)[PF]) {
    const int e0 = b * E;
    const int ecount = (n - e0 < E) ? n - e0 : E;
    if (vec_in) {
#pragma unroll
      for (int i = 0; i < PFV; ++i) {
        const int idx = tid + i * NT;
        const int vr = idx % VPR, el = idx / VPR;
        if (el < ecount) {
          const TIn* src = in + static_cast<int64_t>(e0 + el) * args.estride + vr * VIN_;
          if constexpr (sizeof(TIn) == 8) {
            const nt_dbl2 t = __builtin_nontemporal_load(reinterpret_cast<const nt_dbl2*>(src));
            pf[i * 2] = t.x;
            pf[i * 2 + 1] = t.y;
          } else {
            const nt_flt4 t = __builtin_nontemporal_load(reinterpret_cast<const nt_flt4*>(src));
            pf[i * 4] = t.x;
            pf[i * 4 + 1] = t.y;
            pf[i * 4 + 2] = t.z;
            pf[i * 4 + 3] = t.w;
          }
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < PFS; ++i) {
        const int idx = tid + i * NT;
        const int r = idx % ROWS, el = idx / ROWS;
        if (el < ecount && r < rows_valid)
          pf[i] = static_cast<double>(in[r + static_cast<int64_t>(e0 + el) * args.estride]);
      }
    }
  };
  auto deposit = [&](int b, const double (&pf)[PF]) {
    const int e0 = b * E;
    const int buf = b & 1;
    const int ecount = (n - e0 < E) ? n - e0 : E;
    if (vec_in) {
#pragma unroll
      for (int i = 0; i < PFV; ++i) {
        const int idx = tid + i * NT;
        const int vr = idx % VPR, el = idx / VPR;
        if (el < ecount) {
#pragma unroll
          for (int c = 0; c < VIN_; ++c) lds[at(buf, vr * VIN_ + c, el)] = pf[i * VIN_ + c];
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < PFS; ++i) {
        const int idx = tid + i * NT;
        const int r = idx % ROWS, el = idx / ROWS;
        if (el < ecount && r < rows_valid) lds[at(buf, r, el)] = pf[i];
      }
    }
  };

  auto consume = [&](int b) {
    const int e0 = b * E;
    const int buf = b & 1;
    if (my_blk == b && active) {
      const int el0 = my_e0 - e0;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int e = my_e0 + k;
        if (FULL || e < n) {
          const double c = lds[at(buf, ql + HALO, el0 + k)];
          if (PROD == P_ADI) {
            const double up = lds[at(buf, ql, el0 + k)];
            const double dn = lds[at(buf, ql + 2, el0 + k)];
            f[k] = (e == 0 || e == n - 1) ? c : args.s1 * (up + dn) + args.s2 * c;
          } else {
            f[k] = c;
          }
        }
      }
    }
    if (PROD == P_COMPACT && active) {
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        const int e = (i < 3) ? i : n + 2 - i;
        if (e >= e0 && e < e0 + E) bnd[i] = lds[at(buf, ql + HALO, e - e0)];
      }
    }
  };

#pragma unroll
  for (int j = 0; j < DEPTH; ++j)
    if (j < nblk) fetch(j, ring[j]);
  for (int b0 = 0; b0 < nblk; b0 += DEPTH) {
#pragma unroll
    for (int j = 0; j < DEPTH; ++j) {
      const int b = b0 + j;
      if (b < nblk) {                            // uniform across the workgroup
        tid = (static_cast<int>(threadIdx.x) + opaque_vzero()) & (NT - 1);   // the mask keeps its range known
        deposit(b, ring[j]);
        if (b + DEPTH < nblk) fetch(b + DEPTH, ring[j]);
        __syncthreads();
        consume(b);
        // block b+1 goes to the other LDS buffer; this one is rewritten (block
        // b+2) only after the next barrier, which every reader of block b has passed
      }
    }
  }

  double d[R];
  if (PROD == P_COMPACT) {
    const double fl = shup<W>(f[R - 1], 1);
    const double fr = shdn<W>(f[0], 1);
    compact_rhs_chunk<R, FULL>(f, fl, fr, bnd, my_e0, n, args.s1, args.s2, d);
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k) d[k] = f[k];
  }
  double x[R];
  solve_chunk<R, W, PREC, FULL>(d, x, P, lane, args.sweeps);

  for (int b = 0; b < nblk; ++b) {
    const int e0 = b * E;
    const int buf = b & 1;
    const int ecount = (n - e0 < E) ? n - e0 : E;
    // block 0 follows the load phase; later blocks reuse a buffer whose last
    // readers all passed the previous block's barrier
    if (b == 0) __syncthreads();
    if (my_blk == b && active) {
      const int el0 = my_e0 - e0;
#pragma unroll
      for (int k = 0; k < R; ++k)
        if (FULL || my_e0 + k < n) lds[at(buf, ql + HALO, el0 + k)] = x[k];
    }
    __syncthreads();
    if (vec_out) {
      constexpr int VPR = TS / VOUT;
      for (int idx = tid; idx < VPR * ecount; idx += NT) {
        const int vr = idx % VPR, el = idx / VPR;
        TOut* dst = out + static_cast<int64_t>(e0 + el) * args.estride + vr * VOUT;
        if constexpr (sizeof(TOut) == 8) {
          nt_dbl2 t;
          t.x = lds[at(buf, vr * 2, el)];
          t.y = lds[at(buf, vr * 2 + 1, el)];
          __builtin_nontemporal_store(t, reinterpret_cast<nt_dbl2*>(dst));
        } else {
          nt_flt4 t;
          t.x = static_cast<float>(lds[at(buf, vr * 4, el)]);
          t.y = static_cast<float>(lds[at(buf, vr * 4 + 1, el)]);
          t.z = static_cast<float>(lds[at(buf, vr * 4 + 2, el)]);
          t.w = static_cast<float>(lds[at(buf, vr * 4 + 3, el)]);
          __builtin_nontemporal_store(t, reinterpret_cast<nt_flt4*>(dst));
        }
      }
    } else {
      for (int idx = tid; idx < TS * ecount; idx += NT) {
        const int r = idx % TS, el = idx / TS;
        if (p0 + r < nsys)
          out[(r + HALO) + static_cast<int64_t>(e0 + el) * args.estride] =
              static_cast<TOut>(lds[at(buf, r + HALO, el)]);
      }
    }
  }
}

}  // namespace nt
