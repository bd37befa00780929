// Launch-shape selection for the line kernels.  Each (input type, output
// type, layout, producer, precision) combination is instantiated in its own
// translation unit so the library builds in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include "nt_line_kernel.h"
#include "nt_tile_kernel.h"
#include "nt_lane_kernel.h"
#include "nt_xw_kernel.h"

namespace nt {

// Launch the line kernel over `outer` slabs of `a.phys` staged lines each.
// Fills a.tiles.  Returns hipErrorInvalidValue for a size the kernel family
// does not cover.
template <typename TIn, typename TOut, int LAYOUT, int PROD, int PREC>
hipError_t launch_line(LineArgs& a, int64_t outer, hipStream_t st);

// Launch-shape variants, chosen per (layout, precision, ceil(log2 n), n a power of two)
// by the tuning table
// (nt_tune_set; defaults measured on MI355X).  The variants agree to within a few
// roundings (the lane-per-line solves round differently from cyclic reduction).
//   CONTIG  0: one line per wavefront, wave-coalesced nontemporal loads and
//              stores, wave-private LDS transpose (needs 16-byte aligned lines)
//           1: one line per wavefront, each lane loads its own chunk directly
//           2: workgroup tile of 4 lines staged through LDS (any n, alignment)
//           3: one line per workgroup over two wavefronts, n/128 rows per lane
//              (power-of-two n from 256 to 2048, 16-byte aligned lines; other
//              lines take variant 0's path)
//           4: persistent form of 0 with the next line prefetched, a few
//              wavefronts per CU (power-of-two n from 128 (FP32 input: 256) to
//              2048 outside ADI; other lines take variant 0's path)
//   STRIDED 0: tile of 16 adjacent lines, 64 lanes per line (1024 threads)
//           1: tile of 32 adjacent lines, 32 lanes per line (1024 threads);
//              lines longer than 512: tile of 8 lines, 64 lanes per line (512 threads);
//              ADI above 1024 points: whole tiles of the solved lines, RHS formed at load
//           2: persistent whole-tile kernel, 16 lines (8 above 1024 points, and
//              above 512 for ADI), next tile's loads in flight during this one;
//              FP32-stored compact derivatives: 32 lines, two per wavefront
//           3: as 2, with two lines per wavefront, for lines up to 512 points
//              outside ADI; otherwise as 2.  ADI lines above 1024 points (2, 3):
//              whole tiles of the solved lines only, right-hand side formed at
//              load (measured no faster than the halo tiles at 1024 and below)
//           4: as 3, the two lines of a wavefront solved at once on 32-lane
//              segments (4..512-point lines outside ADI and FP32 storage;
//              otherwise as 2)
//           5: one lane per line, 64 lines per workgroup of 16 wavefronts,
//              block-partitioned FP64 solve (49- to 512-point lines with
//              adjacent lines adjacent in memory, FP64 or FP32 storage, outside
//              ADI); with FP32 + refinement, the FP32 block solve of FP32-stored
//              compact derivatives with one sweep (49..496 and 512 points);
//              other lines take lane_fallback()'s shape
// Contiguous lines that variants 0/1 cannot take (n <= 64, unaligned) use 2.
// launch-shape variants per tuning layout: 0 contiguous, 1 strided, 2 and 3 the
// same for the ADI half steps (their right-hand side reads neighbouring lines,
// which favours other shapes than the derivative's)
constexpr int kNumVariants[4] = {5, 6, 4, 4};
int tune_get(int layout, int prec, int n);
void tune_set(int layout, int prec, int n, int variant);

// n == W*R with n a power of two: the level structure is known at compile time
inline bool is_full(int n, int W, int R) { return n == W * R && (n & (n - 1)) == 0; }

#define NT_LAUNCH_(KERNEL, TS_, NT_)                                                       \
  do {                                                                                     \
    const int64_t nsys_ = a.phys - 2 * ((PROD == P_ADI) ? 1 : 0);                          \
    a.tiles = (nsys_ + (TS_)-1) / (TS_);                                                   \
    const int64_t grid_ = outer * a.tiles;                                                 \
    if (grid_ <= 0) return hipSuccess;                                                     \
    if (grid_ > 0x7fffffffLL) return hipErrorInvalidValue;                                 \
    hipLaunchKernelGGL(KERNEL, dim3(static_cast<unsigned>(grid_)), dim3(NT_), 0, st, a);   \
    return hipGetLastError();                                                              \
  } while (0)

// LDS-staged tile kernel (CONTIG variant 2, all STRIDED variants)
#define NT_GO(R_, W_, WAVES_, E_)                                                          \
  do {                                                                                     \
    if ((R_) > (1 << a.plan.nlev)) return hipErrorInvalidValue;                            \
    if (is_full(n, W_, R_))                                                                \
      NT_LAUNCH_((line_kernel<TIn, TOut, R_, W_, WAVES_, E_, LAYOUT, PROD, PREC, true>),   \
                 (WAVES_) * (64 / (W_)), (WAVES_)*64);                                     \
    NT_LAUNCH_((line_kernel<TIn, TOut, R_, W_, WAVES_, E_, LAYOUT, PROD, PREC, false>),    \
               (WAVES_) * (64 / (W_)), (WAVES_)*64);                                       \
  } while (0)

// direct chunk-load kernel (CONTIG variant 1)
#define NT_GOD(R_, W_, WAVES_)                                                             \
  do {                                                                                     \
    if ((R_) > (1 << a.plan.nlev)) return hipErrorInvalidValue;                            \
    if (is_full(n, W_, R_))                                                                \
      NT_LAUNCH_((line_kernel_direct<TIn, TOut, R_, W_, WAVES_, PROD, PREC, true>),        \
                 (WAVES_) * (64 / (W_)), (WAVES_)*64);                                     \
    NT_LAUNCH_((line_kernel_direct<TIn, TOut, R_, W_, WAVES_, PROD, PREC, false>),         \
               (WAVES_) * (64 / (W_)), (WAVES_)*64);                                       \
  } while (0)

// Resident one-wavefront workgroups per CU of the wave-coalesced kernel,
// capped through dynamic LDS (0: no cap).  Fewer lines in flight stream
// faster: a copy with the kernel's access pattern (one 4 KB line per
// wavefront) moves a 512^3 FP64 field at 6.2 TB/s with 32 workgroups per CU
// and 6.7-6.8 TB/s with 6-8.  The kernel itself needs more to hide its solve;
// measured on MI355X, 512^3 d/dx over four interleaved runs: no cap 343-349 us
// (FP32 + refinement 333-361), 16: 333-335 (357-360), 22: 333.7-334.0
// (335.6-337.4), 24: 332-335 (336-340), 8: 411 (514).
//
// 1024-point lines (four one-line wavefronts per workgroup) with the FP64 solve
// run capped at 3 workgroups per CU: 2^27-point FP64 field 388 -> 373 us, the
// 1024^3 FP32-stored field 2.394 -> 2.246 ms.  With FP32 + refinement the same
// cap is slower (2.40 -> 2.71 ms), and 128- and 256-point lines gain nothing.
#ifndef NT_WCT_WGS_PER_CU
#define NT_WCT_WGS_PER_CU 22
#endif
#ifndef NT_WCT_WGS_PER_CU_1024_F64
#define NT_WCT_WGS_PER_CU_1024_F64 3
#endif
constexpr int wct_wgs_per_cu(int r, int waves, int prec) {
  return (waves == 1) ? NT_WCT_WGS_PER_CU
                      : ((r == 16 && waves == 4 && prec == PREC_F64) ? NT_WCT_WGS_PER_CU_1024_F64 : 0);
}
inline size_t wct_lds_pad(int cap, size_t static_lds) {
  if (cap <= 0) return 0;
  const size_t per = (static_cast<size_t>(160 * 1024) / cap) & ~static_cast<size_t>(1023);
  return per > static_lds ? per - static_lds : 0;
}

#define NT_LAUNCH_LDS_(KERNEL, TS_, NT_, LDS_)                                             \
  do {                                                                                     \
    const int64_t nsys_ = a.phys - 2 * ((PROD == P_ADI) ? 1 : 0);                          \
    a.tiles = (nsys_ + (TS_)-1) / (TS_);                                                   \
    const int64_t grid_ = outer * a.tiles;                                                 \
    if (grid_ <= 0) return hipSuccess;                                                     \
    if (grid_ > 0x7fffffffLL) return hipErrorInvalidValue;                                 \
    hipLaunchKernelGGL(KERNEL, dim3(static_cast<unsigned>(grid_)), dim3(NT_), (LDS_), st, a); \
    return hipGetLastError();                                                              \
  } while (0)

// wave-coalesced streaming kernel (CONTIG variant 0)
#define NT_GOW(R_, W_, WAVES_)                                                             \
  do {                                                                                     \
    if ((R_) > (1 << a.plan.nlev)) return hipErrorInvalidValue;                            \
    const size_t lds_ = wct_lds_pad(wct_wgs_per_cu(R_, WAVES_, PREC),                      \
                                    static_cast<size_t>(WAVES_) * 64 * ((R_) + 2) * 8);    \
    if (is_full(n, W_, R_))                                                                \
      NT_LAUNCH_LDS_((line_kernel_wct<TIn, TOut, R_, W_, WAVES_, PROD, PREC, true>),       \
                     (WAVES_) * (64 / (W_)), (WAVES_)*64, lds_);                           \
    NT_LAUNCH_LDS_((line_kernel_wct<TIn, TOut, R_, W_, WAVES_, PROD, PREC, false>),        \
                   (WAVES_) * (64 / (W_)), (WAVES_)*64, lds_);                             \
  } while (0)

// strided-line tile kernel (STRIDED variants)
#define NT_GOS(R_, W_, WAVES_, E_)                                                         \
  do {                                                                                     \
    if ((R_) > (1 << a.plan.nlev)) return hipErrorInvalidValue;                            \
    if (is_full(n, W_, R_))                                                                \
      NT_LAUNCH_((line_kernel_strided<TIn, TOut, R_, W_, WAVES_, E_, PROD, PREC, true>),   \
                 (WAVES_) * (64 / (W_)), (WAVES_)*64);                                     \
    NT_LAUNCH_((line_kernel_strided<TIn, TOut, R_, W_, WAVES_, E_, PROD, PREC, false>),    \
               (WAVES_) * (64 / (W_)), (WAVES_)*64);                                       \
  } while (0)

// wavefronts per workgroup of the wave-private-transpose kernel for 257..512-point
// lines (one: measured 1-3% faster than four on MI355X at 512^3)
#ifndef NT_WCT_WAVES_512
#define NT_WCT_WAVES_512 1
#endif

// Wavefronts per workgroup of the two-lines-per-wavefront tile kernel (variant 3)
#ifndef NT_TILE2_WAVES
#define NT_TILE2_WAVES 8
#endif

// Fast path of the whole-tile kernel (FP64 in and out, power-of-two lines,
// whole aligned tiles; nt_tile_kernel.h), for the shapes whose registers stay
// within four wavefronts per SIMD without spilling: a scratch reload during the
// solve would wait, in order, for every prefetched load (measured at 1024
// points: 40% slower).  0 turns it off.
#ifndef NT_TILE_FAST
#define NT_TILE_FAST 1
#endif
// Largest rows per lane of the FP64 fast path.  At 1024 points (R = 16) it is
// taken for rows less than NT_TILE_FAST_R16_ROW_LIMIT apart only: on 1024^3
// d/dz (rows 8 MiB apart) it ran 3.9-4.0 ms at three of five output placements
// and 5.6-6.0 ms at the other two, where the general kernel with contiguous tile
// ranges stays at 4.2-4.7 ms.  1024^3 d/dy: 3.47 -> 3.40 ms.
#ifndef NT_TILE_FAST_MAX_R
#define NT_TILE_FAST_MAX_R 16
#endif

// At 16 rows per lane the FP64 fast path takes only lines whose rows are closer
// than this many bytes (y-lines); z-lines keep the general kernel and its tile
// order.  FP32 + refinement takes the fast path for any row distance: on 1024^3
// d/dz (rows 8 MiB apart) it ran 4.59-5.59 ms over nine output placements where
// the general kernel, whose refinement spills at 16 rows per lane, ran 6.34-6.42.
#ifndef NT_TILE_FAST_R16_ROW_LIMIT
#define NT_TILE_FAST_R16_ROW_LIMIT (2 << 20)
#endif
// Largest rows per lane of the FP32 + refinement fast path (one line per
// wavefront).  At 16 rows per lane the next tile is fetched after the solve and
// the solve's LDS offsets are re-derived after it, which keeps the loop free of
// spills: FP64-stored 1024^3 d/dy with FP32 + one sweep 5.61 -> 3.45 ms.
#ifndef NT_TILE_FAST_F32R_MAX_R
#define NT_TILE_FAST_F32R_MAX_R 16
#endif

// Whole 16-byte aligned tiles of `ts` adjacent strided lines only: the fast
// tile kernel's loads and stores are all unconditional 16-byte accesses.
inline bool tile_fast_ok(const LineArgs& a, int ts) {
  const uintptr_t i = reinterpret_cast<uintptr_t>(a.in), o = reinterpret_cast<uintptr_t>(a.out);
  return a.pstride == 1 && a.phys % ts == 0 && i % 16 == 0 && o % 16 == 0 && a.estride % 2 == 0 &&
         a.ostride % 2 == 0;
}

// Persistent whole-tile kernel for strided lines (nt_tile_kernel.h): a grid of
// at most as many workgroups as fit on the device at once.
int64_t persistent_grid(const void* kernel, int threads);

#define NT_GOT(R_, WAVES_, LPW_)                                                           \
  do {                                                                                     \
    if ((R_) > (1 << a.plan.nlev)) return hipErrorInvalidValue;                            \
    const int64_t nsys_ = a.phys - 2 * ((PROD == P_ADI) ? 1 : 0);                          \
    a.tiles = (nsys_ + (WAVES_) * (LPW_)-1) / ((WAVES_) * (LPW_));                         \
    a.total = outer * a.tiles;                                                             \
    if (a.total <= 0) return hipSuccess;                                                   \
    const bool full_ = is_full(n, 64, R_);                                                 \
    if constexpr (NT_TILE_FAST && std::is_same_v<TIn, double> && std::is_same_v<TOut, double> &&     \
                  PROD != P_ADI && ((R_) * (LPW_)) % 2 == 0 &&                             \
                  (R_) <= (PREC == PREC_F64 ? NT_TILE_FAST_MAX_R : ((LPW_) == 1 ? NT_TILE_FAST_F32R_MAX_R : 4))) { \
      if (full_ && tile_fast_ok(a, (WAVES_) * (LPW_)) &&                                   \
          ((R_) < 16 || PREC != PREC_F64 ||                                                \
           a.estride * static_cast<int64_t>(sizeof(TIn)) < NT_TILE_FAST_R16_ROW_LIMIT)) {    \
        const void* kf_ = reinterpret_cast<const void*>(&line_kernel_tile_fast<R_, WAVES_, LPW_, PROD, PREC>); \
        const int64_t capf_ = persistent_grid(kf_, (WAVES_)*64);                           \
        const int64_t gridf_ = a.total < capf_ ? a.total : capf_;                          \
        hipLaunchKernelGGL((line_kernel_tile_fast<R_, WAVES_, LPW_, PROD, PREC>),           \
                           dim3(static_cast<unsigned>(gridf_)), dim3((WAVES_)*64), 0, st, a); \
        return hipGetLastError();                                                          \
      }                                                                                    \
    }                                                                                      \
    const void* k_ = full_                                                                 \
        ? reinterpret_cast<const void*>(&line_kernel_tile<TIn, TOut, R_, WAVES_, LPW_, PROD, PREC, true>)  \
        : reinterpret_cast<const void*>(&line_kernel_tile<TIn, TOut, R_, WAVES_, LPW_, PROD, PREC, false>); \
    const int64_t cap_ = persistent_grid(k_, (WAVES_)*64);                                 \
    const int64_t grid_ = a.total < cap_ ? a.total : cap_;                                 \
    if (full_)                                                                             \
      hipLaunchKernelGGL((line_kernel_tile<TIn, TOut, R_, WAVES_, LPW_, PROD, PREC, true>),      \
                         dim3(static_cast<unsigned>(grid_)), dim3((WAVES_)*64), 0, st, a); \
    else                                                                                   \
      hipLaunchKernelGGL((line_kernel_tile<TIn, TOut, R_, WAVES_, LPW_, PROD, PREC, false>),     \
                         dim3(static_cast<unsigned>(grid_)), dim3((WAVES_)*64), 0, st, a); \
    return hipGetLastError();                                                              \
  } while (0)

// One lane per line, block-partitioned (STRIDED variant 5, nt_lane_kernel.h):
// a workgroup per 64 adjacent lines of a slab, the instance for the plan's
// block length m (4..32 rows).
template <typename TIn, int PROD, int M = 4>
hipError_t launch_lanes(const LaneArgs& la, int m, unsigned grid, hipStream_t st) {
  if constexpr (M > NT_LANE_MAX_M) {
    return hipErrorInvalidValue;
  } else {
    if (m == M) {
      if (la.t.j == NT_LANE_P)
        hipLaunchKernelGGL((line_kernel_lanes<TIn, M, PROD, true>), dim3(grid), dim3(NT_LANE_P * 64), 0, st, la);
      else
        hipLaunchKernelGGL((line_kernel_lanes<TIn, M, PROD, false>), dim3(grid), dim3(NT_LANE_P * 64), 0, st, la);
      return hipGetLastError();
    }
    return launch_lanes<TIn, PROD, M + 1>(la, m, grid, st);
  }
}

#define NT_GOL()                                                                           \
  do {                                                                                     \
    LaneArgs la_;                                                                          \
    la_.in = a.in;                                                                         \
    la_.out = static_cast<double*>(a.out);                                                 \
    la_.estride = a.estride;                                                               \
    la_.ostride = a.ostride;                                                               \
    la_.phys = a.phys;                                                                     \
    la_.groups = (a.phys + 63) / 64;                                                       \
    la_.s1 = a.s1;                                                                         \
    la_.s2 = a.s2;                                                                         \
    la_.t = *a.lane;                                                                       \
    const int64_t grid_ = outer * la_.groups;                                              \
    if (grid_ <= 0) return hipSuccess;                                                     \
    if (grid_ > 0x7fffffffLL) return hipErrorInvalidValue;                                 \
    return launch_lanes<TIn, PROD>(la_, a.lane->m, static_cast<unsigned>(grid_), st);      \
  } while (0)

// FP32 + refinement lane solve of FP32-stored compact derivatives (STRIDED
// variant 5, nt_lane_kernel.h): the instance for m rows.
template <int M = 4>
hipError_t launch_lanes_f32r(const LaneArgs32& la, int m, unsigned grid, hipStream_t st) {
  if constexpr (M > NT_LANE_MAX_M) {
    return hipErrorInvalidValue;
  } else {
    if (m == M) {
      if (la.t.j == NT_LANE_P) {
        hipLaunchKernelGGL((line_kernel_lanes_f32r<M, true>), dim3(grid), dim3(NT_LANE_P * 64), 0, st, la);
      } else {
        // 32 rows with a padding row spill 106 VGPRs: not instantiated (lane32_ok)
        if constexpr (M == NT_LANE_MAX_M) return hipErrorInvalidValue;
        else hipLaunchKernelGGL((line_kernel_lanes_f32r<M, false>), dim3(grid), dim3(NT_LANE_P * 64), 0, st, la);
      }
      return hipGetLastError();
    }
    return launch_lanes_f32r<M + 1>(la, m, grid, st);
  }
}

// The FP32 lane solve takes the plan's blocks unless they are 32 rows with
// padding rows (497..511-point lines).
inline bool lane32_ok(const NtLaneTab32* t) {
  return t && (t->j == NT_LANE_P || t->m < NT_LANE_MAX_M);
}

#define NT_GOL32()                                                                         \
  do {                                                                                     \
    LaneArgs32 la_;                                                                        \
    la_.in = static_cast<const float*>(a.in);                                              \
    la_.out = static_cast<double*>(a.out);                                                 \
    la_.estride = a.estride;                                                               \
    la_.ostride = a.ostride;                                                               \
    la_.phys = a.phys;                                                                     \
    la_.groups = (a.phys + 63) / 64;                                                       \
    la_.s1 = a.s1;                                                                         \
    la_.s2 = a.s2;                                                                         \
    la_.b1 = a.plan.b1;                                                                    \
    la_.c1 = a.plan.c1;                                                                    \
    la_.ai = a.plan.ai;                                                                    \
    la_.bi = a.plan.bi;                                                                    \
    la_.ci = a.plan.ci;                                                                    \
    la_.an = a.plan.an;                                                                    \
    la_.bn = a.plan.bn;                                                                    \
    la_.t = *a.lane32;                                                                     \
    const int64_t grid_ = outer * la_.groups;                                              \
    if (grid_ <= 0) return hipSuccess;                                                     \
    if (grid_ > 0x7fffffffLL) return hipErrorInvalidValue;                                 \
    return launch_lanes_f32r(la_, a.lane32->m, static_cast<unsigned>(grid_), st);          \
  } while (0)

// Variant 5 for lines the lane kernels do not take: the shapes that are the
// defaults without it (FP64: 32-line tiles up to 48 points, else the
// segmented tiles; FP32 + refinement: the segmented tiles up to 256 points,
// at 257..512 two lines per wavefront, or the whole-tile kernel for 512)
inline int lane_fallback(int prec, int n) {
  if (prec == PREC_F64) return (n <= 3 * NT_LANE_P) ? 1 : 4;
  if (n <= 64) return (n == 64) ? 1 : 4;
  if (n <= 256) return 4;
  return ((n & (n - 1)) == 0) ? 2 : 3;
}

// Persistent whole-tile kernel, two lines per wavefront solved at once on
// 32-lane segments (STRIDED variant 4).
#define NT_GOTS(R_, WAVES_)                                                                \
  do {                                                                                     \
    if (2 * (R_) > (1 << a.plan.nlev)) return hipErrorInvalidValue;                        \
    a.tiles = (a.phys + (WAVES_) * 2 - 1) / ((WAVES_) * 2);                                \
    a.total = outer * a.tiles;                                                             \
    if (a.total <= 0) return hipSuccess;                                                   \
    const bool full_ = is_full(n, 64, R_);                                                 \
    const void* k_ = full_                                                                 \
        ? reinterpret_cast<const void*>(&line_kernel_tile<TIn, TOut, R_, WAVES_, 2, PROD, PREC, true, 2>)  \
        : reinterpret_cast<const void*>(&line_kernel_tile<TIn, TOut, R_, WAVES_, 2, PROD, PREC, false, 2>); \
    const int64_t cap_ = persistent_grid(k_, (WAVES_)*64);                                 \
    const int64_t grid_ = a.total < cap_ ? a.total : cap_;                                 \
    if (full_)                                                                             \
      hipLaunchKernelGGL((line_kernel_tile<TIn, TOut, R_, WAVES_, 2, PROD, PREC, true, 2>),      \
                         dim3(static_cast<unsigned>(grid_)), dim3((WAVES_)*64), 0, st, a); \
    else                                                                                   \
      hipLaunchKernelGGL((line_kernel_tile<TIn, TOut, R_, WAVES_, 2, PROD, PREC, false, 2>),     \
                         dim3(static_cast<unsigned>(grid_)), dim3((WAVES_)*64), 0, st, a); \
    return hipGetLastError();                                                              \
  } while (0)

// 16-byte row accesses in the ADI tile kernel (line_kernel_tile_adi_vec): lines of
// exactly 64 R points, aligned 8-line tiles covering the staged lines, 16-byte
// aligned rows.  0 turns it off.
#ifndef NT_TILE_ADI_VEC
#define NT_TILE_ADI_VEC 1
#endif

inline bool tile_adi_vec_ok(const LineArgs& a) {
  const uintptr_t i = reinterpret_cast<uintptr_t>(a.in), o = reinterpret_cast<uintptr_t>(a.out);
  return a.pstride == 1 && a.phys % 8 == 0 && i % 16 == 0 && o % 16 == 0 && a.estride % 2 == 0 &&
         a.ostride % 2 == 0;
}

// ADI along strided lines with the right-hand side formed at load (nt_tile_kernel.h)
#define NT_GOTA(R_, WAVES_)                                                                \
  do {                                                                                     \
    if ((R_) > (1 << a.plan.nlev)) return hipErrorInvalidValue;                            \
    if constexpr (NT_TILE_ADI_VEC && (WAVES_) == 8) {                                      \
      if (is_full(n, 64, R_) && tile_adi_vec_ok(a)) {                                      \
        a.tiles = a.phys / 8;                                                              \
        a.total = outer * a.tiles;                                                         \
        if (a.total <= 0) return hipSuccess;                                               \
        const void* kv_ = reinterpret_cast<const void*>(&line_kernel_tile_adi_vec<R_, PREC>); \
        const int64_t capv_ = persistent_grid(kv_, 512);                                   \
        const int64_t gridv_ = a.total < capv_ ? a.total : capv_;                          \
        hipLaunchKernelGGL((line_kernel_tile_adi_vec<R_, PREC>), dim3(static_cast<unsigned>(gridv_)), \
                           dim3(512), 0, st, a);                                           \
        return hipGetLastError();                                                          \
      }                                                                                    \
    }                                                                                      \
    const int64_t nsys_ = a.phys - 2;                                                      \
    a.tiles = (nsys_ + (WAVES_)-1) / (WAVES_);                                             \
    a.total = outer * a.tiles;                                                             \
    if (a.total <= 0) return hipSuccess;                                                   \
    const bool full_ = is_full(n, 64, R_);                                                 \
    const void* k_ = full_ ? reinterpret_cast<const void*>(&line_kernel_tile_adi<R_, WAVES_, PREC, true>)   \
                           : reinterpret_cast<const void*>(&line_kernel_tile_adi<R_, WAVES_, PREC, false>); \
    const int64_t cap_ = persistent_grid(k_, (WAVES_)*64);                                 \
    const int64_t grid_ = a.total < cap_ ? a.total : cap_;                                 \
    if (full_)                                                                             \
      hipLaunchKernelGGL((line_kernel_tile_adi<R_, WAVES_, PREC, true>), dim3(static_cast<unsigned>(grid_)), \
                         dim3((WAVES_)*64), 0, st, a);                                     \
    else                                                                                   \
      hipLaunchKernelGGL((line_kernel_tile_adi<R_, WAVES_, PREC, false>), dim3(static_cast<unsigned>(grid_)), \
                         dim3((WAVES_)*64), 0, st, a);                                     \
    return hipGetLastError();                                                              \
  } while (0)

// Fast path of the FP32-row tile kernel (nt_tile_kernel.h): power-of-two lines of
// 512 or 1024 points, whole 32-line tiles, FP32 rows and FP64 output 16-byte
// aligned, rows less than 2 MiB apart.  Measured on MI355X (FP32 + one sweep;
// FP64 solve of the same values):
//   1024^3 d/dy (rows 4 KiB apart): 2.89 -> 2.48 ms; 2.68 -> 2.49 ms
//   512^3 d/dy (2 KiB):             0.50 -> 0.43-0.47 ms; 0.43 -> 0.35 ms
//   512^3 d/dz (1 MiB):             0.55 -> 0.50-0.51 ms; 0.47 -> 0.41 ms
//   1024^3 d/dz (4 MiB):            3.79 -> 3.80-3.82 ms; 3.32 -> 3.43-3.44 ms
// so rows 2 MiB or more apart keep the general kernel.  0 turns it off.
#ifndef NT_TILE_F32_FAST
#define NT_TILE_F32_FAST 1
#endif
#ifndef NT_TILE_F32_FAST_ROW_LIMIT
#define NT_TILE_F32_FAST_ROW_LIMIT (2 << 20)
#endif

inline bool tile_f32_fast_ok(const LineArgs& a) {
  const uintptr_t i = reinterpret_cast<uintptr_t>(a.in), o = reinterpret_cast<uintptr_t>(a.out);
  return a.pstride == 1 && a.phys % 32 == 0 && i % 16 == 0 && o % 16 == 0 && a.estride % 4 == 0 &&
         a.ostride % 4 == 0 && a.estride * static_cast<int64_t>(sizeof(float)) < NT_TILE_F32_FAST_ROW_LIMIT;
}

// FP32-stored compact derivative along strided lines: 32-line tiles (nt_tile_kernel.h)
#define NT_GOTF(R_)                                                                        \
  do {                                                                                     \
    if ((R_) > (1 << a.plan.nlev)) return hipErrorInvalidValue;                            \
    a.tiles = (a.phys + 31) / 32;                                                          \
    a.total = outer * a.tiles;                                                             \
    if (a.total <= 0) return hipSuccess;                                                   \
    const bool full_ = is_full(n, 64, R_);                                                 \
    if constexpr (NT_TILE_F32_FAST && ((R_) == 8 || (R_) == 16)) {                         \
      if (full_ && tile_f32_fast_ok(a)) {                                                  \
        const void* kf_ = reinterpret_cast<const void*>(&line_kernel_tile_f32_fast<R_, PREC>); \
        const int64_t capf_ = persistent_grid(kf_, 1024);                                  \
        const int64_t gridf_ = a.total < capf_ ? a.total : capf_;                          \
        hipLaunchKernelGGL((line_kernel_tile_f32_fast<R_, PREC>), dim3(static_cast<unsigned>(gridf_)), \
                           dim3(1024), 0, st, a);                                          \
        return hipGetLastError();                                                          \
      }                                                                                    \
    }                                                                                      \
    const void* k_ = full_ ? reinterpret_cast<const void*>(&line_kernel_tile_f32<R_, PREC, true>)   \
                           : reinterpret_cast<const void*>(&line_kernel_tile_f32<R_, PREC, false>); \
    const int64_t cap_ = persistent_grid(k_, 1024);                                        \
    const int64_t grid_ = a.total < cap_ ? a.total : cap_;                                 \
    if (full_)                                                                             \
      hipLaunchKernelGGL((line_kernel_tile_f32<R_, PREC, true>), dim3(static_cast<unsigned>(grid_)), \
                         dim3(1024), 0, st, a);                                            \
    else                                                                                   \
      hipLaunchKernelGGL((line_kernel_tile_f32<R_, PREC, false>), dim3(static_cast<unsigned>(grid_)), \
                         dim3(1024), 0, st, a);                                            \
    return hipGetLastError();                                                              \
  } while (0)

// one line per workgroup over NW wavefronts of R rows per lane (nt_xw_kernel.h);
// n == 64 R NW
#define NT_GOX(R_, NW_)                                                                    \
  do {                                                                                     \
    if (a.plan.nlev != Log2<R_>::value + 5 + Log2<NW_>::value) return hipErrorInvalidValue; \
    const int64_t nsys_ = a.phys - 2 * ((PROD == P_ADI) ? 1 : 0);                          \
    a.tiles = nsys_;                                                                       \
    const int64_t grid_ = outer * nsys_;                                                   \
    if (grid_ <= 0) return hipSuccess;                                                     \
    if (grid_ > 0x7fffffffLL) return hipErrorInvalidValue;                                 \
    hipLaunchKernelGGL((line_kernel_xw<TIn, R_, NW_, PROD, PREC>), dim3(static_cast<unsigned>(grid_)), \
                       dim3((NW_)*64), 0, st, a);                                          \
    return hipGetLastError();                                                              \
  } while (0)

// Wavefronts per CU of the persistent wave-coalesced kernel (contiguous variant
// 4), by rows per lane.  Measured on MI355X, compact d/dx of 2^27-point fields
// (us at 4 / 8 / 16 wavefronts per CU; the default non-persistent kernel in
// brackets), FP64: 128 points 1191 / 688 / 589 (555), 256 points 770 / 492 /
// 431 (376), 512 points 474 / 395 / 399 (334), 1024 points 371 / 402 / 404
// (385), 2048 points 385 / 395 / 394 (442); FP32 + refinement at 2048 points
// 385 / 384 / 387 (536).  The persistent form is the default for 2048-point
// lines only.
#ifndef NT_WCP_PER_CU
#define NT_WCP_PER_CU(R) (((R) <= 4) ? 16 : (((R) <= 8) ? 8 : 4))
#endif
int64_t device_cus();
int occupancy_per_cu(const void* kernel, int threads, size_t dyn_lds);
inline int wcp_per_cu(int r) {
  const int c = NT_WCP_PER_CU(r);
  return c < 1 ? 1 : c;
}

// The grid never exceeds the workgroups that fit at once (a persistent
// workgroup that waits for a slot would start its whole share at the end).
#define NT_GOWP(R_)                                                                        \
  if constexpr ((R_) >= 16 / static_cast<int>(sizeof(TIn))) do {                          \
    if ((R_) > (1 << a.plan.nlev)) return hipErrorInvalidValue;                            \
    const int64_t nsys_ = a.phys;                                                          \
    a.tiles = nsys_;                                                                       \
    if (nsys_ <= 0) return hipSuccess;                                                     \
    const void* k_ = reinterpret_cast<const void*>(&line_kernel_wcp<TIn, R_, PROD, PREC>); \
    const int want_ = wcp_per_cu(R_);                                                      \
    const size_t slds_ = static_cast<size_t>(64) * ((R_) + 2) * 8;                         \
    const size_t per_lds_ = (static_cast<size_t>(160 * 1024) / want_) & ~static_cast<size_t>(1023); \
    const size_t lds_ = per_lds_ > slds_ ? per_lds_ - slds_ : 0;                           \
    int per_ = occupancy_per_cu(k_, 64, lds_);                                             \
    per_ = per_ < want_ ? per_ : want_;                                                    \
    int64_t grid_ = device_cus() * per_;                                                   \
    grid_ = grid_ < nsys_ ? grid_ : nsys_;                                                 \
    hipLaunchKernelGGL((line_kernel_wcp<TIn, R_, PROD, PREC>), dim3(static_cast<unsigned>(grid_)), \
                       dim3(64), lds_, st, a);                                             \
    return hipGetLastError();                                                              \
  } while (0)

// 16-byte chunk accesses need 16-byte aligned line starts
inline bool direct_ok(const LineArgs& a, size_t in_size, size_t out_size) {
  const uintptr_t i = reinterpret_cast<uintptr_t>(a.in), o = reinterpret_cast<uintptr_t>(a.out);
  return (i % 16 == 0) && (o % 16 == 0) && ((a.pstride * in_size) % 16 == 0) &&
         ((a.pstride * out_size) % 16 == 0) && a.estride == 1;
}

template <typename TIn, typename TOut, int LAYOUT, int PROD, int PREC>
hipError_t launch_line(LineArgs& a, int64_t outer, hipStream_t st) {
  const int n = a.plan.n;
  int variant = tune_get(LAYOUT + (PROD == P_ADI ? 2 : 0), PREC, n);
  if constexpr (LAYOUT == CONTIG) {
    const int vin = 16 / static_cast<int>(sizeof(TIn)), vout = 16 / static_cast<int>(sizeof(TOut));
    const bool aligned = n > 64 && direct_ok(a, sizeof(TIn), sizeof(TOut));
    if constexpr (std::is_same_v<TOut, double>) {
      if (variant == 3 && aligned && a.pstride == n && (n & (n - 1)) == 0 && n >= 256) {
        if (n == 256) NT_GOX(2, 2);
        if (n == 512) NT_GOX(4, 2);
        if (n == 1024) NT_GOX(8, 2);
        if (n == 2048) NT_GOX(16, 2);
      }
    }
    if constexpr (std::is_same_v<TOut, double> && PROD != P_ADI) {
      if (variant == 4 && aligned && a.pstride == n && outer == 1 && (n & (n - 1)) == 0 &&
          n >= 64 * vin) {
        if (n == 128) NT_GOWP(2);
        if (n == 256) NT_GOWP(4);
        if (n == 512) NT_GOWP(8);
        if (n == 1024) NT_GOWP(16);
        if (n == 2048) NT_GOWP(32);
      }
    }
    if ((variant == 0 || variant >= 3) && aligned && a.pstride == n && n % vin == 0 && n % vout == 0) {
      if constexpr (PROD != P_ADI) {
        // short lines on 16- or 32-lane segments, several lines per wavefront
        // (fewer cross-lane levels per line): 2^26-point fields, FP64,
        // 128 points 338 -> 178 us, 256 points 224 -> 181 us, 384 points
        // 267 -> 229 us; 512 points keep 64 lanes (171 against 189 us)
        if (n <= 128) NT_GOW(8, 16, 4);
        if (n <= 256) NT_GOW(8, 32, 4);
        if (n < 512) NT_GOW(16, 32, 1);
      }
      if (n <= 128) NT_GOW(2, 64, 4);
      if (n <= 256) NT_GOW(4, 64, 4);
      if (n <= 512) NT_GOW(8, 64, NT_WCT_WAVES_512);
      if (n <= 1024) NT_GOW(16, 64, 4);
      if (n <= 2048) NT_GOW(32, 64, 2);
    }
    if (variant <= 1 && aligned) {
      if (n <= 128) NT_GOD(2, 64, 4);
      if (n <= 256) NT_GOD(4, 64, 4);
      if (n <= 512) NT_GOD(8, 64, 4);
      if (n <= 1024) NT_GOD(16, 64, 4);
      if (n <= 2048) NT_GOD(32, 64, 4);
    }
    // LDS-staged workgroup tiles: one system per wavefront (16 lanes for tiny
    // systems); short lines outside ADI on 16-lane segments, four lines per
    // wavefront (2^26-point fields, FP64: 64 points 619 -> 233 us, 32 points
    // 1059 -> 401 us)
    if (n <= 16) NT_GO(1, 16, 4, 16);
    if constexpr (PROD != P_ADI) {
      if (n <= 32) NT_GO(2, 16, 4, 32);
      if (n <= 64) NT_GO(4, 16, 4, 64);
    }
    if (n <= 64) NT_GO(1, 64, 4, 64);
    if (n <= 128) NT_GO(2, 64, 4, 128);
    if (n <= 256) NT_GO(4, 64, 4, 256);
    if (n <= 512) NT_GO(8, 64, 4, 512);
    if (n <= 1024) NT_GO(16, 64, 4, 1024);
    if (n <= 2048) NT_GO(32, 64, 4, 2048);
  } else {
    if constexpr (std::is_same_v<TOut, double> && PROD != P_ADI && PREC == PREC_F64) {
      if (variant == 5 && a.lane && a.lane->m >= 4 && a.pstride == 1) NT_GOL();
    }
    if constexpr (std::is_same_v<TIn, float> && std::is_same_v<TOut, double> && PROD == P_COMPACT &&
                  PREC == PREC_F32R) {
      if (variant == 5 && lane32_ok(a.lane32) && a.pstride == 1 && a.sweeps == 1) NT_GOL32();
    }
    // lines the lane kernels do not take (FP64 below 49 points: 32-line tiles,
    // 33..48 points on 2^26-point fields 441-613 us against 523-744 for the
    // segments)
    if (variant == 5) variant = lane_fallback(PREC, n);
    if constexpr (PROD != P_ADI && !(std::is_same_v<TIn, float> && std::is_same_v<TOut, double> &&
                                     PROD == P_COMPACT)) {
      if (variant == 3) {
        // persistent whole-tile kernel, two lines per wavefront
        if (n <= 64) NT_GOT(1, NT_TILE2_WAVES, 2);
        if (n <= 128) NT_GOT(2, NT_TILE2_WAVES, 2);
        if (n <= 256) NT_GOT(4, NT_TILE2_WAVES, 2);
        if (n <= 512) NT_GOT(8, NT_TILE2_WAVES, 2);
      }
    }
    if constexpr (PROD != P_ADI && !(std::is_same_v<TIn, float> && std::is_same_v<TOut, double> &&
                                     PROD == P_COMPACT)) {
      if (variant == 4) {
        // as 3, both lines of a wavefront solved at once on 32-lane segments
        if (n >= 4 && n <= 64) NT_GOTS(1, NT_TILE2_WAVES);
        if (n > 64 && n <= 128) NT_GOTS(2, NT_TILE2_WAVES);
        if (n > 128 && n <= 256) NT_GOTS(4, NT_TILE2_WAVES);
        if (n > 256 && n <= 512) NT_GOTS(8, NT_TILE2_WAVES);
      }
    }
    if (variant >= 2) {
      // persistent whole-tile kernel, next tile prefetched into registers
      if constexpr (std::is_same_v<TIn, float> && std::is_same_v<TOut, double> && PROD == P_COMPACT) {
        // FP32 rows: 32-line tiles, two lines per wavefront (lines up to 1024 points)
        if (n <= 64) NT_GOTF(1);
        if (n <= 128) NT_GOTF(2);
        if (n <= 256) NT_GOTF(4);
        if (n <= 512) NT_GOTF(8);
        if (n <= 1024) NT_GOTF(16);
        if (n <= 2048) NT_GOT(32, 8, 1);
      } else if constexpr (PROD != P_ADI) {
        if (n <= 64) NT_GOT(1, 16, 1);
        if (n <= 128) NT_GOT(2, 16, 1);
        if (n <= 256) NT_GOT(4, 16, 1);
        if (n <= 512) NT_GOT(8, 16, 1);
        if (n <= 1024) NT_GOT(16, 16, 1);
        if (n <= 2048) NT_GOT(32, 8, 1);
      } else {
        if (n <= 64) NT_GOT(1, 16, 1);
        if (n <= 128) NT_GOT(2, 16, 1);
        if (n <= 256) NT_GOT(4, 16, 1);
        if (n <= 512) NT_GOT(8, 16, 1);
        if (n <= 1024) NT_GOT(16, 8, 1);
        // 2048-point ADI lines with their halo lines do not fit a whole tile in LDS:
        // the right-hand side is formed while loading instead
        if (n <= 2048) NT_GOTA(32, 8);
      }
    }
    if (variant == 1) {
      // ADI on 1025..2048-point lines: whole 8-line tiles with the right-hand side
      // formed at load (ADI 2048^2 step 67.8 -> 60.0 us against the row-block kernel)
      if constexpr (PROD == P_ADI) {
        if (n > 1024 && n <= 2048) NT_GOTA(32, 8);
      }
      // 32 adjacent lines per tile (256-byte FP64 rows), 32 lanes per line
      if (n <= 32) NT_GOS(1, 32, 16, 32);
      if (n <= 64) NT_GOS(2, 32, 16, 64);
      if (n <= 128) NT_GOS(4, 32, 16, 128);
      if (n <= 256) NT_GOS(8, 32, 16, 128);
      if (n <= 512) NT_GOS(16, 32, 16, 128);
      // long lines: 8 adjacent lines, 64 lanes each, 512 threads (room for 32-row chunks)
      if (n <= 1024) NT_GOS(16, 64, 8, 512);
      if (n <= 2048) NT_GOS(32, 64, 8, 512);
    }
    // 16 adjacent lines per tile (128-byte FP64 rows), 64 lanes per line
    if (n <= 64) NT_GOS(1, 64, 16, 64);
    if (n <= 128) NT_GOS(2, 64, 16, 128);
    if (n <= 256) NT_GOS(4, 64, 16, 128);
    if (n <= 512) NT_GOS(8, 64, 16, 128);
    if (n <= 1024) NT_GOS(16, 64, 16, 128);
    if (n <= 2048) NT_GOS(32, 64, 16, 128);
  }
  return hipErrorInvalidValue;
}

#define NT_INSTANTIATE(TIn, TOut, LAYOUT, PROD, PREC) \
  template hipError_t launch_line<TIn, TOut, LAYOUT, PROD, PREC>(LineArgs&, int64_t, hipStream_t);

// Each combination is instantiated (NT_INSTANTIATE) in exactly one nt_k_*.hip
// file.  Everywhere else these declarations stop the implicit instantiation
// that would compile every kernel a second time into the caller's code object.
#define NT_DECLARE(TIn, TOut, LAYOUT, PROD, PREC) \
  extern template hipError_t launch_line<TIn, TOut, LAYOUT, PROD, PREC>(LineArgs&, int64_t, hipStream_t);
#define NT_DECLARE_BOTH(TIn, TOut, LAYOUT, PROD) \
  NT_DECLARE(TIn, TOut, LAYOUT, PROD, PREC_F64) NT_DECLARE(TIn, TOut, LAYOUT, PROD, PREC_F32R)
NT_DECLARE_BOTH(double, double, CONTIG, P_IDENT)
NT_DECLARE_BOTH(double, double, STRIDED, P_IDENT)
NT_DECLARE_BOTH(double, double, CONTIG, P_COMPACT)
NT_DECLARE_BOTH(double, double, STRIDED, P_COMPACT)
NT_DECLARE_BOTH(float, double, CONTIG, P_COMPACT)
NT_DECLARE_BOTH(float, double, STRIDED, P_COMPACT)
NT_DECLARE_BOTH(double, double, CONTIG, P_ADI)
NT_DECLARE_BOTH(double, double, STRIDED, P_ADI)
#undef NT_DECLARE_BOTH
#undef NT_DECLARE

}  // namespace nt
