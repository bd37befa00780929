// Mixed-size persistent batch: many fields of different shapes, derivative
// directions, precisions and storage types solved by one launch.
//
// The host packs a binary image (job table + level tables of the plans the
// jobs use) into a caller-provided host buffer; the caller copies it to the
// device and zeroes a work counter; one persistent launch then pulls chunks
// (consecutive 16-line tiles of one job, about 128 K points) from the counter
// until every job is done.
//
// Each job has a class: layout (contiguous x-lines or strided y/z-lines),
// rows per lane R = 1, 2, 4 or 8 (64 lanes x R >= the line length, up to 512
// points), whether the line length is exactly 64 R (a power of two, so the
// level structure is known at compile time), precision, and producer with its
// storage type.  Every class has its own compile-time specialised loop, with
// the loop shapes of the per-shape kernels: contiguous lines are worked
// through by each wavefront on its own (next line prefetched into registers,
// 16-byte coalesced accesses, LDS transpose for R >= 4); strided lines go
// through whole 16-line tiles in LDS, two lines per wavefront, with the next
// tile prefetched into registers during the solve.  The level factors are
// read with scalar loads.
// Per-line arithmetic depends only on the job, so a field gives the same bits
// alone and in any mix.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "../../include/ntoeplitz.h"
#include "nt_error.h"
#include "nt_line_kernel.h"
#include "nt_plan.h"

namespace nt {

constexpr int BATCH_MAXN = 512;
constexpr int BATCH_TS = 16;                // lines per tile
constexpr int BATCH_WAVES = 8;              // wavefronts per workgroup
constexpr int BATCH_NT = BATCH_WAVES * 64;
// A chunk is the unit handed out by the work counter: consecutive tiles of one
// job (at least one).  Its size in points is the batch's total divided by
// BATCH_CHUNKS_PER_SLOT x 512 (the resident workgroups of a full MI355X),
// kept within [BATCH_CHUNK_MIN_PTS, BATCH_CHUNK_MAX_PTS]: large enough that a
// chunk's fixed cost (finding its job, the first tile's loads) is small, small
// enough that the last chunks do not leave most of the GPU idle.
constexpr int64_t BATCH_CHUNK_MAX_PTS = int64_t(1) << 17;
constexpr int64_t BATCH_CHUNK_MIN_PTS = int64_t(1) << 13;
constexpr int64_t BATCH_CHUNKS_PER_SLOT = 8;

// LDS doubles per staged line of a strided tile (two lines per wavefront):
// lane l's chunk of R points starts R + 1 doubles after lane l-1's, an odd
// number of 8-byte bank pairs, so the solve's chunk reads do not collide.
__host__ __device__ constexpr int batch_ep(int R) { return 64 * (R + (R >= 2 ? 1 : 0)) + 1; }
// LDS doubles per contiguous line (one per wavefront): chunks R + 2 apart, so
// every chunk starts 16-byte aligned.
__host__ __device__ constexpr int batch_ls(int R) { return 64 * (R + 2); }
constexpr int BATCH_LDS = (BATCH_TS * batch_ep(8) > BATCH_WAVES * batch_ls(8)) ? BATCH_TS * batch_ep(8)
                                                                               : BATCH_WAVES * batch_ls(8);

// What a job computes and from which storage type.
enum BatchKind : int { K_SOLVE = 0, K_CD64 = 1, K_CD32 = 2 };
// Class of a job: every combination has a compile-time specialised loop.
constexpr int NUM_CLASSES = 2 * 4 * 2 * 2 * 3;
__host__ __device__ constexpr int batch_class(int layout, int rsel, int full, int prec, int kind) {
  return (((layout * 4 + rsel) * 2 + full) * 2 + prec) * 3 + kind;
}

struct BatchJob {
  const void* in;
  void* out;
  int64_t pstride, estride, ostride;  // line geometry (as LineArgs)
  int64_t phys;                       // lines per slab
  int64_t tiles;                      // tiles per slab
  int64_t ntiles;                     // tiles of this job (slabs x tiles)
  int64_t chunk0;                     // first global chunk of this job
  int64_t chunk_tiles;                // tiles per chunk
  double s1, s2;                      // compact scalars 3/(4h), 1/(2h)
  int32_t plan;                       // index into the plan table
  int32_t cls;                        // batch_class(...)
  int32_t sweeps;
  int32_t vec;                        // buffers and strides allow 16-byte accesses
};

struct BatchArgs {
  const BatchJob* jobs;
  const NtPlanK* plans;
  unsigned int* counter;
  int64_t total;                      // chunks over all jobs
  int32_t njobs;
  int32_t pad_;
};

// The job and plan tables are read through the constant address space, i.e.
// with scalar loads: written before the launch, only read by the kernel.  As
// vector loads they would wait, in order, behind the previous tile's stores;
// and the level factors stay scalar operands of the solve.
using CJob = const __attribute__((address_space(4))) BatchJob;
using CPlan = const __attribute__((address_space(4))) NtPlanK;

template <int KIND>
using BatchIn = typename std::conditional<KIND == K_CD32, float, double>::type;

// ---- contiguous lines (x-lines): every wavefront works through its own lines ----
// Wavefront w takes lines q0 + w, q0 + w + 8, ... of the chunk.  The next
// line's loads are issued into registers before the current line is solved.
// Up to 2 rows per lane a lane loads its own R consecutive points directly
// (64 x 8 R contiguous bytes per wavefront); longer lines are loaded with
// coalesced 16-byte accesses into the wavefront's LDS rows and transposed into
// chunks there.  Loads are unconditional (indexes clamped into the line), so
// the compiler can count them.
template <typename TIn, int R, bool FULL>
__device__ __forceinline__ void cline_fetch(const TIn* __restrict__ src, int n, int lane, bool vec, TIn (&pf)[R]) {
  if (R <= 2) {
    if (R == 2 && vec) {                       // a lane's two points as one access
      const int j = FULL ? lane : min(lane, n / 2 - 1);
      if constexpr (sizeof(TIn) == 8) {
        const nt_dbl2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_dbl2*>(src) + j);
        pf[0] = v.x;
        pf[R - 1] = v.y;
      } else {
        typedef float nt_flt2 __attribute__((ext_vector_type(2)));
        const nt_flt2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_flt2*>(src) + j);
        pf[0] = v.x;
        pf[R - 1] = v.y;
      }
    } else {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int e = lane * R + k;
        pf[k] = __builtin_nontemporal_load(src + (FULL ? e : min(e, n - 1)));
      }
    }
  } else if (vec) {
    constexpr int V = 16 / static_cast<int>(sizeof(TIn));
    const int nvec = n / V;
#pragma unroll
    for (int i = 0; i < R / V; ++i) {
      const int j = FULL ? lane + 64 * i : min(lane + 64 * i, nvec - 1);
      if constexpr (sizeof(TIn) == 8) {
        const nt_dbl2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_dbl2*>(src) + j);
        pf[2 * i] = v.x;
        pf[2 * i + 1] = v.y;
      } else {
        const nt_flt4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_flt4*>(src) + j);
        pf[4 * i] = v.x;
        pf[4 * i + 1] = v.y;
        pf[4 * i + 2] = v.z;
        pf[4 * i + 3] = v.w;
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int e = lane + 64 * i;
      pf[i] = __builtin_nontemporal_load(src + (FULL ? e : min(e, n - 1)));
    }
  }
}

// LDS position of point e of a contiguous line (chunks R + 2 apart)
template <int R>
__device__ __forceinline__ int cl_at(int e) {
  return (e / R) * (R + 2) + (e % R);
}

// the fetched values into the wavefront's LDS rows (R >= 4); clamped indexes
// write a point's own value a second time
template <typename TIn, int R, bool FULL>
__device__ __forceinline__ void cline_deposit(double* __restrict__ L, int n, int lane, bool vec, const TIn (&pf)[R]) {
  if (vec) {
    constexpr int V = 16 / static_cast<int>(sizeof(TIn));
    const int nvec = n / V;
#pragma unroll
    for (int i = 0; i < R / V; ++i) {
      const int j = FULL ? lane + 64 * i : min(lane + 64 * i, nvec - 1);
#pragma unroll
      for (int c = 0; c < V; ++c) L[cl_at<R>(V * j + c)] = static_cast<double>(pf[V * i + c]);
    }
  } else {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int e = lane + 64 * i;
      L[cl_at<R>(FULL ? e : min(e, n - 1))] = static_cast<double>(pf[i]);
    }
  }
}

// right-hand side (compact) or the values themselves (solve), then the solve
// on segments of W lanes
template <int R, bool FULL, int PREC, int KIND, int W = 64>
__device__ __forceinline__ void batch_solve(const double (&f)[R], const double (&bnd)[6], int my_e0, int n,
                                            double s1, double s2, const NtPlanK& P, int lane, int sweeps,
                                            double (&x)[R]) {
  double d[R];
  if constexpr (KIND != K_SOLVE) {
    const double fl = shup<W>(f[R - 1], 1);
    const double fr = shdn<W>(f[0], 1);
    compact_rhs_chunk<R, FULL>(f, fl, fr, bnd, my_e0, n, s1, s2, d);
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k) d[k] = f[k];
  }
  solve_chunk<R, W, PREC, FULL>(d, x, P, lane, sweeps);
}

template <int R, bool FULL, int PREC, int KIND>
__device__ __forceinline__ void contig_chunk(CJob& J, const NtPlanK& P, int64_t t0, int64_t t1, double* lds) {
  using TIn = BatchIn<KIND>;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int n = FULL ? 64 * R : P.n;
  const int64_t phys = J.phys;
  const int64_t pstride = J.pstride;
  const int64_t q_end = (t1 * BATCH_TS < phys) ? t1 * BATCH_TS : phys;
  int64_t q = t0 * BATCH_TS + wave;
  if (q >= q_end) return;                       // uniform across the wavefront
  const TIn* __restrict__ in = static_cast<const TIn*>(J.in);
  double* __restrict__ out = static_cast<double*>(J.out);
  const bool vec = J.vec != 0;
  const double s1 = J.s1, s2 = J.s2;
  const int sweeps = J.sweeps;
  const int my_e0 = lane * R;
  double* __restrict__ L = lds + wave * batch_ls(R);
  double* __restrict__ mine = L + lane * (R + 2);

  TIn pf[R];
  // an opaque lane index here too: offsets derived from threadIdx.x alone
  // would be hoisted out of the loop over chunks
  cline_fetch<TIn, R, FULL>(in + q * pstride, n, (static_cast<int>(threadIdx.x) + opaque_vzero()) & 63, vec, pf);
  for (;;) {
    // the lane index through an opaque move per line: the offsets derived from
    // it are recomputed, not held across the loop
    const int ln = (static_cast<int>(threadIdx.x) + opaque_vzero()) & 63;
    const int64_t qn = q + BATCH_WAVES;
    const bool more = qn < q_end;               // uniform
    double f[R];
    double bnd[6];
    if constexpr (R <= 2) {
#pragma unroll
      for (int k = 0; k < R; ++k) f[k] = (FULL || my_e0 + k < n) ? static_cast<double>(pf[k]) : 0.0;
      if constexpr (KIND != K_SOLVE) {
#pragma unroll
        for (int i = 0; i < 6; ++i) {           // point e sits in lane e / R, register e % R
          const int e = (i < 3) ? i : n + 2 - i;
          const double v = (R == 1 || (e & 1) == 0) ? f[0] : f[R - 1];
          bnd[i] = shidx<64>(v, e / R);
        }
      }
    } else {
      wave_lds_sync();                          // the previous line's stores have read the rows
      cline_deposit<TIn, R, FULL>(L, n, ln, vec, pf);
      wave_lds_sync();
      if (FULL || my_e0 + R <= n) {
#pragma unroll
        for (int k = 0; k < R; k += 2) {
          const nt_dbl2 t = *reinterpret_cast<const nt_dbl2*>(mine + k);
          f[k] = t.x;
          f[k + 1] = t.y;
        }
      } else {
#pragma unroll
        for (int k = 0; k < R; ++k) f[k] = (my_e0 + k < n) ? mine[k] : 0.0;
      }
      if constexpr (KIND != K_SOLVE) {
#pragma unroll
        for (int i = 0; i < 6; ++i) bnd[i] = L[cl_at<R>((i < 3) ? i : n + 2 - i)];
      }
    }
    if (more) cline_fetch<TIn, R, FULL>(in + qn * pstride, n, ln, vec, pf);     // in flight during the solve
    double x[R];
    batch_solve<R, FULL, PREC, KIND>(f, bnd, my_e0, n, s1, s2, P, lane, sweeps, x);
    double* __restrict__ dst = out + q * pstride;
    if constexpr (R <= 2) {
      if (R == 2 && vec && (FULL || my_e0 + 2 <= n)) {
        nt_dbl2 t;
        t.x = x[0];
        t.y = x[R - 1];
        __builtin_nontemporal_store(t, reinterpret_cast<nt_dbl2*>(dst) + lane);
      } else {
#pragma unroll
        for (int k = 0; k < R; ++k)
          if (FULL || my_e0 + k < n) __builtin_nontemporal_store(x[k], dst + my_e0 + k);
      }
    } else {
      wave_lds_sync();                          // every lane has read its chunk and the closures
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
      const int ls = (static_cast<int>(threadIdx.x) + opaque_vzero()) & 63;   // fresh, as ln
      if (vec) {
#pragma unroll
        for (int i = 0; i < R / 2; ++i) {
          const int j = ls + 64 * i;
          if (FULL || j < n / 2) {
            nt_dbl2 t;
            t.x = L[cl_at<R>(2 * j)];
            t.y = L[cl_at<R>(2 * j + 1)];
            __builtin_nontemporal_store(t, reinterpret_cast<nt_dbl2*>(dst) + j);
          }
        }
      } else {
#pragma unroll
        for (int i = 0; i < R; ++i) {
          const int e = ls + 64 * i;
          if (FULL || e < n) __builtin_nontemporal_store(L[cl_at<R>(e)], dst + e);
        }
      }
    }
    if (!more) break;
    q = qn;
  }
}

// ---- contiguous lines of up to 256 points: several lines per wavefront -------------------
// A wavefront takes SPW consecutive lines at once (4 lines on 16-lane segments
// up to 128 points, 2 on 32-lane segments up to 256) and solves them together
// with RS = R x SPW rows per lane: the same padded length as one line on 64
// lanes, but fewer cross-lane levels per line.  Each line is read with the
// whole wavefront's coalesced accesses into the wavefront's LDS rows (chunks
// RS + 2 apart) and the next group of lines is fetched into registers during
// the solve.  Lines past the chunk's end are fetched clamped to its last line
// and solved as zeros (no shuffle then carries a non-finite value into a live
// segment); their results are not stored.
template <int R, bool FULL, int PREC, int KIND>
__device__ __forceinline__ void contig_chunk_seg(CJob& J, const NtPlanK& P, int64_t t0, int64_t t1, double* lds) {
  using TIn = BatchIn<KIND>;
  constexpr int SPW = (R <= 2) ? 4 : 2;       // lines per wavefront
  constexpr int W = 64 / SPW;                 // lanes per line
  constexpr int RS = R * SPW;                 // rows per lane
  constexpr int CS = RS + 2;                  // LDS doubles per chunk
  constexpr int LSEG = W * CS;                // LDS doubles per line
  static_assert(SPW * LSEG <= batch_ls(8), "the wavefront's LDS rows");
  constexpr int V = 16 / static_cast<int>(sizeof(TIn));
  constexpr int NV = (R >= V) ? R / V : 0;    // 16-byte loads per lane and line (0: scalar loads)
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int n = FULL ? 64 * R : P.n;
  const int64_t phys = J.phys;
  const int64_t pstride = J.pstride;
  const int64_t q_end = (t1 * BATCH_TS < phys) ? t1 * BATCH_TS : phys;
  int64_t q = t0 * BATCH_TS + static_cast<int64_t>(wave) * SPW;
  if (q >= q_end) return;                     // uniform across the wavefront
  const TIn* __restrict__ in = static_cast<const TIn*>(J.in);
  double* __restrict__ out = static_cast<double*>(J.out);
  const bool vec = (NV > 0) && J.vec != 0;
  const double s1 = J.s1, s2 = J.s2;
  const int sweeps = J.sweeps;
  const int seg = lane / W;
  const int sl = lane & (W - 1);
  double* __restrict__ L = lds + wave * batch_ls(8);
  auto at = [](int e) -> int { return (e / RS) * CS + (e % RS); };

  TIn pf[SPW][R];
  auto fetch = [&](int64_t q0, int ln) {
#pragma unroll
    for (int s = 0; s < SPW; ++s) {
      const int64_t qs = (q0 + s < q_end) ? q0 + s : q_end - 1;
      const TIn* __restrict__ src = in + qs * pstride;
      if (vec) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
          const int j = FULL ? ln + 64 * i : min(ln + 64 * i, n / V - 1);
          if constexpr (sizeof(TIn) == 8) {
            const nt_dbl2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_dbl2*>(src) + j);
            pf[s][2 * i] = v.x;
            pf[s][2 * i + 1] = v.y;
          } else {
            const nt_flt4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_flt4*>(src) + j);
            pf[s][4 * i] = v.x;
            pf[s][4 * i + 1] = v.y;
            pf[s][4 * i + 2] = v.z;
            pf[s][4 * i + 3] = v.w;
          }
        }
      } else {
#pragma unroll
        for (int i = 0; i < R; ++i) {
          const int e = ln + 64 * i;
          pf[s][i] = __builtin_nontemporal_load(src + (FULL ? e : min(e, n - 1)));
        }
      }
    }
  };
  auto deposit = [&](int ln) {
#pragma unroll
    for (int s = 0; s < SPW; ++s) {
      double* __restrict__ Ls = L + s * LSEG;
      if (vec) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
          const int j = FULL ? ln + 64 * i : min(ln + 64 * i, n / V - 1);
#pragma unroll
          for (int c = 0; c < V; ++c) Ls[at(V * j + c)] = static_cast<double>(pf[s][V * i + c]);
        }
      } else {
#pragma unroll
        for (int i = 0; i < R; ++i) {
          const int e = ln + 64 * i;
          Ls[at(FULL ? e : min(e, n - 1))] = static_cast<double>(pf[s][i]);
        }
      }
    }
  };

  fetch(q, (static_cast<int>(threadIdx.x) + opaque_vzero()) & 63);
  for (;;) {
    const int ln = (static_cast<int>(threadIdx.x) + opaque_vzero()) & 63;   // fresh per group
    const int64_t qn = q + BATCH_WAVES * SPW;
    const bool more = qn < q_end;               // uniform
    wave_lds_sync();                            // the previous group's stores have read the rows
    deposit(ln);
    wave_lds_sync();
    const bool on = q + seg < q_end;            // my segment holds a line
    const double* __restrict__ mine = L + seg * LSEG + sl * CS;
    const int e0 = sl * RS;
    double f[RS];
    if (FULL || e0 + RS <= n) {
#pragma unroll
      for (int k = 0; k < RS; k += 2) {
        const nt_dbl2 t = *reinterpret_cast<const nt_dbl2*>(mine + k);
        f[k] = on ? t.x : 0.0;
        f[k + 1] = on ? t.y : 0.0;
      }
    } else {
#pragma unroll
      for (int k = 0; k < RS; ++k) f[k] = (on && e0 + k < n) ? mine[k] : 0.0;
    }
    double bnd[6];
    if constexpr (KIND != K_SOLVE) {
#pragma unroll
      for (int i = 0; i < 6; ++i) bnd[i] = on ? L[seg * LSEG + at((i < 3) ? i : n + 2 - i)] : 0.0;
    } else {
#pragma unroll
      for (int i = 0; i < 6; ++i) bnd[i] = 0.0;
    }
    if (more) fetch(qn, ln);                    // in flight during the solve
    double x[RS];
    batch_solve<RS, FULL, PREC, KIND, W>(f, bnd, e0, n, s1, s2, P, sl, sweeps, x);
    wave_lds_sync();                            // every lane has read its chunk and the closures
    double* __restrict__ mw = L + seg * LSEG + sl * CS;
    if (FULL || e0 + RS <= n) {
#pragma unroll
      for (int k = 0; k < RS; k += 2) {
        nt_dbl2 t;
        t.x = x[k];
        t.y = x[k + 1];
        *reinterpret_cast<nt_dbl2*>(mw + k) = t;
      }
    } else {
#pragma unroll
      for (int k = 0; k < RS; ++k)
        if (e0 + k < n) mw[k] = x[k];
    }
    wave_lds_sync();
    const int ls = (static_cast<int>(threadIdx.x) + opaque_vzero()) & 63;   // fresh, as ln
#pragma unroll
    for (int s = 0; s < SPW; ++s) {
      if (q + s < q_end) {                      // uniform
        const double* __restrict__ Ls = L + s * LSEG;
        double* __restrict__ dst = out + (q + s) * pstride;
        if (J.vec != 0 && (n % 2) == 0) {
          // 16-byte stores: 32 R of them cover the longest line, so at R = 1
          // only half the lanes store even for a full line
          constexpr int NVS = (32 * R + 63) / 64;
          constexpr bool ALL = FULL && (32 * R) % 64 == 0;
#pragma unroll
          for (int i = 0; i < NVS; ++i) {
            const int j = ls + 64 * i;
            if (ALL || j < n / 2) {
              nt_dbl2 t;
              t.x = Ls[at(2 * j)];
              t.y = Ls[at(2 * j + 1)];
              __builtin_nontemporal_store(t, reinterpret_cast<nt_dbl2*>(dst) + j);
            }
          }
        } else {
#pragma unroll
          for (int i = 0; i < R; ++i) {
            const int e = ls + 64 * i;
            if (FULL || e < n) __builtin_nontemporal_store(Ls[at(e)], dst + e);
          }
        }
      }
    }
    if (!more) break;
    q = qn;
  }
}

// ---- strided lines (y- and z-lines): 16-line whole tiles in LDS ------------------------
// A tile is 16 adjacent lines, whose rows are runs of 16 consecutive elements
// in memory; the whole tile sits in LDS and each of the 8 wavefronts solves
// two of its lines, at once on two 32-lane segments with 2 R rows per lane
// (up to 4 rows per lane; at 8 the doubled rows would spill in this kernel's
// 128 VGPRs) or one after the other on 64 lanes.  The next tile's values are loaded into registers while the
// current tile is solved and stored.
#ifndef NT_BATCH_SEG_MAX_R
#define NT_BATCH_SEG_MAX_R(PREC) 4
#endif
template <int R, bool FULL, int PREC, int KIND>
__device__ __forceinline__ void strided_chunk(CJob& J, const NtPlanK& P, int64_t t0, int64_t t1, double* lds) {
  using TIn = BatchIn<KIND>;
  constexpr int CP = (R >= 2) ? 1 : 0;
  constexpr int EP = batch_ep(R);
  constexpr int NMAX = 64 * R;
  constexpr int VIN = 16 / static_cast<int>(sizeof(TIn));
  constexpr int VPR = BATCH_TS / VIN;                          // input vectors per row
  constexpr int PFV = (VPR * NMAX + BATCH_NT - 1) / BATCH_NT;  // input vectors per thread
  constexpr int PFS = BATCH_TS * NMAX / BATCH_NT;              // scalars per thread (2 R)
  constexpr int PF = (PFV * VIN > PFS) ? PFV * VIN : PFS;
  // A thread's fetch vectors reach past the tile's last point when PFV
  // threads-worth exceed the tile (FP32 rows at R = 1): clamp there too.
  constexpr bool CLAMP_V = !FULL || (PFV * BATCH_NT > VPR * NMAX);
  constexpr int VPO = BATCH_TS / 2;                            // output vectors per row
  constexpr int PFO = VPO * NMAX / BATCH_NT;                   // output vectors per thread (R)
  const int n = FULL ? NMAX : P.n;
  const int64_t estride = J.estride, ostride = J.ostride, phys = J.phys, tiles = J.tiles;
  const bool jvec = J.vec != 0;
  const double s1 = J.s1, s2 = J.s2;
  const int sweeps = J.sweeps;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int my_e0 = lane * R;
  // Reset at the top of every tile through an opaque move, so the per-thread
  // load, store and LDS offsets derived from it are recomputed for each tile
  // instead of being hoisted out of the tile loop (or, for the chunk's first
  // fetch, out of the loop over chunks) and held across it.
  int tid = (static_cast<int>(threadIdx.x) + opaque_vzero()) & (BATCH_NT - 1);
  const TIn* __restrict__ in = static_cast<const TIn*>(J.in);
  double* __restrict__ out = static_cast<double*>(J.out);
  auto at = [&](int row, int el) -> int { return row * EP + el + (el / R) * CP; };

  struct Tile {
    int64_t base;
    int32_t rows;
    int32_t vec;
  };
  auto geometry = [&](int64_t t) -> Tile {
    Tile g;
    const int64_t o = t / tiles;
    const int64_t p0 = (t - o * tiles) * BATCH_TS;
    g.base = o * ostride + p0;
    const int64_t rv = phys - p0;
    g.rows = rv < BATCH_TS ? static_cast<int>(rv) : BATCH_TS;
    g.vec = (jvec && g.rows == BATCH_TS) ? 1 : 0;
    return g;
  };

  TIn pf[PF];
  auto fetch = [&](const Tile& g) {
    const TIn* __restrict__ src = in + g.base;
    if (g.vec) {
#pragma unroll
      for (int i = 0; i < PFV; ++i) {
        const int idx = tid + i * BATCH_NT;
        const int vr = idx % VPR;
        const int el = CLAMP_V ? min(idx / VPR, n - 1) : idx / VPR;
        const TIn* p = src + static_cast<int64_t>(el) * estride + vr * VIN;
        if constexpr (sizeof(TIn) == 8) {
          const nt_dbl2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_dbl2*>(p));
          pf[2 * i] = v.x;
          pf[2 * i + 1] = v.y;
        } else {
          const nt_flt4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_flt4*>(p));
          pf[4 * i] = v.x;
          pf[4 * i + 1] = v.y;
          pf[4 * i + 2] = v.z;
          pf[4 * i + 3] = v.w;
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < PFS; ++i) {
        const int idx = tid + i * BATCH_NT;
        const int r = idx % BATCH_TS, el = idx / BATCH_TS;
        if (el < n && r < g.rows) pf[i] = src[r + static_cast<int64_t>(el) * estride];
      }
    }
  };
  auto deposit = [&](const Tile& g) {
    if (g.vec) {
#pragma unroll
      for (int i = 0; i < PFV; ++i) {
        const int idx = tid + i * BATCH_NT;
        const int vr = idx % VPR;
        const int el = CLAMP_V ? min(idx / VPR, n - 1) : idx / VPR;
#pragma unroll
        for (int c = 0; c < VIN; ++c) lds[at(vr * VIN + c, el)] = static_cast<double>(pf[VIN * i + c]);
      }
    } else {
#pragma unroll
      for (int i = 0; i < PFS; ++i) {
        const int idx = tid + i * BATCH_NT;
        const int r = idx % BATCH_TS, el = idx / BATCH_TS;
        if (el < n && r < g.rows) lds[at(r, el)] = static_cast<double>(pf[i]);
      }
    }
  };
  auto solve_tile = [&](const Tile& g) {
    if constexpr (R <= NT_BATCH_SEG_MAX_R(PREC)) {
      // lanes 0-31 line `wave`, lanes 32-63 line `wave + 8`; a segment past the
      // tile's last line computes on zeros, so every value a shuffle brings
      // across the segment boundary is finite
      constexpr int RS = 2 * R;
      const int sl = lane & 31;
      const int ql = wave + (lane >> 5) * BATCH_WAVES;
      const bool active = ql < g.rows;           // per segment
      const int e0 = sl * RS;
      double f[RS];
#pragma unroll
      for (int k = 0; k < RS; ++k) f[k] = (active && (FULL || e0 + k < n)) ? lds[at(ql, e0 + k)] : 0.0;
      double bnd[6];
      if constexpr (KIND != K_SOLVE) {
#pragma unroll
        for (int i = 0; i < 6; ++i) bnd[i] = active ? lds[at(ql, (i < 3) ? i : n + 2 - i)] : 0.0;
      }
      double x[RS];
      batch_solve<RS, FULL, PREC, KIND, 32>(f, bnd, e0, n, s1, s2, P, sl, sweeps, x);
      wave_lds_sync();                           // the solution overwrites the lines just read
      if (active) {
#pragma unroll
        for (int k = 0; k < RS; ++k)
          if (FULL || e0 + k < n) lds[at(ql, e0 + k)] = x[k];
      }
    } else {
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      const int ql = wave + h * BATCH_WAVES;     // tile-local line of this wavefront
      const bool active = ql < g.rows;           // uniform across the wavefront
      double x[R];
      if (active) {
        double f[R];
#pragma unroll
        for (int k = 0; k < R; ++k) f[k] = (FULL || my_e0 + k < n) ? lds[at(ql, my_e0 + k)] : 0.0;
        double bnd[6];
        if constexpr (KIND != K_SOLVE) {
#pragma unroll
          for (int i = 0; i < 6; ++i) bnd[i] = lds[at(ql, (i < 3) ? i : n + 2 - i)];
        }
        batch_solve<R, FULL, PREC, KIND>(f, bnd, my_e0, n, s1, s2, P, lane, sweeps, x);
      }
      wave_lds_sync();                           // the solution overwrites the line just read
      if (active) {
#pragma unroll
        for (int k = 0; k < R; ++k)
          if (FULL || my_e0 + k < n) lds[at(ql, my_e0 + k)] = x[k];
      }
    }
    }
  };
  auto store = [&](const Tile& g) {
    double* __restrict__ dst = out + g.base;
    if (g.vec) {
#pragma unroll
      for (int i = 0; i < PFO; ++i) {
        const int idx = tid + i * BATCH_NT;
        const int vr = idx % VPO, el = idx / VPO;
        if (FULL || el < n) {
          nt_dbl2 v;
          v.x = lds[at(vr * 2, el)];
          v.y = lds[at(vr * 2 + 1, el)];
          __builtin_nontemporal_store(v, reinterpret_cast<nt_dbl2*>(dst + static_cast<int64_t>(el) * estride + vr * 2));
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < PFS; ++i) {
        const int idx = tid + i * BATCH_NT;
        const int r = idx % BATCH_TS, el = idx / BATCH_TS;
        if (el < n && r < g.rows)
          __builtin_nontemporal_store(lds[at(r, el)], dst + r + static_cast<int64_t>(el) * estride);
      }
    }
  };

  int64_t t = t0;
  Tile g = geometry(t);
  fetch(g);
  for (;;) {
    tid = (static_cast<int>(threadIdx.x) + opaque_vzero()) & (BATCH_NT - 1);   // the mask keeps its range known
    __syncthreads();                             // the LDS is free (previous tile stored, previous chunk done)
    deposit(g);
    const bool more = t + 1 < t1;                // uniform
    Tile gn = g;
    if (more) {
      gn = geometry(t + 1);
      fetch(gn);                                 // the next tile's loads fly during this one
    }
    __syncthreads();
    solve_tile(g);
    __syncthreads();
    // the store offsets from a fresh opaque index: not computed before the
    // solve and held through it
    tid = (static_cast<int>(threadIdx.x) + opaque_vzero()) & (BATCH_NT - 1);
    store(g);
    if (!more) break;
    ++t;
    g = gn;
  }
}

template <int CLS>
__device__ __forceinline__ void run_class(CJob& J, const NtPlanK& P, int64_t t0, int64_t t1, double* lds) {
  constexpr int KIND = CLS % 3;
  constexpr int PREC = (CLS / 3) % 2;
  constexpr bool FULL = ((CLS / 6) % 2) != 0;
  constexpr int R = 1 << ((CLS / 12) % 4);
  constexpr int LAYOUT = CLS / 48;
  if constexpr (LAYOUT == CONTIG) {
    if constexpr (R <= 4) contig_chunk_seg<R, FULL, PREC, KIND>(J, P, t0, t1, lds);
    else contig_chunk<R, FULL, PREC, KIND>(J, P, t0, t1, lds);
  }
  else strided_chunk<R, FULL, PREC, KIND>(J, P, t0, t1, lds);
}

// class `cls` (uniform) of [LO, HI): a tree of scalar branches
template <int LO, int HI>
__device__ __forceinline__ void dispatch_class(int cls, CJob& J, const NtPlanK& P, int64_t t0, int64_t t1,
                                               double* lds) {
  if constexpr (HI - LO == 1) {
    run_class<LO>(J, P, t0, t1, lds);
  } else {
    constexpr int MID = (LO + HI) / 2;
    if (cls < MID) dispatch_class<LO, MID>(cls, J, P, t0, t1, lds);
    else dispatch_class<MID, HI>(cls, J, P, t0, t1, lds);
  }
}

// Persistent: each workgroup takes chunks from the work counter until none is
// left.  The next chunk's index is requested when a chunk starts and read at
// its end (one workgroup barrier per chunk), so the counter's round trip is
// hidden.  Chunks are handed out in increasing order, so the job of a chunk is
// found by a binary search that starts at the previous chunk's job.
#ifndef NT_BATCH_WAVES_PER_EU
#define NT_BATCH_WAVES_PER_EU 4
#endif
__global__ void __launch_bounds__(BATCH_NT) __attribute__((amdgpu_waves_per_eu(NT_BATCH_WAVES_PER_EU)))
batch_kernel(const BatchArgs args) {
  __shared__ __attribute__((aligned(16))) double lds[BATCH_LDS];
  __shared__ unsigned int s_chunk[2];          // chunk indices, alternating slots
  CJob* jobs = (CJob*)(args.jobs);
  CPlan* plans = (CPlan*)(args.plans);
  if (threadIdx.x == 0) s_chunk[0] = atomicAdd(args.counter, 1u);
  __syncthreads();
  int slot = 0;
  int j = 0;
  for (;;) {
    const int64_t c = static_cast<int64_t>(static_cast<unsigned>(uni(static_cast<int>(s_chunk[slot]))));
    if (c >= args.total) break;                // every wavefront reaches this
    unsigned int nxt = 0;
    if (threadIdx.x == 0) nxt = atomicAdd(args.counter, 1u);
    {                                            // the last job with chunk0 <= c, at or after j
      int lo = j, hi = args.njobs - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].chunk0 <= c) lo = mid;
        else hi = mid - 1;
      }
      j = lo;
    }
    CJob& J = jobs[j];
    const int64_t t0 = (c - J.chunk0) * J.chunk_tiles;
    const int64_t t1e = t0 + J.chunk_tiles;
    const int64_t t1 = t1e < J.ntiles ? t1e : J.ntiles;
    const NtPlanK& P = *(const NtPlanK*)(&plans[J.plan]);
    dispatch_class<0, NUM_CLASSES>(J.cls, J, P, t0, t1, lds);
    if (threadIdx.x == 0) s_chunk[slot ^ 1] = nxt;
    __syncthreads();
    slot ^= 1;
  }
}

}  // namespace nt

// ---- the allocation-free batch (nt_batch_run_noalloc) ---------------------------
// Its job image and work counter live in `scratch`, a file-scope device array
// of fixed size: each GPU the library runs on holds its own copy until the
// process ends.  In host code `scratch` names a host-side stand-in, never the
// device memory: kernels get the per-device address hipGetSymbolAddress gives.
// Layout: bytes 0..255 the work counter (and padding), then the image.
constexpr size_t NOALLOC_HEAD = 256;
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_HEAD + NT_BATCH_NOALLOC_IMAGE_MAX;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace nt {

// One chunk of the host-packed image, carried to the device by value in the
// kernel arguments of a fill launch.
constexpr size_t FILL_BYTES = 2048;
struct FillChunk {
  unsigned int w[FILL_BYTES / 4];
};

__global__ void __launch_bounds__(256) scratch_fill_kernel(unsigned int* dst, const FillChunk c, int nwords,
                                                           unsigned int* counter) {
  for (int i = threadIdx.x; i < nwords; i += 256) dst[i] = c.w[i];
  if (counter != nullptr && threadIdx.x == 0) *counter = 0u;
}

}  // namespace nt

namespace {

size_t image_bytes(int njobs, int nplans) {
  const size_t jb = sizeof(nt::BatchJob) * static_cast<size_t>(njobs);
  const size_t jb_al = (jb + 255) & ~static_cast<size_t>(255);
  return jb_al + sizeof(NtPlanK) * static_cast<size_t>(nplans);
}

int launch_batch(const void* dev_image, int njobs, int nplans, int64_t total_tiles, unsigned int* dev_counter,
                 int grid, hipStream_t stream, const char* who) {
  int g = grid;
  if (g <= 0) {
    // persistent: as many workgroups as are resident at once
    int dev = 0, cus = 256, per_cu = 2;
    if (hipGetDevice(&dev) == hipSuccess) {
      hipDeviceProp_t prop;
      if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
    }
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, nt::batch_kernel, nt::BATCH_NT, 0) == hipSuccess && nb > 0)
      per_cu = nb;
    g = per_cu * cus;
  }
  if (g > total_tiles) g = static_cast<int>(total_tiles);
  // every chunk and every workgroup's last request take one counter increment
  if (total_tiles + g >= 0xffffffffLL) return nt::set_error(NT_EINVAL, "%s: too many chunks", who);
  nt::BatchArgs a;
  a.jobs = static_cast<const nt::BatchJob*>(dev_image);
  const size_t jb = sizeof(nt::BatchJob) * static_cast<size_t>(njobs);
  a.plans = reinterpret_cast<const NtPlanK*>(static_cast<const char*>(dev_image) +
                                             ((jb + 255) & ~static_cast<size_t>(255)));
  a.counter = dev_counter;
  a.total = total_tiles;
  a.njobs = njobs;
  a.pad_ = 0;
  hipLaunchKernelGGL(nt::batch_kernel, dim3(g), dim3(nt::BATCH_NT), 0, stream, a);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return nt::set_error(NT_EHIP, "%s: %s", who, hipGetErrorString(e));
  return nt::set_error(NT_OK, "%s", "");
}

// Host staging of the allocation-free path's image and the per-device address
// of `scratch`, both guarded by one mutex (calls from several host threads
// serialize their packing and enqueueing).
std::mutex g_noalloc_mu;
constexpr int NOALLOC_MAX_DEVICES = 64;
unsigned char* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};
alignas(256) unsigned char g_noalloc_host[NT_BATCH_NOALLOC_IMAGE_MAX];

// The address of this device's `scratch`, looked up on its first use here with
// that device current (caller holds g_noalloc_mu).
int scratch_address(unsigned char** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return nt::set_error(NT_EHIP, "nt_batch_run_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return nt::set_error(NT_EUNSUPPORTED, "nt_batch_run_noalloc: device %d beyond %d", dev, NOALLOC_MAX_DEVICES);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return nt::set_error(NT_EHIP, "nt_batch_run_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_dev[dev];
  return NT_OK;
}

}  // namespace

extern "C" {

size_t nt_batch_image_bytes(int njobs, int nplans) {
  if (njobs < 0 || nplans < 0) return 0;
  return image_bytes(njobs, nplans);
}

int nt_batch_pack(const nt_batch_job* jobs, int njobs, const nt_plan* const* plans, int nplans,
                  void* host_image, size_t bytes, int64_t* total_tiles) {
  if (!jobs || !plans || !host_image || !total_tiles || njobs <= 0 || nplans <= 0)
    return nt::set_error(NT_EINVAL, "%s", "nt_batch_pack: null or empty argument");
  if (bytes < image_bytes(njobs, nplans))
    return nt::set_error(NT_EINVAL, "nt_batch_pack: image needs %zu bytes", image_bytes(njobs, nplans));
  std::memset(host_image, 0, image_bytes(njobs, nplans));
  auto* J = static_cast<nt::BatchJob*>(host_image);
  const size_t jb = sizeof(nt::BatchJob) * static_cast<size_t>(njobs);
  auto* Pk = reinterpret_cast<NtPlanK*>(static_cast<char*>(host_image) + ((jb + 255) & ~static_cast<size_t>(255)));
  for (int i = 0; i < nplans; ++i) {
    if (!plans[i]) return nt::set_error(NT_EINVAL, "nt_batch_pack: plan %d is null", i);
    Pk[i] = plans[i]->k;
  }
  double total_pts = 0.0;                       // (a double: extents are validated below)
  for (int i = 0; i < njobs; ++i)
    if (jobs[i].nz > 0 && jobs[i].ny > 0 && jobs[i].nx > 0)
      total_pts += static_cast<double>(jobs[i].nz) * static_cast<double>(jobs[i].ny) * static_cast<double>(jobs[i].nx);
  const double want = total_pts / static_cast<double>(nt::BATCH_CHUNKS_PER_SLOT * 512);
  int64_t chunk_pts = (want >= static_cast<double>(nt::BATCH_CHUNK_MAX_PTS)) ? nt::BATCH_CHUNK_MAX_PTS
                                                                                : static_cast<int64_t>(want);
  if (chunk_pts > nt::BATCH_CHUNK_MAX_PTS) chunk_pts = nt::BATCH_CHUNK_MAX_PTS;
  if (chunk_pts < nt::BATCH_CHUNK_MIN_PTS) chunk_pts = nt::BATCH_CHUNK_MIN_PTS;
  int64_t chunk0 = 0;
  for (int i = 0; i < njobs; ++i) {
    const nt_batch_job& q = jobs[i];
    if (!q.in || !q.out) return nt::set_error(NT_EINVAL, "nt_batch_pack: job %d: null buffer", i);
    if (q.plan_slot < 0 || q.plan_slot >= nplans)
      return nt::set_error(NT_EINVAL, "nt_batch_pack: job %d: plan slot %d", i, q.plan_slot);
    if (q.axis < 0 || q.axis > 2) return nt::set_error(NT_EINVAL, "nt_batch_pack: job %d: axis", i);
    if (q.nz <= 0 || q.ny <= 0 || q.nx <= 0) return nt::set_error(NT_EINVAL, "nt_batch_pack: job %d: extents", i);
    if (q.producer != 0 && q.producer != 1)
      return nt::set_error(NT_EINVAL, "nt_batch_pack: job %d: producer must be 0 (solve) or 1 (compact)", i);
    if (q.in_dtype != NT_F64 && q.in_dtype != NT_F32)
      return nt::set_error(NT_EINVAL, "nt_batch_pack: job %d: in_dtype", i);
    if (q.producer == 0 && (q.in != q.out || q.in_dtype != NT_F64))
      return nt::set_error(NT_EINVAL, "nt_batch_pack: job %d: solves are in place on FP64 fields", i);
    if (q.producer == 1 && q.in == q.out)
      return nt::set_error(NT_EINVAL, "nt_batch_pack: job %d: derivative input and output alias", i);
    if (q.precision != NT_PREC_F64 && q.precision != NT_PREC_F32_REFINE)
      return nt::set_error(NT_EINVAL, "nt_batch_pack: job %d: precision", i);
    if (q.sweeps < 0 || q.sweeps > 8) return nt::set_error(NT_EINVAL, "nt_batch_pack: job %d: sweeps", i);
    const int64_t n = (q.axis == 2) ? q.nx : (q.axis == 1 ? q.ny : q.nz);
    if (n != plans[q.plan_slot]->n)
      return nt::set_error(NT_EINVAL, "nt_batch_pack: job %d: line length %lld != plan n", i, (long long)n);
    // 64 lanes x 8 rows: the solve needs at least 3 reductions (n >= 16)
    if (n > nt::BATCH_MAXN || n < 16)
      return nt::set_error(NT_EUNSUPPORTED, "nt_batch_pack: job %d: batched lines need 16 <= n <= 512", i);
    if (q.producer == 1 && !(std::isfinite(q.dx) && q.dx != 0.0))
      return nt::set_error(NT_EINVAL, "nt_batch_pack: job %d: dx", i);
    nt::BatchJob& b = J[i];
    b.in = q.in;
    b.out = q.out;
    int64_t outer;
    int layout;
    if (q.axis == 2) {
      layout = nt::CONTIG; outer = 1; b.phys = q.nz * q.ny; b.pstride = q.nx; b.estride = 1; b.ostride = 0;
    } else if (q.axis == 1) {
      layout = nt::STRIDED; outer = q.nz; b.phys = q.nx; b.pstride = 1; b.estride = q.nx; b.ostride = q.ny * q.nx;
    } else {
      layout = nt::STRIDED; outer = 1; b.phys = q.ny * q.nx; b.pstride = 1; b.estride = q.ny * q.nx; b.ostride = 0;
    }
    b.tiles = (b.phys + nt::BATCH_TS - 1) / nt::BATCH_TS;
    b.ntiles = outer * b.tiles;
    int64_t ct = chunk_pts / (nt::BATCH_TS * n);
    if (ct < 1) ct = 1;
    b.chunk_tiles = ct;
    b.chunk0 = chunk0;
    chunk0 += (b.ntiles + ct - 1) / ct;
    b.s1 = (q.producer == 1) ? 3.0 / (4.0 * q.dx) : 0.0;
    b.s2 = (q.producer == 1) ? 1.0 / (2.0 * q.dx) : 0.0;
    b.plan = q.plan_slot;
    const int kind = (q.producer == 0) ? nt::K_SOLVE : ((q.in_dtype == NT_F32) ? nt::K_CD32 : nt::K_CD64);
    const int rsel = (n <= 64) ? 0 : (n <= 128) ? 1 : (n <= 256) ? 2 : 3;
    const int full = (n == (int64_t(64) << rsel)) ? 1 : 0;
    const int prec = (q.precision == NT_PREC_F64) ? nt::PREC_F64 : nt::PREC_F32R;
    b.cls = nt::batch_class(layout, rsel, full, prec, kind);
    b.sweeps = q.sweeps;
    // 16-byte accesses: aligned buffers, and every line (contiguous) or tile
    // row (strided) starting on a 16-byte boundary
    const int64_t isz = (q.in_dtype == NT_F32) ? 4 : 8;
    const bool al = reinterpret_cast<uintptr_t>(q.in) % 16 == 0 && reinterpret_cast<uintptr_t>(q.out) % 16 == 0;
    if (layout == nt::CONTIG)
      b.vec = al && (b.pstride * isz) % 16 == 0 && (b.pstride * 8) % 16 == 0;
    else
      b.vec = al && (b.estride * isz) % 16 == 0 && (b.ostride * isz) % 16 == 0 && (b.estride * 8) % 16 == 0 &&
              (b.ostride * 8) % 16 == 0;
  }
  *total_tiles = chunk0;
  return nt::set_error(NT_OK, "%s", "");
}

int nt_batch_launch(const void* dev_image, int njobs, int nplans, int64_t total_tiles,
                    unsigned int* dev_counter, int grid, void* stream) {
  if (!dev_image || !dev_counter || njobs <= 0 || nplans <= 0)
    return nt::set_error(NT_EINVAL, "%s", "nt_batch_launch: null or empty argument");
  if (total_tiles <= 0) return nt::set_error(NT_OK, "%s", "");
  if (total_tiles >= 0xffffffffLL) return nt::set_error(NT_EINVAL, "%s", "nt_batch_launch: too many tiles");
  if (reinterpret_cast<uintptr_t>(dev_image) % 256 != 0)
    return nt::set_error(NT_EINVAL, "%s", "nt_batch_launch: image must be 256-byte aligned");
  return launch_batch(dev_image, njobs, nplans, total_tiles, dev_counter, grid,
                      static_cast<hipStream_t>(stream), "nt_batch_launch");
}

int nt_batch_run_noalloc(const nt_batch_job* jobs, int njobs, const nt_plan* const* plans, int nplans, int grid,
                         void* stream) {
  if (!jobs || !plans || njobs <= 0 || nplans <= 0)
    return nt::set_error(NT_EINVAL, "%s", "nt_batch_run_noalloc: null or empty argument");
  const size_t need = image_bytes(njobs, nplans);
  if (need > NT_BATCH_NOALLOC_IMAGE_MAX)
    return nt::set_error(NT_EINVAL,
                         "nt_batch_run_noalloc: %d jobs and %d plans need a %zu-byte image; the static "
                         "scratch holds at most %zu bytes (use nt_batch_pack + nt_batch_launch)",
                         njobs, nplans, need, static_cast<size_t>(NT_BATCH_NOALLOC_IMAGE_MAX));
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  unsigned char* base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  int64_t total_tiles = 0;
  if (int rc = nt_batch_pack(jobs, njobs, plans, nplans, g_noalloc_host, sizeof(g_noalloc_host), &total_tiles))
    return rc;
  if (total_tiles <= 0) return nt::set_error(NT_OK, "%s", "");
  if (total_tiles >= 0xffffffffLL)
    return nt::set_error(NT_EINVAL, "%s", "nt_batch_run_noalloc: too many tiles");
  const hipStream_t s = static_cast<hipStream_t>(stream);
  unsigned int* counter = reinterpret_cast<unsigned int*>(base);
  unsigned char* image = base + NOALLOC_HEAD;
  // the image travels in the kernel arguments of fill launches, one chunk each;
  // the first also zeroes the work counter (a kernel, not a memset)
  for (size_t off = 0; off < need; off += nt::FILL_BYTES) {
    const size_t len = (need - off < nt::FILL_BYTES) ? need - off : nt::FILL_BYTES;
    nt::FillChunk c;
    std::memcpy(c.w, g_noalloc_host + off, len);
    hipLaunchKernelGGL(nt::scratch_fill_kernel, dim3(1), dim3(256), 0, s,
                       reinterpret_cast<unsigned int*>(image + off), c, static_cast<int>(len / 4),
                       off == 0 ? counter : nullptr);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return nt::set_error(NT_EHIP, "nt_batch_run_noalloc: fill: %s", hipGetErrorString(e));
  }
  return launch_batch(image, njobs, nplans, total_tiles, counter, grid, s, "nt_batch_run_noalloc");
}

}  // extern "C"
