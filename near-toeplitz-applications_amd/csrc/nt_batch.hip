// Mixed-size persistent batch: many fields of different shapes, derivative
// directions, precisions and storage types solved by one launch.
//
// The host packs a binary image (job table + level tables of the plans the
// jobs use) into a caller-provided host buffer; the caller copies it to the
// device and zeroes a work counter; one persistent launch then pulls tiles of
// 16 lines from the counter until every job is done.  A tile is 16 lines
// adjacent in memory (16 rows of an x-line field, or 16 adjacent y/z lines),
// staged through LDS with all of a thread's loads in flight at once; each of
// the 8 wavefronts then solves two of the lines with the same register CR as
// the other kernels, 64 lanes x R rows with R = 1, 2, 4 or 8 by line length
// (up to 512 points).
// Per-line arithmetic does not depend on which other jobs share the launch, so
// a field gives the same bits alone and in any mix.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

#include "../../include/ntoeplitz.h"
#include "nt_error.h"
#include "nt_line_kernel.h"
#include "nt_plan.h"

namespace nt {

constexpr int BATCH_MAXN = 512;
constexpr int BATCH_TS = 16;                // lines per tile
constexpr int BATCH_WAVES = 8;              // wavefronts per workgroup
constexpr int BATCH_LPW = BATCH_TS / BATCH_WAVES;   // lines solved by each wavefront
constexpr int BATCH_NT = BATCH_WAVES * 64;
// LDS row of a line: point e of a line with R = 2^rsel rows per lane sits at
// e + (e >> rsel), i.e. one double of padding after each lane's chunk, so a
// lane's chunk starts R + 1 doubles after its neighbour's (an odd number of
// 8-byte bank pairs for R >= 2): the solve's chunk reads and writes do not
// collide.  A row then holds at most 64 (R + 1) doubles.
constexpr int BATCH_EP = BATCH_MAXN + BATCH_MAXN / 8 + 2;   // LDS stride between lines (doubles)
// every thread moves BATCH_TS * BATCH_MAXN / BATCH_NT elements of a tile
constexpr int BATCH_PER = BATCH_TS * BATCH_MAXN / BATCH_NT;

struct BatchJob {
  const void* in;
  void* out;
  int64_t pstride, estride, ostride;  // line geometry (as LineArgs)
  int64_t phys;                       // lines per slab
  int64_t tiles;                      // tiles per slab
  int64_t tile0;                      // first global tile of this job
  int64_t ntiles;                     // tiles of this job (slabs x tiles)
  double s1, s2;                      // compact scalars 3/(4h), 1/(2h)
  int32_t plan;                       // index into the plan table
  int32_t layout;                     // CONTIG / STRIDED
  int32_t prod;                       // P_IDENT / P_COMPACT
  int32_t in_f32;                     // input stored as FP32
  int32_t prec;                       // PREC_F64 / PREC_F32R
  int32_t sweeps;
};

struct BatchArgs {
  const BatchJob* jobs;
  const NtPlanK* plans;
  unsigned int* counter;
  int64_t total_tiles;
  int32_t njobs;
  int32_t pad_;
};

// Element k of this thread's share of a tile: (line, position) in the tile.
// Contiguous lines: wavefront w moves lines w, w + WAVES, ...; lane l takes
// positions l, l + 64, ... (each wavefront reads whole 512 B row runs).
// Strided lines: the TS lines are adjacent in memory, so TS consecutive
// threads read one TS * 8 B run of a row; thread t takes line t % TS,
// positions t / TS + k * (NT / TS).
__device__ __forceinline__ void tile_slot(int layout, int tid, int k, int& line, int& e) {
  if (layout == CONTIG) {
    constexpr int PER_LINE = BATCH_MAXN / 64;        // positions per lane per line
    line = (tid >> 6) + (k / PER_LINE) * BATCH_WAVES;
    e = (tid & 63) + (k % PER_LINE) * 64;
  } else {
    line = tid % BATCH_TS;
    e = tid / BATCH_TS + k * (BATCH_NT / BATCH_TS);
  }
}

template <typename TIn>
__device__ __forceinline__ void stage_tile(double* lds, const TIn* __restrict__ in, int layout, int tid, int rows,
                                           int n, int rsel, int64_t base, int64_t pstride, int64_t estride) {
  double v[BATCH_PER];
  // issue every load of the tile before the first LDS write so they overlap
#pragma unroll
  for (int k = 0; k < BATCH_PER; ++k) {
    int line, e;
    tile_slot(layout, tid, k, line, e);
    v[k] = (line < rows && e < n)
               ? static_cast<double>(__builtin_nontemporal_load(in + base + line * pstride + e * estride))
               : 0.0;
  }
#pragma unroll
  for (int k = 0; k < BATCH_PER; ++k) {
    int line, e;
    tile_slot(layout, tid, k, line, e);
    if (line < rows && e < n) lds[line * BATCH_EP + e + (e >> rsel)] = v[k];
  }
}

// One line held in LDS row `row` (n points, padded as described at BATCH_EP):
// lane l takes points l R .. l R + R - 1, builds the right-hand side (compact
// producer) or takes the values as they are (solve), runs the register CR and
// writes the solution back over the row.
template <int R, bool FULL>
__device__ __forceinline__ void solve_line(double* row, int n, int prod, int prec, const NtPlanK& P, double s1,
                                           double s2, int sweeps, int lane) {
  constexpr int RS = (R == 1) ? 0 : (R == 2) ? 1 : (R == 4) ? 2 : 3;
  const int my_e0 = lane * R;
  double* chunk = row + lane * (R + 1);
  double f[R];
#pragma unroll
  for (int k = 0; k < R; ++k) f[k] = (FULL || my_e0 + k < n) ? chunk[k] : 0.0;
  double d[R];
  if (prod == P_COMPACT) {
    double bnd[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const int e = (i < 3) ? i : n + 2 - i;
      bnd[i] = row[e + (e >> RS)];
    }
    const double fl = shup<64>(f[R - 1], 1);
    const double fr = shdn<64>(f[0], 1);
    compact_rhs_chunk<R, FULL>(f, fl, fr, bnd, my_e0, n, s1, s2, d);
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k) d[k] = f[k];
  }
  double x[R];
  if (prec == PREC_F64) solve_chunk<R, 64, PREC_F64, FULL>(d, x, P, lane, sweeps);
  else solve_chunk<R, 64, PREC_F32R, FULL>(d, x, P, lane, sweeps);
#pragma unroll
  for (int k = 0; k < R; ++k)
    if (FULL || my_e0 + k < n) chunk[k] = x[k];
}

__global__ void __launch_bounds__(BATCH_NT) batch_kernel(const BatchArgs args) {
  __shared__ __attribute__((aligned(16))) double lds[BATCH_TS * BATCH_EP];
  __shared__ __attribute__((aligned(16))) NtPlanK s_plan;   // level table of the current job
  __shared__ int64_t s_tile[2];     // tile indices, alternating slots
  // reset through an opaque move at every tile, so the per-thread offsets
  // derived from it are recomputed per tile, not hoisted and held across the loop
  int tid = threadIdx.x;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  // The job and plan tables are read through the constant address space, i.e.
  // with scalar loads: as vector loads they would wait, in order, for the
  // previous tile's stores before the next tile could be located and staged.
  // The tables are written before the launch and only read by the kernel.
  using CJob = const __attribute__((address_space(4))) BatchJob;
  using CPlan = const __attribute__((address_space(4))) NtPlanK;
  CJob* jobs = (CJob*)(args.jobs);
  CPlan* plans = (CPlan*)(args.plans);
  int j = 0;                        // current job: tiles come in increasing order
  int plan_in_lds = -1;
  int slot = 0;

  // One workgroup barrier per tile, at the top: every wavefront has finished the
  // previous tile (its stores have read the LDS, its solve is done with s_plan),
  // and the new tile index is visible.  The index alternates between two slots,
  // so writing the next one cannot race with late readers of this one.  Strided
  // tiles spread each line over all threads and add two barriers; contiguous
  // tiles are staged, solved and stored by each wavefront on its own lines.
  for (;;) {
    tid = (static_cast<int>(threadIdx.x) + opaque_vzero()) & (BATCH_NT - 1);   // the mask keeps its range known
    if (tid == 0) s_tile[slot] = static_cast<int64_t>(atomicAdd(args.counter, 1u));
    __syncthreads();
    // the tile index is the same for the whole workgroup: keep it (and the
    // job index derived from it) in scalar registers
    const int64_t t = static_cast<int64_t>(static_cast<unsigned>(uni(static_cast<int>(s_tile[slot]))));
    slot ^= 1;
    if (t >= args.total_tiles) break;            // every workgroup reaches this
    while (j + 1 < args.njobs && jobs[j + 1].tile0 <= t) ++j;
    const CJob& J = jobs[j];
    const int plan = J.plan;
    if (plan != plan_in_lds) {                   // copy the level table into LDS (uniform branch)
      const int* src = reinterpret_cast<const int*>(args.plans + plan);
      int* dst = reinterpret_cast<int*>(&s_plan);
      for (int i = tid; i < static_cast<int>(sizeof(NtPlanK) / 4); i += BATCH_NT) dst[i] = src[i];
      plan_in_lds = plan;
      __syncthreads();
    }
    const int64_t local = t - J.tile0;
    const int64_t o = local / J.tiles;
    const int64_t p0 = (local - o * J.tiles) * BATCH_TS;
    const int64_t base = o * J.ostride + p0 * J.pstride;
    const int64_t rows64 = J.phys - p0;
    const int rows = rows64 < BATCH_TS ? static_cast<int>(rows64) : BATCH_TS;
    const int layout = J.layout, prod = J.prod, prec = J.prec;
    const int n = plans[plan].n;
    const int64_t pstride = J.pstride, estride = J.estride;
    double* __restrict__ out = static_cast<double*>(J.out);

    // rows per lane R = 2^rsel follow the line length (64 lanes x R >= n)
    const int rsel = (n <= 64) ? 0 : (n <= 128) ? 1 : (n <= 256) ? 2 : 3;

    // ---- stage the tile (TS lines x n) through LDS, coalesced along memory ------
    if (J.in_f32)
      stage_tile(lds, static_cast<const float*>(J.in), layout, tid, rows, n, rsel, base, pstride, estride);
    else
      stage_tile(lds, static_cast<const double*>(J.in), layout, tid, rows, n, rsel, base, pstride, estride);
    if (layout == CONTIG) wave_lds_sync();       // uniform: a wavefront staged its own lines
    else __syncthreads();

    // ---- each wavefront solves its lines; a line's LDS row is read and then
    // overwritten by the same wavefront only, so no barrier in between.  The
    // compile-time level structure applies when n == 64 R. -----------------------
    const NtPlanK& P = s_plan;
    const bool full = (n == (64 << rsel));
    // the job's scalars in scalar registers: loaded inside the solve they would
    // be vector loads, and each would wait for the previous tile's stores (the
    // vector-memory counter is in order)
    const double s1 = J.s1, s2 = J.s2;
    const int sweeps = J.sweeps;
#pragma unroll 1
    for (int h = 0; h < BATCH_LPW; ++h) {
      const int ql = wave + h * BATCH_WAVES;
      if (ql >= rows) break;                     // uniform across the wavefront
      double* row = lds + ql * BATCH_EP;
      if (full) {
        switch (rsel) {
          case 0: solve_line<1, true>(row, n, prod, prec, P, s1, s2, sweeps, lane); break;
          case 1: solve_line<2, true>(row, n, prod, prec, P, s1, s2, sweeps, lane); break;
          case 2: solve_line<4, true>(row, n, prod, prec, P, s1, s2, sweeps, lane); break;
          default: solve_line<8, true>(row, n, prod, prec, P, s1, s2, sweeps, lane); break;
        }
      } else {
        switch (rsel) {
          case 0: solve_line<1, false>(row, n, prod, prec, P, s1, s2, sweeps, lane); break;
          case 1: solve_line<2, false>(row, n, prod, prec, P, s1, s2, sweeps, lane); break;
          case 2: solve_line<4, false>(row, n, prod, prec, P, s1, s2, sweeps, lane); break;
          default: solve_line<8, false>(row, n, prod, prec, P, s1, s2, sweeps, lane); break;
        }
      }
    }
    if (layout == CONTIG) wave_lds_sync();
    else __syncthreads();

    // ---- store ----------------------------------------------------------------------
#pragma unroll
    for (int k = 0; k < BATCH_PER; ++k) {
      int line, e;
      tile_slot(layout, tid, k, line, e);
      if (line < rows && e < n)
        __builtin_nontemporal_store(lds[line * BATCH_EP + e + (e >> rsel)], out + base + line * pstride + e * estride);
    }
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
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
      hipDeviceProp_t prop;
      if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
    }
    g = 4 * cus;   // several workgroups per CU keep loads and solves overlapping
  }
  if (g > total_tiles) g = static_cast<int>(total_tiles);
  nt::BatchArgs a;
  a.jobs = static_cast<const nt::BatchJob*>(dev_image);
  const size_t jb = sizeof(nt::BatchJob) * static_cast<size_t>(njobs);
  a.plans = reinterpret_cast<const NtPlanK*>(static_cast<const char*>(dev_image) +
                                             ((jb + 255) & ~static_cast<size_t>(255)));
  a.counter = dev_counter;
  a.total_tiles = total_tiles;
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
  int64_t tile0 = 0;
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
    if (q.axis == 2) {
      b.layout = nt::CONTIG; outer = 1; b.phys = q.nz * q.ny; b.pstride = q.nx; b.estride = 1; b.ostride = 0;
    } else if (q.axis == 1) {
      b.layout = nt::STRIDED; outer = q.nz; b.phys = q.nx; b.pstride = 1; b.estride = q.nx; b.ostride = q.ny * q.nx;
    } else {
      b.layout = nt::STRIDED; outer = 1; b.phys = q.ny * q.nx; b.pstride = 1; b.estride = q.ny * q.nx; b.ostride = 0;
    }
    b.tiles = (b.phys + nt::BATCH_TS - 1) / nt::BATCH_TS;
    b.ntiles = outer * b.tiles;
    b.tile0 = tile0;
    tile0 += b.ntiles;
    b.s1 = (q.producer == 1) ? 3.0 / (4.0 * q.dx) : 0.0;
    b.s2 = (q.producer == 1) ? 1.0 / (2.0 * q.dx) : 0.0;
    b.plan = q.plan_slot;
    b.prod = (q.producer == 1) ? nt::P_COMPACT : nt::P_IDENT;
    b.in_f32 = (q.in_dtype == NT_F32);
    b.prec = (q.precision == NT_PREC_F64) ? nt::PREC_F64 : nt::PREC_F32R;
    b.sweeps = q.sweeps;
  }
  *total_tiles = tile0;
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
