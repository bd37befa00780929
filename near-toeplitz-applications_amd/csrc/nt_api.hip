// C ABI of libntoeplitz.so (declared in include/ntoeplitz.h).
//
// Validation happens here, before any launch: sizes, pointers, dtypes and the
// plan/field agreement.  The reference performs none of these checks
// (compact/rhs.py:16-18 derives the extents from the launch grid).
#include <hip/hip_runtime.h>

#include <atomic>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

#include "../../include/ntoeplitz.h"
#include "nt_aux.h"
#include "nt_error.h"
#include "nt_launch.h"
#include "nt_plan.h"

namespace {
thread_local char g_err[512] = "";
}  // namespace

int nt::set_error(int code, const char* fmt, ...) {
  if (code == NT_OK) {
    g_err[0] = '\0';
    return NT_OK;
  }
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

namespace {
// launch-variant table: [layout][precision][ceil(log2 n)][n is a power of two].
// Powers of two take the kernels' compile-time level structure, other lengths
// the general path, and the two prefer different shapes.  The defaults are the
// shapes measured fastest on MI355X: compact derivatives on aligned grids at
// n = 2^k and n = 3 * 2^(k-2) (the other two extents multiples of 64), with
// the 512- and 1024-point strided slots taken from whole 512^3 / 1024^3 fields
// and the 2048-point strided slot from the ADI y half step (whose lines do not
// fit the whole-tile kernel).  FP32 + refinement on 512-point strided lines takes
// the one-line-per-wavefront whole-tile kernel: against two lines per wavefront,
// 512^3 d/dz 569 -> 531 and 585 -> 554 us in two processes, d/dy 461 -> 436 in
// the first and 483 -> 538 us in the second (these shapes vary with where the
// buffers land).
int g_tune[4][2][13][2] = {
    // contiguous lines: 0 wave-private transpose, 1 direct chunk loads, 2 workgroup tile,
    // 3 one line over two wavefronts, 4 persistent form of 0 (2048-point lines: FP64
    // 442 -> 385 us on 2^27 points; FP32 + refinement 536 -> 384 us, and 3 takes 366
    // against 428 us for 4 in one run)
    // 65..255-point lines take 0 on 16- or 32-lane segments (2^26 points, FP64: 96 points
    // 191 us against 425 us for 1, 160 points 222 against 324)
    {{{2, 2}, {2, 2}, {2, 2}, {2, 2}, {2, 2}, {2, 2}, {2, 2}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {1, 4}, {1, 1}},
     {{2, 2}, {2, 2}, {2, 2}, {2, 2}, {2, 2}, {2, 2}, {2, 2}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 3}, {0, 1}}},
    // strided lines: 0 16-line tiles of 64 lanes per line, 1 32-line tiles (8 lines above 512),
    // 2 persistent whole-tile kernel, 3 the same with two lines per wavefront (two workgroups per CU),
    // 4 as 3 with the two lines solved at once on 32-lane segments (2^25-2^26-point fields,
    // y and z: FP64 96 points 519-582 -> 331-336 us, 128: 163 -> 120-126, 160: 393-396 ->
    // 286-298, 480: 243-244 -> 216-227; FP32 + refinement 96: 673-713 -> 478-480, 256: 267-285 ->
    // 238-284, but 288-480 points slower: 2 x 8 rows per lane take 173 VGPRs),
    // 5 one lane per line, block-partitioned (FP64 solve of 49..512-point lines; shorter
    // ones take 1).  2^26-point fields, d/dz and d/dy against the previous default:
    // 50 points 543 -> 305, 537 -> 258 us; 64: 327 -> 217, 329 -> 214; 96: 336 -> 236,
    // 331 -> 199; 160: 301 -> 211, 288 -> 190; 300: 265 -> 204, 259 -> 188; 384: 241 ->
    // 182, 213 -> 181; 480: 192 -> 168, 179 -> 158; 512^3: 475-557 -> 427-443, 419-432 ->
    // 395-412.  FP32-stored fields (FP64 solve, 2^27 points): 128 points 866 -> 328,
    // 827 -> 286; 384: 581 -> 331, 541 -> 294; 512^3: 372 -> 321, 340 -> 312
    // FP32 + refinement: 5 takes the FP32 lane solve of FP32-stored compact
    // derivatives with one sweep (2^27-point fields, d/dz and d/dy against the FP32
    // tile kernels: 128 points 992 -> 376, 956 -> 358 us; 256: 645 -> 351, 616 -> 334;
    // 512: 425 -> 367, 393 -> 349; 64: 678 -> 590, 642 -> 547; 100: 1796 -> 888,
    // 1782 -> 831; 300: 931 -> 357, 916 -> 341; 450: 677 -> 320, 618 -> 317) and
    // gives everything else the shapes before it (lane_fallback in nt_launch.h)
    {{{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {5, 5}, {5, 5}, {5, 5}, {5, 5}, {2, 2}, {2, 1}, {2, 1}},
     {{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {5, 5}, {5, 5}, {5, 5}, {5, 5}, {2, 2}, {2, 1}, {1, 1}}},
    // ADI x half step (contiguous lines), variants as layout 0: at 2048 points the
    // wave-private transpose, which forms the stencil from coalesced reads of the
    // neighbouring lines (ADI 2048^2 step 58.7 -> 52.3 us; the compact derivative
    // at 2048 points stays faster with direct chunk loads, 37.4 against 46.5 us)
    {{{2, 2}, {2, 2}, {2, 2}, {2, 2}, {2, 2}, {2, 2}, {2, 2}, {1, 0}, {1, 0}, {0, 0}, {0, 0}, {1, 0}, {1, 1}},
     {{2, 2}, {2, 2}, {2, 2}, {2, 2}, {2, 2}, {2, 2}, {2, 2}, {1, 0}, {1, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 1}}},
    // ADI y half step (strided lines, with their halo lines), variants as layout 1
    {{{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {0, 3}, {3, 3}, {3, 2}, {2, 2}, {2, 1}, {2, 1}},
     {{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 3}, {3, 3}, {3, 3}, {2, 2}, {2, 1}, {1, 1}}},
};

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

int log2_bucket(int n) {
  int b = 0;
  while ((1 << b) < n && b < 12) ++b;
  return b;
}
}  // namespace

int64_t nt::device_cus() {
  // compute units per device, cached (set once per device; a race only repeats the query)
  static std::atomic<int> cus_cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) dev = 0;
  int cus = (dev < 64) ? cus_cache[dev].load(std::memory_order_relaxed) : 0;
  if (cus <= 0) {
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) {
      (void)hipGetLastError();
      cus = 256;
    }
    if (dev < 64) cus_cache[dev].store(cus, std::memory_order_relaxed);
  }
  return cus;
}

int nt::occupancy_per_cu(const void* kernel, int threads, size_t dyn_lds) {
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, dyn_lds) != hipSuccess || per < 1) {
    (void)hipGetLastError();
    per = 1;
  }
  return per;
}

int64_t nt::persistent_grid(const void* kernel, int threads) {
  return device_cus() * occupancy_per_cu(kernel, threads, 0);
}

int nt::tune_get(int layout, int prec, int n) {
  return g_tune[layout & 3][prec & 1][log2_bucket(n)][is_pow2(n) ? 1 : 0];
}
void nt::tune_set(int layout, int prec, int n, int variant) {
  g_tune[layout & 3][prec & 1][log2_bucket(n)][is_pow2(n) ? 1 : 0] = variant;
}

namespace {

#define fail nt::set_error

int ok() { return nt::set_error(NT_OK, "%s", ""); }

int hip_status(hipError_t e, const char* what) {
  if (e == hipSuccess) return ok();
  if (e == hipErrorInvalidValue)
    return fail(NT_EUNSUPPORTED, "%s: launch shape not supported (%s)", what, hipGetErrorString(e));
  return fail(NT_EHIP, "%s: %s", what, hipGetErrorString(e));
}

using nt::LineArgs;

LineArgs make_args(const nt_plan* p) {
  LineArgs a;
  std::memset(&a, 0, sizeof(a));
  a.plan = p->k;
  a.lane = p->lane.m ? &p->lane : nullptr;
  a.lane32 = p->lane32.m ? &p->lane32 : nullptr;
  return a;
}

// geometry of the lines along `axis` of a C-contiguous (nz, ny, nx) field
struct AxisGeom {
  int layout;
  int64_t n, outer, phys, pstride, estride, ostride;
};

AxisGeom axis_geom(int64_t nz, int64_t ny, int64_t nx, int axis) {
  AxisGeom g{};
  if (axis == 2) {
    g = {nt::CONTIG, nx, 1, nz * ny, nx, 1, 0};
  } else if (axis == 1) {
    g = {nt::STRIDED, ny, nz, nx, 1, nx, ny * nx};
  } else {
    g = {nt::STRIDED, nz, 1, ny * nx, 1, ny * nx, 0};
  }
  return g;
}

template <typename TIn, typename TOut, int PROD>
hipError_t dispatch(LineArgs& a, int layout, int precision, int64_t outer, hipStream_t st) {
  if (layout == nt::CONTIG) {
    if (precision == NT_PREC_F64)
      return nt::launch_line<TIn, TOut, nt::CONTIG, PROD, nt::PREC_F64>(a, outer, st);
    return nt::launch_line<TIn, TOut, nt::CONTIG, PROD, nt::PREC_F32R>(a, outer, st);
  }
  if (precision == NT_PREC_F64)
    return nt::launch_line<TIn, TOut, nt::STRIDED, PROD, nt::PREC_F64>(a, outer, st);
  return nt::launch_line<TIn, TOut, nt::STRIDED, PROD, nt::PREC_F32R>(a, outer, st);
}

int check_precision(int precision, int sweeps) {
  if (precision != NT_PREC_F64 && precision != NT_PREC_F32_REFINE)
    return fail(NT_EINVAL, "unknown precision mode %d", precision);
  if (sweeps < 0 || sweeps > 8) return fail(NT_EINVAL, "sweeps=%d out of range [0, 8]", sweeps);
  return NT_OK;
}

int check_dims(int64_t nz, int64_t ny, int64_t nx) {
  if (nz <= 0 || ny <= 0 || nx <= 0) return fail(NT_EINVAL, "field extents must be positive");
  if (nz > (int64_t(1) << 40) / ny / nx) return fail(NT_EINVAL, "field too large");
  return NT_OK;
}

}  // namespace

extern "C" {

const char* nt_last_error(void) { return g_err; }

int nt_version(void) { return 105; }

int nt_tune_set(int layout, int precision, int64_t n, int variant) {
  if (layout < 0 || layout > 3) return fail(NT_EINVAL, "nt_tune_set: layout must be 0..3");
  if (precision != NT_PREC_F64 && precision != NT_PREC_F32_REFINE)
    return fail(NT_EINVAL, "nt_tune_set: precision %d", precision);
  if (n < 2 || n > NT_MAX_N) return fail(NT_EINVAL, "nt_tune_set: n out of range");
  if (variant < 0 || variant >= nt::kNumVariants[layout])
    return fail(NT_EINVAL, "nt_tune_set: variant %d out of range [0, %d)", variant,
                nt::kNumVariants[layout]);
  nt::tune_set(layout, precision == NT_PREC_F64 ? nt::PREC_F64 : nt::PREC_F32R, static_cast<int>(n), variant);
  return ok();
}

int nt_tune_get(int layout, int precision, int64_t n, int* variant, int* nvariants) {
  if (layout < 0 || layout > 3) return fail(NT_EINVAL, "nt_tune_get: layout must be 0..3");
  if (precision != NT_PREC_F64 && precision != NT_PREC_F32_REFINE)
    return fail(NT_EINVAL, "nt_tune_get: precision %d", precision);
  if (n < 2 || n > NT_MAX_N) return fail(NT_EINVAL, "nt_tune_get: n out of range");
  if (variant)
    *variant = nt::tune_get(layout, precision == NT_PREC_F64 ? nt::PREC_F64 : nt::PREC_F32R, static_cast<int>(n));
  if (nvariants) *nvariants = nt::kNumVariants[layout];
  return ok();
}

int nt_plan_create(int64_t n, const double coeffs[7], nt_plan** out) {
  if (!out || !coeffs) return fail(NT_EINVAL, "nt_plan_create: null argument");
  *out = nullptr;
  nt_plan* p = new (std::nothrow) nt_plan;
  if (!p) return fail(NT_EINVAL, "nt_plan_create: out of host memory");
  char msg[256];
  const int rc = nt_plan_build(p, n, coeffs, msg, sizeof(msg));
  if (rc != 0) {
    delete p;
    return fail(rc == -2 ? NT_ESINGULAR : NT_EINVAL, "nt_plan_create: %s", msg);
  }
  *out = p;
  return ok();
}

int nt_plan_destroy(nt_plan* plan) {
  delete plan;
  return ok();
}

int nt_plan_info(const nt_plan* plan, int64_t* n, int* nlevels) {
  if (!plan) return fail(NT_EINVAL, "nt_plan_info: null plan");
  if (n) *n = plan->n;
  if (nlevels) *nlevels = plan->nlevels;
  return ok();
}

int nt_plan_levels(const nt_plan* plan, double* out, int cap) {
  if (!plan || !out) return fail(NT_EINVAL, "nt_plan_levels: null argument");
  if (cap < 8 * plan->nlevels) return fail(NT_EINVAL, "nt_plan_levels: need %d doubles", 8 * plan->nlevels);
  for (int l = 0; l < plan->nlevels; ++l)
    for (int k = 0; k < 8; ++k) out[8 * l + k] = plan->lev[l][k];
  return ok();
}

int nt_plan_reference_arrays(const nt_plan* plan, double* out, int cap) {
  if (!plan || !out) return fail(NT_EINVAL, "nt_plan_reference_arrays: null argument");
  const int64_t n = plan->n;
  if (n < 4 || (n & (n - 1)) != 0)
    return fail(NT_EINVAL, "nt_plan_reference_arrays: n=%lld is not a power of 2 >= 4", (long long)n);
  int L = 0;
  while ((int64_t(1) << L) < n) ++L;
  if (cap < 8 * L) return fail(NT_EINVAL, "nt_plan_reference_arrays: need %d doubles", 8 * L);
  // Array i of the reference holds the factors of reduction i+1 and the system
  // after it; the last slot of a and b holds the final last-row values.
  double* a = out;
  double* b = out + L;
  double* c = out + 2 * L;
  double* k1 = out + 3 * L;
  double* k2 = out + 4 * L;
  double* bf = out + 5 * L;
  double* k1f = out + 6 * L;
  double* k1l = out + 7 * L;
  for (int i = 0; i < 8 * L; ++i) out[i] = 0.0;
  for (int i = 0; i + 1 < plan->nlevels; ++i) {
    const double* s = plan->lev[i];       // level i: m, bF, cF, a, b, c, aL, bL
    const double* t = plan->lev[i + 1];   // after reduction i+1
    k1[i] = s[3] / s[4];
    k2[i] = s[5] / s[4];
    k1f[i] = s[3] / s[1];
    k1l[i] = s[6] / s[4];
    a[i] = t[3];
    b[i] = t[4];
    c[i] = t[5];
    bf[i] = t[1];
  }
  const double* top = plan->lev[plan->nlevels - 1];
  a[L - 1] = top[6];
  b[L - 1] = top[7];
  return ok();
}

int nt_solve(const nt_plan* plan, void* x, int dtype, int64_t nrhs, int precision, int sweeps,
             void* stream) {
  if (!plan) return fail(NT_EINVAL, "nt_solve: null plan");
  if (!x && nrhs > 0) return fail(NT_EINVAL, "nt_solve: null x");
  if (nrhs < 0) return fail(NT_EINVAL, "nt_solve: nrhs=%lld < 0", (long long)nrhs);
  if (dtype != NT_F64) return fail(NT_EUNSUPPORTED, "nt_solve: only NT_F64 systems are supported");
  if (int rc = check_precision(precision, sweeps)) return rc;
  if (plan->n > NT_MAX_N) return fail(NT_EUNSUPPORTED, "nt_solve: n=%lld > %d", (long long)plan->n, NT_MAX_N);
  if (nrhs == 0) return ok();
  LineArgs a = make_args(plan);
  a.in = x;
  a.out = x;
  a.pstride = plan->n;
  a.estride = 1;
  a.ostride = 0;
  a.phys = nrhs;
  a.sweeps = sweeps;
  return hip_status(dispatch<double, double, nt::P_IDENT>(a, nt::CONTIG, precision, 1,
                                                          static_cast<hipStream_t>(stream)),
                    "nt_solve");
}

int nt_solve_axis(const nt_plan* plan, void* x, int dtype, int64_t nz, int64_t ny, int64_t nx,
                  int axis, int precision, int sweeps, void* stream) {
  if (!plan || !x) return fail(NT_EINVAL, "nt_solve_axis: null argument");
  if (dtype != NT_F64) return fail(NT_EUNSUPPORTED, "nt_solve_axis: only NT_F64 fields are supported");
  if (axis < 0 || axis > 2) return fail(NT_EINVAL, "nt_solve_axis: axis=%d", axis);
  if (int rc = check_dims(nz, ny, nx)) return rc;
  if (int rc = check_precision(precision, sweeps)) return rc;
  const AxisGeom g = axis_geom(nz, ny, nx, axis);
  if (g.n != plan->n)
    return fail(NT_EINVAL, "nt_solve_axis: line length %lld != plan n %lld", (long long)g.n,
                (long long)plan->n);
  if (plan->n > NT_MAX_N) return fail(NT_EUNSUPPORTED, "nt_solve_axis: n > %d", NT_MAX_N);
  LineArgs a = make_args(plan);
  a.in = x;
  a.out = x;
  a.pstride = g.pstride;
  a.estride = g.estride;
  a.ostride = g.ostride;
  a.phys = g.phys;
  a.sweeps = sweeps;
  return hip_status(dispatch<double, double, nt::P_IDENT>(a, g.layout, precision, g.outer,
                                                          static_cast<hipStream_t>(stream)),
                    "nt_solve_axis");
}

int nt_compact_derivative(const nt_plan* plan, const void* f, int f_dtype, void* out,
                          int out_dtype, int64_t nz, int64_t ny, int64_t nx, int axis, double dx,
                          int precision, int sweeps, void* stream) {
  if (!plan || !f || !out) return fail(NT_EINVAL, "nt_compact_derivative: null argument");
  if (f == out) return fail(NT_EINVAL, "nt_compact_derivative: f and out must not alias");
  if (axis < 0 || axis > 2) return fail(NT_EINVAL, "nt_compact_derivative: axis=%d", axis);
  if (f_dtype != NT_F64 && f_dtype != NT_F32)
    return fail(NT_EINVAL, "nt_compact_derivative: f_dtype=%d", f_dtype);
  if (out_dtype != NT_F64)
    return fail(NT_EUNSUPPORTED, "nt_compact_derivative: out_dtype must be NT_F64");
  if (!(std::isfinite(dx) && dx != 0.0)) return fail(NT_EINVAL, "nt_compact_derivative: dx=%g", dx);
  if (int rc = check_dims(nz, ny, nx)) return rc;
  if (int rc = check_precision(precision, sweeps)) return rc;
  {
    // the kernels read one tile while storing another, so any shared byte of f
    // and out (not only equal base pointers) would corrupt the result
    const uintptr_t npts = static_cast<uintptr_t>(nz) * static_cast<uintptr_t>(ny) * static_cast<uintptr_t>(nx);
    const uintptr_t f0 = reinterpret_cast<uintptr_t>(f), o0 = reinterpret_cast<uintptr_t>(out);
    const uintptr_t f1 = f0 + npts * (f_dtype == NT_F64 ? 8u : 4u), o1 = o0 + npts * 8u;
    if (f0 < o1 && o0 < f1) return fail(NT_EINVAL, "nt_compact_derivative: f and out overlap");
  }
  const AxisGeom g = axis_geom(nz, ny, nx, axis);
  if (g.n != plan->n)
    return fail(NT_EINVAL, "nt_compact_derivative: line length %lld != plan n %lld",
                (long long)g.n, (long long)plan->n);
  if (g.n < 3) return fail(NT_EINVAL, "nt_compact_derivative: lines need >= 3 points");
  if (plan->n > NT_MAX_N) return fail(NT_EUNSUPPORTED, "nt_compact_derivative: n > %d", NT_MAX_N);
  LineArgs a = make_args(plan);
  a.in = f;
  a.out = out;
  a.pstride = g.pstride;
  a.estride = g.estride;
  a.ostride = g.ostride;
  a.phys = g.phys;
  a.s1 = 3.0 / (4.0 * dx);
  a.s2 = 1.0 / (2.0 * dx);
  a.sweeps = sweeps;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e = (f_dtype == NT_F64)
                     ? dispatch<double, double, nt::P_COMPACT>(a, g.layout, precision, g.outer, st)
                     : dispatch<float, double, nt::P_COMPACT>(a, g.layout, precision, g.outer, st);
  return hip_status(e, "nt_compact_derivative");
}

int nt_compact_rhs(const void* f, void* rhs, int64_t nz, int64_t ny, int64_t nx, double dx,
                   void* stream) {
  if (!f || !rhs) return fail(NT_EINVAL, "nt_compact_rhs: null argument");
  if (int rc = check_dims(nz, ny, nx)) return rc;
  if (nx < 3) return fail(NT_EINVAL, "nt_compact_rhs: nx=%lld < 3", (long long)nx);
  if (!(std::isfinite(dx) && dx != 0.0)) return fail(NT_EINVAL, "nt_compact_rhs: dx=%g", dx);
  return hip_status(nt::launch_compact_rhs(static_cast<const double*>(f), static_cast<double*>(rhs),
                                           nz * ny * nx, nx, dx, static_cast<hipStream_t>(stream)),
                    "nt_compact_rhs");
}

int nt_permute(const void* a, void* b, const int64_t shape[3], const int perm[3], void* stream) {
  if (!a || !b || !shape || !perm) return fail(NT_EINVAL, "nt_permute: null argument");
  if (a == b) return fail(NT_EINVAL, "nt_permute: in-place permute is not supported");
  if (int rc = check_dims(shape[0], shape[1], shape[2])) return rc;
  int seen = 0;
  for (int i = 0; i < 3; ++i) {
    if (perm[i] < 0 || perm[i] > 2) return fail(NT_EINVAL, "nt_permute: bad permutation");
    seen |= 1 << perm[i];
  }
  if (seen != 7) return fail(NT_EINVAL, "nt_permute: not a permutation");
  return hip_status(nt::launch_permute(static_cast<const double*>(a), static_cast<double*>(b), shape,
                                       perm, static_cast<hipStream_t>(stream)),
                    "nt_permute");
}

int nt_adi_rhs(const void* u, void* d, double dx, double dt, int64_t N, void* stream) {
  if (!u || !d) return fail(NT_EINVAL, "nt_adi_rhs: null argument");
  if (u == d) return fail(NT_EINVAL, "nt_adi_rhs: u and d must not alias");
  if (N < 3) return fail(NT_EINVAL, "nt_adi_rhs: N=%lld < 3", (long long)N);
  if (!(dx != 0.0 && dt != 0.0)) return fail(NT_EINVAL, "nt_adi_rhs: zero dx or dt");
  return hip_status(nt::launch_adi_rhs(static_cast<const double*>(u), static_cast<double*>(d), N, dx,
                                       dt, static_cast<hipStream_t>(stream)),
                    "nt_adi_rhs");
}

int nt_transpose(const void* a, void* b, int64_t N, void* stream) {
  if (!a || !b) return fail(NT_EINVAL, "nt_transpose: null argument");
  if (a == b) return fail(NT_EINVAL, "nt_transpose: in-place transpose is not supported");
  if (N < 1) return fail(NT_EINVAL, "nt_transpose: N=%lld", (long long)N);
  const int64_t shape[3] = {1, N, N};
  const int perm[3] = {0, 2, 1};
  return hip_status(nt::launch_permute(static_cast<const double*>(a), static_cast<double*>(b), shape,
                                       perm, static_cast<hipStream_t>(stream)),
                    "nt_transpose");
}

int nt_adi_prepare(const void* u, void* work, int64_t N, void* stream) {
  if (!u || !work) return fail(NT_EINVAL, "nt_adi_prepare: null argument");
  if (N < 3) return fail(NT_EINVAL, "nt_adi_prepare: N=%lld < 3", (long long)N);
  return hip_status(hipMemcpyAsync(work, u, sizeof(double) * N * N, hipMemcpyDeviceToDevice,
                                   static_cast<hipStream_t>(stream)),
                    "nt_adi_prepare");
}

int nt_adi_steps(const nt_plan* plan, void* u, void* work, int64_t N, double r, int nsteps,
                 int precision, int sweeps, void* stream) {
  if (!plan || !u || !work) return fail(NT_EINVAL, "nt_adi_steps: null argument");
  if (u == work) return fail(NT_EINVAL, "nt_adi_steps: u and work must not alias");
  if (N < 3 || plan->n != N)
    return fail(NT_EINVAL, "nt_adi_steps: N=%lld does not match plan n=%lld", (long long)N,
                (long long)plan->n);
  if (N > NT_MAX_N) return fail(NT_EUNSUPPORTED, "nt_adi_steps: N > %d", NT_MAX_N);
  if (nsteps < 0) return fail(NT_EINVAL, "nt_adi_steps: nsteps < 0");
  if (!std::isfinite(r)) return fail(NT_EINVAL, "nt_adi_steps: r not finite");
  if (int rc = check_precision(precision, sweeps)) return rc;
  const double* c = plan->coeffs;
  if (c[0] != 1.0 || c[1] != 0.0 || c[5] != 0.0 || c[6] != 1.0)
    return fail(NT_EINVAL, "nt_adi_steps: plan must have identity boundary rows (1, 0, ., ., ., 0, 1)");
  hipStream_t st = static_cast<hipStream_t>(stream);
  LineArgs ax = make_args(plan);   // x-implicit: rows 1..N-2 of u -> work
  ax.pstride = N;
  ax.estride = 1;
  ax.ostride = 0;
  ax.phys = N;
  ax.s1 = 0.5 * r;
  ax.s2 = 1.0 - r;
  ax.sweeps = sweeps;
  LineArgs ay = ax;                // y-implicit: columns 1..N-2 of work -> u
  ay.pstride = 1;
  ay.estride = N;
  for (int s = 0; s < nsteps; ++s) {
    ax.in = u;
    ax.out = work;
    hipError_t e = dispatch<double, double, nt::P_ADI>(ax, nt::CONTIG, precision, 1, st);
    if (e != hipSuccess) return hip_status(e, "nt_adi_steps (x half step)");
    ay.in = work;
    ay.out = u;
    e = dispatch<double, double, nt::P_ADI>(ay, nt::STRIDED, precision, 1, st);
    if (e != hipSuccess) return hip_status(e, "nt_adi_steps (y half step)");
  }
  return ok();
}

}  // extern "C"
