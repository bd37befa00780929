// Partitioned-solve plans and launches (C ABI part; see nt_part_kernel.h).
//
// Plan time (host, FP64, once): the three subdomain level tables, the
// secondary solutions xU/xL of parallelcompact/misc.py:21-46 (here by a Thomas
// sweep on the host instead of SciPy), the reduced system of
// parallelcompact/reduced.py:163-178 and its inverse.  Everything is
// copied to one device allocation owned by the plan; the reference instead
// recomputes the secondary solutions and reallocates buffers on every call
// (run-parallel-compact.py:82, reduced.py:136-145).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <new>
#include <vector>

#include "../../include/ntoeplitz.h"
#include "nt_error.h"
#include "nt_part_kernel.h"
#include "nt_plan.h"

struct nt_part_plan {
  int64_t n;
  int P, m;
  double coeffs[7];
  std::vector<double> xu, xl;    // host copies [3][m]
  nt::NtPartDev host;            // host image of the device header
  void* dev;                     // device allocation: header + xU + xL
  int device;
  size_t bytes;
};

namespace {

// Thomas solve of a tridiagonal system (a sub, b diag, c super), no pivoting.
bool thomas(int n, const double* a, const double* b, const double* c, const double* d, double* x) {
  std::vector<double> cp(n), dp(n);
  if (b[0] == 0.0) return false;
  cp[0] = c[0] / b[0];
  dp[0] = d[0] / b[0];
  for (int i = 1; i < n; ++i) {
    const double den = b[i] - a[i] * cp[i - 1];
    if (den == 0.0 || !std::isfinite(den)) return false;
    cp[i] = c[i] / den;
    dp[i] = (d[i] - a[i] * dp[i - 1]) / den;
  }
  x[n - 1] = dp[n - 1];
  for (int i = n - 2; i >= 0; --i) x[i] = dp[i] - cp[i] * x[i + 1];
  return true;
}

}  // namespace

namespace {

// Segment of W lanes x R rows per subdomain (W R >= m).  Many subdomains per
// line (P >= 12) or short ones (m <= 64): 16-lane segments (four subdomains per
// wavefront, R <= 8) up to 128 points, 32 lanes up to 256, fewer wavefronts per
// line.  Otherwise one 64-lane wavefront per subdomain: in the one-line kernel
// the first and last subdomains have a wavefront each, so at small P packed
// segments would sit idle.  Measured on MI355X: P = 16, m = 128: 0.296 ms
// (segments) against 0.400 ms on 128^2 lines; P = 4, m = 128: 0.132 ms against
// 0.104 ms.
void pick_segments(int P, int m, int* R, int* W) {
  *R = 1;
  *W = 64;
  if ((P >= 12 || m <= 64) && m <= 32 * 8) *W = (m <= 16 * 8) ? 16 : 32;
  while (*W * *R < m) *R *= 2;
}

}  // namespace

extern "C" {

int nt_part_plan_create(int64_t n, int P, const double coeffs[7], nt_part_plan** out) {
  if (!out || !coeffs) return nt::set_error(NT_EINVAL, "%s", "nt_part_plan_create: null argument");
  *out = nullptr;
  if (P < 1 || P > NT_PART_MAX_P) return nt::set_error(NT_EINVAL, "%s", "nt_part_plan_create: P must be in [1, 16]");
  if (n < 2 || n > NT_PART_MAX_N || n % P != 0)
    return nt::set_error(NT_EINVAL, "%s", "nt_part_plan_create: need 2 <= n <= 2048 and n divisible by P");
  const int m = static_cast<int>(n / P);
  if (m < 4 || m > 64 * 32)
    return nt::set_error(NT_EINVAL, "%s", "nt_part_plan_create: subdomains need 4 <= m <= 2048 points");
  nt_part_plan* pp = new (std::nothrow) nt_part_plan;
  if (!pp) return nt::set_error(NT_EINVAL, "%s", "nt_part_plan_create: out of host memory");
  pp->n = n;
  pp->P = P;
  pp->m = m;
  pp->dev = nullptr;
  std::memcpy(pp->coeffs, coeffs, sizeof(pp->coeffs));
  std::memset(&pp->host, 0, sizeof(pp->host));
  const double b1 = coeffs[0], c1 = coeffs[1], ai = coeffs[2], bi = coeffs[3], ci = coeffs[4],
               an = coeffs[5], bn = coeffs[6];
  // subdomain kinds: 0 first (keeps b1, c1), 1 interior (principal block), 2 last (keeps an, bn)
  double kinds[3][7] = {{b1, c1, ai, bi, ci, ai, bi},
                        {bi, ci, ai, bi, ci, ai, bi},
                        {bi, ci, ai, bi, ci, an, bn}};
  if (P == 1) std::memcpy(kinds[0], coeffs, sizeof(kinds[0]));
  pp->xu.assign(3 * m, 0.0);
  pp->xl.assign(3 * m, 0.0);
  char msg[256];
  for (int k = 0; k < 3; ++k) {
    nt_plan tmp;
    const int rc = nt_plan_build(&tmp, m, kinds[k], msg, sizeof(msg));
    if (rc != 0) {
      delete pp;
      char buf[400];
      std::snprintf(buf, sizeof(buf), "nt_part_plan_create: subdomain kind %d: %s", k, msg);
      return nt::set_error(rc == -2 ? NT_ESINGULAR : NT_EINVAL, "%s", buf);
    }
    pp->host.plan[k] = tmp.k.d;
    pp->host.nlev = tmp.k.nlev;
    // secondary solutions (misc.py:21-46): rU = -a[0] e_0, rL = -c[-1] e_{m-1}
    std::vector<double> a(m, ai), b(m, bi), c(m, ci), rU(m, 0.0), rL(m, 0.0);
    a[0] = (k == 0) ? 0.0 : ai;
    b[0] = kinds[k][0];
    c[0] = kinds[k][1];
    a[m - 1] = kinds[k][5];
    b[m - 1] = kinds[k][6];
    c[m - 1] = (k == 2 || P == 1) ? 0.0 : ci;
    if (P == 1) a[0] = 0.0;
    rU[0] = -a[0];
    rL[m - 1] = -c[m - 1];
    if (!thomas(m, a.data(), b.data(), c.data(), rU.data(), &pp->xu[k * m]) ||
        !thomas(m, a.data(), b.data(), c.data(), rL.data(), &pp->xl[k * m])) {
      delete pp;
      return nt::set_error(NT_ESINGULAR, "%s", "nt_part_plan_create: zero pivot in a secondary solve");
    }
  }
  // reduced system (reduced.py:163-178) from [x[0], x[-1]] of every subdomain
  const int N2 = 2 * P;
  std::vector<double> ra(N2, 0.0), rb(N2, 0.0), rc(N2, 0.0);
  for (int p = 0; p < P; ++p) {
    const int k = (P == 1) ? 0 : (p == 0 ? 0 : (p == P - 1 ? 2 : 1));
    const double xu0 = pp->xu[k * m], xum = pp->xu[k * m + m - 1];
    const double xl0 = pp->xl[k * m], xlm = pp->xl[k * m + m - 1];
    ra[2 * p] = -1.0;
    rb[2 * p] = xu0;
    rc[2 * p] = xl0;
    ra[2 * p + 1] = xum;
    rb[2 * p + 1] = xlm;
    rc[2 * p + 1] = -1.0;
  }
  ra[0] = rc[0] = 0.0;
  rb[0] = 1.0;
  ra[N2 - 1] = rc[N2 - 1] = 0.0;
  rb[N2 - 1] = 1.0;
  ra[1] = 0.0;
  rc[N2 - 2] = 0.0;
  // its inverse, column by column with the reference's Thomas sweep (reduced.py:11-49)
  std::vector<double> ej(N2), col(N2);
  for (int j = 0; j < N2; ++j) {
    std::fill(ej.begin(), ej.end(), 0.0);
    ej[j] = 1.0;
    if (!thomas(N2, ra.data(), rb.data(), rc.data(), ej.data(), col.data())) {
      delete pp;
      return nt::set_error(NT_ESINGULAR, "%s", "nt_part_plan_create: zero pivot in the reduced system");
    }
    for (int i = 0; i < N2; ++i) pp->host.red_inv[i * N2 + j] = col[i];
  }
  pp->host.m = m;
  pp->host.P = P;
  pp->host.n = static_cast<int32_t>(n);
  pp->host.b1 = b1; pp->host.c1 = c1; pp->host.ai = ai; pp->host.bi = bi;
  pp->host.ci = ci; pp->host.an = an; pp->host.bn = bn;
  // device copy (skipped without a HIP device: the host data stays queryable,
  // solves then fail with a clear error)
  pp->bytes = sizeof(nt::NtPartDev) + 6 * static_cast<size_t>(m) * sizeof(double);
  pp->device = -1;
  int devcount = 0;
  if (hipGetDeviceCount(&devcount) == hipSuccess && devcount > 0) {
    hipError_t e = hipGetDevice(&pp->device);
    if (e == hipSuccess) e = hipMalloc(&pp->dev, pp->bytes);
    if (e == hipSuccess)
      e = hipMemcpy(pp->dev, &pp->host, sizeof(nt::NtPartDev), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      double* tail = reinterpret_cast<double*>(static_cast<char*>(pp->dev) + sizeof(nt::NtPartDev));
      e = hipMemcpy(tail, pp->xu.data(), 3 * m * sizeof(double), hipMemcpyHostToDevice);
      if (e == hipSuccess)
        e = hipMemcpy(tail + 3 * m, pp->xl.data(), 3 * m * sizeof(double), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
      if (pp->dev) (void)hipFree(pp->dev);
      delete pp;
      char buf[300];
      std::snprintf(buf, sizeof(buf), "nt_part_plan_create: %s", hipGetErrorString(e));
      return nt::set_error(NT_EHIP, "%s", buf);
    }
  } else {
    (void)hipGetLastError();
  }
  *out = pp;
  return nt::set_error(NT_OK, "%s", "");
}

int nt_part_plan_destroy(nt_part_plan* pp) {
  if (pp) {
    if (pp->dev) (void)hipFree(pp->dev);
    delete pp;
  }
  return nt::set_error(NT_OK, "%s", "");
}

int nt_part_plan_secondary(const nt_part_plan* pp, int p, double* xU, double* xL, int cap) {
  if (!pp || !xU || !xL) return nt::set_error(NT_EINVAL, "%s", "nt_part_plan_secondary: null argument");
  if (p < 0 || p >= pp->P) return nt::set_error(NT_EINVAL, "%s", "nt_part_plan_secondary: bad subdomain");
  if (cap < pp->m) return nt::set_error(NT_EINVAL, "%s", "nt_part_plan_secondary: buffer too small");
  const int k = (pp->P == 1) ? 0 : (p == 0 ? 0 : (p == pp->P - 1 ? 2 : 1));
  std::memcpy(xU, &pp->xu[k * pp->m], pp->m * sizeof(double));
  std::memcpy(xL, &pp->xl[k * pp->m], pp->m * sizeof(double));
  return nt::set_error(NT_OK, "%s", "");
}

int nt_part_solve_lines(const nt_part_plan* pp, int producer, const void* in, void* out,
                        int64_t lines, double p1, double p2, void* stream) {
  (void)p2;
  if (!pp || !in || !out) return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_lines: null argument");
  if (lines < 0) return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_lines: lines < 0");
  int cur = -1;
  if (!pp->dev) return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_lines: plan has no device copy (no HIP device at creation)");
  if (hipGetDevice(&cur) != hipSuccess || cur != pp->device)
    return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_lines: plan was created on another device");
  nt::PartArgs a;
  std::memset(&a, 0, sizeof(a));
  a.dev = static_cast<const nt::NtPartDev*>(pp->dev);
  const double* tail =
      reinterpret_cast<const double*>(static_cast<const char*>(pp->dev) + sizeof(nt::NtPartDev));
  a.xu = tail;
  a.xl = tail + 3 * pp->m;
  a.in = in;
  a.out = out;
  a.lstride = pp->n;
  int64_t grid = lines;
  if (producer == 0) {
    a.lines = lines;
  } else if (producer == 1) {
    if (!(std::isfinite(p1) && p1 != 0.0)) return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_lines: dx");
    if (in == out) return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_lines: f and out must not alias");
    a.lines = lines;
    a.s1 = 3.0 / (4.0 * p1);
    a.s2 = 1.0 / (2.0 * p1);
  } else if (producer == 2) {
    // ADI x-implicit half step: `lines` = N rows of an N x N field, rows 1..N-2 solved
    if (lines != pp->n) return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_lines: ADI needs lines == n");
    if (in == out) return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_lines: u and work must not alias");
    const double* c = pp->coeffs;
    if (c[0] != 1.0 || c[1] != 0.0 || c[5] != 0.0 || c[6] != 1.0)
      return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_lines: ADI plan needs identity end rows");
    a.lines = lines;
    a.s1 = 0.5 * p1;
    a.s2 = 1.0 - p1;
    grid = lines - 2;
  } else {
    return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_lines: producer must be 0, 1 or 2");
  }
  if (grid <= 0) return nt::set_error(NT_OK, "%s", "");
  if (grid > 0x7fffffffLL) return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_lines: too many lines");
  int R = 1, W = 64;
  pick_segments(pp->P, pp->m, &R, &W);
  const dim3 g(static_cast<unsigned>(grid)), b(64 * nt::part_waves(pp->P, W));
  hipStream_t st = static_cast<hipStream_t>(stream);
#define NT_PGO(RR, WW)                                                                   \
  if (R == (RR) && W == (WW)) {                                                          \
    if (static_cast<int>(b.x) > nt::part_max_threads<RR>())                          \
      return nt::set_error(NT_EUNSUPPORTED, "%s", "nt_part_solve_lines: too many subdomains"); \
    if (producer == 0)                                                                   \
      hipLaunchKernelGGL((nt::part_kernel<RR, WW, nt::P_IDENT>), g, b, 0, st, a);        \
    else if (producer == 1)                                                              \
      hipLaunchKernelGGL((nt::part_kernel<RR, WW, nt::P_COMPACT>), g, b, 0, st, a);      \
    else                                                                                 \
      hipLaunchKernelGGL((nt::part_kernel<RR, WW, nt::P_ADI>), g, b, 0, st, a);          \
  } else
  NT_PGO(1, 16)
  NT_PGO(2, 16)
  NT_PGO(4, 16)
  NT_PGO(8, 16)
  NT_PGO(8, 32)
  NT_PGO(1, 64)
  NT_PGO(2, 64)
  NT_PGO(4, 64)
  NT_PGO(8, 64)
  NT_PGO(16, 64)
  NT_PGO(32, 64) {
    return nt::set_error(NT_EUNSUPPORTED, "%s", "nt_part_solve_lines: subdomain too large");
  }
#undef NT_PGO
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    char buf[300];
    std::snprintf(buf, sizeof(buf), "nt_part_solve_lines: %s", hipGetErrorString(e));
    return nt::set_error(NT_EHIP, "%s", buf);
  }
  return nt::set_error(NT_OK, "%s", "");
}

int nt_part_solve_axis(const nt_part_plan* pp, int producer, const void* in, void* out, int64_t nz, int64_t ny,
                       int64_t nx, int axis, double p1, void* stream) {
  if (!pp || !in || !out) return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_axis: null argument");
  if (axis < 0 || axis > 2) return nt::set_error(NT_EINVAL, "nt_part_solve_axis: axis=%d", axis);
  if (nz <= 0 || ny <= 0 || nx <= 0 || nz > (int64_t(1) << 40) / ny / nx)
    return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_axis: bad extents");
  if (producer < 0 || producer > 2)
    return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_axis: producer must be 0, 1 or 2");
  const int64_t n = (axis == 2) ? nx : (axis == 1 ? ny : nz);
  if (n != pp->n)
    return nt::set_error(NT_EINVAL, "nt_part_solve_axis: line length %lld != plan n %lld", (long long)n,
                         (long long)pp->n);
  if (producer == 2 && !(nz == 1 && ny == pp->n && nx == pp->n && axis >= 1))
    return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_axis: ADI half steps take an n x n field (nz = 1), axis 2 or 1");
  if (producer != 0 && in == out)
    return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_axis: in and out must not alias");
  if (producer == 0 && in != out)
    return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_axis: solves are in place (in == out)");
  if (!pp->dev) return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_axis: plan has no device copy (no HIP device at creation)");
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != pp->device)
    return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_axis: plan was created on another device");
  nt::PartTileArgs a;
  std::memset(&a, 0, sizeof(a));
  a.dev = static_cast<const nt::NtPartDev*>(pp->dev);
  const double* tail =
      reinterpret_cast<const double*>(static_cast<const char*>(pp->dev) + sizeof(nt::NtPartDev));
  a.xu = tail;
  a.xl = tail + 3 * pp->m;
  a.in = in;
  a.out = out;
  int64_t outer;
  const int layout = (axis == 2) ? nt::CONTIG : nt::STRIDED;
  if (axis == 2) {
    outer = 1; a.phys = nz * ny; a.lstride = nx; a.estride = 1; a.ostride = 0;
  } else if (axis == 1) {
    outer = nz; a.phys = nx; a.lstride = 1; a.estride = nx; a.ostride = ny * nx;
  } else {
    outer = 1; a.phys = ny * nx; a.lstride = 1; a.estride = ny * nx; a.ostride = 0;
  }
  if (producer == 1) {
    if (!(std::isfinite(p1) && p1 != 0.0)) return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_axis: dx");
    a.s1 = 3.0 / (4.0 * p1);
    a.s2 = 1.0 / (2.0 * p1);
  } else if (producer == 2) {
    const double* c = pp->coeffs;
    if (c[0] != 1.0 || c[1] != 0.0 || c[5] != 0.0 || c[6] != 1.0)
      return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_axis: ADI plan needs identity end rows");
    a.s1 = 0.5 * p1;
    a.s2 = 1.0 - p1;
  }
  const int64_t nsys = a.phys - ((producer == 2) ? 2 : 0);
  if (nsys <= 0) return nt::set_error(NT_OK, "%s", "");
  int R = 1, W = 64;
  pick_segments(pp->P, pp->m, &R, &W);
  const bool full = (pp->m == W * R);
  const int CP = (R >= 2) ? 1 : 0;
  const int LR = (R == 1) ? 0 : (R == 2) ? 1 : (R == 4) ? 2 : (R == 8) ? 3 : (R == 16) ? 4 : 5;
  const int lp = static_cast<int>(pp->n + (pp->n >> LR) * CP + 2);
  auto lds_bytes = [&](int tl) { return static_cast<size_t>(tl) * (lp + 4 * pp->P) * sizeof(double); };
  // 16 lines per tile (128-byte rows) where two workgroups still share a CU, else 8,
  // else 4 (2048-point lines of 256-point subdomains)
  const int tls = (lds_bytes(16) <= 80 * 1024) ? 4 : ((lds_bytes(8) <= 160 * 1024) ? 3 : 2);
  const int TL = 1 << tls;
  if (lds_bytes(TL) > 160 * 1024)
    return nt::set_error(NT_EUNSUPPORTED, "%s", "nt_part_solve_axis: tile does not fit in LDS");
  a.tl_log2 = tls;
  a.lp = lp;
  // pairs of adjacent strided lines as 16-byte vectors: every row and slab start,
  // and both buffers, 16-byte aligned (tiles start at even lines)
  // contiguous lines: pairs of adjacent points, with n even and every line start
  // 16-byte aligned
  const bool al16 = reinterpret_cast<uintptr_t>(in) % 16 == 0 && reinterpret_cast<uintptr_t>(out) % 16 == 0 &&
                    a.ostride % 2 == 0;
  a.vec = ((layout == nt::STRIDED && al16 && a.estride % 2 == 0) ||
           (layout == nt::CONTIG && al16 && pp->n % 2 == 0 && pp->n >= 4 && a.lstride % 2 == 0))
              ? 1
              : 0;
  // paired strided ADI tiles start at line 0 (see part_tile_kernel): lines 1 .. phys - 2
  // in tiles of TL from line 0; other tiles start at the first solved line
  const bool colal = (producer == 2 && layout == nt::STRIDED && (NT_PART_VEC & 1) && a.vec);
  a.tiles = colal ? (a.phys - 2) / TL + 1 : (nsys + TL - 1) / TL;
  const int64_t grid = outer * a.tiles;
  if (grid > 0x7fffffffLL) return nt::set_error(NT_EINVAL, "%s", "nt_part_solve_axis: too many lines");
  const int S = 64 / W;
  const int ntasks = pp->P * TL;
  int waves = (ntasks + S - 1) / S;
  if (waves > 16) waves = 16;
  const dim3 g(static_cast<unsigned>(grid)), b(64 * waves);
  const size_t shm = lds_bytes(TL);
  hipStream_t st = static_cast<hipStream_t>(stream);
#define NT_PTK(RR, WW, PR, LY, FU) hipLaunchKernelGGL((nt::part_tile_kernel<RR, WW, PR, LY, FU>), g, b, shm, st, a)
#define NT_PTL(RR, WW, PR)                                                                        \
  do {                                                                                            \
    if (layout == nt::CONTIG) {                                                                   \
      if (full) NT_PTK(RR, WW, PR, nt::CONTIG, true); else NT_PTK(RR, WW, PR, nt::CONTIG, false); \
    } else {                                                                                      \
      if (full) NT_PTK(RR, WW, PR, nt::STRIDED, true); else NT_PTK(RR, WW, PR, nt::STRIDED, false); \
    }                                                                                             \
  } while (0)
#define NT_PTGO(RR, WW)                                                                          \
  if (R == (RR) && W == (WW)) {                                                                  \
    if (producer == 0) NT_PTL(RR, WW, nt::P_IDENT);                                              \
    else if (producer == 1) NT_PTL(RR, WW, nt::P_COMPACT);                                       \
    else NT_PTL(RR, WW, nt::P_ADI);                                                              \
  } else
  NT_PTGO(1, 16)
  NT_PTGO(2, 16)
  NT_PTGO(4, 16)
  NT_PTGO(8, 16)
  NT_PTGO(8, 32)
  NT_PTGO(1, 64)
  NT_PTGO(2, 64)
  NT_PTGO(4, 64)
  NT_PTGO(8, 64)
  NT_PTGO(16, 64)
  NT_PTGO(32, 64) {
    return nt::set_error(NT_EUNSUPPORTED, "%s", "nt_part_solve_axis: subdomain too large");
  }
#undef NT_PTGO
#undef NT_PTL
#undef NT_PTK
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return nt::set_error(NT_EHIP, "nt_part_solve_axis: %s", hipGetErrorString(e));
  return nt::set_error(NT_OK, "%s", "");
}

}  // extern "C"
