// Host-side harness for the plan builder (csrc/nt_plan.cpp), built with g++
// under AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_host_sanitize.py.  It builds plans for every n in [2, 4096] and
// several matrix families, feeds the builder invalid sizes, non-finite
// coefficients and singular systems, writes error messages into buffers
// allocated at exactly the length given, and checks the level table's
// structure:
//   * level 0 holds n equations and each reduction keeps floor(m / 2);
//   * the top system has 2 or 3 equations;
//   * every factor is finite and the FP32 table is the FP64 one rounded;
//   * diagonal dominance of the interior rows survives every reduction for
//     the diagonally dominant families (cyclic reduction preserves it);
//   * the lane-solve tables (49 <= n <= 512): every block response solves
//     its block equations, and the lane kernel's arithmetic restated here
//     matches a band elimination with partial pivoting of the whole matrix;
//     other n, and a zero block pivot, leave no lane tables.
// Prints one summary line and exits 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "nt_plan.h"

namespace {

int failures = 0;

void fail(const char* what, long long n) {
  if (failures < 20) std::printf("FAIL %s (n=%lld)\n", what, n);
  ++failures;
}

template <typename T>
bool level_finite(const NtLevel<T>& L) {
  const T v[] = {L.k1, L.k2, L.k1f, L.kl, L.a, L.c, L.ib, L.cf, L.ibf, L.al, L.ibl};
  for (T x : v)
    if (!std::isfinite(static_cast<double>(x))) return false;
  return true;
}

bool rounded(double d, float f) { return static_cast<float>(d) == f; }

// build with an error buffer of exactly `errlen` bytes (none for 0)
int build(nt_plan* p, long long n, const double co[7], int errlen, std::string* msg) {
  std::unique_ptr<char[]> buf(errlen > 0 ? new char[errlen]() : nullptr);   // zeroed: success writes nothing
  const int rc = nt_plan_build(p, n, co, buf.get(), errlen);
  if (msg) *msg = (errlen > 0) ? std::string(buf.get()) : std::string();
  return rc;
}

void check_plan(const nt_plan& p, long long n, bool dominant) {
  if (p.n != n || p.k.n != n) fail("n recorded", n);
  if (p.nlevels < 1 || p.nlevels > NT_MAX_LEVELS || p.k.nlev != p.nlevels - 1) fail("level count", n);
  long long m = n;
  for (int s = 0; s < p.nlevels; ++s) {
    if (static_cast<long long>(p.lev[s][0]) != m) fail("equations per level", n);
    if (p.k.d.lv[s].m != m) fail("m in the kernel table", n);
    if (!level_finite(p.k.d.lv[s]) || !level_finite(p.k.f.lv[s])) fail("finite factors", n);
    const NtLevel<double>& D = p.k.d.lv[s];
    const NtLevel<float>& F = p.k.f.lv[s];
    if (!rounded(D.k1, F.k1) || !rounded(D.k2, F.k2) || !rounded(D.ib, F.ib) || !rounded(D.ibf, F.ibf))
      fail("FP32 factors are the FP64 ones rounded", n);
    const double a = p.lev[s][3], b = p.lev[s][4], c = p.lev[s][5];
    if (dominant && m >= 4 && !(std::fabs(b) >= std::fabs(a) + std::fabs(c))) fail("diagonal dominance", n);
    if (s + 1 < p.nlevels) m /= 2;
  }
  if (m < 2 || m > 3) fail("top system size", n);
  if (p.k.d.top.m != m || p.k.f.top.m != m) fail("top m", n);
  for (int i = 0; i < 7; ++i)
    if (!std::isfinite(p.k.d.top.t[i])) fail("finite top factors", n);
}

// Tridiagonal solve by Gaussian elimination with partial pivoting (row i may
// swap with row i+1; the swap fills one extra super-diagonal, kept in f).
std::vector<double> gepp_solve(long long n, const double co[7], std::vector<double> x) {
  std::vector<double> l(n, co[2]), d(n, co[3]), u(n, co[4]), f(n, 0.0);
  d[0] = co[0];
  u[0] = co[1];
  l[n - 1] = co[5];   // l[i] couples row i+1 to column i: the last row's is an
  d[n - 1] = co[6];
  for (long long i = 0; i + 1 < n; ++i) {
    const double li = (i + 1 == n - 1) ? co[5] : co[2];
    if (std::fabs(d[i]) >= std::fabs(li)) {
      const double t = li / d[i];
      d[i + 1] -= t * u[i];
      x[i + 1] -= t * x[i];
    } else {
      const double t = d[i] / li;
      d[i] = li;
      const double dn = d[i + 1];
      d[i + 1] = u[i] - t * dn;
      if (i + 2 < n) {
        f[i] = u[i + 1];
        u[i + 1] = -t * f[i];
      }
      u[i] = dn;
      const double xi = x[i];
      x[i] = x[i + 1];
      x[i + 1] = xi - t * x[i + 1];
    }
  }
  x[n - 1] /= d[n - 1];
  if (n > 1) x[n - 2] = (x[n - 2] - u[n - 2] * x[n - 1]) / d[n - 2];
  for (long long i = n - 3; i >= 0; --i) x[i] = (x[i] - u[i] * x[i + 1] - f[i] * x[i + 2]) / d[i];
  return x;
}

// The lane kernel's arithmetic for one line: block sweeps (blocks 0..j-1 of
// m rows, the others of m - 1), the reduced system's forward and back sweeps,
// the correction.
std::vector<double> lane_kernel_host(const NtLaneTab& t, const std::vector<double>& rhs) {
  const int P = NT_LANE_P, m = t.m, NR = NT_LANE_NR;
  std::vector<double> y(rhs), red(NR);
  std::vector<int> start(P + 1, 0), len(P);
  for (int w = 0; w < P; ++w) {
    len[w] = (w < t.j) ? m : m - 1;
    start[w + 1] = start[w] + len[w];
  }
  for (int w = 0; w < P; ++w) {
    double* v = &y[start[w]];
    const int L = len[w];
    const double* ip = t.ip[w == 0 ? 0 : 1];
    const double* cp = t.cp[w == 0 ? 0 : 1];
    v[0] *= ip[0];
    for (int r = 1; r < L - 1; ++r) v[r] = (v[r] - t.ai * v[r - 1]) * ip[r];
    v[L - 1] = (v[L - 1] - (w == P - 1 ? t.an : t.ai) * v[L - 2]) * (w == P - 1 ? t.ip_last : ip[L - 1]);
    for (int r = L - 2; r >= 0; --r) v[r] -= cp[r] * v[r + 1];
    if (w >= 1) red[2 * w - 1] = -v[0];
    if (w <= P - 2) red[2 * w] = -v[L - 1];
  }
  red[0] *= t.rip[0];
  for (int i = 1; i < NR; ++i) red[i] = (red[i] - t.rsub[i] * red[i - 1]) * t.rip[i];
  for (int i = NR - 2; i >= 0; --i) red[i] -= t.rcp[i] * red[i + 1];
  for (int w = 0; w < P; ++w) {
    const double al = (w >= 1) ? red[2 * w - 1] : 0.0, be = (w <= P - 2) ? red[2 * w] : 0.0;
    const double* uh = t.uh[w == P - 1 ? 2 : (len[w] == m ? 0 : 1)];
    const double* lh = t.lh[w == 0 ? 0 : (len[w] == m ? 1 : 2)];
    for (int r = 0; r < len[w]; ++r) y[start[w] + r] += al * uh[r] + be * lh[r];
  }
  return y;
}

// residual of one block's equations for the response v: first row (bf, cf),
// interior (ai, bi, ci), last row (al, bl); rhs e0 on row 0, em on row m-1
double block_residual(int m, double bf, double cf, double ai, double bi, double ci, double al, double bl,
                      const double* v, double e0, double em) {
  double worst = 0.0;
  for (int r = 0; r < m; ++r) {
    double ax;
    if (r == 0) ax = bf * v[0] + cf * v[1];
    else if (r == m - 1) ax = al * v[m - 2] + bl * v[m - 1];
    else ax = ai * v[r - 1] + bi * v[r] + ci * v[r + 1];
    const double want = (r == 0 ? e0 : 0.0) + (r == m - 1 ? em : 0.0);
    worst = std::fmax(worst, std::fabs(ax - want));
  }
  return worst;
}

void check_lane(const nt_plan& p, long long n, const double co[7], bool expect) {
  const NtLaneTab& t = p.lane;
  const bool size_ok = (n > 3 * NT_LANE_P && n <= NT_LANE_P * NT_LANE_MAX_M);
  if (!(size_ok && expect)) {
    if (t.m != 0) fail("no lane tables expected", n);
    return;
  }
  const int m = static_cast<int>((n + NT_LANE_P - 1) / NT_LANE_P);
  if (t.m != m || t.j != n - NT_LANE_P * (m - 1) || t.j < 1 || t.j > NT_LANE_P) {
    fail("lane tables built with m = ceil(n / 16) and j long blocks", n);
    return;
  }
  const double b1 = co[0], c1 = co[1], ai = co[2], bi = co[3], ci = co[4], an = co[5], bn = co[6];
  const double scale = std::fabs(ai) + std::fabs(ci) + 1e-300;
  const int mlast = (t.j == NT_LANE_P) ? m : m - 1;
  // uh: the interior (m and m - 1 rows) and last blocks' response to -ai on
  // row 0; lh: the first and interior blocks' response to -ci on their last row
  if (block_residual(m, bi, ci, ai, bi, ci, ai, bi, t.uh[0], -ai, 0.0) > 1e-14 * scale ||
      block_residual(m - 1, bi, ci, ai, bi, ci, ai, bi, t.uh[1], -ai, 0.0) > 1e-14 * scale ||
      block_residual(mlast, bi, ci, ai, bi, ci, an, bn, t.uh[2], -ai, 0.0) > 1e-14 * scale ||
      block_residual(m, b1, c1, ai, bi, ci, ai, bi, t.lh[0], 0.0, -ci) > 1e-14 * scale ||
      block_residual(m, bi, ci, ai, bi, ci, ai, bi, t.lh[1], 0.0, -ci) > 1e-14 * scale ||
      block_residual(m - 1, bi, ci, ai, bi, ci, ai, bi, t.lh[2], 0.0, -ci) > 1e-14 * scale)
    fail("block responses solve their block equations", n);
  std::vector<double> rhs(n);
  for (long long i = 0; i < n; ++i) rhs[i] = std::cos(1.3 * i) - 0.5 * std::sin(0.37 * i * i);
  const std::vector<double> want = gepp_solve(n, co, rhs), got = lane_kernel_host(t, rhs);
  double e = 0.0, s = 0.0;
  for (long long i = 0; i < n; ++i) {
    e = std::fmax(e, std::fabs(got[i] - want[i]));
    s = std::fmax(s, std::fabs(want[i]));
  }
  if (!(e <= 1e-12 * s)) fail("lane solve equals elimination with pivoting", n);
}

}  // namespace

int main() {
  const double compact[7] = {1.0, 2.0, 0.25, 1.0, 0.25, 2.0, 1.0};
  const double general[7] = {2.0, 0.3, 0.5, 3.0, 0.7, 0.2, 1.7};
  std::vector<std::vector<double>> dominant_sets = {
      {2.0, 0.3, 0.5, 3.0, 0.7, 0.2, 1.7},
      {1.0, 0.25, 0.25, 1.0, 0.25, 0.25, 1.0}};
  for (double r : {0.1, 0.5, 2.0, 50.0}) dominant_sets.push_back({1.0, 0.0, -r, 1.0 + 2.0 * r, -r, 0.0, 1.0});
  auto plan = std::make_unique<nt_plan>();
  long long built = 0;
  // every size, every family
  for (long long n = 2; n <= 4096; ++n) {
    std::string msg;
    // the compact matrix is singular at n = 3: [[1, 2, 0], [1/4, 1, 1/4], [0, 2, 1]]
    // has determinant 1 * (1 - 1/2) - 2 * (1/4) = 0
    if (n == 3) {
      if (build(plan.get(), n, compact, 256, &msg) != -2 || msg.find("singular") == std::string::npos)
        fail("singular compact n = 3 refused", n);
    } else if (build(plan.get(), n, compact, 256, &msg) != 0) {
      fail(("compact: " + msg).c_str(), n);
    } else {
      check_plan(*plan, n, false);
      check_lane(*plan, n, compact, true);
      ++built;
    }
    for (const auto& co : dominant_sets) {
      if (build(plan.get(), n, co.data(), 256, &msg) != 0) fail(("dominant: " + msg).c_str(), n);
      else { check_plan(*plan, n, true); check_lane(*plan, n, co.data(), true); ++built; }
    }
  }
  // sizes out of range
  for (long long n : {-5LL, 0LL, 1LL, 4097LL, 1LL << 40}) {
    std::string msg;
    if (build(plan.get(), n, compact, 128, &msg) != -1 || msg.find("out of range") == std::string::npos)
      fail("size out of range refused", n);
  }
  // non-finite coefficients, each position
  for (int i = 0; i < 7; ++i) {
    for (double bad : {NAN, INFINITY, -INFINITY}) {
      double co[7];
      std::memcpy(co, general, sizeof(co));
      co[i] = bad;
      std::string msg;
      if (build(plan.get(), 64, co, 64, &msg) != -1 || msg.find("not finite") == std::string::npos)
        fail("non-finite coefficient refused", i);
    }
  }
  // zero pivots and singular top systems
  {
    const double zero_b1[7] = {0.0, 1.0, 0.25, 1.0, 0.25, 2.0, 1.0};
    const double zero_bn[7] = {1.0, 2.0, 0.25, 1.0, 0.25, 2.0, 0.0};
    const double zero_b[7] = {1.0, 2.0, 0.25, 0.0, 0.25, 2.0, 1.0};
    const double sing2[7] = {1.0, 1.0, 0.5, 2.0, 0.5, 1.0, 1.0};   // n = 2: [[1, 1], [1, 1]]
    std::string msg;
    if (build(plan.get(), 64, zero_b1, 128, &msg) != -2) fail("zero first pivot refused", 64);
    if (build(plan.get(), 64, zero_bn, 128, &msg) != -2) fail("zero last pivot refused", 64);
    if (build(plan.get(), 64, zero_b, 128, &msg) != -2) fail("zero interior pivot refused", 64);
    if (build(plan.get(), 2, sing2, 128, &msg) != -2 || msg.find("singular") == std::string::npos)
      fail("singular 2x2 refused", 2);
  }
  // error buffers of every small length, and none at all
  for (int len = 0; len <= 300; ++len) {
    if (build(plan.get(), 0, compact, len, nullptr) != -1) fail("refusal with short buffer", len);
    double co[7];
    std::memcpy(co, general, sizeof(co));
    co[3] = NAN;
    if (build(plan.get(), 16, co, len, nullptr) != -1) fail("refusal with short buffer", len);
  }
  if (nt_plan_build(plan.get(), 1, compact, nullptr, 0) != -1) fail("null error buffer", 1);
  // a refused build leaves no stale table behind
  build(plan.get(), 512, compact, 64, nullptr);
  if (build(plan.get(), 5000, compact, 64, nullptr) != -1 || plan->nlevels != 0 || plan->n != 0)
    fail("refused build clears the plan", 5000);
  // the general (non-dominant ends) family still builds at the sizes the kernels take
  for (long long n : {2LL, 3LL, 5LL, 64LL, 100LL, 512LL, 2048LL}) {
    if (build(plan.get(), n, general, 64, nullptr) != 0) fail("general family", n);
    else check_plan(*plan, n, false);
  }
  // a zero pivot in the first block's sweep (b_i - a_i c1 / b1 = 0) leaves no lane tables
  {
    const double zero_block[7] = {1.0, 2.0, 0.5, 1.0, 0.25, 2.0, 1.0};
    NtLaneTab t;
    if (nt_lane_build(&t, 512, zero_block) || t.m != 0) fail("zero block pivot refused", 512);
    if (!nt_lane_build(&t, 256, general) || t.m != 16) fail("general family lane tables", 256);
  }
  std::printf("plans built %lld, failures %d\n", built, failures);
  return failures == 0 ? 0 : 1;
}
