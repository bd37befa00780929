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
//     the diagonally dominant families (cyclic reduction preserves it).
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
      ++built;
    }
    for (const auto& co : dominant_sets) {
      if (build(plan.get(), n, co.data(), 256, &msg) != 0) fail(("dominant: " + msg).c_str(), n);
      else { check_plan(*plan, n, true); ++built; }
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
  std::printf("plans built %lld, failures %d\n", built, failures);
  return failures == 0 ? 0 : 1;
}
