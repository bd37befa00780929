// Host-side construction of the cyclic-reduction level table (FP64).
//
// Level recurrences: reference neato.py:241-272 (power-of-2 n).  For an odd
// equation count the last equation is eliminated into the last kept one
// (see nt_plan.h); for power-of-2 n the table reproduces the reference's
// eight arrays, which nt_plan_reference_arrays() rebuilds from it.
#include "nt_plan.h"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace {

struct Sys {
  double m, bF, cF, a, b, c, aL, bL;
};

bool finite_nonzero(double v) { return std::isfinite(v) && v != 0.0; }

template <typename T>
void fill_level(NtLevel<T>& L, const Sys& s) {
  const double m = s.m;
  L.m = static_cast<int32_t>(m);
  L.pad_ = 0;
  L.k1 = static_cast<T>(s.a / s.b);
  L.k2 = static_cast<T>(s.c / s.b);
  L.k1f = static_cast<T>(s.a / s.bF);
  const bool even = (static_cast<int64_t>(m) % 2) == 0;
  L.kl = static_cast<T>(even ? s.aL / s.b : s.c / s.bL);
  L.a = static_cast<T>(s.a);
  L.c = static_cast<T>(s.c);
  L.ib = static_cast<T>(1.0 / s.b);
  L.cf = static_cast<T>(s.cF);
  L.ibf = static_cast<T>(1.0 / s.bF);
  L.al = static_cast<T>(s.aL);
  L.ibl = static_cast<T>(1.0 / s.bL);
}

template <typename T>
void fill_top(NtTop<T>& t, const Sys& s) {
  t.m = static_cast<int32_t>(s.m);
  t.pad_ = 0;
  for (auto& v : t.t) v = T(0);
  if (t.m == 1) {
    t.t[0] = static_cast<T>(1.0 / s.bF);
  } else if (t.m == 2) {
    const double idet = 1.0 / (s.bF * s.bL - s.cF * s.aL);
    t.t[0] = static_cast<T>(s.bL * idet);
    t.t[1] = static_cast<T>(s.cF * idet);
    t.t[2] = static_cast<T>(s.bF * idet);
    t.t[3] = static_cast<T>(s.aL * idet);
  } else {
    const double k1f = s.a / s.bF, k2l = s.c / s.bL;
    t.t[0] = static_cast<T>(k1f);
    t.t[1] = static_cast<T>(k2l);
    t.t[2] = static_cast<T>(1.0 / (s.b - k1f * s.cF - k2l * s.aL));
    t.t[3] = static_cast<T>(s.cF);
    t.t[4] = static_cast<T>(1.0 / s.bF);
    t.t[5] = static_cast<T>(s.aL);
    t.t[6] = static_cast<T>(1.0 / s.bL);
  }
}

}  // namespace

int nt_plan_build(nt_plan* p, int64_t n, const double coeffs[7], char* err, int errlen) {
  std::memset(p, 0, sizeof(*p));
  if (n < 2 || n > 4096) {
    std::snprintf(err, errlen, "system size n=%lld out of range [2, 4096]", (long long)n);
    return -1;
  }
  for (int i = 0; i < 7; ++i) {
    if (!std::isfinite(coeffs[i])) {
      std::snprintf(err, errlen, "coefficient %d is not finite", i);
      return -1;
    }
  }
  p->n = n;
  std::memcpy(p->coeffs, coeffs, sizeof(p->coeffs));
  Sys s{static_cast<double>(n), coeffs[0], coeffs[1], coeffs[2], coeffs[3], coeffs[4],
        coeffs[5], coeffs[6]};
  int nl = 0;
  for (;;) {
    if (nl >= NT_MAX_LEVELS) {
      std::snprintf(err, errlen, "too many CR levels for n=%lld", (long long)n);
      return -1;
    }
    double* row = p->lev[nl];
    row[0] = s.m; row[1] = s.bF; row[2] = s.cF; row[3] = s.a;
    row[4] = s.b; row[5] = s.c; row[6] = s.aL; row[7] = s.bL;
    const int64_t m = static_cast<int64_t>(s.m);
    // every pivot the solve divides by must be usable (no pivoting, as the reference)
    if (!finite_nonzero(s.bF) || !finite_nonzero(s.bL) || (m >= 4 && !finite_nonzero(s.b))) {
      std::snprintf(err, errlen, "zero or non-finite pivot at CR level %d (m=%lld)", nl,
                    (long long)m);
      return -2;
    }
    fill_level(p->k.d.lv[nl], s);
    fill_level(p->k.f.lv[nl], s);
    ++nl;
    if (m <= 3) break;
    const double k1 = s.a / s.b, k2 = s.c / s.b, k1f = s.a / s.bF;
    Sys t;
    t.m = static_cast<double>(m / 2);
    t.a = -k1 * s.a;
    t.b = s.b - k1 * s.c - k2 * s.a;
    t.c = -k2 * s.c;
    t.bF = s.b - k1f * s.cF - k2 * s.a;
    t.cF = t.c;
    if (m % 2 == 0) {
      const double k1l = s.aL / s.b;
      t.aL = -k1l * s.a;
      t.bL = s.bL - k1l * s.c;
    } else {
      const double k2l = s.c / s.bL;
      t.aL = -k1 * s.a;
      t.bL = s.b - k1 * s.c - k2l * s.aL;
    }
    s = t;
  }
  // top-system checks
  if (s.m == 2 && !finite_nonzero(s.bF * s.bL - s.cF * s.aL)) {
    std::snprintf(err, errlen, "singular 2x2 top system");
    return -2;
  }
  if (s.m == 3) {
    const double piv = s.b - (s.a / s.bF) * s.cF - (s.c / s.bL) * s.aL;
    if (!finite_nonzero(piv)) {
      std::snprintf(err, errlen, "singular 3x3 top system");
      return -2;
    }
  }
  p->nlevels = nl;
  p->k.n = static_cast<int32_t>(n);
  p->k.nlev = nl - 1;
  p->k.b1 = coeffs[0]; p->k.c1 = coeffs[1]; p->k.ai = coeffs[2]; p->k.bi = coeffs[3];
  p->k.ci = coeffs[4]; p->k.an = coeffs[5]; p->k.bn = coeffs[6];
  fill_top(p->k.d.top, s);
  fill_top(p->k.f.top, s);
  return 0;
}
