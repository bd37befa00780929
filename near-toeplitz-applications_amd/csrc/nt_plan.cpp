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
  nt_lane_build(&p->lane, n, coeffs);
  nt_lane_build32(&p->lane32, p->lane);
  return 0;
}

namespace {

// One block of len rows of the lane solve, as the kernel runs it: forward
// y_0 = d_0 ip_0, y_r = (d_r - a_r y_{r-1}) ip_r, then y_r -= cp_r y_{r+1}.
// kind: 0 first block, 1 interior, 2 last (its row len-1 has an and ip_last).
void lane_block(const NtLaneTab& t, int kind, int len, const double* d, double* y) {
  const double* ip = t.ip[kind == 0 ? 0 : 1];
  const double* cp = t.cp[kind == 0 ? 0 : 1];
  y[0] = d[0] * ip[0];
  for (int r = 1; r < len; ++r) {
    const bool last = (kind == 2 && r == len - 1);
    y[r] = (d[r] - (last ? t.an : t.ai) * y[r - 1]) * (last ? t.ip_last : ip[r]);
  }
  for (int r = len - 2; r >= 0; --r) y[r] -= cp[r] * y[r + 1];
}

int lane_len(const NtLaneTab& t, int p) { return p < t.j ? t.m : t.m - 1; }
int lane_kind(int p) { return p == 0 ? 0 : (p == NT_LANE_P - 1 ? 2 : 1); }
const double* lane_uh(const NtLaneTab& t, int p) {
  return t.uh[p == NT_LANE_P - 1 ? 2 : (lane_len(t, p) == t.m ? 0 : 1)];
}
const double* lane_lh(const NtLaneTab& t, int p) {
  return t.lh[p == 0 ? 0 : (lane_len(t, p) == t.m ? 1 : 2)];
}

// The whole lane solve of one line on the host (the kernel's arithmetic).
void lane_solve(const NtLaneTab& t, const double* d, double* x) {
  const int P = NT_LANE_P, NR = NT_LANE_NR;
  double y[NT_LANE_P * NT_LANE_MAX_M], g[NT_LANE_NR];
  int start[NT_LANE_P + 1];
  start[0] = 0;
  for (int p = 0; p < P; ++p) start[p + 1] = start[p] + lane_len(t, p);
  for (int p = 0; p < P; ++p) lane_block(t, lane_kind(p), lane_len(t, p), d + start[p], y + start[p]);
  for (int i = 0; i < NR; ++i) {
    const int p = (i + 1) / 2;
    const double r = (i % 2 == 1) ? -y[start[p]] : -y[start[p + 1] - 1];
    g[i] = (r - (i ? t.rsub[i] * g[i - 1] : 0.0)) * t.rip[i];
  }
  for (int i = NR - 2; i >= 0; --i) g[i] -= t.rcp[i] * g[i + 1];
  for (int p = 0; p < P; ++p) {
    const double al = (p >= 1) ? g[2 * p - 1] : 0.0, be = (p <= P - 2) ? g[2 * p] : 0.0;
    const double* uh = lane_uh(t, p);
    const double* lh = lane_lh(t, p);
    for (int r = 0; r < lane_len(t, p); ++r) x[start[p] + r] = y[start[p] + r] + al * uh[r] + be * lh[r];
  }
}

bool lane_off(NtLaneTab* t) {
  t->m = 0;
  t->j = 0;
  return false;
}

}  // namespace

bool nt_lane_build(NtLaneTab* t, int64_t n, const double coeffs[7]) {
  std::memset(t, 0, sizeof(*t));
  const int P = NT_LANE_P, NR = NT_LANE_NR;
  if (n <= 3 * P || n > P * NT_LANE_MAX_M) return false;
  const int m = static_cast<int>((n + P - 1) / P);
  const double b1 = coeffs[0], c1 = coeffs[1], ai = coeffs[2], bi = coeffs[3], ci = coeffs[4],
               an = coeffs[5], bn = coeffs[6];
  t->m = m;
  t->j = static_cast<int>(n - static_cast<int64_t>(P) * (m - 1));
  t->ai = ai;
  t->an = an;
  // block sweeps: the first block starts with (b1, c1), the others with (bi, ci);
  // a block of m - 1 rows uses the first m - 1 pivots of its kind
  for (int k = 0; k < 2; ++k) {
    double cprev = 0.0;
    for (int r = 0; r < m; ++r) {
      const bool first = (k == 0 && r == 0);
      const double piv = (first ? b1 : bi) - (r ? ai * cprev : 0.0);
      if (!finite_nonzero(piv)) return lane_off(t);
      t->ip[k][r] = 1.0 / piv;
      t->cp[k][r] = (first ? c1 : ci) / piv;
      cprev = t->cp[k][r];
    }
  }
  const int llast = lane_len(*t, P - 1);
  const double pl = bn - an * t->cp[1][llast - 2];
  if (!finite_nonzero(pl)) return lane_off(t);
  t->ip_last = 1.0 / pl;
  // responses to the couplings: uh = A^-1 (-ai e_0), lh = A^-1 (-ci e_{len-1})
  double e[NT_LANE_MAX_M] = {};
  e[0] = -ai;
  lane_block(*t, 1, m, e, t->uh[0]);
  lane_block(*t, 1, m - 1, e, t->uh[1]);
  lane_block(*t, 2, llast, e, t->uh[2]);
  e[0] = 0.0;
  e[m - 1] = -ci;
  lane_block(*t, 0, m, e, t->lh[0]);
  lane_block(*t, 1, m, e, t->lh[1]);
  e[m - 1] = 0.0;
  e[m - 2] = -ci;
  lane_block(*t, 1, m - 1, e, t->lh[2]);
  // reduced system, row i for z_{i+1}: odd i: first row of block (i+1)/2,
  // x_p[0] = z_{2p-1}; even i: last row of block i/2, x_p[len-1] = z_{2p+2}
  double cprev = 0.0;
  for (int i = 0; i < NR; ++i) {
    double sub, dia, sup;
    if (i % 2 == 1) {
      const int p = (i + 1) / 2;
      sub = -1.0;
      dia = lane_uh(*t, p)[0];
      sup = (p <= P - 2) ? lane_lh(*t, p)[0] : 0.0;
    } else {
      const int p = i / 2, len = lane_len(*t, p);
      sub = (p >= 1) ? lane_uh(*t, p)[len - 1] : 0.0;
      dia = lane_lh(*t, p)[len - 1];
      sup = -1.0;
    }
    const double piv = dia - (i ? sub * cprev : 0.0);
    if (!finite_nonzero(piv)) return lane_off(t);
    t->rsub[i] = sub;
    t->rip[i] = 1.0 / piv;
    t->rcp[i] = sup / piv;
    cprev = t->rcp[i];
  }
  // self-check against a whole-system Thomas solve of a test vector
  double d[NT_LANE_P * NT_LANE_MAX_M], x[NT_LANE_P * NT_LANE_MAX_M], cw[NT_LANE_P * NT_LANE_MAX_M],
      dw[NT_LANE_P * NT_LANE_MAX_M];
  for (int i = 0; i < n; ++i) d[i] = std::sin(0.7 * i) + 0.3 * std::cos(2.1 * i + 0.4);
  for (int i = 0; i < n; ++i) {
    const double a = (i == n - 1) ? an : ai, b = (i == 0) ? b1 : (i == n - 1 ? bn : bi),
                 c = (i == 0) ? c1 : ci;
    const double piv = b - (i ? a * cw[i - 1] : 0.0);
    if (!finite_nonzero(piv)) return lane_off(t);
    cw[i] = c / piv;
    dw[i] = (d[i] - (i ? a * dw[i - 1] : 0.0)) / piv;
  }
  for (int i = static_cast<int>(n) - 2; i >= 0; --i) dw[i] -= cw[i] * dw[i + 1];
  lane_solve(*t, d, x);
  double emax = 0.0, xmax = 0.0;
  for (int i = 0; i < n; ++i) {
    if (!std::isfinite(x[i]) || !std::isfinite(dw[i])) return lane_off(t);
    emax = std::fmax(emax, std::fabs(x[i] - dw[i]));
    xmax = std::fmax(xmax, std::fabs(dw[i]));
  }
  if (!(emax <= 1e-11 * xmax)) return lane_off(t);
  return true;
}

bool nt_lane_build32(NtLaneTab32* t, const NtLaneTab& d) {
  std::memset(t, 0, sizeof(*t));
  if (d.m < 4) return false;
  const int m = d.m;
  t->ai = static_cast<float>(d.ai);
  t->an = static_cast<float>(d.an);
  t->ip_last = static_cast<float>(d.ip_last);
  for (int r = 0; r < m; ++r) {
    for (int k = 0; k < 2; ++k) {
      t->ip[k][r] = static_cast<float>(d.ip[k][r]);
      t->cp[k][r] = static_cast<float>(d.cp[k][r]);
    }
    for (int k = 0; k < 3; ++k) {
      t->uh[k][r] = static_cast<float>(d.uh[k][r]);
      t->lh[k][r] = static_cast<float>(d.lh[k][r]);
    }
  }
  for (int i = 0; i < NT_LANE_NR; ++i) {
    t->rsub[i] = static_cast<float>(d.rsub[i]);
    t->rip[i] = static_cast<float>(d.rip[i]);
    t->rcp[i] = static_cast<float>(d.rcp[i]);
  }
  bool ok = std::isfinite(t->ai) && std::isfinite(t->an) && std::isfinite(t->ip_last);
  for (int r = 0; r < m; ++r) {
    for (int k = 0; k < 2; ++k) ok = ok && std::isfinite(t->ip[k][r]) && std::isfinite(t->cp[k][r]);
    for (int k = 0; k < 3; ++k) ok = ok && std::isfinite(t->uh[k][r]) && std::isfinite(t->lh[k][r]);
  }
  for (int i = 0; i < NT_LANE_NR; ++i)
    ok = ok && std::isfinite(t->rsub[i]) && std::isfinite(t->rip[i]) && std::isfinite(t->rcp[i]);
  if (!ok) {
    std::memset(t, 0, sizeof(*t));
    return false;
  }
  t->m = m;
  t->j = d.j;
  return true;
}
