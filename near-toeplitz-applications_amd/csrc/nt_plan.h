// Near-Toeplitz cyclic-reduction plan: the level table computed once on the
// host in FP64 and handed to every kernel by value (it lives in the kernel
// argument segment, i.e. constant memory read through the scalar cache).
//
// A near-Toeplitz matrix is the 7-tuple (b1, c1, ai, bi, ci, an, bn): first row
// [b1, c1], constant interior rows (ai, bi, ci), last row [an, bn]
// (reference neato.py:136-138).  One CR reduction keeps the odd equations of a
// level; the kept equations again form a near-Toeplitz system, so each level
// is described by seven numbers and its equation count m.  The reference
// precomputes these per level for power-of-2 n (neato.py:199-274); this plan
// does the same for any n >= 2 (an odd m eliminates the level's last
// equation into the last kept one, which then carries the factor c/bL).
#pragma once
#include <cstdint>

#define NT_MAX_LEVELS 12   // levels 0..11: enough for any n <= 4096
#define NT_MAX_N 2048      // the register-resident CR holds at most 64 lanes x 32 rows

// Factors of one level, in the arithmetic type T of the solve.
template <typename T>
struct NtLevel {
  // forward reduction that starts from this level (reduction s uses level s-1)
  T k1;    // a/b      left factor of an interior kept equation
  T k2;    // c/b      right factor of an interior kept equation
  T k1f;   // a/bF     left factor of the first kept equation (its left neighbour is row 0)
  T kl;    // m even: aL/b (left factor of the kept last row); m odd: c/bL (right factor of row m-2)
  // back substitution of this level's eliminated (even) equations
  T a, c, ib;     // interior row: x = (d - a*xl - c*xr) * ib
  T cf, ibf;      // row 0:        x = (d - cf*xr) * ibf
  T al, ibl;      // row m-1 (m odd only): x = (d - al*xl) * ibl
  int32_t m;      // equations at this level
  int32_t pad_;
};

// Direct solve of the top system (m <= 3 equations).
template <typename T>
struct NtTop {
  int32_t m;
  int32_t pad_;
  // m == 1: x0 = d0*t[0]
  // m == 2: x0 = d0*t[0] - d1*t[1];  x1 = d1*t[2] - d0*t[3]   (Cramer, neato.py:78-93)
  // m == 3: x1 = (d1 - t[0]*d0 - t[1]*d2)*t[2]; x0 = (d0 - t[3]*x1)*t[4]; x2 = (d2 - t[5]*x1)*t[6]
  T t[7];
};

template <typename T>
struct NtPlanT {
  NtLevel<T> lv[NT_MAX_LEVELS];
  NtTop<T> top;
};

// Everything a kernel needs about one matrix.  Passed by value.
struct NtPlanK {
  int32_t n;       // system size
  int32_t nlev;    // number of reductions (top system is level nlev)
  double b1, c1, ai, bi, ci, an, bn;   // level-0 matrix (FP64 residuals)
  NtPlanT<double> d;                   // FP64 factors
  NtPlanT<float> f;                    // the same factors rounded to FP32
};

// Block-partitioned solve with one lane per line (strided lines whose adjacent
// lines are adjacent in memory): the n rows are cut into NT_LANE_P blocks, one
// wavefront per block, each lane a line.  With m = ceil(n / NT_LANE_P), the
// first j = n - NT_LANE_P (m - 1) blocks have m rows and the others m - 1.  Each
// block is solved on its own by a Thomas sweep (y), then
//   x = y + alpha * uh + beta * lh
// with alpha the previous block's last value and beta the next block's first
// (uh, lh: the block's responses to its two couplings, RHS-independent).  The
// unknowns z = (alpha_0, beta_0, alpha_1, ...) satisfy the tridiagonal reduced
// system of the reference's partition method (parallelcompact/reduced.py:163-178;
// alpha_0 = beta_{P-1} = 0 are dropped, leaving 2P - 2 rows).  All factors are
// RHS-independent and computed here once, in FP64.
#define NT_LANE_P 16
#define NT_LANE_MAX_M 32
#define NT_LANE_NR (2 * NT_LANE_P - 2)
struct NtLaneTab {
  int32_t m;                        // rows of the long blocks; 0: no lane solve for this plan
  int32_t j;                        // blocks 0..j-1 have m rows, blocks j..P-1 have m - 1
  double ai, an;                    // sub-diagonal of block rows >= 1 (an: last row of the last block)
  double ip[2][NT_LANE_MAX_M];      // 1 / pivot of the block sweep: [0] first block, [1] the others
  double cp[2][NT_LANE_MAX_M];      // super-diagonal / pivot (the last block shares [1])
  double ip_last;                   // 1 / pivot of the last block's last row
  // response to alpha: [0] interior blocks of m rows, [1] of m - 1 rows, [2] the last block
  double uh[3][NT_LANE_MAX_M];
  // response to beta: [0] the first block, [1] interior blocks of m rows, [2] of m - 1 rows
  double lh[3][NT_LANE_MAX_M];
  double rsub[NT_LANE_NR];          // reduced system: sub-diagonal,
  double rip[NT_LANE_NR];           //   1 / pivot,
  double rcp[NT_LANE_NR];           //   super-diagonal / pivot
};

// The same factors rounded to FP32, for the FP32 + refinement lane solve of
// FP32-stored fields (same m, j and array meanings as NtLaneTab).
struct NtLaneTab32 {
  int32_t m;                        // rows of the long blocks; 0: no FP32 lane solve for this plan
  int32_t j;                        // blocks 0..j-1 have m rows, blocks j..P-1 have m - 1
  float ai, an, ip_last, pad_;
  float ip[2][NT_LANE_MAX_M];       // [0] first block, [1] the others
  float cp[2][NT_LANE_MAX_M];
  float uh[3][NT_LANE_MAX_M];       // [0] interior blocks of m rows, [1] of m - 1 rows, [2] the last block
  float lh[3][NT_LANE_MAX_M];       // [0] the first block, [1] interior blocks of m rows, [2] of m - 1 rows
  float rsub[NT_LANE_NR], rip[NT_LANE_NR], rcp[NT_LANE_NR];
};

// Host-side plan object behind the opaque C handle.
struct nt_plan {
  int64_t n;
  double coeffs[7];
  int nlevels;                       // number of CR levels including level 0
  double lev[NT_MAX_LEVELS][8];      // m, bF, cF, a, b, c, aL, bL per level
  NtPlanK k;
  NtLaneTab lane;                    // block-partitioned lane solve (lane.m == 0: none)
  NtLaneTab32 lane32;                // its FP32 form (lane32.m == 0: none)
};

// Fill `p` for an n x n near-Toeplitz matrix.  Returns 0, or a negative code
// with a message in `err` (zero pivot, n out of range, non-finite input).
int nt_plan_build(nt_plan* p, int64_t n, const double coeffs[7], char* err, int errlen);

// Fill t for 3 NT_LANE_P < n <= NT_LANE_P NT_LANE_MAX_M (blocks of at least 3
// rows) and the matrix coeffs; returns false (t->m = 0) for other n, a zero
// pivot, or when the block solve of a test vector differs from a whole-system
// Thomas solve by more than 1e-11 relative.
bool nt_lane_build(NtLaneTab* t, int64_t n, const double coeffs[7]);

// t = d rounded to FP32; false and t->m = 0 when d holds no tables or a
// factor overflows FP32.
bool nt_lane_build32(NtLaneTab32* t, const NtLaneTab& d);
