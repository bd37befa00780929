/*
 * ntoeplitz.h - C ABI of the MI355X near-Toeplitz tridiagonal library
 * (libntoeplitz.so, HIP kernels for gfx950).
 *
 * Device pointers are plain HIP device addresses (e.g. torch tensor
 * data_ptr()), streams are hipStream_t passed as void* (NULL = the null
 * stream).  Every entry point returns NT_OK (0) or a negative NT_E* code and
 * then leaves a message for nt_last_error() (per host thread).  No entry point
 * allocates device memory on a solve path; launches are asynchronous on the
 * given stream, as the reference's PyCUDA launches are on the default stream.
 *
 * The reference (ashwinsrnth/near-toeplitz-applications) is Python 2 +
 * PyCUDA; each entry below names the reference interface it replaces.
 */
#ifndef NTOEPLITZ_H
#define NTOEPLITZ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NT_OK 0
#define NT_EINVAL (-1)      /* bad argument: size, shape, pointer, stride */
#define NT_ESINGULAR (-2)   /* zero pivot in the cyclic reduction */
#define NT_EHIP (-3)        /* HIP runtime error (launch, graph, device) */
#define NT_EUNSUPPORTED (-4)

/* element types of device buffers */
#define NT_F64 0
#define NT_F32 1

/* arithmetic of the bulk solve */
#define NT_PREC_F64 0         /* FP64 cyclic reduction */
#define NT_PREC_F32_REFINE 1  /* FP32 cyclic reduction + FP64-residual refinement sweeps */

typedef struct nt_plan nt_plan;

/* Message of the last failing call on this host thread ("" if none). */
const char* nt_last_error(void);
/* Library ABI version (major*100 + minor). */
int nt_version(void);

/* ---- launch tuning (no reference counterpart: the reference hard-codes one
 * block per system, neato.py:181-183) --------------------------------------
 * Launch-shape variant used for lines of length n (bucketed by ceil(log2 n)
 * and by whether n is a power of two) in layout 0 (contiguous lines) or 1
 * (strided lines) of the solves and derivatives, or 2 / 3 (the same for the
 * ADI half steps, nt_adi_steps) at the given precision mode (NT_PREC_*).  Process-wide; the defaults are the shapes measured
 * fastest on MI355X, and the autotuner (ntamd/autotune.py) can re-measure and
 * set them from its JSON cache.  Every variant gives the same results to
 * within a few roundings; only the launch shape changes.  Contiguous layout:
 * 0 one line per wavefront, 1 direct chunk loads, 2 workgroup tiles, 3 one
 * line over several wavefronts, 4 persistent with the next line prefetched
 * (not for ADI, whose layout 2 has variants 0-3).  Strided layouts: 0 16-line
 * tiles, 1 32-line tiles, 2 persistent whole tiles, 3 the same with two lines
 * per wavefront, 4 the two lines on 32-lane segments, 5 one lane per line with
 * a block-partitioned FP64 solve (49..512-point lines, FP64 or FP32 storage),
 * or at NT_PREC_F32_REFINE the FP32 block solve with one sweep of FP32-stored
 * compact derivatives (49..496 and 512 points); lines variant 5
 * does not take run the shapes that are the defaults without it (not for ADI,
 * whose layout 3 has variants 0-3). */
int nt_tune_set(int layout, int precision, int64_t n, int variant);
int nt_tune_get(int layout, int precision, int64_t n, int* variant, int* nvariants);

/* ---- plans ---------------------------------------------------------------
 * Replaces NearToeplitzSolver.__init__(n, nrhs, coeffs) (neato.py:131-166):
 * coeffs = (b1, c1, ai, bi, ci, an, bn) as in neato.py:136-138.  The level
 * coefficients (neato.py:199-274, here for any 2 <= n <= 4096) are computed
 * once on the host in FP64; the plan is immutable and usable from any stream.
 * No device memory is held: kernels receive the table as a launch argument. */
int nt_plan_create(int64_t n, const double coeffs[7], nt_plan** out);
int nt_plan_destroy(nt_plan* plan);
int nt_plan_info(const nt_plan* plan, int64_t* n, int* nlevels);
/* Level table: out[8*l + {0..7}] = (m, bF, cF, a, b, c, aL, bL) of level l,
 * l = 0 .. nlevels-1.  cap = number of doubles available in out. */
int nt_plan_levels(const nt_plan* plan, double* out, int cap);
/* The reference's eight arrays (a, b, c, k1, k2, b_first, k1_first, k1_last),
 * each of length log2(n), concatenated in that order; power-of-2 n >= 4 only
 * (neato.py:199-274). */
int nt_plan_reference_arrays(const nt_plan* plan, double* out, int cap);

/* ---- solves ---------------------------------------------------------------
 * Replaces NearToeplitzSolver.solve(x_d) (neato.py:168-197): solves nrhs
 * systems of length n stored contiguously from x (x may point into a larger
 * array, as run-adi.py:65 passes d_x_gpu[1,0]), in place.
 * dtype NT_F64 with precision NT_PREC_F64 or NT_PREC_F32_REFINE (sweeps >= 0). */
int nt_solve(const nt_plan* plan, void* x, int dtype, int64_t nrhs, int precision, int sweeps,
             void* stream);

/* In-place solve of every line along `axis` (0 = z, 1 = y, 2 = x) of a
 * C-contiguous (nz, ny, nx) field: the reference's permute + solve + permute
 * (run-compact.py:66-90) without the permutes. */
int nt_solve_axis(const nt_plan* plan, void* x, int dtype, int64_t nz, int64_t ny, int64_t nx,
                  int axis, int precision, int sweeps, void* stream);

/* ---- compact derivative (fused) ------------------------------------------
 * out = d f / d axis by the 4th-order compact scheme: RHS stencil
 * (compact/rhs.py:21-36) + solve (neato.py) + the permutes of
 * run-compact.py:66-90, in one pass over HBM.  The plan must be the compact
 * matrix of size n = extent of `axis`; dx is the grid spacing.
 * f_dtype NT_F64 or NT_F32 (FP32 storage), out_dtype NT_F64. */
int nt_compact_derivative(const nt_plan* plan, const void* f, int f_dtype, void* out,
                          int out_dtype, int64_t nz, int64_t ny, int64_t nx, int axis,
                          double dx, int precision, int sweeps, void* stream);

/* compact.rhs.compute_rhs(f, rhs, dx) (compact/rhs.py:47-53): RHS along x of
 * a (nz, ny, nx) FP64 field. */
int nt_compact_rhs(const void* f, void* rhs, int64_t nz, int64_t ny, int64_t nx, double dx,
                   void* stream);

/* compact.permute.permute(a_d, b_d, permutation) (compact/permute.py:32-52):
 * b = a.transpose(perm) for a C-contiguous 3-D FP64 array of shape `shape`
 * (any shape, not only cubes). */
int nt_permute(const void* a, void* b, const int64_t shape[3], const int perm[3], void* stream);

/* ---- ADI ----------------------------------------------------------------
 * adi.rhs.compute_rhs(u, d, dx, dt, N) (adi/rhs.py:35-55): interior-only
 * explicit stencil along the slow axis of an N x N FP64 array. */
int nt_adi_rhs(const void* u, void* d, double dx, double dt, int64_t N, void* stream);

/* adi.transpose.transpose(a, b, N) (adi/transpose.py:40-50): b = a^T, N x N FP64. */
int nt_transpose(const void* a, void* b, int64_t N, void* stream);

/* Fused Peaceman-Rachford steps (run-adi.py:60-71 without the transposes):
 * each step is an x-implicit half step u -> work and a y-implicit half step
 * work -> u, lines (-r/2, 1+r, -r/2) with identity (Dirichlet) ends.  `plan`
 * must be built from (1, 0, -r/2, 1+r, -r/2, 0, 1) with n = N.  `work` is an
 * N x N FP64 buffer whose boundary rows/columns must equal u's (Dirichlet
 * values); nt_adi_prepare fills it. */
int nt_adi_prepare(const void* u, void* work, int64_t N, void* stream);
int nt_adi_steps(const nt_plan* plan, void* u, void* work, int64_t N, double r, int nsteps,
                 int precision, int sweeps, void* stream);

/* ---- one-GPU partitioned (superposition) solve ------------------------------
 * The reference's distributed solver cuts each line into P subdomains over P
 * MPI ranks: local solves, secondary solutions (parallelcompact/misc.py:21-46),
 * a 2P-unknown reduced system gathered to one rank and solved there
 * (parallelcompact/reduced.py:11-196), a correction sum (parallelcompact/sum.py).
 * Here the same method runs on one GPU as one launch per batch of lines.
 * A plan fixes n = P*m, P <= 16 and the 7-tuple; it precomputes the subdomain
 * level tables, xU/xL and the reduced system's factors on the host in FP64
 * and keeps them in one device allocation (made at creation, on the current
 * device; none if no device is present). */
typedef struct nt_part_plan nt_part_plan;
int nt_part_plan_create(int64_t n, int P, const double coeffs[7], nt_part_plan** out);
int nt_part_plan_destroy(nt_part_plan* plan);
/* Host copies of xU, xL of subdomain p (misc.py:21-46 solve_secondary_systems). */
int nt_part_plan_secondary(const nt_part_plan* plan, int p, double* xU, double* xL, int cap);
/* producer 0: solve `lines` contiguous FP64 lines of length n in place (in == out allowed);
 * producer 1: compact derivative along contiguous lines, p1 = dx
 *             (parallelcompact rhs.py + solve + get_params + sum_solutions of
 *             run-parallel-compact.py:79-85, one GPU);
 * producer 2: ADI x-implicit half step of an n x n field in -> out (lines == n),
 *             p1 = r; rows 0 and n-1 of out are not written.
 * p2 is reserved (pass 0). */
int nt_part_solve_lines(const nt_part_plan* plan, int producer, const void* in, void* out,
                        int64_t lines, double p1, double p2, void* stream);

/* The partitioned solve along `axis` (0 = z, 1 = y, 2 = x) of a C-contiguous
 * (nz, ny, nx) FP64 field: the reference's permute + partitioned solve +
 * permute back for y and z (run-parallel-compact.py:92-164) without the
 * permutes.  Strided lines are staged in tiles of 8 or 16 adjacent lines.
 * producer 0: solve in place (in == out); 1: compact derivative, p1 = dx;
 * 2: an ADI half step of an n x n field (nz = 1) in -> out, p1 = r: axis 2 is
 * the x-implicit half step (rows), axis 1 the y-implicit one (columns); lines
 * 0 and n-1 are not solved. */
int nt_part_solve_axis(const nt_part_plan* plan, int producer, const void* in, void* out, int64_t nz,
                       int64_t ny, int64_t nx, int axis, double p1, void* stream);

/* The partition method one step at a time, as the reference's ranks run it
 * (run-parallel-compact.py:72-76).  Subdomain arrays are FP64, C-contiguous.
 *
 * parallelcompact.rhs.compute_RHS(f_local, x, dx, shape, mx, npx)
 * (parallelcompact/rhs.py:8-73): x (nz, ny, nx) = compact RHS along x of the
 * width-1 ghosted f_local (nz+2, ny+2, nx+2); the one-sided closures only on
 * rank mx == 0 (first point) and mx == npx-1 (last point).  nx >= 3. */
int nt_part_rhs_ghosted(const void* f_local, void* x, int64_t nz, int64_t ny, int64_t nx, double dx,
                        int mx, int npx, void* stream);

/* parallelcompact.reduced.get_params(line_da, xU_d, xL_d, xR_d)
 * (parallelcompact/reduced.py:135-196) for the `nsub` subdomains of a line
 * held on one device, so the reference's gather / root solve / scatter
 * (reduced.py:147-192) is one launch: xU, xL (nsub, m) secondary solutions,
 * xR (nsub, nz, ny, m) subdomain solutions; writes alpha, beta (nsub, nz, ny).
 * 1 <= nsub <= 16. */
int nt_part_params(const void* xU, const void* xL, const void* xR, void* alpha, void* beta, int nsub,
                   int64_t nz, int64_t ny, int64_t m, void* stream);

/* parallelcompact.sum.sum_solutions(x_UH, x_LH, x_R, alpha, beta, shape)
 * (parallelcompact/sum.py:7-51): x_R[z,y,x] += alpha[z,y] x_UH[x] + beta[z,y] x_LH[x]. */
int nt_part_sum(const void* xU, const void* xL, void* xR, const void* alpha, const void* beta, int64_t nz,
                int64_t ny, int64_t nx, void* stream);

/* ---- mixed-size persistent batch --------------------------------------------
 * No reference counterpart (the reference launches one grid per field and
 * direction, run-compact.py:57-90).  Many fields of different extents,
 * directions, storage types and precisions in one persistent launch that
 * pulls tiles of 16 lines from a device work counter.  Lines of 16..512 points.
 * The library allocates nothing: nt_batch_pack writes a binary image into a
 * caller-provided host buffer (nt_batch_image_bytes long); the caller copies it
 * to a 256-byte aligned device buffer, zeroes a device unsigned counter and
 * calls nt_batch_launch on the stream of the copy.  Each field's result does
 * not depend on the other jobs in the launch. */
typedef struct nt_batch_job {
  const void* in;     /* field (producer 1), or the systems to solve (producer 0; == out) */
  void* out;          /* FP64 result */
  int64_t nz, ny, nx; /* C-contiguous extents */
  double dx;          /* grid spacing along axis (producer 1) */
  int axis;           /* 0 = z, 1 = y, 2 = x */
  int producer;       /* 0 = solve in place, 1 = compact derivative */
  int in_dtype;       /* NT_F64 or NT_F32 (producer 1) */
  int plan_slot;      /* index into the plans array of nt_batch_pack */
  int precision;      /* NT_PREC_F64 or NT_PREC_F32_REFINE */
  int sweeps;
} nt_batch_job;
size_t nt_batch_image_bytes(int njobs, int nplans);
int nt_batch_pack(const nt_batch_job* jobs, int njobs, const nt_plan* const* plans, int nplans,
                  void* host_image, size_t bytes, int64_t* total_tiles);
/* grid <= 0: four workgroups per compute unit */
int nt_batch_launch(const void* dev_image, int njobs, int nplans, int64_t total_tiles,
                    unsigned int* dev_counter, int grid, void* stream);

/* ---- allocation-free batch ----------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device or
 * pinned host memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same jobs and kernel as
 * nt_batch_pack + nt_batch_launch, with the job image and work counter kept in
 * a fixed 1 MiB device array of the library (one per GPU, held until the process
 * ends).  The image is packed on the host into a static buffer and carried to
 * the device in the arguments of small fill kernels on `stream`, so nothing is
 * allocated and no host memory needs pinning; the call returns once the work is
 * enqueued.  Results are identical to nt_batch_launch's.
 * Limits: nt_batch_image_bytes(njobs, nplans) <= NT_BATCH_NOALLOC_IMAGE_MAX
 * (NT_EINVAL beyond it).  `stream` must belong to the current device.  Every
 * call on one GPU shares that GPU's array: calls on the same stream are
 * ordered, but calls that overlap on one GPU (other streams, other host
 * threads enqueueing while earlier work still runs) race on it. */
#define NT_BATCH_NOALLOC_IMAGE_MAX ((size_t)1048320)   /* 1 MiB minus the 256-byte counter block */
int nt_batch_run_noalloc(const nt_batch_job* jobs, int njobs, const nt_plan* const* plans, int nplans,
                         int grid, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NTOEPLITZ_H */
