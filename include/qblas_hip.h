/*
 * qblas_hip.h -- C ABI of libqblas_hip.so, the MI355X (gfx950) backend for the QuadBLAS binary128
 * hot path: dot, gemv and gemm.
 *
 * Part 1 keeps the reference's C entry points byte-for-byte (same names, argument order and
 * meaning, same return types); they take HOST pointers to Sleef_quad / __float128 arrays, copy
 * them to the GPU, run the kernels and copy results back before returning.  A pinned buffer
 * (hipHostMalloc) is copied by one DMA; a pageable one of 2 MiB or more is staged by up to
 * QBLAS_STAGE_WORKERS (default 8) library threads through pinned slots of QBLAS_STAGE_CHUNK_MB
 * (default 8) MiB, two per thread, allocated at a calling thread's first staged copy and kept.
 * quadblas_qgemm copies C in on a second stream while the kernels that do not read C run.
 * Part 2 carries the C++ template API's full-precision entry points (binary128 alpha/beta and a
 * binary128 dot result), which include/quadblas/quadblas.hpp wraps as QuadBLAS::dot/gemv/gemm
 * and Vector/Matrix.
 * Part 3 is new: the same routines on DEVICE pointers, asynchronous on a caller stream, with an
 * explicit evaluation mode.  It serves data already resident in HBM and kernel timing.
 *
 * Binary128 values cross the ABI as 16-byte little-endian images of __float128 ({lo, hi} 64-bit
 * words), 16-byte aligned.  No function throws; failures are reported through
 * quadblas_hip_last_error() (thread-local) and leave outputs unchanged, except that the
 * functions returning a value return NaN (quadblas_qdot, quadblas_qnrm2; quadblas_dot_q and
 * quadblas_nrm2_q write a binary128 NaN).
 */
#ifndef QBLAS_HIP_H
#define QBLAS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Part 1: the reference C ABI (include/quadblas/interface/c_interface.hpp) ------------- */

/* c_interface.hpp:21-31.  Returns the binary128 dot product rounded to double (round to nearest
   even).  n <= 0 returns 0.  Evaluation order follows quadblas_get_num_threads() like the
   reference's OpenMP chunking (level1.hpp:38-137) unless the mode is changed. */
double quadblas_qdot(int n, void *x, int incx, void *y, int incy);

/* c_interface.hpp:64-90.  layout 'C'/'c' = column-major, anything else row-major; trans in
   {T,t,C,c} swaps m and n and flips the layout.  alpha, beta are converted exactly. */
void quadblas_qgemv(char layout, char trans, int m, int n, double alpha, void *A, int lda, void *x, int incx,
                    double beta, void *y, int incy);

/* c_interface.hpp:95-117.  transa/transb are accepted and ignored, as in the reference. */
void quadblas_qgemm(char layout, char transa, char transb, int m, int n, int k, double alpha, void *A, int lda,
                    void *B, int ldb, double beta, void *C, int ldc);

/* c_interface.hpp:122-131.  The thread count fixes the reference's dot chunking (T); the
   default is the host's OpenMP default (OMP_NUM_THREADS, else the available cores). */
/* quadblas_qnrm2 c_interface.hpp:34-46: (double) RN(sqrt(dot(x, x))), the dot in the current
   mode (reference order by default), the square root correctly rounded (Sleef_sqrtq1_u05). */
double quadblas_qnrm2(int n, void *x, int incx);

/* quadblas_qaxpy c_interface.hpp:49-58 (QuadBLAS::axpy, level1.hpp:139-223):
   y_i = fma(alpha, x_i, y_i), one correctly rounded FMA per element; n <= 0 is a no-op. */
void quadblas_qaxpy(int n, double alpha, void *x, int incx, void *y, int incy);

void quadblas_set_num_threads(int num_threads);
int quadblas_get_num_threads(void);

/* c_interface.hpp:134-137 and :140-144 (32-byte alignment test, constants.hpp:18). */
const char *quadblas_get_version(void);
int quadblas_is_aligned(const void *ptr);

/* ---- Part 2: full-precision entry points behind the C++ template API --------------------- */

/* QuadBLAS::dot (level1.hpp:80-137; Vector<L>::dot, cpp_classes.hpp:66-69): writes the
   binary128 result to *result. */
void quadblas_dot_q(size_t n, const void *x, size_t incx, const void *y, size_t incy, void *result);

/* QuadBLAS::axpy (level1.hpp:190-223; Vector<L>::axpy, cpp_classes.hpp:72-76). */
void quadblas_axpy_q(size_t n, const void *alpha, const void *x, size_t incx, void *y, size_t incy);

/* Vector<L>::norm (cpp_classes.hpp:78-81): *result = RN(sqrt(QuadBLAS::dot(x, x))). */
void quadblas_nrm2_q(size_t n, const void *x, size_t incx, void *result);

/* QuadBLAS::gemv (level2.hpp:85-99; Matrix<L>::gemv, cpp_classes.hpp:143-147).
   layout: 0 = RowMajor, 1 = ColMajor.  alpha, beta point to binary128 values. */
void quadblas_gemv_q(int layout, size_t m, size_t n, const void *alpha, const void *A, size_t lda,
                     const void *x, size_t incx, const void *beta, void *y, size_t incy);

/* QuadBLAS::gemm (level3.hpp:215-336; Matrix<L>::gemm, cpp_classes.hpp:150-154). */
void quadblas_gemm_q(int layout, size_t m, size_t n, size_t k, const void *alpha, const void *A, size_t lda,
                     const void *B, size_t ldb, const void *beta, void *C, size_t ldc);

/* ---- Part 3: device-pointer API (new; no reference counterpart) --------------------------- */

/* Evaluation modes.
   QBLAS_MODE_REFERENCE: the reference's summation order and blocking, every operation correctly
                         rounded -> bit-identical to the reference (dot additionally needs the
                         same T).  Default for Parts 1 and 2.
   QBLAS_MODE_EXACT:     dot/gemv: the exact sum of products rounded once to binary128 (correctly
                         rounded result, independent of order and T).
   QBLAS_MODE_DD:        double-double arithmetic on FP64 FMA (inputs rounded to 106 bits); faster,
                         ~2^-104 relative accuracy on well-conditioned data.
   QBLAS_MODE_OZAKI:     gemm only: each operand scaled per row / per column by a power of two and
                         rounded to an integer, the integer products formed on the int8 matrix
                         cores with exact int32 accumulation, recombined exactly and rounded once.
                         Two methods (qblas_set_ozaki_method): the integers split into S int8
                         slices (slice-pair products), or carried as residues modulo up to 47
                         pairwise coprime moduli <= 256 (one int8 GEMM per modulus, Chinese
                         remainder reconstruction).  Every entry is certified to <= 2^-104
                         relative error by an a-priori bound, or recomputed exactly (correctly
                         rounded) when the bound cannot certify it. */
enum {
  QBLAS_MODE_REFERENCE = 0,
  QBLAS_MODE_EXACT = 1,
  QBLAS_MODE_DD = 2,
  QBLAS_MODE_OZAKI = 3
};

/* Status codes returned by Part 3 (0 = success). */
enum {
  QBLAS_OK = 0,
  QBLAS_ERR_ARG = 1,      /* invalid argument (negative size/stride, misaligned pointer, ...) */
  QBLAS_ERR_HIP = 2,      /* a HIP runtime call failed */
  QBLAS_ERR_MODE = 3,     /* mode not available for this routine */
  QBLAS_ERR_RANGE = 4     /* EXACT form and the OZAKI form's exact repair: a term fell outside the
                             full-range accumulator.  Cannot happen for binary128 inputs: products
                             too far apart for the fast accumulation windows are recomputed on a
                             full-range path (qblas_hip_exact_wide_items counts them) */
};

/* Mode used by Parts 1 and 2 (default QBLAS_MODE_REFERENCE; environment QBLAS_MODE overrides at
   load: "reference", "exact", "dd", "ozaki").  Per routine: with QBLAS_MODE_OZAKI, dot / gemv /
   nrm2 use QBLAS_MODE_EXACT; with QBLAS_MODE_EXACT, gemm uses QBLAS_MODE_OZAKI.  Part 3 takes the
   mode as given and reports QBLAS_ERR_MODE where a routine lacks it. */
int qblas_set_mode(int mode);
int qblas_get_mode(void);

/* k-block length of the reference's blocked GEMM (BlockingParams::kc, blocking.hpp:39-41):
   126 on x86-64 Linux (the default), 256 on Apple silicon. */
void qblas_set_kc(int kc);
int qblas_get_kc(void);

/* Slice count S of QBLAS_MODE_OZAKI (8..28); 0 (the default) lets the library choose per call:
   it computes 256 sampled entries exactly and takes the smallest S whose error bound certifies
   all of them (this probe synchronizes with the stream).  Returns QBLAS_ERR_ARG for other values. */
int qblas_set_ozaki_slices(int slices);
int qblas_get_ozaki_slices(void);

/* Method of QBLAS_MODE_OZAKI.  QBLAS_OZAKI_SLICES: int8 slices (precision 8 S bits).
   QBLAS_OZAKI_CRT: residues modulo the moduli (precision t bits, 48..158; t bits at depth k need
   the fewest moduli whose product is >= k 2^(2t-2)); when the precision the data needs exceeds
   158 bits the slice method is used.  QBLAS_OZAKI_AUTO (the default): a fixed slice count selects
   the slice method and fixed CRT bits the CRT method; otherwise the probe sets both precisions
   and the method with less matrix-core work is used.  Returns QBLAS_ERR_ARG for other values. */
enum { QBLAS_OZAKI_AUTO = 0, QBLAS_OZAKI_SLICES = 1, QBLAS_OZAKI_CRT = 2 };
int qblas_set_ozaki_method(int method);
int qblas_get_ozaki_method(void);
/* Bits t of the CRT method (48..158); 0 (the default): chosen per call by the probe. */
int qblas_set_ozaki_crt_bits(int bits);
int qblas_get_ozaki_crt_bits(void);
/* Method (QBLAS_OZAKI_SLICES or QBLAS_OZAKI_CRT), CRT bits and CRT moduli used by the last
   QBLAS_MODE_OZAKI gemm on this thread (0 where not applicable). */
int qblas_hip_ozaki_last_method(void);
int qblas_hip_ozaki_last_crt_bits(void);
int qblas_hip_ozaki_last_moduli(void);

/* After a QBLAS_MODE_OZAKI gemm on this thread: synchronizes the stream and returns how many
   entries were recomputed exactly because the bound could not certify them (-1 if there was no
   such call or on a HIP error). */
long long qblas_hip_ozaki_repairs(void *stream);
/* Slice count used by the last QBLAS_MODE_OZAKI gemm on this thread (0 if none or CRT). */
int qblas_hip_ozaki_last_slices(void);

/* Phase timing of QBLAS_MODE_OZAKI gemm calls (off by default; costs four event records per
   call).  qblas_hip_phase_times writes the durations in ms of the last such call on this thread
   -- [0] scaling and slicing (or residues), [1] products (matrix cores), [2] finish and
   certification (with the CRT reconstruction), [3] exact repair -- after synchronizing with it;
   returns 4, or -1 if no timed call exists. */
void qblas_hip_set_phase_timing(int on);
int qblas_hip_phase_times(float *ms);

/* Last error message of the calling thread ("" if none); cleared by qblas_clear_error(). */
const char *quadblas_hip_last_error(void);
void qblas_clear_error(void);

/* After a device-API call in QBLAS_MODE_EXACT or QBLAS_MODE_OZAKI on this thread (the last such
   call; any later host-pointer call replaces it): synchronizes the stream and returns 1 if a term
   fell outside the full-range accumulator (the output was set to NaN; cannot happen for binary128
   inputs), 0 if the result is exact or certified, -1 on a HIP error, when a later call on the
   same stream has reused the call's workspace, or when a graph replay of a batched OZAKI call found
   the descriptors' shapes changed (nothing computed).  Host-pointer entry points check this
   themselves. */
int qblas_hip_exact_status(void *stream);

/* After the same kind of call: synchronizes the stream and returns how many items (dots, gemv
   rows, GEMM entries of the exact repair) were recomputed on the full-range path because their
   products lay too far apart for the accumulation windows (about 2^+-124 around an anchor); -1 if
   unknown (as for qblas_hip_exact_status).  A dot counts as one item. */
long long qblas_hip_exact_wide_items(void *stream);

/* dot: *d_result (device, 16 bytes) = x . y.  T: the reference thread count for
   QBLAS_MODE_REFERENCE (ignored otherwise).  stream: hipStream_t or NULL. */
int qblas_hip_dot(int mode, int64_t n, const void *d_x, int64_t incx, const void *d_y, int64_t incy, int T,
                  void *d_result, void *stream);

/* axpy on device pointers; alpha is a HOST pointer to a binary128 value. */
int qblas_hip_axpy(int64_t n, const void *alpha, const void *d_x, int64_t incx, void *d_y, int64_t incy,
                   void *stream);

/* nrm2: *d_result = RN(sqrt(dot(x, x))) with the dot of `mode` (T as for qblas_hip_dot). */
int qblas_hip_nrm2(int mode, int64_t n, const void *d_x, int64_t incx, int T, void *d_result, void *stream);

/* gemv on device pointers; alpha, beta are HOST pointers to binary128 values. */
int qblas_hip_gemv(int mode, int layout, int64_t m, int64_t n, const void *alpha, const void *d_A, int64_t lda,
                   const void *d_x, int64_t incx, const void *beta, void *d_y, int64_t incy, void *stream);

/* gemm on device pointers; alpha, beta are HOST pointers to binary128 values. */
int qblas_hip_gemm(int mode, int layout, int64_t m, int64_t n, int64_t k, const void *alpha, const void *d_A,
                   int64_t lda, const void *d_B, int64_t ldb, const void *beta, void *d_C, int64_t ldc,
                   void *stream);

/* ---- Allocation-free dot (new; no reference counterpart) ------------------------------------
   For callers that cannot allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc): qblas_hip_dot's computation (modes REFERENCE, EXACT and
   DD, same results bit for bit) with its scratch in a static device array of
   QBLAS_NOALLOC_SCRATCH_BYTES per GPU that the library holds until the process ends; it calls no
   allocating routine.  The EXACT and DD forms fit any n; REFERENCE needs 32 T bytes (T <= 65536).
   A call that needs more returns QBLAS_ERR_ARG.  Calls on one GPU share its array: calls that
   overlap on one GPU (two streams or two threads) race and must be ordered by the caller.
   qblas_hip_exact_status / qblas_hip_exact_wide_items do not report on these calls. */
#define QBLAS_NOALLOC_SCRATCH_BYTES (2u << 20)
int qblas_hip_dot_noalloc(int mode, int64_t n, const void *d_x, int64_t incx, const void *d_y, int64_t incy, int T,
                          void *d_result, void *stream);

/* ---- Batched problems (new; no reference counterpart) --------------------------------------
   Many independent problems of different sizes in one call: a device array of descriptors, each
   problem evaluated with the semantics of the single-problem routine in the given mode (REFERENCE:
   level3.hpp:215-336 / level2.hpp:15-99 per problem, bit-identical to a loop of single calls).
   All problems share the layout.  The work is one grouped persistent kernel over the tiles (gemm)
   or rows (gemv) of all problems; the launch sequence allocates nothing when the workspace of an
   earlier call of the same count on the same stream suffices, synchronizes nothing and may be
   captured in a hipGraph (all modes but gemm's QBLAS_MODE_OZAKI, below).  Workspaces are kept per
   (device, stream) of the calling thread; one used while its stream was capturing is never freed
   (so the graph stays valid).  A capture cannot allocate: the capturing stream uses its own
   buffer if large enough, else takes over another stream's (the least recently used large enough
   one), else the call fails with QBLAS_ERR_ARG -- make the same call once, on any stream of the
   thread, before capturing.
   alpha/beta are binary128 {lo, hi} words. */
typedef struct qblas_gemm_problem {
  int64_t m, n, k;
  int64_t lda, ldb, ldc;
  const void *A;
  const void *B;
  void *C;
  uint64_t alpha[2];
  uint64_t beta[2];
  int64_t reserved;
} qblas_gemm_problem;

typedef struct qblas_gemv_problem {
  int64_t m, n;
  int64_t lda, incx, incy;
  const void *A;
  const void *x;
  void *y;
  uint64_t alpha[2];
  uint64_t beta[2];
} qblas_gemv_problem;

/* d_problems: device array of count descriptors; gemm: modes REFERENCE, DD and OZAKI; gemv:
   REFERENCE, DD and EXACT (rows whose products lie too far apart for the accumulation windows are
   found again by a second walk over the batch and computed on the full-range path).
   gemm in QBLAS_MODE_OZAKI: the slice method with one slice count S for the whole batch (the tuned
   or selected configuration; configuration 0 is S = 17), every entry within 2^-104 relative error
   or recomputed exactly, per problem bit-identical to qblas_hip_gemm in QBLAS_MODE_OZAKI with
   method QBLAS_OZAKI_SLICES and slice count S.  An eager call synchronizes the stream once (the
   per-problem work counts are summed on the device and read back to size the workspace) and
   remembers the totals for its (descriptor array, count, layout, configuration).  A call under
   graph capture uses the totals of the last eager call with the same descriptor array, count,
   layout and configuration (QBLAS_ERR_ARG if there was none) and neither synchronizes nor
   allocates; each replay checks on the device that the descriptors still give those totals, and
   if they do not, computes nothing (C untouched) and qblas_hip_exact_status returns QBLAS_ERR_ARG.
   qblas_hip_ozaki_last_slices, qblas_hip_ozaki_repairs and qblas_hip_exact_status report on the
   batch. */
int qblas_hip_gemm_batched(int mode, int layout, int64_t count, const qblas_gemm_problem *d_problems, void *stream);
int qblas_hip_gemv_batched(int mode, int layout, int64_t count, const qblas_gemv_problem *d_problems, void *stream);

/* Autotuner for the batched gemm: times every configuration of the mode (tile shapes; for
   QBLAS_MODE_OZAKI slice counts) on a replica of the batch's first 512 problems (same shapes,
   synthetic data in a scratch workspace; QBLAS_MODE_OZAKI reads the caller's A and B, since the
   share of entries sent to the exact repair depends on the data; the caller's C is not touched),
   keeps the fastest for calls with this (mode, layout, count), and returns its index (>= 0) or an
   error code (< 0).  Synchronizes; call it outside graph capture.  Without a tuned entry the
   batched call uses configuration 0. */
int qblas_hip_gemm_batched_tune(int mode, int layout, int64_t count, const qblas_gemm_problem *d_problems,
                                void *stream);
/* Description of configuration idx of a mode ("TMxTN ...", or "S slices" for QBLAS_MODE_OZAKI),
   or NULL. */
const char *qblas_hip_gemm_batched_config(int mode, int idx);
/* Use configuration idx for batched gemm calls with this (mode, layout, count), as the autotuner
   does with its choice.  QBLAS_ERR_ARG for an unknown mode or index. */
int qblas_hip_gemm_batched_select(int mode, int layout, int64_t count, int idx);

/* Seeded synthetic inputs written to device memory (definitions shared with the test oracle):
   kind 0: a + b*2^-60 summed in binary128, a, b uniform [-1, 1) doubles (full 113-bit mantissa);
   kind 1: random sign, exponent uniform in [lo_exp, hi_exp), 112 random fraction bits;
   kind 2: the qdot cancellation variant's x: +-(1e20*u) pairs on slots 0,1 of every 8, kind 0
           elsewhere; kind 3: the matching y (equal values on each pair); kind 4: kind 2 with
           +-(1e22*u) pairs. */
int qblas_hip_fill(int kind, uint64_t seed, int64_t n, int lo_exp, int hi_exp, void *d_out, void *stream);
/* The same generators over a rows x cols matrix: value number r*cols + c is stored at element
   (r, c) (layout 0: r*ld + c, 1: r + c*ld).  With kind 2 for A (rows = m, cols = k) and kind 3 for
   B^T (rows = n, cols = k) every product pair A_il B_lj, l % 8 in {0, 1}, cancels exactly (exact
   sums ~1e-18 of the largest products); kind 4 instead of 2 gives ~1e-20: the ill-conditioned
   qgemm workload. */
int qblas_hip_fill_2d(int kind, uint64_t seed, int64_t rows, int64_t cols, int64_t ld, int layout, int lo_exp,
                      int hi_exp, void *d_out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* QBLAS_HIP_H */
