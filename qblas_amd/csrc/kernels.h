// kernels.h -- launchers shared between the kernel translation units and the C ABI (capi.hip).
// All pointers are device pointers to packed binary128 values; all launches are asynchronous on
// the given stream and return hipGetLastError().
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "q128.hpp"
#include "../../include/qblas_hip.h"

namespace qbk {
using qb::q128;

// One launch dimension holds at most 2^32 - 1 work-items: the dispatch packet's grid size is a
// 32-bit work-item count, and a larger launch covers only the count modulo 2^32, without an
// error.  Element-wise kernels therefore stride over their elements from a capped grid
// (stride_blocks); the launchers of the other kernels check grid_fits and fail otherwise.
constexpr int64_t QB_GRID_ITEMS_MAX = 0xffffffffll;
inline bool grid_fits(int64_t blocks, int64_t block_size) {
  return blocks >= 0 && blocks <= QB_GRID_ITEMS_MAX / block_size;
}
inline unsigned stride_blocks(int64_t items, int block_size) {
  const int64_t b = (items + block_size - 1) / block_size;
  return (unsigned)(b < 65536 ? b : 65536);
}

// reference order (kernels_ref.hip)
hipError_t launch_dot_ref(const q128 *x, int64_t incx, const q128 *y, int64_t incy, int64_t n, int T,
                          q128 *part /* 2*T */, q128 *out, hipStream_t st);
hipError_t launch_gemv_ref(int colmajor, int64_t m, int64_t n, q128 alpha, const q128 *A, int64_t lda,
                           const q128 *x, int64_t incx, q128 beta, q128 *y, int64_t incy,
                           q128 *ax /* n, colmajor only */, hipStream_t st);
hipError_t launch_gemm_ref(int colmajor, int64_t m, int64_t n, int64_t k, int64_t kc, q128 alpha, const q128 *A,
                           int64_t lda, const q128 *B, int64_t ldb, q128 beta, q128 *C, int64_t ldc,
                           hipStream_t st);
hipError_t launch_zero(q128 *p, int64_t n, hipStream_t st);
hipError_t launch_axpy(int64_t n, q128 alpha, const q128 *x, int64_t incx, q128 *y, int64_t incy, hipStream_t st);
hipError_t launch_sqrt1(q128 *p, hipStream_t st);  // *p = RN(sqrt(*p))
hipError_t launch_fill(int kind, uint64_t seed, int64_t n, int lo_exp, int hi_exp, q128 *out, hipStream_t st);
hipError_t launch_fill_2d(int kind, uint64_t seed, int64_t rows, int64_t cols, int64_t ld, int colmajor, int lo_exp,
                          int hi_exp, q128 *out, hipStream_t st);

// double-double form (kernels_dd.hip); ws sizes from the *_ws functions
size_t dot_dd_ws(int64_t n);
hipError_t launch_dot_dd(const q128 *x, int64_t incx, const q128 *y, int64_t incy, int64_t n, void *ws, q128 *out,
                         hipStream_t st);
size_t gemv_dd_ws(int colmajor, int64_t m, int64_t n);
hipError_t launch_gemv_dd(int colmajor, int64_t m, int64_t n, q128 alpha, const q128 *A, int64_t lda,
                          const q128 *x, int64_t incx, q128 beta, q128 *y, int64_t incy, void *ws,
                          hipStream_t st);

size_t gemm_dd_ws(int64_t m, int64_t n, int64_t k);
hipError_t launch_gemm_dd(int colmajor, int64_t m, int64_t n, int64_t k, q128 alpha, const q128 *A, int64_t lda,
                          const q128 *B, int64_t ldb, q128 beta, q128 *C, int64_t ldc, void *ws, hipStream_t st);

// exact (correctly rounded) form (kernels_exact.hip); ws[0] (u32) is the status word:
// bit 0 set when a product fell outside its window (the output is then NaN)
size_t dot_exact_ws(int64_t n);
hipError_t launch_dot_exact(const q128 *x, int64_t incx, const q128 *y, int64_t incy, int64_t n, void *ws,
                            q128 *out, hipStream_t st);
size_t gemv_exact_ws(int colmajor, int64_t m, int64_t n);
hipError_t launch_gemv_exact(int colmajor, int64_t m, int64_t n, q128 alpha, const q128 *A, int64_t lda,
                             const q128 *x, int64_t incx, q128 beta, q128 *y, int64_t incy, void *ws,
                             hipStream_t st);

// Called by the gemm launchers right before they enqueue the first kernel that reads C: a caller
// still copying C in on another stream makes st wait for that copy there.
struct c_gate {
  hipError_t (*fn)(void *arg, hipStream_t st);
  void *arg;
};

// OZAKI form of gemm (kernels_ozaki.hip): S int8 slices per operand.  ws[0] (u32) is the status
// word as for the exact form; ws[1] counts the entries the call recomputed exactly.
constexpr int OZ_SMIN = 8, OZ_SMAX = 28;
size_t gemm_ozaki_ws(int64_t m, int64_t n, int64_t k, int S);
hipError_t launch_gemm_ozaki(int colmajor, int64_t m, int64_t n, int64_t k, int S, q128 alpha, const q128 *A,
                             int64_t lda, const q128 *B, int64_t ldb, q128 beta, q128 *C, int64_t ldc, void *ws,
                             hipStream_t st, hipEvent_t *ev /* null, or 5 events: start, after slicing, after
                             the slice products, after the finish, after the repair */,
                             const void *probe_ws /* null, or the workspace of a probe of the same operands:
                             its scaling exponents are used (it must stay intact until the call's work ends) */,
                             const c_gate *gate = nullptr);
hipError_t ozaki_repair_count(const void *ws, uint32_t *out, hipStream_t st);
// probe: scales of A and B, then for `samples` (<= 1024) exact entries ilogb|AB_ij| - (E_i + F_j)
// as ints at ws + ozaki_probe_out_offset (INT_MAX where unusable); ozaki_slices_for turns a chosen
// order statistic of them into a slice count
constexpr int OZ_CRT_TMIN = 48, OZ_CRT_TMAX = 158;
// OZAKI form, CRT method (residues modulo nq pairwise coprime moduli <= 256, one int8 GEMM per
// modulus): t bits per scaled operand (48..158), workspace bytes (0 when t needs too many moduli)
size_t gemm_crt_ws(int64_t m, int64_t n, int64_t k, int t);
int crt_moduli_for(int64_t k, int t);
hipError_t launch_gemm_crt(int colmajor, int64_t m, int64_t n, int64_t k, int t, q128 alpha, const q128 *A,
                           int64_t lda, const q128 *B, int64_t ldb, q128 beta, q128 *C, int64_t ldc, void *ws,
                           hipStream_t st, hipEvent_t *ev, const void *probe_ws, const c_gate *gate = nullptr,
                           const void *tab = nullptr /* null, or the tables of crt_moduli_for(k, t) moduli built by
                           launch_crt_tables and complete before the call's work */);
// the CRT method's per-modulus constants for nq moduli (crt_tab_bytes() bytes at out); they depend
// on nq only
size_t crt_tab_bytes();
hipError_t launch_crt_tables(int nq, void *out, hipStream_t st);
// bits t of the CRT method for a probe value (as ozaki_slices_for); 0: use the slice method
int ozaki_crt_bits_for(int64_t k, int r);
size_t ozaki_probe_ws(int64_t m, int64_t n);
size_t ozaki_probe_out_offset(int64_t m, int64_t n);
hipError_t launch_ozaki_probe(int colmajor, int64_t m, int64_t n, int64_t k, const q128 *A, int64_t lda,
                              const q128 *B, int64_t ldb, int samples, void *ws, hipStream_t st,
                              int *out = nullptr /* null: ws + ozaki_probe_out_offset; else e.g. pinned host memory */);
int ozaki_slices_for(int64_t k, int r);

// batched problems (kernels_batched.hip); ws holds count + 1 prefix sums
size_t batched_ws(int64_t count);
int gemm_batched_configs(int dd_mode);
const char *gemm_batched_config_name(int dd_mode, int idx);
hipError_t launch_gemm_batched(int dd_mode, int colmajor, int64_t count, const qblas_gemm_problem *P, int cfg,
                               int64_t kc, void *ws, hipStream_t st);
// form 0 REFERENCE, 1 DD, 2 EXACT (status word at ws + batched_status_offset(count))
size_t batched_status_offset(int64_t count);
hipError_t launch_gemv_batched(int form, int colmajor, int64_t count, const qblas_gemv_problem *P, void *ws,
                               hipStream_t st);
// batched OZAKI form (kernels_ozaki.hip, slice method, one slice count for the batch): the scan
// writes per-problem prefix sums into ws (gemm_ozaki_batched_scan_ws bytes); the caller reads the
// 9 totals at ws + gemm_ozaki_batched_totals_offset, grows ws to gemm_ozaki_batched_ws (running the
// scan again when ws moved) and launches the rest.  ws word 0 is the status word.
constexpr int OZ_BATCHED_TOTALS = 9;
constexpr int OZ_HDR_STALE = 8;  // header word: planned totals did not match (nothing computed)
int ozaki_batched_configs();
int ozaki_batched_slices(int cfg);  // slice count of configuration cfg (0 = default)
size_t gemm_ozaki_batched_scan_ws(int64_t count);
size_t gemm_ozaki_batched_totals_offset(int64_t count);
hipError_t launch_gemm_ozaki_batched_scan(int colmajor, int64_t count, const qblas_gemm_problem *P, void *ws,
                                          hipStream_t st);
size_t gemm_ozaki_batched_ws(int64_t count, const int64_t *tot, int S);
hipError_t launch_gemm_ozaki_batched(int colmajor, int64_t count, const qblas_gemm_problem *P, int S,
                                     const int64_t *tot, void *ws, hipStream_t st, bool planned = false);

// hooks into the C ABI's thread-local state (capi.hip): the last-error message, and forgetting the
// EXACT-form status of earlier calls
void record_error(const char *msg);
void forget_exact_status();
}  // namespace qbk
