// quadblas.hpp -- C++ drop-in for the QuadBLAS hot path on MI355X.
//
// Same names, signatures and semantics as the reference's C++ API for dot / gemv / gemm:
//   enum class Layout                          include/quadblas/core/types.hpp:8-12
//   configuration constants                    include/quadblas/core/constants.hpp:10-18
//   VERSION, VERSION_MAJOR/MINOR/PATCH         include/quadblas/quadblas.hpp:54-59
//   QuadBLAS::dot                              include/quadblas/algorithms/level1.hpp:80-137
//   QuadBLAS::axpy                             include/quadblas/algorithms/level1.hpp:190-223
//   QuadBLAS::gemv                             include/quadblas/algorithms/level2.hpp:85-99
//   QuadBLAS::gemm                             include/quadblas/algorithms/level3.hpp:215-336
//   QuadBLAS::get_num_threads / set_num_threads  include/quadblas/threading/openmp_utils.hpp:10-24
//   QuadBLAS::aligned_alloc / aligned_free     include/quadblas/memory/allocation.hpp:18-41
//   Vector<L>, Matrix<L> and the aliases       include/quadblas/interface/cpp_classes.hpp:18-168
// The arithmetic runs on the GPU through libqblas_hip.so (qblas_hip.h): buffers stay caller-owned
// host memory, calls are synchronous, results are valid on return.  The evaluation form is the
// library's mode (qblas_set_mode / QBLAS_MODE): by default the reference's own order, bit-identical
// to a correctly rounded binary128 evaluation of the reference's loops.
//
// Link with -lqblas_hip (and the HIP runtime it depends on).  Sleef_quad is __float128 on x86-64,
// declared under the same guard as SLEEF's own header so both can be included together.
#ifndef QUADBLAS_MI355X_HPP
#define QUADBLAS_MI355X_HPP

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>

#include "../qblas_hip.h"

#if !defined(Sleef_quad_DEFINED)
#define Sleef_quad_DEFINED
typedef struct {
  uint64_t x, y;
} Sleef_uint64_2t;
typedef __float128 Sleef_quad;
#endif
#ifndef SLEEF_QUAD_C
#if defined(__STRICT_ANSI__)
// strict ISO mode has no Q literal suffix: the literal goes through double, as the reference's
// own fallback (core/platform.hpp:21-23, Sleef_cast_from_doubleq1)
#define SLEEF_QUAD_C(x) ((Sleef_quad)(x))
#else
#define SLEEF_QUAD_C(x) (x##Q)
#endif
#endif

static_assert(sizeof(Sleef_quad) == 16, "Sleef_quad must be IEEE binary128");

namespace QuadBLAS {

enum class Layout { RowMajor, ColMajor };

// ---- configuration constants, with the reference's values ------------------------------------
// Here they describe the REFERENCE evaluation form, which keeps the reference's order:
// VECTOR_SIZE lane chains per dot (two), PARALLEL_THRESHOLD the length from which a dot is cut
// into one chunk per thread, GEMM_BLOCK_SIZE the largest m, n, k of the unblocked GEMM, and
// L1_CACHE_SIZE the cache size whose blocking gives kc = 126 (qblas_set_kc changes kc).
// ALIGNMENT is the byte alignment of aligned_alloc and of quadblas_is_aligned.
constexpr size_t VECTOR_SIZE = 2;
constexpr size_t CACHE_LINE_SIZE = 64;
constexpr size_t L1_CACHE_SIZE = 32768;
constexpr size_t L2_CACHE_SIZE = 262144;
constexpr size_t PARALLEL_THRESHOLD = 500;
constexpr size_t GEMM_BLOCK_SIZE = 64;
constexpr size_t ALIGNMENT = 32;

// ---- version of the API this library implements ------------------------------------------------
constexpr const char *VERSION = "1.0.0";
constexpr int VERSION_MAJOR = 1;
constexpr int VERSION_MINOR = 0;
constexpr int VERSION_PATCH = 0;

template <Layout layout>
class Matrix;
template <Layout layout>
class Vector;

// ---- threading: the thread count fixes the reference's dot chunking (level1.hpp:46, :96) ----
inline int get_num_threads() { return quadblas_get_num_threads(); }
inline void set_num_threads(int num_threads) { quadblas_set_num_threads(num_threads); }

// ---- ALIGNMENT-byte aligned host buffers (nullptr on failure) ---------------------------------
template <typename T>
inline T *aligned_alloc(size_t count) {
  void *p = nullptr;
  const size_t bytes = ((count * sizeof(T) + ALIGNMENT - 1) / ALIGNMENT) * ALIGNMENT;
  if (posix_memalign(&p, ALIGNMENT, bytes ? bytes : ALIGNMENT) != 0) return nullptr;
  return static_cast<T *>(p);
}
template <typename T>
inline void aligned_free(T *ptr) {
  free(ptr);
}

// ---- level 1/2/3 ------------------------------------------------------------------------------
inline Sleef_quad dot(size_t n, const Sleef_quad *x, size_t incx, const Sleef_quad *y, size_t incy) {
  Sleef_quad r = 0;
  if (n == 0) return r;
  quadblas_dot_q(n, x, incx, y, incy, &r);
  return r;
}

inline void axpy(size_t n, Sleef_quad alpha, const Sleef_quad *x, size_t incx, Sleef_quad *y, size_t incy) {
  if (n == 0) return;
  quadblas_axpy_q(n, &alpha, x, incx, y, incy);
}

inline void gemv(Layout layout, size_t m, size_t n, Sleef_quad alpha, const Sleef_quad *A, size_t lda,
                 const Sleef_quad *x, size_t incx, Sleef_quad beta, Sleef_quad *y, size_t incy) {
  quadblas_gemv_q(layout == Layout::ColMajor ? 1 : 0, m, n, &alpha, A, lda, x, incx, &beta, y, incy);
}

inline void gemm(Layout layout, size_t m, size_t n, size_t k, Sleef_quad alpha, const Sleef_quad *A, size_t lda,
                 const Sleef_quad *B, size_t ldb, Sleef_quad beta, Sleef_quad *C, size_t ldc) {
  quadblas_gemm_q(layout == Layout::ColMajor ? 1 : 0, m, n, k, &alpha, A, lda, B, ldb, &beta, C, ldc);
}

// ---- containers -------------------------------------------------------------------------------
// Owning (zero-filled, 32-byte aligned) or viewing a strided host array; move-only.  The sized
// constructor is not explicit, as in the reference (cpp_classes.hpp:28): `VectorRowMajor v = n;`
// compiles.
template <Layout layout>
class Vector {
 public:
  Vector(size_t size) : data_(aligned_alloc<Sleef_quad>(size)), size_(size), stride_(1), owns_(true) {
    if (data_) std::fill(data_, data_ + size, Sleef_quad(0));
  }
  Vector(Sleef_quad *data, size_t size, size_t stride = 1) : data_(data), size_(size), stride_(stride), owns_(false) {}
  ~Vector() {
    if (owns_) aligned_free(data_);
  }
  Vector(Vector &&o) noexcept : data_(o.data_), size_(o.size_), stride_(o.stride_), owns_(o.owns_) { o.owns_ = false; }
  Vector(const Vector &) = delete;
  Vector &operator=(const Vector &) = delete;

  Sleef_quad &operator[](size_t i) { return data_[i * stride_]; }
  const Sleef_quad &operator[](size_t i) const { return data_[i * stride_]; }
  size_t size() const { return size_; }
  size_t stride() const { return stride_; }
  Sleef_quad *data() { return data_; }
  const Sleef_quad *data() const { return data_; }

  Sleef_quad dot(const Vector &other) const { return QuadBLAS::dot(size_, data_, stride_, other.data_, other.stride_); }
  // this = alpha * other + this
  void axpy(Sleef_quad alpha, const Vector &other) {
    QuadBLAS::axpy(size_, alpha, other.data_, other.stride_, data_, stride_);
  }
  // correctly rounded square root of dot(*this)
  Sleef_quad norm() const {
    Sleef_quad r = 0;
    if (size_) quadblas_nrm2_q(size_, data_, stride_, &r);
    return r;
  }

 private:
  Sleef_quad *data_;
  size_t size_, stride_;
  bool owns_;
};

// Element (i, j) at i*ld + j (RowMajor) or j*ld + i (ColMajor).  As in the reference
// (cpp_classes.hpp:96, :104), the leading dimension defaults to cols for both layouts.
template <Layout layout>
class Matrix {
 public:
  Matrix(size_t rows, size_t cols)
      : data_(aligned_alloc<Sleef_quad>(rows * cols)), rows_(rows), cols_(cols), ld_(cols), owns_(true) {
    if (data_) std::fill(data_, data_ + rows * cols, Sleef_quad(0));
  }
  Matrix(Sleef_quad *data, size_t rows, size_t cols, size_t ld = 0)
      : data_(data), rows_(rows), cols_(cols), ld_(ld == 0 ? cols : ld), owns_(false) {}
  ~Matrix() {
    if (owns_) aligned_free(data_);
  }
  Matrix(Matrix &&o) noexcept : data_(o.data_), rows_(o.rows_), cols_(o.cols_), ld_(o.ld_), owns_(o.owns_) {
    o.owns_ = false;
  }
  Matrix(const Matrix &) = delete;
  Matrix &operator=(const Matrix &) = delete;

  Sleef_quad &operator()(size_t i, size_t j) { return data_[index(i, j)]; }
  const Sleef_quad &operator()(size_t i, size_t j) const { return data_[index(i, j)]; }
  size_t rows() const { return rows_; }
  size_t cols() const { return cols_; }
  size_t leading_dimension() const { return ld_; }
  Sleef_quad *data() { return data_; }
  const Sleef_quad *data() const { return data_; }

  void gemv(Sleef_quad alpha, const Vector<layout> &x, Sleef_quad beta, Vector<layout> &y) const {
    QuadBLAS::gemv(layout, rows_, cols_, alpha, data_, ld_, x.data(), x.stride(), beta, y.data(), y.stride());
  }
  void gemm(Sleef_quad alpha, const Matrix &B, Sleef_quad beta, Matrix &C) const {
    QuadBLAS::gemm(layout, rows_, B.cols(), cols_, alpha, data_, ld_, B.data(), B.leading_dimension(), beta,
                   C.data(), C.leading_dimension());
  }

 private:
  size_t index(size_t i, size_t j) const { return layout == Layout::RowMajor ? i * ld_ + j : j * ld_ + i; }
  Sleef_quad *data_;
  size_t rows_, cols_, ld_;
  bool owns_;
};

using VectorRowMajor = Vector<Layout::RowMajor>;
using VectorColMajor = Vector<Layout::ColMajor>;
using MatrixRowMajor = Matrix<Layout::RowMajor>;
using MatrixColMajor = Matrix<Layout::ColMajor>;
template <Layout layout = Layout::RowMajor>
using DefaultVector = Vector<layout>;
template <Layout layout = Layout::RowMajor>
using DefaultMatrix = Matrix<layout>;

}  // namespace QuadBLAS

#endif  // QUADBLAS_MI355X_HPP
