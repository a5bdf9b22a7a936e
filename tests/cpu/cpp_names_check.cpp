// Every public name of the reference's C++ headers that a caller can use -- interface/cpp_classes.hpp,
// core/types.hpp, core/constants.hpp, quadblas.hpp (version) -- used the way a caller of the
// reference would, compiled against include/quadblas/quadblas.hpp with g++ -Wall -Wextra -Werror.
// Runs without a GPU: nothing here computes (containers, constants, threading, alignment).
// Exit status 0 when every value matches the reference's.
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <utility>

#include "quadblas/quadblas.hpp"

static int failures = 0;
static void check(bool ok, const char *what) {
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", what);
  if (!ok) ++failures;
}

template <QuadBLAS::Layout L>
static size_t vec_size(const QuadBLAS::Vector<L> &v) {
  return v.size();
}

int main() {
  using namespace QuadBLAS;
  // quadblas.hpp:54-59
  check(std::strcmp(VERSION, "1.0.0") == 0, "VERSION");
  check(VERSION_MAJOR == 1 && VERSION_MINOR == 0 && VERSION_PATCH == 0, "VERSION_MAJOR/MINOR/PATCH");
  // constants.hpp:10-18
  static_assert(std::is_same<decltype(VECTOR_SIZE), const size_t>::value, "size_t constants");
  check(VECTOR_SIZE == 2 && CACHE_LINE_SIZE == 64 && L1_CACHE_SIZE == 32768 && L2_CACHE_SIZE == 262144,
        "VECTOR_SIZE, CACHE_LINE_SIZE, L1_CACHE_SIZE, L2_CACHE_SIZE");
  check(PARALLEL_THRESHOLD == 500 && GEMM_BLOCK_SIZE == 64 && ALIGNMENT == 32,
        "PARALLEL_THRESHOLD, GEMM_BLOCK_SIZE, ALIGNMENT");
  // types.hpp:8-12
  const Layout ls[2] = {Layout::RowMajor, Layout::ColMajor};
  check(ls[0] != ls[1], "Layout::RowMajor / Layout::ColMajor");
  // cpp_classes.hpp:28: the sized constructor takes copy-initialisation
  VectorRowMajor v = 5;
  VectorColMajor vc = 3;
  DefaultVector<> dv = 4;
  DefaultVector<Layout::ColMajor> dvc = 2;
  check(vec_size<Layout::RowMajor>(7) == 7, "Vector<L> from a size_t argument");
  check(v.size() == 5 && v.stride() == 1 && vc.size() == 3 && dv.size() == 4 && dvc.size() == 2, "sizes, strides");
  bool zero = true;
  for (size_t i = 0; i < v.size(); ++i) zero &= v[i] == 0;
  check(zero, "owning Vector is zero-filled");
  check(quadblas_is_aligned(v.data()) == 1, "owning Vector is ALIGNMENT-aligned");
  const VectorRowMajor &cv = v;
  check(cv.data() == v.data() && &cv[1] == &v[1], "const accessors");
  // views, strides, moves
  Sleef_quad buf[12] = {};
  VectorRowMajor view(buf, 4, 3);
  view[2] = SLEEF_QUAD_C(1.5);
  check(buf[6] == SLEEF_QUAD_C(1.5) && view.stride() == 3, "strided view");
  VectorRowMajor moved(std::move(v));
  check(moved.size() == 5, "move constructor");
  static_assert(!std::is_copy_constructible<VectorRowMajor>::value, "Vector is move-only");
  static_assert(!std::is_copy_assignable<MatrixRowMajor>::value, "Matrix is move-only");
  // Matrix: owning, view with ld, accessors, aliases
  MatrixRowMajor A(3, 4);
  MatrixColMajor B(4, 2);
  DefaultMatrix<> D(2, 2);
  DefaultMatrix<Layout::ColMajor> E(1, 1);
  A(1, 2) = 7;
  check(A.data()[1 * 4 + 2] == 7, "RowMajor element (i, j) at i*ld + j");
  B(3, 1) = 9;
  check(B.data()[1 * 2 + 3] == 9, "ColMajor element (i, j) at j*ld + i, ld = cols as in the reference");
  check(A.rows() == 3 && A.cols() == 4 && A.leading_dimension() == 4 && D.rows() == 2 && E.cols() == 1,
        "rows, cols, leading_dimension");
  Sleef_quad mbuf[20] = {};
  MatrixRowMajor Av(mbuf, 3, 4, 6);
  Av(2, 3) = 1;
  check(mbuf[2 * 6 + 3] == 1 && Av.leading_dimension() == 6, "Matrix view with a leading dimension");
  MatrixRowMajor Am(std::move(A));
  check(Am(1, 2) == 7, "Matrix move constructor");
  // memory/allocation and threading
  Sleef_quad *p = aligned_alloc<Sleef_quad>(10);
  check(p != nullptr && quadblas_is_aligned(p) == 1, "aligned_alloc");
  aligned_free(p);
  const int t0 = get_num_threads();
  set_num_threads(3);
  check(get_num_threads() == 3 && quadblas_get_num_threads() == 3, "set/get_num_threads");
  set_num_threads(t0);
  // the member functions that compute exist with the reference's signatures
  Sleef_quad (VectorRowMajor::*f_dot)(const VectorRowMajor &) const = &VectorRowMajor::dot;
  void (VectorRowMajor::*f_axpy)(Sleef_quad, const VectorRowMajor &) = &VectorRowMajor::axpy;
  Sleef_quad (VectorRowMajor::*f_norm)() const = &VectorRowMajor::norm;
  void (MatrixRowMajor::*f_gemv)(Sleef_quad, const VectorRowMajor &, Sleef_quad, VectorRowMajor &) const =
      &MatrixRowMajor::gemv;
  void (MatrixRowMajor::*f_gemm)(Sleef_quad, const MatrixRowMajor &, Sleef_quad, MatrixRowMajor &) const =
      &MatrixRowMajor::gemm;
  Sleef_quad (*g_dot)(size_t, const Sleef_quad *, size_t, const Sleef_quad *, size_t) = &QuadBLAS::dot;
  void (*g_gemv)(Layout, size_t, size_t, Sleef_quad, const Sleef_quad *, size_t, const Sleef_quad *, size_t,
                 Sleef_quad, Sleef_quad *, size_t) = &QuadBLAS::gemv;
  void (*g_gemm)(Layout, size_t, size_t, size_t, Sleef_quad, const Sleef_quad *, size_t, const Sleef_quad *,
                 size_t, Sleef_quad, Sleef_quad *, size_t) = &QuadBLAS::gemm;
  check(f_dot && f_axpy && f_norm && f_gemv && f_gemm && g_dot && g_gemv && g_gemm, "signatures");
  std::printf("%d failure(s)\n", failures);
  return failures ? 1 : 0;
}
