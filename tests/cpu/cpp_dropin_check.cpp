// C++ drop-in check: a program written against the reference's C++ API (QuadBLAS::dot, gemv,
// gemm, Vector, Matrix) and C ABI (quadblas_qdot, quadblas_qgemv, quadblas_qgemm), built with g++
// against include/quadblas/quadblas.hpp and linked to libqblas_hip.so.  The cases follow the
// reference's own known-answer tests (test_quadblas.cpp:204-224 dot = 55, :327-350 gemv 3x3,
// :446-464 gemm 2x2, :679-688 gemm 1x1, :691-712 identity, :715-739 cancellation to 1), here
// checked bit for bit, plus ColMajor variants and the OZAKI form through the same API.
// Exit status 0 when every case passes.
#include <cstdio>
#include <cstring>
#include <vector>

#include "quadblas/quadblas.hpp"

using QuadBLAS::Layout;

static int failures = 0;
static void check(bool ok, const char *what) {
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", what);
  if (!ok) ++failures;
}
static bool same(Sleef_quad a, Sleef_quad b) { return std::memcmp(&a, &b, 16) == 0; }

// splitmix64-driven full-mantissa values in [-1, 1)
static Sleef_quad rnd(uint64_t &s) {
  auto next = [&]() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  };
  const double a = (double)(next() >> 11) * 0x1p-53 * 2.0 - 1.0;
  const double b = (double)(next() >> 11) * 0x1p-53 * 2.0 - 1.0;
  return (Sleef_quad)a + (Sleef_quad)b * (Sleef_quad)0x1p-60;
}

int main() {
  std::printf("%s\n", quadblas_get_version());
  // dot = 55 (x = 1..10, y = 1)
  {
    std::vector<Sleef_quad> x(10), y(10, 1);
    for (int i = 0; i < 10; ++i) x[i] = i + 1;
    check(same(QuadBLAS::dot(10, x.data(), 1, y.data(), 1), 55), "dot 1..10 . 1 = 55");
    check(quadblas_qdot(10, x.data(), 1, y.data(), 1) == 55.0, "quadblas_qdot = 55");
    check(same(QuadBLAS::dot(0, x.data(), 1, y.data(), 1), 0), "dot n = 0 -> +0");
  }
  // gemv 3x3: [1..9] x [1,1,1] = [6, 15, 24], RowMajor, ColMajor and through Matrix / the C ABI
  {
    std::vector<Sleef_quad> A = {1, 2, 3, 4, 5, 6, 7, 8, 9}, x = {1, 1, 1}, y(3, 0);
    QuadBLAS::gemv(Layout::RowMajor, 3, 3, 1, A.data(), 3, x.data(), 1, 0, y.data(), 1);
    check(same(y[0], 6) && same(y[1], 15) && same(y[2], 24), "gemv RowMajor 3x3 = [6, 15, 24]");
    std::vector<Sleef_quad> At = {1, 4, 7, 2, 5, 8, 3, 6, 9}, yc(3, 0);
    QuadBLAS::MatrixColMajor M(At.data(), 3, 3);
    QuadBLAS::VectorColMajor xv(x.data(), 3), yv(yc.data(), 3);
    M.gemv(1, xv, 0, yv);
    check(same(yc[0], 6) && same(yc[1], 15) && same(yc[2], 24), "Matrix<ColMajor>::gemv 3x3 = [6, 15, 24]");
    std::vector<Sleef_quad> yt(3, 0);
    quadblas_qgemv('R', 'T', 3, 3, 1.0, A.data(), 3, x.data(), 1, 0.0, yt.data(), 1);  // A^T x
    check(same(yt[0], 12) && same(yt[1], 15) && same(yt[2], 18), "quadblas_qgemv trans = A^T x");
  }
  // gemm 2x2 = [[19, 22], [43, 50]] and 1x1 = 12
  {
    std::vector<Sleef_quad> A = {1, 2, 3, 4}, B = {5, 6, 7, 8}, C(4, 0);
    QuadBLAS::gemm(Layout::RowMajor, 2, 2, 2, 1, A.data(), 2, B.data(), 2, 0, C.data(), 2);
    check(same(C[0], 19) && same(C[1], 22) && same(C[2], 43) && same(C[3], 50), "gemm 2x2 = [[19,22],[43,50]]");
    std::vector<Sleef_quad> Cc(4, 0);
    quadblas_qgemm('C', 'N', 'N', 2, 2, 2, 1.0, A.data(), 2, B.data(), 2, 0.0, Cc.data(), 2);
    // column-major: A = [[1,3],[2,4]], B = [[5,7],[6,8]] -> C = [[23,31],[34,46]] stored by columns
    check(same(Cc[0], 23) && same(Cc[1], 34) && same(Cc[2], 31) && same(Cc[3], 46), "quadblas_qgemm ColMajor 2x2");
    Sleef_quad a = 3, b = 4, c = 0;
    QuadBLAS::gemm(Layout::RowMajor, 1, 1, 1, 1, &a, 1, &b, 1, 0, &c, 1);
    check(same(c, 12), "gemm 1x1 = 12");
  }
  // identity: I * A == A exactly (n = 50, full-mantissa A), through Matrix<RowMajor>::gemm
  {
    const size_t n = 50;
    uint64_t s = 42;
    QuadBLAS::MatrixRowMajor I(n, n), A(n, n), C(n, n);
    for (size_t i = 0; i < n; ++i) I(i, i) = 1;
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j < n; ++j) A(i, j) = rnd(s);
    I.gemm(1, A, 0, C);
    bool ok = true;
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j < n; ++j) ok &= same(C(i, j), A(i, j));
    check(ok, "identity 50x50: I * A == A bit for bit");
  }
  // cancellation: [1e20, 1, -1e20, 0, ...] . 1 = 1 exactly (Vector::dot)
  {
    QuadBLAS::DefaultVector<> x(10), y(10);
    x[0] = 1e20;
    x[1] = 1;
    x[2] = -1e20;
    for (size_t i = 0; i < 10; ++i) y[i] = 1;
    check(same(x.dot(y), 1), "cancellation dot = 1");
  }
  // Vector::axpy and Vector::norm (cpp_classes.hpp:72-81), quadblas_qaxpy / quadblas_qnrm2
  {
    QuadBLAS::VectorRowMajor x(3), y(3), v(2);
    for (size_t i = 0; i < 3; ++i) {
      x[i] = i + 1;
      y[i] = 1;
    }
    y.axpy(2, x);
    check(same(y[0], 3) && same(y[1], 5) && same(y[2], 7), "Vector::axpy 2*[1,2,3] + 1 = [3,5,7]");
    v[0] = 3;
    v[1] = 4;
    check(same(v.norm(), 5), "Vector::norm [3,4] = 5");
    check(quadblas_qnrm2(2, v.data(), 1) == 5.0, "quadblas_qnrm2 [3,4] = 5");
    quadblas_qaxpy(3, -1.0, x.data(), 1, y.data(), 1);
    check(same(y[0], 2) && same(y[1], 3) && same(y[2], 4), "quadblas_qaxpy");
  }
  // the OZAKI form through the unchanged C++ API: within 2^-104 of the reference-order result
  {
    const size_t n = 96;
    uint64_t s = 7;
    std::vector<Sleef_quad> A(n * n), B(n * n), Cr(n * n, 0), Co(n * n, 0);
    for (auto &v : A) v = rnd(s);
    for (auto &v : B) v = rnd(s);
    QuadBLAS::gemm(Layout::RowMajor, n, n, n, 1, A.data(), n, B.data(), n, 0, Cr.data(), n);
    const int old = qblas_get_mode();
    qblas_set_mode(QBLAS_MODE_OZAKI);
    QuadBLAS::gemm(Layout::RowMajor, n, n, n, 1, A.data(), n, B.data(), n, 0, Co.data(), n);
    qblas_set_mode(old);
    // both are within a few ulps of the exact value; compare to each other at 2^-100
    bool ok = true;
    for (size_t i = 0; i < n * n; ++i) {
      Sleef_quad d = Co[i] - Cr[i];
      if (d < 0) d = -d;
      Sleef_quad m = Cr[i] < 0 ? -Cr[i] : Cr[i];
      ok &= d <= m * (Sleef_quad)0x1p-100;
    }
    check(ok, "OZAKI form via QuadBLAS::gemm agrees with the reference order to 2^-100");
  }
  std::printf("%d failure(s)\n", failures);
  return failures ? 1 : 0;
}
