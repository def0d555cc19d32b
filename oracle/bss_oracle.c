/*
 * bss_oracle.c -- CPU restatement of the reference's FP64/FP32 block-tridiagonal
 * Cholesky (the `cpu.jl` path).  TEST INFRASTRUCTURE ONLY: this file is the
 * checker for the HIP library and the CPU baseline timed beside it.  Nothing in
 * the product (blockstructuredsolvers_amd/) links, imports or calls it.
 *
 * Matrix convention (reference test/utils.jl:1-30, benchmark/utils.jl:59-103):
 *   M = tridiag(B_{i-1}^T, A_i, B_i), B_i sits at block position (i, i+1).
 * Storage (reference src/gpu_seq.jl:29-43 create_matrix_list): every block is
 * column-major with leading dimension n, blocks contiguous:
 *   A : n x n x N,   B : n x n x (N-1),   d : n x nrhs x N.
 *
 * Operation order follows src/backbone_cpu.jl exactly:
 *   cholesky_factorize!  (backbone_cpu.jl:5-15)
 *     potrf!('U', A_1)
 *     for i = 2..N: trsm!('L','U','T','N', 1, A_{i-1}, B_{i-1})   -> C_{i-1} = U_{i-1}^{-T} B_{i-1}
 *                   gemm!('T','N', -1, B_{i-1}, B_{i-1}, 1, A_i)   -> A_i -= C^T C
 *                   potrf!('U', A_i)
 *   cholesky_solve!      (backbone_cpu.jl:17-33)
 *     trsm!(U_1^T) d_1 ; i=2..N: d_i -= C_{i-1}^T d_{i-1}; d_i = U_i^{-T} d_i
 *     d_N = U_N^{-1} d_N ; i=N-1..1: d_i -= C_i d_{i+1}; d_i = U_i^{-1} d_i
 *
 * LAPACK semantics kept: potrf('U') reads and writes only the upper triangle
 * (the strict lower triangle keeps its input values); gemm updates the full
 * block.  Unlike the reference (which ignores potrf's info), the first
 * non-positive pivot is reported as a 1-based block index.
 *
 * The dense kernels are written out by hand (unblocked, plain loops) so that
 * the oracle needs no BLAS; they are cross-checked against scipy's LAPACK
 * dpotrf/dtrsm and a dense numpy Cholesky in tests/test_oracle.py.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#define IDX(r, c, ld) ((size_t)(c) * (size_t)(ld) + (size_t)(r))

#define DEFINE_ORACLE(T, SFX, SQRT)                                                       \
  /* LAPACK ?potrf('U'): A = U^T U, upper triangle overwritten by U.  Returns 0 or the  \
   * 1-based column of the first non-positive pivot (LAPACK info convention).  */        \
  static int potrf_u_##SFX(int n, T* a) {                                                 \
    for (int j = 0; j < n; ++j) {                                                         \
      T s = a[IDX(j, j, n)];                                                              \
      _Pragma("omp simd reduction(-:s)")                                                  \
      for (int k = 0; k < j; ++k) s -= a[IDX(k, j, n)] * a[IDX(k, j, n)];                 \
      if (!(s > (T)0)) {                                                                  \
        a[IDX(j, j, n)] = s;                                                              \
        return j + 1;                                                                     \
      }                                                                                   \
      T ujj = SQRT(s);                                                                    \
      a[IDX(j, j, n)] = ujj;                                                              \
      for (int c = j + 1; c < n; ++c) {                                                   \
        T t = a[IDX(j, c, n)];                                                            \
        _Pragma("omp simd reduction(-:t)")                                                \
        for (int k = 0; k < j; ++k) t -= a[IDX(k, j, n)] * a[IDX(k, c, n)];               \
        a[IDX(j, c, n)] = t / ujj;                                                        \
      }                                                                                   \
    }                                                                                     \
    return 0;                                                                             \
  }                                                                                       \
  /* ?trsm('L','U','T','N', 1, U, X): X <- U^{-T} X, X is n x m (ld n).  */               \
  static void trsm_lutn_##SFX(int n, int m, const T* u, T* x) {                          \
    for (int c = 0; c < m; ++c) {                                                         \
      T* xc = x + (size_t)c * n;                                                          \
      for (int r = 0; r < n; ++r) {                                                       \
        T t = xc[r];                                                                      \
        _Pragma("omp simd reduction(-:t)")                                                \
        for (int k = 0; k < r; ++k) t -= u[IDX(k, r, n)] * xc[k];                         \
        xc[r] = t / u[IDX(r, r, n)];                                                      \
      }                                                                                   \
    }                                                                                     \
  }                                                                                       \
  /* ?trsm('L','U','N','N', 1, U, X): X <- U^{-1} X.  */                                  \
  static void trsm_lunn_##SFX(int n, int m, const T* u, T* x) {                          \
    for (int c = 0; c < m; ++c) {                                                         \
      T* xc = x + (size_t)c * n;                                                          \
      /* column-oriented back substitution (same maths, contiguous columns of U) */      \
      for (int k = n - 1; k >= 0; --k) {                                                  \
        const T xk = xc[k] / u[IDX(k, k, n)];                                             \
        xc[k] = xk;                                                                       \
        const T* uk = u + (size_t)k * n;                                                  \
        for (int r = 0; r < k; ++r) xc[r] -= uk[r] * xk;                                  \
      }                                                                                   \
    }                                                                                     \
  }                                                                                       \
  /* ?gemm('T','N', -1, X, Y, 1, Z): Z (n x m) -= X^T Y, X is n x n, Y is n x m.  */      \
  static void gemm_tn_sub_##SFX(int n, int m, const T* x, const T* y, T* z) {            \
    for (int c = 0; c < m; ++c)                                                           \
      for (int r = 0; r < n; ++r) {                                                       \
        T t = (T)0;                                                                       \
        _Pragma("omp simd reduction(+:t)")                                                \
        for (int k = 0; k < n; ++k) t += x[IDX(k, r, n)] * y[IDX(k, c, n)];               \
        z[IDX(r, c, n)] -= t;                                                             \
      }                                                                                   \
  }                                                                                       \
  /* ?gemm('N','N', -1, X, Y, 1, Z): Z (n x m) -= X Y.  */                                \
  static void gemm_nn_sub_##SFX(int n, int m, const T* x, const T* y, T* z) {            \
    for (int c = 0; c < m; ++c)                                                           \
      for (int k = 0; k < n; ++k) {                                                       \
        const T yk = y[IDX(k, c, n)];                                                     \
        const T* xk = x + (size_t)k * n;                                                  \
        T* zc = z + (size_t)c * n;                                                        \
        for (int r = 0; r < n; ++r) zc[r] -= xk[r] * yk;                                  \
      }                                                                                   \
  }                                                                                       \
  /* cholesky_factorize! -- src/backbone_cpu.jl:5-15.  Returns 0 or the 1-based index   \
   * of the first block whose potrf fails (the reference ignores this).  */              \
  int oracle_factorize_##SFX(int N, int n, T* A, T* B) {                                  \
    const size_t nn = (size_t)n * n;                                                      \
    if (potrf_u_##SFX(n, A)) return 1;                                                    \
    for (int i = 1; i < N; ++i) {                                                         \
      T* Bp = B + (size_t)(i - 1) * nn;                                                   \
      trsm_lutn_##SFX(n, n, A + (size_t)(i - 1) * nn, Bp);                                \
      gemm_tn_sub_##SFX(n, n, Bp, Bp, A + (size_t)i * nn);                                \
      if (potrf_u_##SFX(n, A + (size_t)i * nn)) return i + 1;                             \
    }                                                                                     \
    return 0;                                                                             \
  }                                                                                       \
  /* cholesky_solve! -- src/backbone_cpu.jl:17-33 (any number of RHS columns).  */       \
  void oracle_solve_##SFX(int N, int n, int nrhs, const T* A, const T* B, T* d) {         \
    const size_t nn = (size_t)n * n, nd = (size_t)n * nrhs;                               \
    trsm_lutn_##SFX(n, nrhs, A, d);                                                       \
    for (int i = 1; i < N; ++i) {                                                         \
      gemm_tn_sub_##SFX(n, nrhs, B + (size_t)(i - 1) * nn, d + (size_t)(i - 1) * nd,      \
                        d + (size_t)i * nd);                                              \
      trsm_lutn_##SFX(n, nrhs, A + (size_t)i * nn, d + (size_t)i * nd);                   \
    }                                                                                     \
    trsm_lunn_##SFX(n, nrhs, A + (size_t)(N - 1) * nn, d + (size_t)(N - 1) * nd);         \
    for (int i = N - 2; i >= 0; --i) {                                                    \
      gemm_nn_sub_##SFX(n, nrhs, B + (size_t)i * nn, d + (size_t)(i + 1) * nd,            \
                        d + (size_t)i * nd);                                              \
      trsm_lunn_##SFX(n, nrhs, A + (size_t)i * nn, d + (size_t)i * nd);                   \
    }                                                                                     \
  }                                                                                       \
  /* Block operator y = M x (y may not alias x): y_i = B_{i-1}^T x_{i-1} + A_i x_i +      \
   * B_i x_{i+1}; the manufactured-solution RHS of test/utils.jl:22-27.  */               \
  void oracle_matvec_##SFX(int N, int n, int nrhs, const T* A, const T* B, const T* x,    \
                           T* y) {                                                        \
    const size_t nn = (size_t)n * n, nd = (size_t)n * nrhs;                               \
    memset(y, 0, sizeof(T) * nd * (size_t)N);                                             \
    for (int i = 0; i < N; ++i) {                                                         \
      T* yi = y + (size_t)i * nd;                                                         \
      /* gemm_*_sub computes z -= ..., so accumulate -y then flip the sign. */            \
      gemm_nn_sub_##SFX(n, nrhs, A + (size_t)i * nn, x + (size_t)i * nd, yi);             \
      if (i > 0)                                                                          \
        gemm_tn_sub_##SFX(n, nrhs, B + (size_t)(i - 1) * nn, x + (size_t)(i - 1) * nd, yi); \
      if (i + 1 < N)                                                                      \
        gemm_nn_sub_##SFX(n, nrhs, B + (size_t)i * nn, x + (size_t)(i + 1) * nd, yi);     \
      for (size_t e = 0; e < nd; ++e) yi[e] = -yi[e];                                     \
    }                                                                                     \
  }

DEFINE_ORACLE(double, f64, sqrt)
DEFINE_ORACLE(float, f32, sqrtf)

/* Normwise backward error per RHS column, computed in FP64 from FP64 data:
 *   berr_j = ||d_j - M x_j||_inf / (||M||_inf ||x_j||_inf + ||d_j||_inf)
 * Writes nrhs values to berr and returns their maximum. */
double oracle_backward_error_f64(int N, int n, int nrhs, const double* A, const double* B,
                                 const double* x, const double* d, double* r_work,
                                 double* berr) {
  const size_t nn = (size_t)n * n, nd = (size_t)n * nrhs;
  oracle_matvec_f64(N, n, nrhs, A, B, x, r_work);
  double mnorm = 0.0;
  for (int i = 0; i < N; ++i)
    for (int r = 0; r < n; ++r) {
      double s = 0.0;
      for (int c = 0; c < n; ++c) s += fabs(A[(size_t)i * nn + IDX(r, c, n)]);
      if (i > 0)
        for (int c = 0; c < n; ++c) s += fabs(B[(size_t)(i - 1) * nn + IDX(c, r, n)]);
      if (i + 1 < N)
        for (int c = 0; c < n; ++c) s += fabs(B[(size_t)i * nn + IDX(r, c, n)]);
      if (s > mnorm) mnorm = s;
    }
  double worst = 0.0;
  for (int j = 0; j < nrhs; ++j) {
    double rn = 0.0, xn = 0.0, dn = 0.0;
    for (int i = 0; i < N; ++i)
      for (int r = 0; r < n; ++r) {
        size_t e = (size_t)i * nd + (size_t)j * n + r;
        double rv = fabs(d[e] - r_work[e]);
        if (rv > rn) rn = rv;
        if (fabs(x[e]) > xn) xn = fabs(x[e]);
        if (fabs(d[e]) > dn) dn = fabs(d[e]);
      }
    double den = mnorm * xn + dn;
    berr[j] = den > 0.0 ? rn / den : rn;
    if (berr[j] > worst) worst = berr[j];
  }
  return worst;
}
