// smoother.hip -- block-tridiagonal normal equations of a linear state-space smoother,
// built on the device.
//
// For the model x_{i+1} = F x_i + w_i (w ~ N(0, Q)), y_i = H x_i + v_i (v ~ N(0, R))
// the smoother's normal equations are block-tridiagonal with, for every block,
//   A_i = Q^{-1} + F^T Q^{-1} F + H^T R^{-1} H,    B_i = -Q^{-1} F
// (the served-stream workload; benchmark/benchmark_kf.jl:79-154 builds the same blocks
// on the host, with A_1 and A_N special-cased, and uploads them).  Here the two n x n
// blocks are formed once in FP64 and broadcast into the reference layout (A: N blocks,
// B: N-1 blocks, column-major), so a request carries only the model matrices.
#include <hip/hip_runtime.h>

#include "../../include/bss.h"

namespace bss {
int api_fail(int code, const char* msg);  // records bss_last_error (bss_api.hip)
}

namespace {

// One thread per entry (r, c) of the n x n blocks: A_0 and B_0 into block 0 of A, B.
// Column-major inputs: F (n x n), H (m x n), Qinv (n x n), Rinv (m x m).
template <typename T>
__global__ void __launch_bounds__(256) smoother_block0_kernel(int n, int m, const double* F,
                                                              const double* H, const double* Qinv,
                                                              const double* Rinv, T* A, T* B) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int r = e % n, c = e / n;
  // entries of Q^{-1} F and R^{-1} H are recomputed where used: n, m <= 128 keep it cheap
  double qf_rc = 0.0;  // (Q^{-1} F)[r][c]
  for (int k = 0; k < n; ++k) qf_rc += Qinv[k * n + r] * F[c * n + k];
  double ftqf = 0.0;  // (F^T Q^{-1} F)[r][c] = sum_k F[k][r] (Q^{-1} F)[k][c]
  for (int k = 0; k < n; ++k) {
    double qf_kc = 0.0;
    for (int j = 0; j < n; ++j) qf_kc += Qinv[j * n + k] * F[c * n + j];
    ftqf += F[r * n + k] * qf_kc;
  }
  double htrh = 0.0;  // (H^T R^{-1} H)[r][c] = sum_k H[k][r] (R^{-1} H)[k][c]
  for (int k = 0; k < m; ++k) {
    double rh_kc = 0.0;
    for (int j = 0; j < m; ++j) rh_kc += Rinv[j * m + k] * H[c * m + j];
    htrh += H[r * m + k] * rh_kc;
  }
  A[e] = (T)(Qinv[c * n + r] + ftqf + htrh);
  if (B) B[e] = (T)(-qf_rc);
}

// Copies block 0 to blocks 1 .. count-1 (16-byte vectors; n*n*sizeof(T) is a multiple
// of 16 for every supported n).
__global__ void __launch_bounds__(256) broadcast_block_kernel(int4* __restrict__ base,
                                                              long long vec_per_block,
                                                              long long count) {
  const long long total = vec_per_block * (count - 1);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x)
    base[vec_per_block + i] = base[i % vec_per_block];
}

}  // namespace

extern "C" int bss_build_smoother(int N, int n, int m, const double* F, const double* H,
                                  const double* Qinv, const double* Rinv, void* A, void* B,
                                  int dtype, void* stream) {
  if (N < 1 || n < 16 || n > 128 || (n % 16) != 0 || m < 1 || m > 128)
    return bss::api_fail(BSS_E_ARG, "smoother sizes: need N >= 1, n in {16, ..., 128} (multiple of 16), 1 <= m <= 128");
  if (!F || !H || !Qinv || !Rinv || !A || (N > 1 && !B))
    return bss::api_fail(BSS_E_ARG, "smoother: null buffer");
  if (dtype != BSS_F32 && dtype != BSS_F64) return bss::api_fail(BSS_E_ARG, "smoother: bad dtype");
  hipStream_t s = (hipStream_t)stream;
  const int nn = n * n;
  void* B0 = N > 1 ? B : nullptr;
  if (dtype == BSS_F64)
    hipLaunchKernelGGL(smoother_block0_kernel<double>, dim3((nn + 255) / 256), dim3(256), 0, s, n,
                       m, F, H, Qinv, Rinv, (double*)A, (double*)B0);
  else
    hipLaunchKernelGGL(smoother_block0_kernel<float>, dim3((nn + 255) / 256), dim3(256), 0, s, n, m,
                       F, H, Qinv, Rinv, (float*)A, (float*)B0);
  if (hipGetLastError() != hipSuccess) return bss::api_fail(BSS_E_HIP, "smoother: block launch failed");
  const long long vpb = (long long)nn * (dtype == BSS_F64 ? 8 : 4) / 16;
  const int grid = 1024;
  if (N > 1) {
    hipLaunchKernelGGL(broadcast_block_kernel, dim3(grid), dim3(256), 0, s, (int4*)A, vpb,
                       (long long)N);
    if (N > 2)
      hipLaunchKernelGGL(broadcast_block_kernel, dim3(grid), dim3(256), 0, s, (int4*)B, vpb,
                         (long long)(N - 1));
  }
  return hipGetLastError() == hipSuccess ? BSS_OK : bss::api_fail(BSS_E_HIP, "smoother: broadcast launch failed");
}
