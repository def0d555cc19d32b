// update_probe.hip -- the Schur-update and trsm products of the fused factor kernel on
// their own, to measure their MFMA rate (test and measurement probe, not the solver).
//
// Per block b (one workgroup, NW waves, NB = 64) it runs, with the same tile helpers
// (common.hpp) and the same per-wave task mapping as chain_factor_kernel
// (blockstructuredsolvers_amd/csrc/chain_factor.hpp, the "[C | W] = Uinv^T [B | Z]" and
// "S_L += W^T W ; Z = -C^T W ; A -= C^T C" phases of its block loop):
//   [C | W] = Uinv^T [B | Z]        (Uinv block upper triangular by 16x16 tiles)
//   S  = W^T W                      (upper tiles)
//   Zn = -C^T W
//   A' = A - C^T C                  (upper tiles)
// Operands come from HBM into LDS once per block; `reps` > 1 repeats the products on the
// LDS-resident operands (the ceiling of the products without their loads).
#include <hip/hip_runtime.h>

#include "../../blockstructuredsolvers_amd/csrc/chain_factor.hpp"

namespace {

using namespace bss;

template <typename T, int NB, int NW>
__global__ void __launch_bounds__(64 * NW)
    update_probe_kernel(const T* __restrict__ U, const T* __restrict__ BZ,
                        const T* __restrict__ A, T* __restrict__ out, int reps) {
  constexpr int TT = NB / 16;
  constexpr int LD = NB + LdsPad<T>::value;
  constexpr int NT = 64 * NW;
  constexpr int NN = NB * NB;
  constexpr int NS = TT * (TT + 1) / 2;
  constexpr int SQ = (NS + NW - 1) / NW;
  constexpr int XG = NW / TT;
  constexpr int XQ = 2 * TT / XG;
  constexpr int ZQ = (TT * TT + NW - 1) / NW;
  constexpr int PQ = (TT * TT + NS + NW - 1) / NW;
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T Abuf[NB * LD];
  __shared__ T Ubuf[NB * LD];
  __shared__ T Rbuf[2 * NB * LD];
  const long long b = blockIdx.x;
  const int w = wave_id(), tid = threadIdx.x;
  for (int e = tid; e < NN; e += NT) {
    const int r = e % NB, c = e / NB;
    Ubuf[c * LD + r] = U[b * NN + e];
    Abuf[c * LD + r] = A[b * NN + e];
  }
  for (int e = tid; e < 2 * NN; e += NT) Rbuf[(e / NB) * LD + e % NB] = BZ[b * 2 * NN + e];
  __syncthreads();
  T* o = out + b * 5 * NN;  // [C | W], S, Zn, A'
  acc_t Sacc[SQ];
#pragma unroll
  for (int q = 0; q < SQ; ++q) Sacc[q] = Mfma<T>::zero();
  acc_t Z[ZQ];
  for (int rep = 0; rep < reps; ++rep) {
    const int xg = w / TT, rho = w % TT;
    acc_t X[XQ];
#pragma unroll
    for (int q = 0; q < XQ; ++q) {
      const int c = xg * XQ + q;
      const int r = ((rho - c) % TT + TT) % TT;
      X[q] = Mfma<T>::zero();
      for (int t = 0; t <= r; ++t)
        tile_mma<T, true, false, false, 16>(X[q], Ubuf + tile_off(t, r, LD), LD,
                                            Rbuf + tile_off(t, c, LD), LD);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < XQ; ++q) {
      const int c = xg * XQ + q;
      const int r = ((rho - c) % TT + TT) % TT;
      tile_store<T>(Rbuf + tile_off(r, c, LD), LD, X[q]);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SQ; ++q) {
      int s = w + NW * q;
      if (s < NS) {
        int a = 0;
        while (s >= TT - a) {
          s -= TT - a;
          ++a;
        }
        const int bb = a + s;
        tile_mma<T, true, false, false, NB>(Sacc[q], Rbuf + (NB + 16 * a) * LD, LD,
                                            Rbuf + (NB + 16 * bb) * LD, LD);
      }
    }
#pragma unroll
    for (int q = 0; q < PQ; ++q) {
      const int task = w + NW * q;
      if (task < TT * TT) {
        if (q < ZQ) {
          const int a = task % TT, bb = task / TT;
          Z[q] = Mfma<T>::zero();
          tile_mma<T, true, false, true, NB>(Z[q], Rbuf + 16 * a * LD, LD,
                                             Rbuf + (NB + 16 * bb) * LD, LD);
        }
      } else if (task < TT * TT + NS) {
        int s = task - TT * TT, a = 0;
        while (s >= TT - a) {
          s -= TT - a;
          ++a;
        }
        const int bb = a + s;
        acc_t acc = tile_load<T>(Abuf + tile_off(a, bb, LD), LD);
        tile_mma<T, true, false, true, NB>(acc, Rbuf + 16 * a * LD, LD, Rbuf + 16 * bb * LD, LD);
        tile_store<T>(Abuf + tile_off(a, bb, LD), LD, acc);
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < 2 * NN; e += NT) o[e] = Rbuf[(e / NB) * LD + e % NB];
#pragma unroll
  for (int q = 0; q < SQ; ++q) {
    int s = w + NW * q;
    if (s < NS) {
      int a = 0;
      while (s >= TT - a) {
        s -= TT - a;
        ++a;
      }
      tile_store<T>(o + 2 * NN + tile_off(a, a + s, NB), NB, Sacc[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < ZQ; ++q) {
    const int task = w + NW * q;
    if (task < TT * TT) tile_store<T>(o + 3 * NN + tile_off(task % TT, task / TT, NB), NB, Z[q]);
  }
  for (int e = tid; e < NN; e += NT) o[4 * NN + e] = Abuf[(e / NB) * LD + e % NB];
}

}  // namespace

// dtype 1 = FP64, 0 = FP32; n = 64, 8 waves per block.  U, A: nblk blocks of n x n;
// BZ: nblk blocks of n x 2n; out: nblk x 5 n^2.  Returns 0 or a HIP error code.
extern "C" int update_probe(int dtype, int nblk, int reps, const void* U, const void* BZ,
                            const void* A, void* out, void* stream) {
  if (nblk < 1 || reps < 1) return -1;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == 1)
    hipLaunchKernelGGL((update_probe_kernel<double, 64, 8>), dim3(nblk), dim3(512), 0, s,
                       (const double*)U, (const double*)BZ, (const double*)A, (double*)out, reps);
  else
    hipLaunchKernelGGL((update_probe_kernel<float, 64, 8>), dim3(nblk), dim3(512), 0, s,
                       (const float*)U, (const float*)BZ, (const float*)A, (float*)out, reps);
  return (int)hipGetLastError();
}
