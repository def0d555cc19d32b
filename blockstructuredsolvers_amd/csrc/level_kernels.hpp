// level_kernels.hpp -- hand-off between recursion levels.
//
// The reference gathers separators and scatters Schur contributions with strided
// copyto!/broadcasts (src/gpu_batched.jl:207-216, 223-224, 251-253;
// src/backbone_rocm.jl:134-136, 152-153).  Here each chain writes its contributions
// to per-separator slots (two writers per separator would otherwise race), and these
// kernels combine them in a fixed order, so results do not depend on scheduling:
//   A'_p = A[sep_p] - SR_p - SL_p         (reduced diagonal)
//   d'_p = d[sep_p] - rR_p - rL_p         (reduced right-hand side)
//   d[sep_p] = x'_p                       (separator solution back into the level)
#pragma once

#include "common.hpp"

namespace bss {

template <typename Tin, typename T>
__global__ void __launch_bounds__(256) assemble_diag_kernel(const Tin* __restrict__ A,
                                                             const T* __restrict__ SR,
                                                             const T* __restrict__ SL,
                                                             const int* __restrict__ seppos,
                                                             T* __restrict__ Anext, int nn,
                                                             long long sysA, long long sysS,
                                                             long long sysN) {
  const int p = blockIdx.x, sys = blockIdx.y;
  const Tin* a = A + sys * sysA + (long long)seppos[p] * nn;
  const T* sr = SR + sys * sysS + (long long)p * nn;
  const T* sl = SL + sys * sysS + (long long)p * nn;
  T* an = Anext + sys * sysN + (long long)p * nn;
  for (int e = threadIdx.x; e < nn; e += blockDim.x) {
    const double v = (double)a[e] - (double)sr[e] - (double)sl[e];
    an[e] = (T)v;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) assemble_rhs_kernel(const T* __restrict__ d,
                                                            const T* __restrict__ rR,
                                                            const T* __restrict__ rL,
                                                            const int* __restrict__ seppos,
                                                            T* __restrict__ dnext, int nd,
                                                            long long sysD, long long sysR,
                                                            long long sysN) {
  const int p = blockIdx.x, sys = blockIdx.y;
  const T* src = d + sys * sysD + (long long)seppos[p] * nd;
  const T* a = rR + sys * sysR + (long long)p * nd;
  const T* b = rL + sys * sysR + (long long)p * nd;
  T* dst = dnext + sys * sysN + (long long)p * nd;
  for (int e = threadIdx.x; e < nd; e += blockDim.x) dst[e] = src[e] - a[e] - b[e];
}

template <typename T>
__global__ void __launch_bounds__(256) scatter_sep_kernel(const T* __restrict__ x,
                                                           const int* __restrict__ seppos,
                                                           T* __restrict__ d, int nd,
                                                           long long sysX, long long sysD) {
  const int p = blockIdx.x, sys = blockIdx.y;
  const T* src = x + sys * sysX + (long long)p * nd;
  T* dst = d + sys * sysD + (long long)seppos[p] * nd;
  for (int e = threadIdx.x; e < nd; e += blockDim.x) dst[e] = src[e];
}

}  // namespace bss
