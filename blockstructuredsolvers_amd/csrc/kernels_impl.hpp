// kernels_impl.hpp -- launcher definitions; included by one translation unit per
// element-type combination so the (long) template instantiations compile in parallel.
#pragma once

#include "launch.hpp"
#include "level_kernels.hpp"

namespace bss {

template <typename Tin, typename T, int NB, int NW>
static hipError_t factor_nb(const FactorLevelArgs<T>& a, int nchains, int nsys, hipStream_t s) {
  hipLaunchKernelGGL((chain_factor_kernel<Tin, T, NB, NW>), dim3(nchains, nsys), dim3(64 * NW),
                     0, s, a);
  return hipGetLastError();
}

template <typename Tin, typename T>
hipError_t launch_factor(int nb, int waves, const FactorLevelArgs<T>& a, int nchains, int nsys,
                         hipStream_t s) {
  if (nchains <= 0) return hipSuccess;
  switch (nb) {
    case 16: return factor_nb<Tin, T, 16, 1>(a, nchains, nsys, s);
    case 32: return factor_nb<Tin, T, 32, 2>(a, nchains, nsys, s);
    case 48: return factor_nb<Tin, T, 48, 3>(a, nchains, nsys, s);
    case 64:
      if (waves == 8) return factor_nb<Tin, T, 64, 8>(a, nchains, nsys, s);
      return factor_nb<Tin, T, 64, 4>(a, nchains, nsys, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t launch_forward(int nb, const SolveLevelArgs<T>& a, int nchains, int nsys,
                          hipStream_t s) {
  if (nchains <= 0) return hipSuccess;
  switch (nb) {
#define BSS_FWD(NB)                                                                         \
  case NB:                                                                                  \
    hipLaunchKernelGGL((chain_forward_kernel<T, NB>), dim3(nchains, nsys), dim3(4 * NB), 0, \
                       s, a);                                                               \
    return hipGetLastError();
    BSS_FWD(16) BSS_FWD(32) BSS_FWD(48) BSS_FWD(64)
#undef BSS_FWD
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t launch_backward(int nb, const SolveLevelArgs<T>& a, int nchains, int nsys,
                           hipStream_t s) {
  if (nchains <= 0) return hipSuccess;
  switch (nb) {
#define BSS_BWD(NB)                                                                          \
  case NB:                                                                                   \
    hipLaunchKernelGGL((chain_backward_kernel<T, NB>), dim3(nchains, nsys), dim3(4 * NB), 0, \
                       s, a);                                                                \
    return hipGetLastError();
    BSS_BWD(16) BSS_BWD(32) BSS_BWD(48) BSS_BWD(64)
#undef BSS_BWD
    default: return hipErrorInvalidValue;
  }
}

template <typename Tin, typename T>
hipError_t launch_assemble_diag(const Tin* A, const T* SR, const T* SL, const int* seppos,
                                T* Anext, int nb, int P, int nsys, long long sysA,
                                long long sysS, long long sysN, hipStream_t s) {
  if (P <= 0) return hipSuccess;
  hipLaunchKernelGGL((assemble_diag_kernel<Tin, T>), dim3(P, nsys), dim3(256), 0, s, A, SR, SL,
                     seppos, Anext, nb * nb, sysA, sysS, sysN);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_assemble_rhs(const T* d, const T* rR, const T* rL, const int* seppos,
                               T* dnext, int nd, int P, int nsys, long long sysD,
                               long long sysR, long long sysN, hipStream_t s) {
  if (P <= 0) return hipSuccess;
  hipLaunchKernelGGL((assemble_rhs_kernel<T>), dim3(P, nsys), dim3(256), 0, s, d, rR, rL,
                     seppos, dnext, nd, sysD, sysR, sysN);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scatter_sep(const T* x, const int* seppos, T* d, int nd, int P, int nsys,
                              long long sysX, long long sysD, hipStream_t s) {
  if (P <= 0) return hipSuccess;
  hipLaunchKernelGGL((scatter_sep_kernel<T>), dim3(P, nsys), dim3(256), 0, s, x, seppos, d, nd,
                     sysX, sysD);
  return hipGetLastError();
}

}  // namespace bss
