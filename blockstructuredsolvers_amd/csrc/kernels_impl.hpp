// kernels_impl.hpp -- launcher definitions; included by one translation unit per
// element-type combination so the (long) template instantiations compile in parallel.
#pragma once

#include "launch.hpp"
#include "chain_factor_staged.hpp"
#include "level_kernels.hpp"

namespace bss {

// Compute units of the current device (cached per device).
static int device_cus() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cached[dev]) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    cached[dev] = cus;
  }
  return cached[dev];
}

template <typename Tin, typename T, int NB, int NW, bool FWD>
static hipError_t factor_nb_f(const FactorLevelArgs& a, int nchains, hipStream_t s) {
  if constexpr (factor_has_latency_variant<T, NB, NW>()) {
    if (nchains <= device_cus()) {  // one chain per CU at most: latency variant
      hipLaunchKernelGGL((chain_factor_kernel<Tin, T, NB, NW, true, FWD>), dim3(nchains),
                         dim3(64 * NW), 0, s, a);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((chain_factor_kernel<Tin, T, NB, NW, false, FWD>), dim3(nchains),
                     dim3(64 * NW), 0, s, a);
  return hipGetLastError();
}

template <typename Tin, typename T, int NB, int NW>
static hipError_t factor_nb(const FactorLevelArgs& a, int nchains, hipStream_t s) {
  if (a.fwd) {  // fused forward sweep: FP32 factors only (factor_fuses_forward)
    if constexpr (sizeof(T) == 4)
      return factor_nb_f<Tin, T, NB, NW, true>(a, nchains, s);
    else
      return hipErrorInvalidValue;
  }
  return factor_nb_f<Tin, T, NB, NW, false>(a, nchains, s);
}

template <typename Tin, typename T, int NB, int NW>
static hipError_t factor_big_nb(const FactorLevelArgs& a, int nchains, hipStream_t s) {
  if (a.fwd) return hipErrorInvalidValue;  // no fused forward sweep (factor_fuses_forward)
  hipLaunchKernelGGL((chain_factor_staged_kernel<Tin, T, NB, NW>), dim3(nchains), dim3(64 * NW), 0, s, a);
  return hipGetLastError();
}

template <typename Tin, typename T>
hipError_t launch_factor(int nb, int waves, const FactorLevelArgs& a, int nchains, hipStream_t s) {
  if (nchains <= 0) return hipSuccess;
  switch (nb) {
    case 16: return factor_nb<Tin, T, 16, 1>(a, nchains, s);
    case 32: return factor_nb<Tin, T, 32, 2>(a, nchains, s);
    case 48: return factor_nb<Tin, T, 48, 3>(a, nchains, s);
    case 64:
      if (waves == 8) return factor_nb<Tin, T, 64, 8>(a, nchains, s);
      return factor_nb<Tin, T, 64, 4>(a, nchains, s);
    case 96:
      // FP32 factors at n = 96: the four blocks of the LDS kernel fit (154 KB, one
      // workgroup per CU); FP64 ones take the staged kernel (triangle in LDS)
      if constexpr (sizeof(T) == 4)
        return factor_nb<Tin, T, 96, 12>(a, nchains, s);
      else
        return factor_big_nb<Tin, T, 96, 8>(a, nchains, s);
    case 128: return factor_big_nb<Tin, T, 128, 8>(a, nchains, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t launch_forward(int nb, const SolveLevelArgs& a, int nchains, hipStream_t s) {
  if (nchains <= 0) return hipSuccess;
  switch (nb) {
#define BSS_FWD(NB)                                                                          \
  case NB:                                                                                   \
    if (a.nrhs == 1)                                                                         \
      hipLaunchKernelGGL((chain_forward1_kernel<T, NB>), dim3(nchains), dim3(4 * NB), 0, s, a); \
    else                                                                                     \
      hipLaunchKernelGGL((chain_forward_kernel<T, NB>), dim3(nchains), dim3(4 * NB), 0, s, a); \
    return hipGetLastError();
    BSS_FWD(16) BSS_FWD(32) BSS_FWD(48) BSS_FWD(64) BSS_FWD(96) BSS_FWD(128)
#undef BSS_FWD
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t launch_backward(int nb, const SolveLevelArgs& a, int nchains, hipStream_t s) {
  if (nchains <= 0) return hipSuccess;
  switch (nb) {
#define BSS_BWD(NB)                                                                           \
  case NB:                                                                                    \
    if (a.nrhs == 1)                                                                          \
      hipLaunchKernelGGL((chain_backward1_kernel<T, NB>), dim3(nchains), dim3(4 * NB), 0, s, a); \
    else                                                                                      \
      hipLaunchKernelGGL((chain_backward_kernel<T, NB>), dim3(nchains), dim3(4 * NB), 0, s, a); \
    return hipGetLastError();
    BSS_BWD(16) BSS_BWD(32) BSS_BWD(48) BSS_BWD(64) BSS_BWD(96) BSS_BWD(128)
#undef BSS_BWD
    default: return hipErrorInvalidValue;
  }
}

template <typename Tin, typename Ts, typename To>
hipError_t launch_assemble_diag(const SepRec* seps, const LevelSys* sys, int nb, int nsep,
                                hipStream_t s) {
  if (nsep <= 0) return hipSuccess;
  hipLaunchKernelGGL((assemble_diag_kernel<Tin, Ts, To>), dim3(nsep), dim3(256), 0, s, seps, sys,
                     nb * nb);
  return hipGetLastError();
}

template <typename Ts, typename To>
hipError_t launch_assemble_rhs(const SepRec* seps, const LevelSys* sys, int nd, int nsep,
                               hipStream_t s) {
  if (nsep <= 0) return hipSuccess;
  hipLaunchKernelGGL((assemble_rhs_kernel<Ts, To>), dim3(nsep), dim3(256), 0, s, seps, sys, nd);
  return hipGetLastError();
}

template <typename Ts, typename To>
hipError_t launch_scatter_sep(const SepRec* seps, const LevelSys* sys, int nd, int nsep,
                              hipStream_t s) {
  if (nsep <= 0) return hipSuccess;
  hipLaunchKernelGGL((scatter_sep_kernel<Ts, To>), dim3(nsep), dim3(256), 0, s, seps, sys, nd);
  return hipGetLastError();
}

}  // namespace bss
