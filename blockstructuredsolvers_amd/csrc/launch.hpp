// launch.hpp -- host-side launchers for the level kernels, instantiated per element
// type in the kernels_*.hip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "chain_factor.hpp"
#include "chain_solve.hpp"

namespace bss {

// returns hipErrorInvalidValue for an unsupported block size
// `waves`: workgroup size in waves for nb = 64 (4 or 8); other sizes use nb/16
template <typename Tin, typename T>
hipError_t launch_factor(int nb, int waves, const FactorLevelArgs<T>& a, int nchains, int nsys,
                         hipStream_t s);
template <typename T>
hipError_t launch_forward(int nb, const SolveLevelArgs<T>& a, int nchains, int nsys,
                          hipStream_t s);
template <typename T>
hipError_t launch_backward(int nb, const SolveLevelArgs<T>& a, int nchains, int nsys,
                           hipStream_t s);
template <typename Tin, typename T>
hipError_t launch_assemble_diag(const Tin* A, const T* SR, const T* SL, const int* seppos,
                                T* Anext, int nb, int P, int nsys, long long sysA,
                                long long sysS, long long sysN, hipStream_t s);
template <typename T>
hipError_t launch_assemble_rhs(const T* d, const T* rR, const T* rL, const int* seppos,
                               T* dnext, int nd, int P, int nsys, long long sysD,
                               long long sysR, long long sysN, hipStream_t s);
template <typename T>
hipError_t launch_scatter_sep(const T* x, const int* seppos, T* d, int nd, int P, int nsys,
                              long long sysX, long long sysD, hipStream_t s);

}  // namespace bss
