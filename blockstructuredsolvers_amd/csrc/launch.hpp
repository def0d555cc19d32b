// launch.hpp -- host-side launchers for the level kernels, instantiated per element
// type in the kernels_*.hip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "chain_factor.hpp"
#include "chain_solve.hpp"

namespace bss {

// `waves`: workgroup size in waves for nb = 64 (4 or 8); other sizes use nb/16.
// All launchers return hipErrorInvalidValue for an unsupported block size.
template <typename Tin, typename T>
hipError_t launch_factor(int nb, int waves, const FactorLevelArgs& a, int nchains, hipStream_t s);
template <typename T>
hipError_t launch_forward(int nb, const SolveLevelArgs& a, int nchains, hipStream_t s);
template <typename T>
hipError_t launch_backward(int nb, const SolveLevelArgs& a, int nchains, hipStream_t s);
// Ts: type of this level's contributions/vectors; To: type of the next level.
template <typename Tin, typename Ts, typename To>
hipError_t launch_assemble_diag(const SepRec* seps, const LevelSys* sys, int nb, int nsep,
                                hipStream_t s);
template <typename Ts, typename To>
hipError_t launch_assemble_rhs(const SepRec* seps, const LevelSys* sys, int nd, int nsep,
                               hipStream_t s);
template <typename Ts, typename To>
hipError_t launch_scatter_sep(const SepRec* seps, const LevelSys* sys, int nd, int nsep,
                              hipStream_t s);
template <typename S, typename D>
hipError_t launch_convert(const ConvRec<S, D>* recs, int nrec, hipStream_t s);

// Whether a level's forward sweep runs in the factor kernel (FactorLevelArgs::fwd): FP32
// factors on the LDS kernel (n <= 64 and n = 96), one right-hand side.  With FP64
// factors the fused kernel is at its register limit and spills; measured at N = 8192,
// n = 64 (16 systems) its level-0 factor took 11.1 ms instead of 9.5, more than the
// separate sweep it saves (FP32: 4.82 instead of 4.74 ms, saving a 1.0 ms sweep).
inline bool factor_fuses_forward(int nb, bool fp32_factors, int nrhs) {
  return nrhs == 1 && fp32_factors && (nb == 16 || nb == 32 || nb == 48 || nb == 64 || nb == 96);
}

}  // namespace bss
