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

}  // namespace bss
