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
#include "plan_tables.hpp"

namespace bss {

// one workgroup per separator (all systems of the level in one launch)
template <typename Tin, typename T>
__global__ void __launch_bounds__(256) assemble_diag_kernel(const SepRec* __restrict__ seps,
                                                             const LevelSys* __restrict__ sys,
                                                             int nn) {
  const SepRec sr = seps[blockIdx.x];
  const LevelSys& S = sys[sr.slot];
  const Tin* a = (const Tin*)S.A + (long long)sr.pos * nn;
  const T* r = (const T*)S.SR + (long long)sr.p * nn;
  const T* l = (const T*)S.SL + (long long)sr.p * nn;
  T* an = (T*)S.An + (long long)sr.p * nn;
  for (int e = threadIdx.x; e < nn; e += blockDim.x) {
    const double v = (double)a[e] - (double)r[e] - (double)l[e];
    an[e] = (T)v;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) assemble_rhs_kernel(const SepRec* __restrict__ seps,
                                                            const LevelSys* __restrict__ sys,
                                                            int nd) {
  const SepRec sr = seps[blockIdx.x];
  const LevelSys& S = sys[sr.slot];
  const T* src = (const T*)S.d + (long long)sr.pos * nd;
  const T* a = (const T*)S.rR + (long long)sr.p * nd;
  const T* b = (const T*)S.rL + (long long)sr.p * nd;
  T* dst = (T*)S.dn + (long long)sr.p * nd;
  for (int e = threadIdx.x; e < nd; e += blockDim.x) dst[e] = src[e] - a[e] - b[e];
}

template <typename T>
__global__ void __launch_bounds__(256) scatter_sep_kernel(const SepRec* __restrict__ seps,
                                                           const LevelSys* __restrict__ sys,
                                                           int nd) {
  const SepRec sr = seps[blockIdx.x];
  const LevelSys& S = sys[sr.slot];
  const T* src = (const T*)S.dn + (long long)sr.p * nd;
  T* dst = (T*)S.d + (long long)sr.pos * nd;
  for (int e = threadIdx.x; e < nd; e += blockDim.x) dst[e] = src[e];
}

}  // namespace bss
