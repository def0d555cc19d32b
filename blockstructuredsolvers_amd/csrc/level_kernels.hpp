// level_kernels.hpp -- hand-off between recursion levels.
//
// The reference gathers separators and scatters Schur contributions with strided
// copyto!/broadcasts (src/gpu_batched.jl:207-216, 223-224, 251-253;
// src/backbone_rocm.jl:134-136, 152-153).  Here each chain writes its contributions
// to per-separator slots (two writers per separator would otherwise race), and these
// kernels combine them in a fixed order, so results do not depend on scheduling:
//   A'_p = A[sep_p] - SR_p - SL_p         (reduced diagonal, combined in FP64)
//   d'_p = d[sep_p] - rR_p - rL_p         (reduced right-hand side)
//   d[sep_p] = x'_p                       (separator solution back into the level)
#pragma once

#include "common.hpp"
#include "plan_tables.hpp"

namespace bss {

// One workgroup per separator (all systems of the level in one launch).  Ts is the
// element type of the chain contributions and of this level's vectors, To that of the
// next level; they differ only at the hand-off from an FP32 level to FP64 separator
// levels (mixed mode after a separator re-factorisation).
template <typename Tin, typename Ts, typename To>
__global__ void __launch_bounds__(256) assemble_diag_kernel(const SepRec* __restrict__ seps,
                                                             const LevelSys* __restrict__ sys,
                                                             int nn) {
  const SepRec sr = seps[blockIdx.x];
  const LevelSys& S = sys[sr.slot];
  const auto a = gptr((const Tin*)S.A) + (long long)sr.pos * nn;
  const auto r = gptr((const Ts*)S.SR) + (long long)sr.p * nn;
  const auto l = gptr((const Ts*)S.SL) + (long long)sr.p * nn;
  const auto an = gptr((To*)S.An) + (long long)sr.p * nn;
  for (int e = threadIdx.x; e < nn; e += blockDim.x) {
    const double v = (double)a[e] - (double)r[e] - (double)l[e];
    an[e] = (To)v;
  }
}

template <typename Ts, typename To>
__global__ void __launch_bounds__(256) assemble_rhs_kernel(const SepRec* __restrict__ seps,
                                                            const LevelSys* __restrict__ sys,
                                                            int nd) {
  const SepRec sr = seps[blockIdx.x];
  const LevelSys& S = sys[sr.slot];
  const auto src = gptr((const Ts*)S.d) + (long long)sr.pos * nd;
  const auto a = gptr((const Ts*)S.rR) + (long long)sr.p * nd;
  const auto b = gptr((const Ts*)S.rL) + (long long)sr.p * nd;
  const auto dst = gptr((To*)S.dn) + (long long)sr.p * nd;
  for (int e = threadIdx.x; e < nd; e += blockDim.x) {
    if constexpr (sizeof(Ts) == sizeof(To))
      dst[e] = src[e] - a[e] - b[e];
    else
      dst[e] = (To)((double)src[e] - (double)a[e] - (double)b[e]);
  }
}

template <typename Ts, typename To>
__global__ void __launch_bounds__(256) scatter_sep_kernel(const SepRec* __restrict__ seps,
                                                           const LevelSys* __restrict__ sys,
                                                           int nd) {
  const SepRec sr = seps[blockIdx.x];
  const LevelSys& S = sys[sr.slot];
  const auto src = gptr((const To*)S.dn) + (long long)sr.p * nd;
  const auto dst = gptr((Ts*)S.d) + (long long)sr.pos * nd;
  for (int e = threadIdx.x; e < nd; e += blockDim.x) dst[e] = (Ts)src[e];
}

// Element-type conversion of whole buffers (one ConvRec per buffer, grid-stride in
// blockIdx.y).
template <typename S, typename D>
__global__ void __launch_bounds__(256) convert_kernel(const ConvRec<S, D>* __restrict__ recs) {
  const ConvRec<S, D> r = recs[blockIdx.x];
  const auto src = gptr(r.src);
  const auto dst = gptr(r.dst);
  for (long long e = (long long)blockIdx.y * blockDim.x + threadIdx.x; e < r.count;
       e += (long long)gridDim.y * blockDim.x)
    dst[e] = (D)src[e];
}

}  // namespace bss
