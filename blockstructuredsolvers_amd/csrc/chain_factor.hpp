// chain_factor.hpp -- fused per-chain factorisation kernel of one recursion level.
//
// A level holds N diagonal blocks A_i and N-1 couplings B_i (M = tridiag(B^T, A, B)).
// The schedule splits it into chains of consecutive interior blocks separated by
// single separator blocks.  For a chain [s .. s+m-1] with left separator L = s-1 and
// right separator R = s+m (either may be absent) the kernel eliminates the chain in
// order, keeping the fill that the left separator creates ("panel" Z):
//
//   Z_s = B_L^T,  for i = s .. s+m-1:
//     A_i  <- A_i - C_{i-1}^T C_{i-1}               (Schur update, MFMA)
//     U_i^T U_i = A_i ; Uinv_i = U_i^{-1}           (blocked potrf + inverse)
//     [C_i | W_i] = Uinv_i^T [B_i | Z_i]            (trsm as MFMA products)
//     Z_{i+1} = -C_i^T W_i ;  S_L += W_i^T W_i      (MFMA)
//   reduced system: A'_L -= S_L, A'_R -= C_last^T C_last, B'_{L,R} = -W_last^T C_last
//
// This is the reference's per-level work -- cholesky_factorize_batched! (factor the
// P-1 interior chains, src/backbone_rocm.jl:62-82), cholesky_solve_batched! with the
// 2n-column G (src/backbone_rocm.jl:84-120) and compute_schur_complement!
// (src/backbone_rocm.jl:122-138) -- re-derived so that only the one-sided panel
// W = U^{-T} G is formed (G is non-zero only in the first and last interior block,
// src/gpu_batched.jl:214-215) and the Schur complement G^T T^{-1} G = W^T W is
// accumulated as the chain is eliminated, in one launch per level with the blocks
// held in LDS.
//
// Stored factors (consumed by the substitution sweeps): Uinv_i (upper 16x16 tiles of
// the A slot), C_i (the B slot), W_i (panel slot, only for chains with a left
// separator).  One workgroup per chain; blockIdx.y is the system in a batch.
#pragma once

#include "common.hpp"
#include "plan_tables.hpp"

namespace bss {

struct FactorLevelArgs {
  const ChainRec* chains;  // one workgroup per chain
  const LevelSys* sys;     // per-system buffers of this level
  int* info;               // per system, atomicMin of the 1-based failing block index
  unsigned* pivmin;        // per system, atomicMin of the float bits of min_j U_jj^2 / A_jj
                           // over the level's blocks (null: not monitored)
  int fwd = 0;             // 1: also run this level's forward sweep for one right-hand
                           // side (chain_forward1_kernel's arithmetic, on the factors
                           // while they are in LDS): d <- y in place, r_R and r_L out
                           // (chain_factor_kernel only; see factor_fuses_forward)
};

// The forward sweep's dot product (XOps::dot of chain_solve.hpp) with the factor's
// column in LDS: (M^T y)[row] over the 16-deep slices [s_lo, s_hi), x[k] = M[k][row],
// lane (lr, lk) holding k = 16 s + 4 lk + q, two partial sums, then the sum over the
// four lanes of the row; the same operations in the same order, so the same bits.
template <typename T>
__device__ inline T fwd_dot(const T* __restrict__ x, const T* __restrict__ y, int s_lo, int s_hi) {
  const int lk = lane_id() >> 4;
  T t0 = T(0), t1 = T(0);
  for (int s = s_lo; s < s_hi; ++s) {
    t0 = fma(x[16 * s + 4 * lk + 0], y[16 * s + 4 * lk + 0], t0);
    t1 = fma(x[16 * s + 4 * lk + 1], y[16 * s + 4 * lk + 1], t1);
    t0 = fma(x[16 * s + 4 * lk + 2], y[16 * s + 4 * lk + 2], t0);
    t1 = fma(x[16 * s + 4 * lk + 3], y[16 * s + 4 * lk + 3], t1);
  }
  return row4_sum(t0 + t1);
}
// ... and its one-accumulator form (the sweep's C_last^T y_last)
template <typename T, int TT>
__device__ inline T fwd_dot1(const T* __restrict__ x, const T* __restrict__ y) {
  const int lk = lane_id() >> 4;
  T t = T(0);
#pragma unroll
  for (int s = 0; s < TT; ++s)
#pragma unroll
    for (int q = 0; q < 4; ++q) t = fma(x[16 * s + 4 * lk + q], y[16 * s + 4 * lk + q], t);
  return row4_sum(t);
}

template <typename T>
struct LdsPad {
  static constexpr int value = sizeof(T) == 8 ? 2 : 4;  // 16-byte column shift
};

__device__ __host__ constexpr int tile_off(int r, int c, int ld) { return 16 * c * ld + 16 * r; }

// The value of x, hidden from the optimiser: offsets derived from it are recomputed
// where they are used instead of being hoisted out of a kernel's block loop and held in
// registers for the whole chain.
__device__ inline int opaque_v(int x) {
  asm volatile("" : "+v"(x));
  return x;
}
__device__ inline int opaque_s(int x) {
  asm volatile("" : "+s"(x));
  return x;
}

// Cross-lane helpers for the diagonal-tile routine.
template <typename T>
__device__ inline T read_lane(T x, int lane);
template <>
__device__ inline double read_lane<double>(double x, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
  return __hiloint2double(hi, lo);
}
template <>
__device__ inline float read_lane<float>(float x, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}

// every lane takes the value of lane J of its 16-lane row (DPP row_newbcast; a
// broadcast never reads outside the row, so bound_ctrl changes nothing).  The FP64 form
// is one 64-bit DPP move (gfx950 has row_newbcast for 64-bit operands), not two 32-bit
// halves that each need a zeroed destination first.
template <int J>
__device__ inline double row_bcast(double x) {
  return __longlong_as_double(
      __builtin_amdgcn_mov_dpp(__double_as_longlong(x), 0x150 + J, 0xf, 0xf, true));
}
template <int J>
__device__ inline float row_bcast(float x) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x150 + J, 0xf, 0xf, false));
}

// 1/sqrt(p) on the serial path, avoiding the long correctly-rounded sqrt + divide
// sequences: f64 refines the hardware estimate with two Newton steps; f32 takes
// v_rsq_f32 as is (1 ulp), the FP32 factor's own rounding level (a Newton step there
// measured 2% of the FP32 factor time and no accuracy change in the refined result).
__device__ inline double rsqrt_nr(double p) {
  double y = __builtin_amdgcn_rsq(p);
  y = y * (1.5 - 0.5 * p * y * y);
  y = y * (1.5 - 0.5 * p * y * y);
  return y;
}
__device__ inline float rsqrt_nr(float p) { return __builtin_amdgcn_rsqf(p); }

// One elimination step J of the 16x16 tile (see diag_chol_inv16).  Lane c (of every
// 16-lane row; the four rows compute the same thing) holds column c: a[r] = A[r][c],
// e[r] = E[r][c].  Everything the step needs is in the lane's own registers or in
// lane J of the same row, so the serial path uses only readlane/DPP and VALU ops.
// `s` = 1/sqrt of pivot J.  Row J+1 is updated first: it holds the next pivot, whose
// 1/sqrt (the long dependent part of a step) then overlaps this step's other rows.
template <typename T, int J>
__device__ __attribute__((always_inline)) inline void diag_step(T (&a)[16], T (&e)[16], T s,
                                                                bool& ok) {
  // keep steps apart in the schedule: interleaving whole steps multiplies the live
  // registers (the pivot of step J+1 is already computed inside step J)
  __builtin_amdgcn_sched_barrier(0);
  const T ujc = a[J] * s;  // U[J][c]
  const T ejc = e[J] * s;  // row J of E, scaled
  a[J] = ujc;
  e[J] = ejc;
  T s_next = T(0);
  // U[J][r] = A[J][r] s = A[r][J] s (the trailing matrix stays symmetric): lane J's a[r].
  // FP32 folds the broadcast into the multiply by s (v_mul_f32 takes a DPP operand, the
  // FP64 multiply does not): 3 VALU ops per row.  FP64 moves s into the pivot-row
  // factors once per step instead, so a row costs one 64-bit DPP move and two FMAs.
  [[maybe_unused]] const T nu = -(ujc * s), ne = -(ejc * s);
  auto update = [&](int r) {
    if constexpr (sizeof(T) == 4) {
      const T ujr = row_bcast<J>(a[r]) * s;
      a[r] -= ujr * ujc;
      e[r] -= ujr * ejc;
    } else {
      const T m = row_bcast<J>(a[r]);
      e[r] = fma(m, ne, e[r]);
      a[r] = fma(m, nu, a[r]);
    }
  };
  if constexpr (J < 15) {
    update(J + 1);
    const T p = read_lane(a[J + 1], J + 1);  // pivot A[J+1][J+1]
    ok = ok && (p > T(0));
    s_next = rsqrt_nr(p);
  }
#pragma unroll
  for (int r = J + 2; r < 16; ++r) update(r);
  // E is consumed only by the final store; without this the compiler defers all of
  // its updates to the end and keeps every multiplier of every step live
#pragma unroll
  for (int r = 0; r < 16; ++r) asm volatile("" : "+v"(e[r]));
  if constexpr (J < 15) diag_step<T, J + 1>(a, e, s_next, ok);
}

// Wave-level Cholesky + inverse of one 16x16 SPD tile held (upper triangle) in LDS.
// Row operations are applied to [A | I]; at the end the right half is E = U^{-T}.
// Writes Uinv = E^T into the upper triangle of the output tile (strict lower part
// exact zeros).  Returns false if a pivot was not positive.
// OPQ: the lane's 16 load offsets are recomputed per call, not held across the caller's
// loop (for kernels whose registers are otherwise exhausted)
template <typename T, bool OPQ = false>
__device__ __attribute__((always_inline)) inline bool diag_chol_inv16(const T* at, int lda, T* ut, int ldu) {
  const int c = OPQ ? opaque_v(lane_id() & 15) : lane_id() & 15;
  T a[16], e[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    a[r] = (r <= c) ? at[c * lda + r] : at[r * lda + c];
    e[r] = (r == c) ? T(1) : T(0);
  }
  const T p0 = read_lane(a[0], 0);
  bool ok = p0 > T(0);
  diag_step<T, 0>(a, e, rsqrt_nr(p0), ok);
  if (lane_id() < 16) {
#pragma unroll
    for (int r = 0; r < 16; ++r) ut[r * ldu + c] = e[r];  // Uinv[c][r] = E[r][c]
  }
  return ok;
}

// Blocked potrf + inverse of the NB x NB block in Abuf (upper triangle used) with
// 16x16 tiles: right-looking elimination on [A | I] where the diagonal tiles go
// through diag_chol_inv16 and every panel/trailing/inverse update is an MFMA tile
// product.  On exit Ubuf's upper tiles hold Uinv = U^{-1}.  Abuf is consumed.
// LA (lookahead): the diagonal step of kk+1 overlaps the elimination of kk (below);
// without it the diagonal step waits for the whole elimination (the variants with a
// 128-register cap other than the slot kernel, where the lookahead's extra live
// registers spill; with the slot kernel's per-call offsets it fits and measured 1.5%
// faster).
// Slots: waves 1..NW-1 run slot(k), k = 0..TT-1, while wave 0 factors diagonal tile k
// (LA: slot k >= 1 follows those waves' share of elimination k-1); a barrier separates
// consecutive slots.  The kernels give them memory work only: an MFMA on the diagonal
// wave's SIMD stalls its serial VALU steps (a 16x16 diagonal tile measured 2.8K cycles
// beside idle, VALU, LDS or other-SIMD MFMA work and 6.8K-11.6K beside MFMAs on its
// own SIMD), and waves w, w + 4 share a SIMD.
// SL: slots in use (the step loops are unrolled so that each slot's work is fixed at
// compile time); without it `slot` is never called.
template <typename T, int NB, int LD, int NW, bool LA, bool SL, typename Slot>
__device__ __attribute__((always_inline)) inline void potrf_inv(T* Abuf, T* Ubuf, int* fail,
                                                                Slot slot) {
  constexpr int TT = NB / 16;
  typedef typename Mfma<T>::acc_t acc_t;
  const int w = wave_id();
  auto diag = [&](int kk) {
    const bool ok = diag_chol_inv16<T, SL>(Abuf + tile_off(kk, kk, LD), LD,
                                           Ubuf + tile_off(kk, kk, LD), LD);
    if (!ok && lane_id() == 0) *fail = 1;
  };
  // scaling: U_{kk,j} = Uinv_kk^T A_{kk,j} (j > kk);  Et_{t,kk} <- Et_{t,kk} Uinv_kk (t < kk)
  auto scale = [&](int kk) {
    const int nrow = TT - 1 - kk;
    for (int task = w; task < TT - 1; task += NW) {
      acc_t acc = Mfma<T>::zero();
      if (task < nrow) {
        const int j = kk + 1 + task;
        tile_mma<T, true, false, false, 16>(acc, Ubuf + tile_off(kk, kk, LD), LD,
                                        Abuf + tile_off(kk, j, LD), LD);
        tile_store<T>(Abuf + tile_off(kk, j, LD), LD, acc);
      } else {
        const int t = task - nrow;
        tile_mma<T, false, false, false, 16>(acc, Ubuf + tile_off(t, kk, LD), LD,
                                         Ubuf + tile_off(kk, kk, LD), LD);
        tile_store<T>(Ubuf + tile_off(t, kk, LD), LD, acc);
      }
    }
  };
  // elimination tasks first, first + stride, ...:
  //   A_ij -= U_{kk,i}^T U_{kk,j} (kk < i <= j);  Et_{t,i} -= Et_{t,kk} U_{kk,i}
  // (task 0 is the next diagonal tile)
  auto eliminate = [&](int kk, int first, int stride) {
    const int nrow = TT - 1 - kk;
    const int nA = nrow * (nrow + 1) / 2;
    const int nE = (kk + 1) * nrow;
    for (int task = first; task < nA + nE; task += stride) {
      if (task < nA) {
        int idx = task, ii = 0;
        while (idx >= nrow - ii) {
          idx -= nrow - ii;
          ++ii;
        }
        const int i = kk + 1 + ii, j = i + idx;
        acc_t acc = tile_load<T>(Abuf + tile_off(i, j, LD), LD);
        tile_mma<T, true, false, true, 16>(acc, Abuf + tile_off(kk, i, LD), LD,
                                       Abuf + tile_off(kk, j, LD), LD);
        tile_store<T>(Abuf + tile_off(i, j, LD), LD, acc);
      } else {
        const int e = task - nA;
        const int t = e % (kk + 1), i = kk + 1 + e / (kk + 1);
        acc_t acc = (t == kk) ? Mfma<T>::zero() : tile_load<T>(Ubuf + tile_off(t, i, LD), LD);
        tile_mma<T, false, false, true, 16>(acc, Ubuf + tile_off(t, kk, LD), LD,
                                        Abuf + tile_off(kk, i, LD), LD);
        tile_store<T>(Ubuf + tile_off(t, i, LD), LD, acc);
      }
    }
  };
  if constexpr (!LA) {
    auto step = [&](int kk) {
      if (w == 0)
        diag(kk);
      else if (SL)
        slot(kk);
      __syncthreads();
      scale(kk);
      __syncthreads();
      if (kk < TT - 1) {
        eliminate(kk, w, NW);
        __syncthreads();
      }
    };
    if constexpr (SL) {
#pragma unroll
      for (int kk = 0; kk < TT; ++kk) step(kk);
    } else {
      for (int kk = 0; kk < TT; ++kk) step(kk);
    }
  } else {
    // Lookahead: during the elimination of step kk, wave 0 updates the next diagonal
    // tile (task 0) and factors it at once while the other waves share the remaining
    // tasks, so the serial diagonal step overlaps the elimination instead of following
    // it.  LDS accesses of one wave complete in order.
    if (w == 0)
      diag(0);
    else if (SL)
      slot(0);
    __syncthreads();
    auto step = [&](int kk) {
      scale(kk);
      __syncthreads();
      if (kk < TT - 1) {
        if (w == 0) {
          const int i = kk + 1;
          acc_t acc = tile_load<T>(Abuf + tile_off(i, i, LD), LD);
          tile_mma<T, true, false, true, 16>(acc, Abuf + tile_off(kk, i, LD), LD,
                                         Abuf + tile_off(kk, i, LD), LD);
          tile_store<T>(Abuf + tile_off(i, i, LD), LD, acc);
          diag(i);
        } else {
          // with eight waves, wave 4 shares wave 0's SIMD, where an MFMA stalls the
          // diagonal steps: the other six waves take the elimination tiles
          if constexpr (NW == 8) {
            if (w != 4) eliminate(kk, w - (w > 4), NW - 2);
          } else {
            eliminate(kk, w, NW - 1);
          }
          if (SL) slot(kk + 1);
        }
        __syncthreads();
      }
    };
    if constexpr (SL) {
#pragma unroll
      for (int kk = 0; kk < TT; ++kk) step(kk);
    } else {
      for (int kk = 0; kk < TT; ++kk) step(kk);
    }
  }
}

// One NB x NB block staged through registers: every thread issues all of its 16-byte
// global loads at once (load), and writes them to LDS later (store), so the HBM
// latency of the next block hides behind the current block's factorisation.
template <typename Tin, int NB, int NT>
struct BlockStage {
  typedef Tin vec_t __attribute__((ext_vector_type(16 / sizeof(Tin))));
  static constexpr int VEC = 16 / sizeof(Tin);
  static constexpr int PER = NB * NB / (NT * VEC);
  static_assert(PER * NT * VEC == NB * NB, "block does not split evenly over the threads");
  vec_t r[PER];
  __device__ inline void load(const Tin* __restrict__ g) {
#pragma unroll
    for (int q = 0; q < PER; ++q) r[q] = *(const vec_t*)(g + (q * NT + (int)threadIdx.x) * VEC);
  }
  template <typename T, int LD, bool TRANS>
  __device__ inline void store(T* s) const {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        const int idx = (q * NT + (int)threadIdx.x) * VEC + e;
        const int row = idx % NB, col = idx / NB;
        if (TRANS)
          s[row * LD + col] = (T)r[q][e];
        else
          s[col * LD + row] = (T)r[q][e];
      }
    }
  }
};

// One NB x NB block staged through the registers of NTB threads (thread t of them):
// 16-byte loads issued at once, written to LDS later.
template <typename Tin, int NB, int NTB>
struct BgStage {
  typedef Tin vec_t __attribute__((ext_vector_type(16 / sizeof(Tin))));
  static constexpr int VEC = 16 / sizeof(Tin);
  static constexpr int NV = NB * NB / VEC;  // 16-byte vectors per block
  static constexpr int PER = (NV + NTB - 1) / NTB;
  vec_t r[PER];
  __device__ inline void load(const BSS_GLOBAL Tin* g, int t) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int v = q * NTB + t;
      if (q * NTB + NTB <= NV || v < NV) r[q] = *(const BSS_GLOBAL vec_t*)(g + v * VEC);
    }
  }
  template <typename T, int LD>
  __device__ inline void store(T* s, int t) const {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int v = q * NTB + t;
      if (q * NTB + NTB <= NV || v < NV) {
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
          const int idx = v * VEC + e;
          s[(idx / NB) * LD + idx % NB] = (T)r[q][e];
        }
      }
    }
  }
};

// FP32 at n = 64 with 8 waves: LDS (70 KB) leaves room for two workgroups per CU, so
// registers are capped at 128 (4 waves per SIMD) to let two chains overlap their
// serial phases; measured faster than 191 registers and one workgroup per CU.
template <typename T, int NB, int NW>
constexpr int factor_min_waves_per_eu() {
  return (sizeof(T) == 4 && NB == 64 && NW == 8) ? 4 : 1;
}
// With two workgroups per CU the other chain hides HBM latency, and the registers a
// prefetch of the next block would hold are better spent not spilling: the next
// blocks are then loaded when needed (measured 17% faster at FP32, n = 64; issuing
// the loads after the potrf instead, in flight through the trsm only, spilled and was
// 10-25% slower).
// At n = 96 (FP32, 12 waves, one workgroup per CU) the prefetch registers spill; loading
// when needed measured 10% faster there.
template <typename T, int NB, int NW>
constexpr bool factor_prefetch() {
  return factor_min_waves_per_eu<T, NB, NW>() == 1 && NB != 96;
}
// LAT: the latency variant for levels with at most one chain per CU, where no second
// workgroup shares the CU: one workgroup per CU, all registers, next blocks prefetched.
// Identical to the default variant wherever that already runs one workgroup per CU.
template <typename T, int NB, int NW>
constexpr bool factor_has_latency_variant() {
  return factor_min_waves_per_eu<T, NB, NW>() > 1;
}

template <typename Tin, typename T, int NB, int NW, bool LAT = false, bool FWD = false>
__global__ void __launch_bounds__(64 * NW)
    __attribute__((amdgpu_waves_per_eu(LAT ? 1 : factor_min_waves_per_eu<T, NB, NW>())))
    chain_factor_kernel(FactorLevelArgs p) {
  constexpr int TT = NB / 16;
  constexpr int LD = NB + LdsPad<T>::value;
  constexpr int NT = 64 * NW;
  constexpr int NN = NB * NB;
  constexpr int NS = TT * (TT + 1) / 2;           // upper tiles of a symmetric block
  // waves with the products that follow the trsm (S_L, Z, Schur, reduced-system blocks):
  // all of them, or with the fused forward sweep all but its wave
  constexpr int NWP = (FWD && NW > 1) ? NW - 1 : NW;
  constexpr int SQ = (NS + NWP - 1) / NWP;        // S accumulators per wave
  static_assert(NW % TT == 0, "waves must be a multiple of the tile rows");
  // SLOTS (FP32 factors, n = 64, eight waves, where the 128-register cap leaves no room
  // to hold a prefetched block through the potrf): the next blocks' HBM loads are
  // issued and stored by waves 1-7 in the potrf's background slots, while wave 0
  // factors diagonal tiles (measured 3% faster than loading when needed; the FP64
  // kernel's prefetch at the top of a block measured faster than slots):
  //   slot 0: B_blk load;  slot 1: B_blk into the coupling buffer;
  //   slot 2: A_{blk+1} load, written to the A buffer once the potrf has freed it
  constexpr bool SLOTS = sizeof(T) == 4 && NB == 64 && NW == 8;
  constexpr int NTB = SLOTS ? 64 * (NW - 1) : 64;  // slot threads (unused without SLOTS)
  constexpr int XG = NW / TT;                      // wave groups splitting the RHS columns
  constexpr int XQ = 2 * TT / XG;                  // trsm output tiles per wave
  constexpr int ZQ = (TT * TT + NWP - 1) / NWP;   // panel tiles per wave
  constexpr int PQ = (TT * TT + NS + NWP - 1) / NWP;
  typedef typename Mfma<T>::acc_t acc_t;

  __shared__ T Abuf[NB * LD];
  __shared__ T Ubuf[NB * LD];
  __shared__ T Rbuf[2 * NB * LD];  // columns [0,NB): B -> C ; [NB,2NB): Z -> W
  __shared__ int fail;
  // fused forward sweep: r and y of the current block, C^T y of the previous block, the
  // running sum W^T y
  __shared__ T Fr[NB], Fy[NB], Ft[NB], Fl[NB], Fd[NB];  // Fd: d of the current block

  const ChainRec cd = p.chains[blockIdx.x];
  const int start = cd.start, len = cd.len, sepL = cd.sepL, sepR = cd.sepR;
  if (len <= 0) return;
  const LevelSys S = p.sys[cd.slot];
  const int sys = S.sys;
  const bool hasL = sepL >= 0, hasR = sepR >= 0;
  const Tin* A = (const Tin*)S.A;
  const Tin* B = (const Tin*)S.B;
  T* U = (T*)S.U;
  T* C = (T*)S.C;
  T* W = (T*)S.W;
  const int w = wave_id();
  const int tid = threadIdx.x;

  acc_t Sacc[SQ];
#pragma unroll
  for (int q = 0; q < SQ; ++q) Sacc[q] = Mfma<T>::zero();

  BlockStage<Tin, NB, NT> sa, sb;  // next diagonal block / next coupling block
  if (tid == 0) fail = 0;
  sa.load(A + (long long)start * NN);
  if (hasL) {
    sb.load(B + (long long)(start - 1) * NN);
    sb.template store<T, LD, true>(Rbuf + NB * LD);  // Z_1 = B_L^T
  }
  int fail_i = len;       // first failed block of the chain (thread 0)
  float piv_run = 1.0f;   // smallest pivot ratio of the chain (wave 0)
  constexpr bool PF = !SLOTS && (LAT || factor_prefetch<T, NB, NW>());
  if (PF && (len > 1 || hasR)) sb.load(B + (long long)start * NN);
  sa.template store<T, LD, false>(Abuf);
  // FWD, the fused forward sweep (one right-hand side, the variant launched for
  // p.fwd): the last wave, which has the fewest product tasks, runs
  // chain_forward1_kernel's steps for every row tile of each block once its factors are
  // in LDS; lane (lr, lk) of row tile t holds what lane (lr, lk) of that kernel's wave
  // t holds
  const bool fwd = FWD && w == NW - 1;
  const auto dg = gptr((T*)S.d);
  const int flr = lane_id() & 15, flk = lane_id() >> 4;
  constexpr int FDQ = (NB + 63) / 64;  // d elements per lane of the sweep's wave
  if (fwd) {
#pragma unroll
    for (int q = 0; q < FDQ; ++q) {
      const int row = lane_id() + 64 * q;
      if (row < NB) Fd[row] = dg[(long long)start * NB + row];
    }
  }
  __syncthreads();

  for (int i = 0; i < len; ++i) {
    const int blk = start + i;
    const bool last = (i == len - 1);
    const bool hasC = !last || hasR;  // block couples to a later block
    if (hasC && !SLOTS) {  // B_blk (prefetched, or loaded now)
      if (!PF) sb.load(B + (long long)blk * NN);
      sb.template store<T, LD, false>(Rbuf);
    }
    // B_blk / A_{blk+1} in flight in the slots.  Declared per block: on wave 0's path
    // they hold no value, so no register is kept for them across its diagonal steps.
    BgStage<Tin, NB, NTB> gb, ga;
    auto slot = [&](int k) {
      if constexpr (SLOTS) {
        const int bt = opaque_v(tid) - 64;
        if (k == 0) {
          if (hasC) gb.load(gptr(B) + (long long)blk * NN, bt);
        } else if (k == 1) {
          if (hasC) gb.template store<T, LD>(Rbuf, bt);
        } else if (k == 2) {
          if (!last) ga.load(gptr(A) + (long long)(blk + 1) * NN, bt);
        }
      }
    };
    // prefetch the next block pair; it lands while this block is factored
    if (PF && !last) sa.load(A + (long long)(blk + 1) * NN);
    if (PF && (i + 1 < len - 1 || (i + 1 == len - 1 && hasR))) sb.load(B + (long long)(blk + 1) * NN);

    // pivot monitor: wave 0, lane j < NB keeps the diagonal entry A_jj of this block
    float adiag = 0.0f;
    if (p.pivmin && w == 0 && lane_id() < NB) adiag = (float)Abuf[lane_id() * (LD + 1)];
    potrf_inv<T, NB, LD, NW, (SLOTS || LAT || factor_min_waves_per_eu<T, NB, NW>() == 1), SLOTS>(
        Abuf, Ubuf, &fail, slot);
    // failures and pivot ratios are kept per workgroup and reported once after the
    // chain: a global atomic inside the loop, on a path some lanes skip, would make the
    // memory waits after it conservative (and drain the prefetches)
    if (tid == 0 && fail) {
      fail_i = min(fail_i, i);
      fail = 0;
    }
    if (p.pivmin && w == 0) {
      // U_jj^2 / A_jj = 1 / (Uinv_jj^2 A_jj): the fraction of the diagonal left after
      // elimination; a failed or non-finite pivot counts as 0
      float ratio = 1.0f;
      if (lane_id() < NB) {
        const float ui = (float)Ubuf[lane_id() * (LD + 1)];
        const float q = 1.0f / (ui * ui * adiag);
        ratio = (q >= 0.0f && q <= 3.0e38f) ? q : 0.0f;
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) ratio = fminf(ratio, __shfl_xor(ratio, o));
      piv_run = fminf(piv_run, ratio);
    }

    // [C | W] = Uinv^T [B | Z].  Output tile (r, c) costs r+1 tile products; wave w
    // (group g = w / TT, residue rho = w % TT) takes columns g*XQ .. g*XQ+XQ-1 with
    // r = (rho - c) mod TT, so every wave does the same amount of MFMA work.
    const int c0 = hasC ? 0 : TT;
    const int c1 = hasL ? 2 * TT : TT;
    const int wb = SLOTS ? opaque_s(w) : w;  // task tiles recomputed per block
    const int xg = wb / TT, rho = wb % TT;
    acc_t X[XQ];
#pragma unroll
    for (int q = 0; q < XQ; ++q) {
      const int c = xg * XQ + q;
      const int r = ((rho - c) % TT + TT) % TT;
      X[q] = Mfma<T>::zero();
      if (c >= c0 && c < c1) {
        for (int t = 0; t <= r; ++t)
          tile_mma<T, true, false, false, 16>(X[q], Ubuf + tile_off(t, r, LD), LD,
                                              Rbuf + tile_off(t, c, LD), LD);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < XQ; ++q) {
      const int c = xg * XQ + q;
      const int r = ((rho - c) % TT + TT) % TT;
      if (c >= c0 && c < c1) tile_store<T>(Rbuf + tile_off(r, c, LD), LD, X[q]);
    }
    if (!last) {  // A_{blk+1} (prefetched, loaded in a slot, or loaded now)
      if constexpr (SLOTS) {
        if (w > 0) ga.template store<T, LD>(Abuf, opaque_v(tid) - 64);
      } else {
        if (!PF) sa.load(A + (long long)(blk + 1) * NN);
        sa.template store<T, LD, false>(Abuf);
      }
    }
    __syncthreads();
    if (fwd) {
      const int blk1 = start + i;
      // lane offsets recomputed per block (held across the chain they would spill)
      const int flr = opaque_v(lane_id() & 15), flk = lane_id() >> 4;
      // d of the next block, one element per lane, in flight through this step's
      // products (registers held within this step only)
      T nd[FDQ];
      const int fl = opaque_v(lane_id());  // (its addresses held across the chain spill)
#pragma unroll
      for (int q = 0; q < FDQ; ++q) {
        const int row = fl + 64 * q;
        nd[q] = (!last && row < NB) ? dg[(long long)(blk1 + 1) * NB + row] : T(0);
      }
      // r = d_blk - C_{blk-1}^T y_{blk-1}
#pragma unroll
      for (int t = 0; t < TT; ++t) {
        const int row = 16 * t + flr;
        T v = Fd[row];
        if (i > 0) v -= Ft[row];
        if (flk == 0) Fr[row] = v;
      }
      // y = Uinv^T r (tiles t' <= t of column tile t), in place of d; the row-tile loops
      // stay rolled, which keeps this step's registers few (the accumulators of the
      // other waves' products share the budget)
#pragma unroll 1
      for (int t = 0; t < TT; ++t) {
        const int row = 16 * t + flr;
        const T y = fwd_dot(Ubuf + row * LD, Fr, 0, t + 1);
        if (flk == 0) {
          dg[(long long)blk1 * NB + row] = y;
          Fy[row] = y;
        }
      }
      // r_L += W^T y;  C^T y for the next block, or r_R after the last one
#pragma unroll 1
      for (int t = 0; t < TT; ++t) {
        const int row = 16 * t + flr;
        if (hasL) {
          const T v = fwd_dot(Rbuf + (NB + row) * LD, Fy, 0, TT);
          if (flk == 0) Fl[row] = (i > 0 ? Fl[row] : T(0)) + v;
        }
        if (!last) {
          const T v = fwd_dot(Rbuf + row * LD, Fy, 0, TT);
          if (flk == 0) Ft[row] = v;
        } else if (hasR) {
          const T v = fwd_dot1<T, TT>(Rbuf + row * LD, Fy);
          if (flk == 0) gptr((T*)S.rR)[(long long)sepR * NB + row] = v;
        }
      }
      if (!last) {
#pragma unroll
        for (int q = 0; q < FDQ; ++q) {
          const int row = fl + 64 * q;
          if (row < NB) Fd[row] = nd[q];
        }
      }
    }
    // factors to HBM as flat 16-byte stores: C (B slot), W (panel slot), Uinv upper tiles
    {
      typedef T vt __attribute__((ext_vector_type(16 / sizeof(T))));
      constexpr int V = 16 / sizeof(T);
      for (int idx = tid * V; idx < NN; idx += NT * V) {
        const int row = idx % NB, col = idx / NB;
        vt u, cc, ww;
#pragma unroll
        for (int e = 0; e < V; ++e) {
          u[e] = Ubuf[col * LD + row + e];
          cc[e] = Rbuf[col * LD + row + e];
          ww[e] = Rbuf[(NB + col) * LD + row + e];
        }
        if ((row >> 4) <= (col >> 4)) *(vt*)(U + (long long)blk * NN + idx) = u;
        if (hasC) *(vt*)(C + (long long)blk * NN + idx) = cc;
        if (hasL) *(vt*)(W + (long long)blk * NN + idx) = ww;
      }
    }

    // S_L += W^T W (upper tiles, persistent per wave)
    if (hasL && (NWP == NW || wb < NWP)) {
#pragma unroll
      for (int q = 0; q < SQ; ++q) {
        int s = wb + NWP * q;
        if (s < NS) {
          int a = 0;
          while (s >= TT - a) {
            s -= TT - a;
            ++a;
          }
          const int b = a + s;
          tile_mma<T, true, false, false, NB>(Sacc[q], Rbuf + (NB + 16 * a) * LD, LD,
                                          Rbuf + (NB + 16 * b) * LD, LD);
        }
      }
    }
    if (!last) {
      // Z_{i+1} = -C^T W (held until all reads of Rbuf are done); A_{i+1} -= C^T C
      acc_t Z[ZQ];
#pragma unroll
      for (int q = 0; q < PQ; ++q) {
        const int task = (NWP == NW || wb < NWP) ? wb + NWP * q : TT * TT + NS;
        if (task < TT * TT) {
          if (hasL && q < ZQ) {
            const int a = task % TT, b = task / TT;
            Z[q] = Mfma<T>::zero();
            tile_mma<T, true, false, true, NB>(Z[q], Rbuf + 16 * a * LD, LD,
                                           Rbuf + (NB + 16 * b) * LD, LD);
          }
        } else if (task < TT * TT + NS) {
          int s = task - TT * TT, a = 0;
          while (s >= TT - a) {
            s -= TT - a;
            ++a;
          }
          const int b = a + s;
          acc_t acc = tile_load<T>(Abuf + tile_off(a, b, LD), LD);
          tile_mma<T, true, false, true, NB>(acc, Rbuf + 16 * a * LD, LD, Rbuf + 16 * b * LD, LD);
          tile_store<T>(Abuf + tile_off(a, b, LD), LD, acc);
        }
      }
      __syncthreads();
      if (hasL) {
#pragma unroll
        for (int q = 0; q < ZQ; ++q) {
          const int task = (NWP == NW || wb < NWP) ? wb + NWP * q : TT * TT;
          if (task < TT * TT) {
            const int a = task % TT, b = task / TT;
            tile_store<T>(Rbuf + tile_off(a, TT + b, LD), LD, Z[q]);
          }
        }
      }
    } else {
      // reduced-system contributions of the last block
      if (hasR) {
        T* sr = (T*)S.SR + (long long)sepR * NN;
        for (int task = w; (NWP == NW || w < NWP) && task < NS; task += NWP) {
          int s = task, a = 0;
          while (s >= TT - a) {
            s -= TT - a;
            ++a;
          }
          const int b = a + s;
          acc_t acc = Mfma<T>::zero();
          tile_mma<T, true, false, false, NB>(acc, Rbuf + 16 * a * LD, LD, Rbuf + 16 * b * LD, LD);
          tile_store<T>(sr + tile_off(a, b, NB), NB, acc);
        }
        if (hasL) {
          T* bn = (T*)S.Bn + (long long)sepL * NN;
          for (int task = w; (NWP == NW || w < NWP) && task < TT * TT; task += NWP) {
            const int a = task % TT, b = task / TT;
            acc_t acc = Mfma<T>::zero();
            tile_mma<T, true, false, true, NB>(acc, Rbuf + (NB + 16 * a) * LD, LD,
                                           Rbuf + 16 * b * LD, LD);
            tile_store<T>(bn + tile_off(a, b, NB), NB, acc);
          }
        }
      }
    }
  }
  if (hasL) {
    T* sl = (T*)S.SL + (long long)sepL * NN;
#pragma unroll
    for (int q = 0; q < SQ; ++q) {
      int s = w + NWP * q;
      if ((NWP == NW || w < NWP) && s < NS) {
        int a = 0;
        while (s >= TT - a) {
          s -= TT - a;
          ++a;
        }
        const int b = a + s;
        tile_store<T>(sl + tile_off(a, b, NB), NB, Sacc[q]);
      }
    }
  }
  if (fwd && hasL && flk == 0) {
#pragma unroll
    for (int t = 0; t < TT; ++t) gptr((T*)S.rL)[(long long)sepL * NB + 16 * t + flr] = Fl[16 * t + flr];
  }
  // the chain's first failed block (original indices increase along a chain, so the
  // first local failure is the smallest original index) and its smallest pivot ratio
  if (tid == 0 && fail_i < len) {
    const int blk = start + fail_i;
    atomicMin(p.info + sys, (S.orig ? S.orig[blk] : blk) + 1);
  }
  if (p.pivmin && tid == 0) atomicMin(p.pivmin + sys, __float_as_uint(piv_run));
}

}  // namespace bss
