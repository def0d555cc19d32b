// common.hpp -- shared device helpers for the gfx950 block-tridiagonal Cholesky kernels.
//
// Every dense product the solver needs is a 16x16 output tile of op(X) * op(Y) where
// X and Y are column-major blocks (the reference layout, src/gpu_seq.jl:29-43).
// They map onto the CDNA MFMA 16x16x4 instructions (f64 and f32, wave64):
//   A operand: lane l supplies P[l%16][k] for k-slot l/16
//   B operand: lane l supplies Q[k][l%16] for k-slot l/16
//   D        : lane l, register v (0..3) holds D[row][l%16] with
//              row = 4*(l/16) + v for f32  and  row = l/16 + 4*v for f64
//              (the f64 mapping was measured on gfx950 by tests/test_gpu_tiles.py)
// The k ordering inside a product is free, so k = k0 + 4*(l/16) + s over the four
// instructions s = 0..3 of each 16-deep step.  With that ordering a lane's A operand
// of X^T (column i of X) and its B operand of Y (column j of Y) are 4 consecutive
// elements of one column; D tiles are stored column by column (32-byte runs of 4
// rows across the lanes of a column), so loads and stores stay contiguous in the
// column-major layout.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bss {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  typedef d4_t acc_t;
  // row of the 16x16 result held by register v of a lane with l/16 == lk
  __device__ static constexpr int row(int lk, int v) { return lk + 4 * v; }
  __device__ static inline acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
  __device__ static inline acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

template <>
struct Mfma<float> {
  typedef f4_t acc_t;
  __device__ static constexpr int row(int lk, int v) { return 4 * lk + v; }
  __device__ static inline acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
  __device__ static inline acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

__device__ inline int lane_id() { return threadIdx.x & 63; }

// Sum over the four 16-lane rows of the wave (lanes l, l ^ 16, l ^ 32, l ^ 48), in
// every lane, as (t0 + t1) + (t2 + t3).  gfx950's row swaps (v_permlane16_swap and
// v_permlane32_swap) are VALU operations; a shuffle would be two round trips through
// the LDS crossbar (ds_bpermute) on the critical path of every block step.
__device__ inline float row4_sum(float t) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(t), __float_as_uint(t),
                                                  false, false);
  t = __uint_as_float(r[0]) + __uint_as_float(r[1]);
  const auto q = __builtin_amdgcn_permlane32_swap(__float_as_uint(t), __float_as_uint(t),
                                                  false, false);
  return __uint_as_float(q[0]) + __uint_as_float(q[1]);
}
__device__ inline double row4_sum(double t) {
  unsigned lo = (unsigned)__double2loint(t), hi = (unsigned)__double2hiint(t);
  auto rl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  auto rh = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  t = __hiloint2double((int)rh[0], (int)rl[0]) + __hiloint2double((int)rh[1], (int)rl[1]);
  lo = (unsigned)__double2loint(t);
  hi = (unsigned)__double2hiint(t);
  rl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  rh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  return __hiloint2double((int)rh[0], (int)rl[0]) + __hiloint2double((int)rh[1], (int)rl[1]);
}
// wave index, in a scalar register: branches on it are uniform, so loads under them stay
// unpredicated and the wait counters can be tracked exactly
__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

template <typename T>
__device__ inline T rsqrt_t(T x);
template <>
__device__ inline double rsqrt_t<double>(double x) {
  return 1.0 / sqrt(x);
}
template <>
__device__ inline float rsqrt_t<float>(float x) {
  return 1.0f / sqrtf(x);
}

// Pointers read from device tables are generic, and loads through them compile to FLAT
// memory instructions, which count against the LDS wait counter as well: every wait for
// an LDS read then also waits for all HBM loads in flight, defeating prefetches.
// Kernels reach HBM through global-address-space pointers instead (gptr).
#define BSS_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __attribute__((always_inline)) inline BSS_GLOBAL T* gptr(T* p) {
  return (BSS_GLOBAL T*)p;
}

// acc += op(X) * op(Y) for one 16x16 output tile, K a multiple of 16 (compile-time, so
// the whole product is unrolled and every operand load issues before the MFMAs).
//   TA: op(X) = X^T  -> P[i][k] = X[k][i] at xb[i*ldx + k]; else X[i][k] at xb[k*ldx + i]
//   TB: op(Y) = Y^T  -> Q[k][j] = Y[j][k] at yb[k*ldy + j]; else Y[k][j] at yb[j*ldy + k]
//   NEG: subtract instead of add (the A operand is negated).
// xb/yb point at element (0,0) of the operands' 16-row/col panels; they may live in
// LDS or global memory.
template <typename T, bool TA, bool TB, bool NEG, int K, typename PX, typename PY>
__device__ inline void tile_mma(typename Mfma<T>::acc_t& acc, PX __restrict__ xb, int ldx,
                                PY __restrict__ yb, int ldy) {
  static_assert(K % 16 == 0, "K must be a multiple of 16");
  const int l = lane_id();
  const int lr = l & 15, lk = l >> 4;
#pragma unroll
  for (int k0 = 0; k0 < K; k0 += 16) {
    T a[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = k0 + 4 * lk + q;
      a[q] = TA ? xb[lr * ldx + k] : xb[k * ldx + lr];
      b[q] = TB ? yb[k * ldy + lr] : yb[lr * ldy + k];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) acc = Mfma<T>::mma(NEG ? -a[q] : a[q], b[q], acc);
  }
}

// acc += op(X) op(Y) over K in 16-deep steps unrolled by two (not fully), which keeps
// the operand registers of a full-depth product bounded
template <typename T, bool TA, bool TB, bool NEG, int K, typename PX, typename PY>
__device__ inline void tile_mma_k(typename Mfma<T>::acc_t& acc, PX __restrict__ xb, int ldx,
                                  PY __restrict__ yb, int ldy) {
#pragma unroll 2
  for (int k0 = 0; k0 < K; k0 += 16)
    tile_mma<T, TA, TB, NEG, 16>(acc, TA ? xb + k0 : xb + (long long)k0 * ldx, ldx,
                                 TB ? yb + (long long)k0 * ldy : yb + k0, ldy);
}

// Load / store a D-layout tile from/to a column-major buffer (element (0,0) at p).
template <typename T>
__device__ inline typename Mfma<T>::acc_t tile_load(const T* p, int ld) {
  const int l = lane_id();
  const int lr = l & 15, lk = l >> 4;
  typename Mfma<T>::acc_t a;
#pragma unroll
  for (int v = 0; v < 4; ++v) a[v] = p[lr * ld + Mfma<T>::row(lk, v)];
  return a;
}

template <typename T>
__device__ inline void tile_store(T* p, int ld, const typename Mfma<T>::acc_t& a) {
  const int l = lane_id();
  const int lr = l & 15, lk = l >> 4;
#pragma unroll
  for (int v = 0; v < 4; ++v) p[lr * ld + Mfma<T>::row(lk, v)] = a[v];
}

// Order-independent max of non-negative doubles through their IEEE bit patterns
// (non-negative doubles order like their uint64 encodings), so device-wide norm
// reductions are deterministic.
__device__ inline void atomic_max_nonneg(unsigned long long* addr, double v) {
  atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

// Whether a level's forward sweep runs in the factor kernel (FactorLevelArgs::fwd): FP32
// factors on the LDS kernel (n <= 64 and n = 96), one right-hand side.  With FP64
// factors the fused kernel is at its register limit and spills; measured at N = 8192,
// n = 64 (16 systems) its level-0 factor took 11.1 ms instead of 9.5, more than the
// separate sweep it saves (FP32: 4.82 instead of 4.74 ms, saving a 1.0 ms sweep).
inline bool factor_fuses_forward(int nb, bool fp32_factors, int nrhs) {
  return nrhs == 1 && fp32_factors && (nb == 16 || nb == 32 || nb == 48 || nb == 64 || nb == 96);
}

}  // namespace bss
