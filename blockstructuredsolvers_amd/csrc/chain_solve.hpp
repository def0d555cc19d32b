// chain_solve.hpp -- per-chain forward and backward substitution sweeps of one level.
//
// With the factors of chain_factor.hpp the level's solve is, per chain,
//   forward : y_i = Uinv_i^T (d_i - C_{i-1}^T y_{i-1})          (first block: no C term)
//             r_R = C_last^T y_last,   r_L = sum_i W_i^T y_i     (reduced right-hand side)
//   [reduced system over the separators solved by the next level: x_sep]
//   backward: x_i = Uinv_i (y_i - C_i x_{i+1} - W_i x_L),  x_{last+1} = x_R
// which replaces the reference's compute_schur_RHS! (src/backbone_rocm.jl:140-154),
// update_boundary! (src/backbone_rocm.jl:156-170) and cholesky_solve_batched! with
// nd = 1 (src/backbone_rocm.jl:84-120).  Each wave owns one 16-row tile of the block;
// the right-hand sides (up to 16 columns) ride in the MFMA's 16-wide N dimension, so
// one code path serves nrhs = 1 and nrhs = 16.  Factor blocks are read straight from
// HBM into MFMA operand registers (contiguous 32-byte column segments); only the
// n x 16 vectors live in LDS.
#pragma once

#include "common.hpp"
#include "plan_tables.hpp"

namespace bss {

struct SolveLevelArgs {
  const ChainRec* chains;
  const LevelSys* sys;  // U, C, W, d; forward writes rR/rL, backward reads dn (x_sep)
  int nrhs;
  // forward sweeps: when set and *skip == 0, the sweep already ran (inside the factor
  // kernel, on the same right-hand side) and the launch does nothing
  const int* skip = nullptr;
};

template <typename T>
struct VecPad {
  static constexpr int value = sizeof(T) == 8 ? 2 : 4;
};

// zero an LDS vector buffer (NB rows x 16 columns, leading dimension LDV)
template <typename T, int LDV, int NT>
__device__ inline void zero_vec(T* v) {
  for (int i = threadIdx.x; i < 16 * LDV; i += NT) v[i] = T(0);
}

// d-tile (rows 16w.., all rhs columns) from a global n x nrhs column-major block
// Every lane loads (columns past nrhs read column 0): a predicated load would make the
// in-order wait counters conservative for every later load.  load_rhs_raw leaves those
// columns as loaded, so a tile loaded ahead is not waited for until mask_rhs uses it.
template <typename T, int NB, typename P>
__device__ inline typename Mfma<T>::acc_t load_rhs_raw(P g, int w, int nrhs) {
  const int l = lane_id();
  const int lr = l & 15, lk = l >> 4;
  const int c = lr < nrhs ? lr : 0;
  typename Mfma<T>::acc_t a;
#pragma unroll
  for (int v = 0; v < 4; ++v) a[v] = g[c * NB + 16 * w + Mfma<T>::row(lk, v)];
  return a;
}
template <typename T>
__device__ inline typename Mfma<T>::acc_t mask_rhs(typename Mfma<T>::acc_t a, int nrhs) {
  const bool in = (lane_id() & 15) < nrhs;
#pragma unroll
  for (int v = 0; v < 4; ++v) a[v] = in ? a[v] : T(0);
  return a;
}
template <typename T, int NB, typename P>
__device__ inline typename Mfma<T>::acc_t load_rhs_tile(P g, int w, int nrhs) {
  return mask_rhs<T>(load_rhs_raw<T, NB>(g, w, nrhs), nrhs);
}

template <typename T, int NB, typename P>
__device__ inline void store_rhs_tile(P g, int w, int nrhs, const typename Mfma<T>::acc_t& a) {
  const int l = lane_id();
  const int lr = l & 15, lk = l >> 4;
  if (lr < nrhs) {
#pragma unroll
    for (int v = 0; v < 4; ++v) g[lr * NB + 16 * w + Mfma<T>::row(lk, v)] = a[v];
  }
}

// global n x nrhs block -> LDS vector buffer (columns >= nrhs left untouched)
template <typename T, int NB, int LDV, int NT, typename P>
__device__ inline void load_vec(P g, T* v, int nrhs) {
  for (int i = threadIdx.x; i < NB * nrhs; i += NT) v[(i / NB) * LDV + (i % NB)] = g[i];
}

// The HBM operand of a 16-row product, loaded ahead into registers: op(X) rows
// i = 0..15 over the 16-deep k slices [s_lo, s_hi) (TA: X[k][i] at xb[i*ldx + k], else
// X[i][k] at xb[k*ldx + i]), in tile_mma's k order.
template <typename T, int NB, bool TA>
struct XOps {
  T a[NB / 16][4];
  template <typename P>
  __device__ inline void load(P __restrict__ xb, int ldx, int s_lo = 0,
                              int s_hi = NB / 16) {
    const int l = lane_id();
    const int lr = l & 15, lk = l >> 4;
    // slices outside [s_lo, s_hi) re-read an in-range slice (a cache hit) and are zeroed:
    // a fixed number of loads per call keeps the in-order wait counters exact
#pragma unroll
    for (int s = 0; s < NB / 16; ++s)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int sc = s < s_lo ? s_lo : (s >= s_hi ? s_hi - 1 : s);
        const int k = 16 * sc + 4 * lk + q;
        const T x = TA ? xb[lr * ldx + k] : xb[k * ldx + lr];
        a[s][q] = (s >= s_lo && s < s_hi) ? x : T(0);
      }
  }
  // acc (+/-)= op(X) Y over the slices [s_lo, s_hi), Y[k][j] at yb[j*ldy + k] (LDS)
  template <bool NEG>
  __device__ inline void mma(typename Mfma<T>::acc_t& acc, const T* __restrict__ yb, int ldy,
                             int s_lo = 0, int s_hi = NB / 16) const {
    const int l = lane_id();
    const int lr = l & 15, lk = l >> 4;
#pragma unroll
    for (int s = 0; s < NB / 16; ++s) {
      if (s < s_lo || s >= s_hi) continue;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const T b = yb[lr * ldy + 16 * s + 4 * lk + q];
        acc = Mfma<T>::mma(NEG ? -a[s][q] : a[s][q], b, acc);
      }
    }
  }
  // One right-hand side: (op(X) y)_i for row i = lane % 16 over the slices [s_lo, s_hi),
  // y[k] at yb[k] (LDS), in every lane of the row.  Vector FMAs over the lane's own k
  // slots, then a sum over the four lanes that share the row.
  __device__ inline T dot(const T* __restrict__ yb, int s_lo = 0, int s_hi = NB / 16) const {
    const int lk = lane_id() >> 4;
    T t0 = T(0), t1 = T(0);
#pragma unroll
    for (int s = 0; s < NB / 16; ++s) {
      if (s < s_lo || s >= s_hi) continue;
      t0 = fma(a[s][0], yb[16 * s + 4 * lk + 0], t0);
      t1 = fma(a[s][1], yb[16 * s + 4 * lk + 1], t1);
      t0 = fma(a[s][2], yb[16 * s + 4 * lk + 2], t0);
      t1 = fma(a[s][3], yb[16 * s + 4 * lk + 3], t1);
    }
    return row4_sum(t0 + t1);
  }
};

template <typename T, int NB>
__global__ void __launch_bounds__(4 * NB) chain_forward_kernel(SolveLevelArgs p) {
  if (p.skip && *p.skip == 0) return;  // the whole workgroup leaves (uniform)
  constexpr int TT = NB / 16;
  constexpr int NT = 64 * TT;
  constexpr int LDV = NB + VecPad<T>::value;
  constexpr long long NN = (long long)NB * NB;
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T Y[16 * LDV];
  __shared__ T R[16 * LDV];

  const ChainRec cd = p.chains[blockIdx.x];
  const int start = cd.start, len = cd.len, sepL = cd.sepL, sepR = cd.sepR;
  const LevelSys S = p.sys[cd.slot];
  const int nrhs = p.nrhs;
  const long long ND = (long long)NB * nrhs;
  const auto U = gptr((const T*)S.U);
  const auto C = gptr((const T*)S.C);
  const auto W = gptr((const T*)S.W);
  const auto d = gptr((T*)S.d);
  const int w = wave_id();
  // right-hand side of block q: below the top level (fused hand-off) it is assembled
  // from the previous level, d[q] = dp[pp[q]] - rRp[q] - rLp[q]
  const bool assemble = S.fused && S.dp;
  // the loaded parts of a right-hand side tile, combined only where the tile is used
  // (so a tile loaded one block ahead does not make this block wait for it)
  struct RhsParts {
    acc_t a, r, l;
  };
  auto rhs_parts = [&](long long q) -> RhsParts {
    RhsParts t;
    if (!assemble) {
      t.a = load_rhs_raw<T, NB>(d + q * ND, w, nrhs);
    } else {
      t.a = load_rhs_raw<T, NB>(gptr((const T*)S.dp) + (long long)gptr(S.pp)[q] * ND, w, nrhs);
      t.r = load_rhs_raw<T, NB>(gptr((const T*)S.rRp) + q * ND, w, nrhs);
      t.l = load_rhs_raw<T, NB>(gptr((const T*)S.rLp) + q * ND, w, nrhs);
    }
    return t;
  };
  auto rhs_combine = [&](const RhsParts& t) -> acc_t {
    if (!assemble) return mask_rhs<T>(t.a, nrhs);
    acc_t a;
#pragma unroll
    for (int v = 0; v < 4; ++v) a[v] = t.a[v] - t.r[v] - t.l[v];
    return mask_rhs<T>(a, nrhs);
  };
  auto rhs_tile = [&](long long q) -> acc_t { return rhs_combine(rhs_parts(q)); };
  // the separators this chain owns get their assembled right-hand side, for the next
  // level: the one on its right, and a separator at block 0 (which has no chain on its
  // left; the reference layout starts a level with one)
  if (assemble && sepR >= 0) {
    const long long q = start + len;
    store_rhs_tile<T, NB>(d + q * ND, w, nrhs, rhs_tile(q));
  }
  if (assemble && sepL >= 0 && start == 1) store_rhs_tile<T, NB>(d, w, nrhs, rhs_tile(0));
  if (len <= 0) return;

  zero_vec<T, LDV, NT>(Y);
  zero_vec<T, LDV, NT>(R);
  __syncthreads();

  // Each HBM operand panel of block i+1 is loaded into the registers of block i's panel
  // once that panel's product has been consumed by a later instruction (so the
  // products have retired before their source registers are overwritten); the load
  // latency then hides behind the rest of block i.
  XOps<T, NB, true> oc, ou, ow;  // C_{i-1} panel, Uinv column tiles t <= w, W_i panel
  ou.load(U + (long long)start * NN + 16 * w * NB, NB, 0, w + 1);
  if (sepL >= 0) ow.load(W + (long long)start * NN + 16 * w * NB, NB);
  acc_t accL = Mfma<T>::zero();
  // the right-hand side tile of the next block is loaded one block ahead (below the
  // top level it is three dependent loads through the hand-off index)
  RhsParts rnext = rhs_parts(start);
  for (int i = 0; i < len; ++i) {
    const long long blk = start + i;
    acc_t acc = rhs_combine(rnext);
    if (i + 1 < len) rnext = rhs_parts(blk + 1);
    if (i > 0) oc.template mma<true>(acc, Y, LDV);
    tile_store<T>(R + 16 * w, LDV, acc);
    if (i + 1 < len) oc.load(C + blk * NN + 16 * w * NB, NB);  // C_blk, used by row blk + 1
    // W_blk, into registers whose products (row blk - 1) retired before the C product
    // that tile_store above waited on
    if (sepL >= 0 && i > 0) ow.load(W + blk * NN + 16 * w * NB, NB);
    __syncthreads();
    acc = Mfma<T>::zero();
    ou.template mma<false>(acc, R, LDV, 0, w + 1);  // Uinv_blk^T r
    store_rhs_tile<T, NB>(d + blk * ND, w, nrhs, acc);
    tile_store<T>(Y + 16 * w, LDV, acc);
    if (i + 1 < len) ou.load(U + (blk + 1) * NN + 16 * w * NB, NB, 0, w + 1);
    __syncthreads();
    if (sepL >= 0) ow.template mma<false>(accL, Y, LDV);
  }
  if (sepR >= 0) {
    const long long last = start + len - 1;
    acc_t acc = Mfma<T>::zero();
    tile_mma<T, true, false, false, NB>(acc, C + last * NN + 16 * w * NB, NB, Y, LDV);
    store_rhs_tile<T, NB>(gptr((T*)S.rR) + sepR * ND, w, nrhs, acc);
  }
  if (sepL >= 0) store_rhs_tile<T, NB>(gptr((T*)S.rL) + sepL * ND, w, nrhs, accL);
}

// The forward sweep for one right-hand side (nrhs = 1).  Same schedule, operand loads
// and hand-off as chain_forward_kernel; the products are vector FMAs (XOps::dot), since
// an MFMA would compute 16 columns to keep one.  Lane (lr, lk) of wave w holds element
// 16w + lr of a block's vector; the LDS vectors hold only column 0.
template <typename T, int NB>
__global__ void __launch_bounds__(4 * NB) chain_forward1_kernel(SolveLevelArgs p) {
  if (p.skip && *p.skip == 0) return;  // the whole workgroup leaves (uniform)
  constexpr int TT = NB / 16;
  constexpr long long NN = (long long)NB * NB;
  __shared__ T Y[NB];
  __shared__ T R[NB];

  const ChainRec cd = p.chains[blockIdx.x];
  const int start = cd.start, len = cd.len, sepL = cd.sepL, sepR = cd.sepR;
  const LevelSys S = p.sys[cd.slot];
  const auto U = gptr((const T*)S.U);
  const auto C = gptr((const T*)S.C);
  const auto W = gptr((const T*)S.W);
  const auto d = gptr((T*)S.d);
  const int w = wave_id();
  const int l = lane_id(), lr = l & 15, lk = l >> 4;
  const int row = 16 * w + lr;
  const bool assemble = S.fused && S.dp;
  struct RhsParts {
    T a, r, l;
  };
  auto rhs_parts = [&](long long q) -> RhsParts {
    RhsParts t{T(0), T(0), T(0)};
    if (!assemble) {
      t.a = d[q * NB + row];
    } else {
      t.a = gptr((const T*)S.dp)[(long long)gptr(S.pp)[q] * NB + row];
      t.r = gptr((const T*)S.rRp)[q * NB + row];
      t.l = gptr((const T*)S.rLp)[q * NB + row];
    }
    return t;
  };
  auto rhs_combine = [&](const RhsParts& t) -> T { return assemble ? t.a - t.r - t.l : t.a; };
  if (assemble && sepR >= 0) {
    const long long q = start + len;
    const T v = rhs_combine(rhs_parts(q));
    if (lk == 0) d[q * NB + row] = v;
  }
  if (assemble && sepL >= 0 && start == 1) {
    const T v = rhs_combine(rhs_parts(0));
    if (lk == 0) d[row] = v;
  }
  if (len <= 0) return;

  XOps<T, NB, true> oc, ou, ow;  // C_{i-1} panel, Uinv column tiles t <= w, W_i panel
  ou.load(U + (long long)start * NN + 16 * w * NB, NB, 0, w + 1);
  if (sepL >= 0) ow.load(W + (long long)start * NN + 16 * w * NB, NB);
  T accL = T(0);
  RhsParts rnext = rhs_parts(start);
  for (int i = 0; i < len; ++i) {
    const long long blk = start + i;
    T v = rhs_combine(rnext);
    if (i + 1 < len) rnext = rhs_parts(blk + 1);
    if (i > 0) v -= oc.dot(Y);
    if (lk == 0) R[row] = v;
    if (i + 1 < len) oc.load(C + blk * NN + 16 * w * NB, NB);
    if (sepL >= 0 && i > 0) ow.load(W + blk * NN + 16 * w * NB, NB);
    __syncthreads();
    v = ou.dot(R, 0, w + 1);  // Uinv_blk^T r
    if (lk == 0) {
      d[blk * NB + row] = v;
      Y[row] = v;
    }
    if (i + 1 < len) ou.load(U + (blk + 1) * NN + 16 * w * NB, NB, 0, w + 1);
    __syncthreads();
    if (sepL >= 0) accL += ow.dot(Y);
  }
  if (sepR >= 0) {
    // C_last^T y_last, straight from HBM (one use)
    const auto cb = C + (long long)(start + len - 1) * NN + (long long)row * NB;
    T t = T(0);
#pragma unroll
    for (int s = 0; s < TT; ++s)
#pragma unroll
      for (int q = 0; q < 4; ++q) t = fma(cb[16 * s + 4 * lk + q], Y[16 * s + 4 * lk + q], t);
    t = row4_sum(t);
    if (lk == 0) gptr((T*)S.rR)[(long long)sepR * NB + row] = t;
  }
  if (sepL >= 0 && lk == 0) gptr((T*)S.rL)[(long long)sepL * NB + row] = accL;
}

template <typename T, int NB>
__global__ void __launch_bounds__(4 * NB) chain_backward_kernel(SolveLevelArgs p) {
  constexpr int TT = NB / 16;
  constexpr int NT = 64 * TT;
  constexpr int LDV = NB + VecPad<T>::value;
  constexpr long long NN = (long long)NB * NB;
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T Xn[16 * LDV];
  __shared__ T XL[16 * LDV];
  __shared__ T R[16 * LDV];

  const ChainRec cd = p.chains[blockIdx.x];
  const int start = cd.start, len = cd.len, sepL = cd.sepL, sepR = cd.sepR;
  const LevelSys S = p.sys[cd.slot];
  const int nrhs = p.nrhs;
  const long long ND = (long long)NB * nrhs;
  const auto U = gptr((const T*)S.U);
  const auto C = gptr((const T*)S.C);
  const auto W = gptr((const T*)S.W);
  const auto d = gptr((T*)S.d);
  const auto xs = gptr((const T*)S.dn);
  const int w = wave_id();
  // fused hand-off: the solutions of the separators this chain owns (as in the forward
  // sweep) into this level's vector, for the level above or the caller at the top
  if (S.fused && sepR >= 0)
    store_rhs_tile<T, NB>(d + (long long)(start + len) * ND, w, nrhs,
                          load_rhs_tile<T, NB>(xs + sepR * ND, w, nrhs));
  if (S.fused && sepL >= 0 && start == 1)
    store_rhs_tile<T, NB>(d, w, nrhs, load_rhs_tile<T, NB>(xs + sepL * ND, w, nrhs));
  if (len <= 0) return;

  zero_vec<T, LDV, NT>(Xn);
  zero_vec<T, LDV, NT>(XL);
  zero_vec<T, LDV, NT>(R);
  __syncthreads();
  if (sepR >= 0) load_vec<T, NB, LDV, NT>(xs + sepR * ND, Xn, nrhs);
  if (sepL >= 0) load_vec<T, NB, LDV, NT>(xs + sepL * ND, XL, nrhs);
  __syncthreads();

  // operand panels of block i-1 are loaded right after block i consumed them
  XOps<T, NB, false> oc, ow, ou;  // rows 16w.. of C_i, W_i, and of Uinv_i (tiles t >= w)
  {
    const long long blk = start + len - 1;
    // C of the last block exists only when a right separator follows it (a level's
    // last block has no coupling block)
    if (sepR >= 0) oc.load(C + blk * NN + 16 * w, NB);
    if (sepL >= 0) ow.load(W + blk * NN + 16 * w, NB);
    ou.load(U + blk * NN + 16 * w, NB, w, TT);
  }
  acc_t rnext = load_rhs_raw<T, NB>(d + (long long)(start + len - 1) * ND, w, nrhs);
  for (int i = len - 1; i >= 0; --i) {
    const long long blk = start + i;
    acc_t acc = mask_rhs<T>(rnext, nrhs);  // y_blk; block blk - 1's tile loads one ahead
    if (i > 0) rnext = load_rhs_raw<T, NB>(d + (blk - 1) * ND, w, nrhs);
    if (i < len - 1 || sepR >= 0) oc.template mma<true>(acc, Xn, LDV);
    if (sepL >= 0) ow.template mma<true>(acc, XL, LDV);
    tile_store<T>(R + 16 * w, LDV, acc);  // consumes the products above
    if (i > 0) {
      oc.load(C + (blk - 1) * NN + 16 * w, NB);
      if (sepL >= 0) ow.load(W + (blk - 1) * NN + 16 * w, NB);
    }
    __syncthreads();
    acc = Mfma<T>::zero();
    ou.template mma<false>(acc, R, LDV, w, TT);  // Uinv_blk t
    store_rhs_tile<T, NB>(d + blk * ND, w, nrhs, acc);
    tile_store<T>(Xn + 16 * w, LDV, acc);
    if (i > 0) ou.load(U + (blk - 1) * NN + 16 * w, NB, w, TT);
    __syncthreads();
  }
}

// The backward sweep for one right-hand side (layout as chain_forward1_kernel).
template <typename T, int NB>
__global__ void __launch_bounds__(4 * NB) chain_backward1_kernel(SolveLevelArgs p) {
  constexpr int TT = NB / 16;
  constexpr long long NN = (long long)NB * NB;
  __shared__ T Xn[NB];
  __shared__ T XL[NB];
  __shared__ T R[NB];

  const ChainRec cd = p.chains[blockIdx.x];
  const int start = cd.start, len = cd.len, sepL = cd.sepL, sepR = cd.sepR;
  const LevelSys S = p.sys[cd.slot];
  const auto U = gptr((const T*)S.U);
  const auto C = gptr((const T*)S.C);
  const auto W = gptr((const T*)S.W);
  const auto d = gptr((T*)S.d);
  const auto xs = gptr((const T*)S.dn);
  const int w = wave_id();
  const int l = lane_id(), lr = l & 15, lk = l >> 4;
  const int row = 16 * w + lr;
  if (S.fused && sepR >= 0) {
    const T v = xs[(long long)sepR * NB + row];
    if (lk == 0) d[(long long)(start + len) * NB + row] = v;
  }
  if (S.fused && sepL >= 0 && start == 1) {
    const T v = xs[(long long)sepL * NB + row];
    if (lk == 0) d[row] = v;
  }
  if (len <= 0) return;

  // the separator solutions (read only when the separator exists)
  if (threadIdx.x < NB) {
    const int k = threadIdx.x;
    if (sepR >= 0) Xn[k] = xs[(long long)sepR * NB + k];
    if (sepL >= 0) XL[k] = xs[(long long)sepL * NB + k];
  }
  __syncthreads();

  XOps<T, NB, false> oc, ow, ou;  // rows 16w.. of C_i, W_i, and of Uinv_i (tiles t >= w)
  {
    const long long blk = start + len - 1;
    if (sepR >= 0) oc.load(C + blk * NN + 16 * w, NB);
    if (sepL >= 0) ow.load(W + blk * NN + 16 * w, NB);
    ou.load(U + blk * NN + 16 * w, NB, w, TT);
  }
  T rnext = d[(long long)(start + len - 1) * NB + row];
  for (int i = len - 1; i >= 0; --i) {
    const long long blk = start + i;
    T v = rnext;
    if (i > 0) rnext = d[(blk - 1) * NB + row];
    if (i < len - 1 || sepR >= 0) v -= oc.dot(Xn);
    if (sepL >= 0) v -= ow.dot(XL);
    if (lk == 0) R[row] = v;
    if (i > 0) {
      oc.load(C + (blk - 1) * NN + 16 * w, NB);
      if (sepL >= 0) ow.load(W + (blk - 1) * NN + 16 * w, NB);
    }
    __syncthreads();
    v = ou.dot(R, w, TT);  // Uinv_blk t
    if (lk == 0) {
      d[blk * NB + row] = v;
      Xn[row] = v;
    }
    if (i > 0) ou.load(U + (blk - 1) * NN + 16 * w, NB, w, TT);
    __syncthreads();
  }
}

}  // namespace bss
