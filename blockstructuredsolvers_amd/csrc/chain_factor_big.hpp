// chain_factor_big.hpp -- chain factorisation for block sizes n = 96 and 128.
//
// Same algorithm and outputs as chain_factor_kernel (chain_factor.hpp): per chain
//   Z_s = B_L^T,  for each block i:  A_i -= C_{i-1}^T C_{i-1};  Uinv_i = chol(A_i)^{-1};
//   [C_i | W_i] = Uinv_i^T [B_i | Z_i];  Z_{i+1} = -C_i^T W_i;  S_L += W_i^T W_i
// but at these sizes four n x n blocks no longer fit in the 160 KB of LDS (one FP64
// block at n = 128 is 128 KB).  Only the upper triangle of the block being factored
// stays in LDS, as packed 16x16 tiles; it is factored in place (potrf, diagonal tiles
// kept inverted) and then inverted in place (trtri), giving Uinv.  The right-hand
// sides B and Z and the products C, W live in their HBM output slots: the triangular
// solve reads them tile by tile (L2-resident, just written), holds its output tiles in
// registers across a barrier and writes them back in place, and the Schur products
// read C and W from there.  Global data handed from one wave to another needs only the
// workgroup barrier: its workgroup-scope release/acquire orders the stores and the
// workgroup's waves share one CU's vector L1.
#pragma once

#include "chain_factor.hpp"

namespace bss {

template <typename T, int NB>
struct BigTri {
  static constexpr int TT = NB / 16;
  // no padding for FP64 at n = 128 so that two workgroups fit in one CU's LDS
  static constexpr int PAD = (sizeof(T) == 8 && NB == 128) ? 0 : LdsPad<T>::value;
  static constexpr int LDT = 16 + PAD;
  static constexpr int TS = 16 * LDT;  // elements per stored tile
  static constexpr int NS = TT * (TT + 1) / 2;
  // upper tile (r, c), r <= c, packed column by column
  __device__ static constexpr int idx(int r, int c) { return (c * (c + 1) / 2 + r) * TS; }
};

// (r, c) of the s-th upper tile in packed column order
__device__ inline void upper_tile(int s, int& r, int& c) {
  c = 0;
  while (s > c) {
    s -= c + 1;
    ++c;
  }
  r = s;
}

// D-layout tile of a column-major Tin buffer, converted to T
template <typename Tin, typename T>
__device__ inline typename Mfma<T>::acc_t tile_load_conv(const Tin* p, int ld) {
  const int l = lane_id();
  const int lr = l & 15, lk = l >> 4;
  typename Mfma<T>::acc_t a;
#pragma unroll
  for (int v = 0; v < 4; ++v) a[v] = (T)p[lr * ld + Mfma<T>::row(lk, v)];
  return a;
}

// global writes of this workgroup visible to all of its waves: the barrier's
// workgroup-scope fences (a device-scope fence would write back the whole L2)
__device__ inline void global_handoff() { __syncthreads(); }

template <typename Tin, typename T, int NB, int NW>
__global__ void __launch_bounds__(64 * NW) chain_factor_big_kernel(FactorLevelArgs p) {
  typedef BigTri<T, NB> Tri;
  constexpr int TT = Tri::TT, NS = Tri::NS, LDT = Tri::LDT;
  constexpr int NT = 64 * NW;
  constexpr long long NN = (long long)NB * NB;
  constexpr int SQ = (NS + NW - 1) / NW;              // S accumulators per wave
  constexpr int XQ = (TT * TT + NW - 1) / NW;         // triangular-solve tiles per wave
  static_assert(NW >= TT - 1, "the in-place inverse gives one tile row per wave");
  typedef typename Mfma<T>::acc_t acc_t;

  __shared__ T tri[NS * Tri::TS];
  __shared__ int fail;

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
  const bool copyB = (const void*)C != (const void*)B;
  const int w = wave_id();
  const int tid = threadIdx.x;

  acc_t Sacc[SQ];
#pragma unroll
  for (int q = 0; q < SQ; ++q) Sacc[q] = Mfma<T>::zero();
  if (tid == 0) fail = 0;
  int fail_i = len;      // first failed block of the chain (thread 0)
  float piv_run = 1.0f;  // smallest pivot ratio of the chain (wave 0)

  if (hasL) {  // Z_1 = B_L^T into the first block's panel slot
    const Tin* bl = B + (long long)(start - 1) * NN;
    T* z = W + (long long)start * NN;
    for (int e = tid; e < NN; e += NT) {
      const int row = e % NB, col = e / NB;
      z[e] = (T)bl[row * NB + col];
    }
  }

  for (int i = 0; i < len; ++i) {
    const long long blk = start + i;
    const bool last = (i == len - 1);
    const bool hasC = !last || hasR;
    T* Cb = C + blk * NN;
    T* Wb = W + blk * NN;
    if (hasC && copyB) {  // B_blk into the C slot (converted), solved there in place
      const Tin* b = B + blk * NN;
      for (int e = tid; e < NN; e += NT) Cb[e] = (T)b[e];
    }
    // 1. upper tiles of A_blk - C_{blk-1}^T C_{blk-1} into the LDS triangle
    for (int s = w; s < NS; s += NW) {
      int a, b;
      upper_tile(s, a, b);
      acc_t acc = tile_load_conv<Tin, T>(A + blk * NN + 16 * b * NB + 16 * a, NB);
      if (i > 0)
        tile_mma<T, true, false, true, NB>(acc, C + (blk - 1) * NN + 16 * a * NB, NB,
                                           C + (blk - 1) * NN + 16 * b * NB, NB);
      tile_store<T>(tri + Tri::idx(a, b), LDT, acc);
    }
    global_handoff();  // triangle complete; B copy and Z visible
    float adiag[2] = {0.0f, 0.0f};
    if (p.pivmin && w == 0) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int j = lane_id() + 64 * h;
        if (j < NB) adiag[h] = (float)tri[Tri::idx(j >> 4, j >> 4) + (j & 15) * (LDT + 1)];
      }
    }
    // 2. potrf in place: diagonal tiles become Uinv_kk, row panels U_kj.  Lookahead:
    //    during the trailing update of step kk, wave 0 updates tile (kk+1, kk+1) first
    //    (task 0) and factors it at once while the other waves share the remaining
    //    tasks (LDS accesses of one wave complete in order).
    auto diag = [&](int kk) {
      const bool ok = diag_chol_inv16<T>(tri + Tri::idx(kk, kk), LDT, tri + Tri::idx(kk, kk), LDT);
      if (!ok && lane_id() == 0) fail = 1;
    };
    auto trailing = [&](int kk, int task) {
      int ii, jj;
      upper_tile(task, ii, jj);
      const int r = kk + 1 + ii, c = kk + 1 + jj;
      acc_t acc = tile_load<T>(tri + Tri::idx(r, c), LDT);
      tile_mma<T, true, false, true, 16>(acc, tri + Tri::idx(kk, r), LDT, tri + Tri::idx(kk, c), LDT);
      tile_store<T>(tri + Tri::idx(r, c), LDT, acc);
    };
    if (w == 0) diag(0);
    __syncthreads();
    for (int kk = 0; kk < TT; ++kk) {
      for (int j = kk + 1 + w; j < TT; j += NW) {
        acc_t acc = Mfma<T>::zero();
        tile_mma<T, true, false, false, 16>(acc, tri + Tri::idx(kk, kk), LDT, tri + Tri::idx(kk, j), LDT);
        tile_store<T>(tri + Tri::idx(kk, j), LDT, acc);
      }
      __syncthreads();
      const int nrow = TT - 1 - kk;
      const int nA = nrow * (nrow + 1) / 2;
      if (nrow > 0) {
        if (w == 0) {
          trailing(kk, 0);
          diag(kk + 1);
        } else {
          for (int task = w; task < nA; task += NW - 1) trailing(kk, task);
        }
        __syncthreads();
      }
    }
    // reported once after the chain (see chain_factor.hpp)
    if (tid == 0 && fail) {
      fail_i = min(fail_i, i);
      fail = 0;
    }
    if (p.pivmin && w == 0) {
      float ratio = 1.0f;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int j = lane_id() + 64 * h;
        if (j < NB) {
          const float ui = (float)tri[Tri::idx(j >> 4, j >> 4) + (j & 15) * (LDT + 1)];
          const float q = 1.0f / (ui * ui * adiag[h]);
          ratio = fminf(ratio, (q >= 0.0f && q <= 3.0e38f) ? q : 0.0f);
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) ratio = fminf(ratio, __shfl_xor(ratio, o));
      piv_run = fminf(piv_run, ratio);
    }
    // 3. trtri in place, column by column: Uinv(0:j, j) = -Uinv(0:j, 0:j) U(0:j, j) Uinv_jj
    for (int j = 1; j < TT; ++j) {
      for (int t = w; t < j; t += NW) {  // Y_t = U(t, j) Uinv_jj
        acc_t acc = Mfma<T>::zero();
        tile_mma<T, false, false, false, 16>(acc, tri + Tri::idx(t, j), LDT, tri + Tri::idx(j, j), LDT);
        tile_store<T>(tri + Tri::idx(t, j), LDT, acc);
      }
      __syncthreads();
      acc_t acc = Mfma<T>::zero();
      const int r = w;  // Uinv(r, j) = -sum_{t=r}^{j-1} Uinv(r, t) Y_t
      if (r < j) {
        for (int t = r; t < j; ++t)
          tile_mma<T, false, false, true, 16>(acc, tri + Tri::idx(r, t), LDT, tri + Tri::idx(t, j), LDT);
      }
      __syncthreads();
      if (r < j) tile_store<T>(tri + Tri::idx(r, j), LDT, acc);
      __syncthreads();
    }
    // 4. [C | W] = Uinv^T [B | Z] in place in the HBM slots, C columns then W columns;
    //    Uinv tiles to U
    for (int half = 0; half < 2; ++half) {
      if (half == 0 ? !hasC : !hasL) continue;
      T* base = half == 0 ? Cb : Wb;
      acc_t X[XQ];
#pragma unroll
      for (int q = 0; q < XQ; ++q) {
        // output tile (r, c) costs r + 1 products: rotating r with c spreads the work
        const int task = w + NW * q;
        const int c = task / TT, r = (task % TT + c) % TT;
        X[q] = Mfma<T>::zero();
        if (task < TT * TT) {
          for (int t = 0; t <= r; ++t)
            tile_mma<T, true, false, false, 16>(X[q], tri + Tri::idx(t, r), LDT,
                                                base + 16 * c * NB + 16 * t, NB);
        }
      }
      __syncthreads();  // every read of these right-hand-side columns is done
#pragma unroll
      for (int q = 0; q < XQ; ++q) {
        const int task = w + NW * q;
        const int c = task / TT, r = (task % TT + c) % TT;
        if (task < TT * TT) tile_store<T>(base + 16 * c * NB + 16 * r, NB, X[q]);
      }
    }
    for (int s = w; s < NS; s += NW) {
      int a, b;
      upper_tile(s, a, b);
      tile_store<T>(U + blk * NN + 16 * b * NB + 16 * a, NB, tile_load<T>(tri + Tri::idx(a, b), LDT));
    }
    global_handoff();  // C and W visible to every wave
    // 5. Schur products
    if (hasL) {
#pragma unroll
      for (int q = 0; q < SQ; ++q) {
        const int s = w + NW * q;
        if (s < NS) {
          int a, b;
          upper_tile(s, a, b);
          tile_mma<T, true, false, false, NB>(Sacc[q], Wb + 16 * a * NB, NB, Wb + 16 * b * NB, NB);
        }
      }
    }
    if (!last) {
      if (hasL) {  // Z_{i+1} = -C^T W into the next block's panel slot
        T* z = W + (blk + 1) * NN;
        for (int task = w; task < TT * TT; task += NW) {
          const int a = task % TT, b = task / TT;
          acc_t acc = Mfma<T>::zero();
          tile_mma<T, true, false, true, NB>(acc, Cb + 16 * a * NB, NB, Wb + 16 * b * NB, NB);
          tile_store<T>(z + 16 * b * NB + 16 * a, NB, acc);
        }
      }
    } else if (hasR) {
      T* sr = (T*)S.SR + (long long)sepR * NN;
      for (int s = w; s < NS; s += NW) {
        int a, b;
        upper_tile(s, a, b);
        acc_t acc = Mfma<T>::zero();
        tile_mma<T, true, false, false, NB>(acc, Cb + 16 * a * NB, NB, Cb + 16 * b * NB, NB);
        tile_store<T>(sr + 16 * b * NB + 16 * a, NB, acc);
      }
      if (hasL) {
        T* bn = (T*)S.Bn + (long long)sepL * NN;
        for (int task = w; task < TT * TT; task += NW) {
          const int a = task % TT, b = task / TT;
          acc_t acc = Mfma<T>::zero();
          tile_mma<T, true, false, true, NB>(acc, Wb + 16 * a * NB, NB, Cb + 16 * b * NB, NB);
          tile_store<T>(bn + 16 * b * NB + 16 * a, NB, acc);
        }
      }
    }
    global_handoff();  // Z_{i+1} and C_blk visible before the next block reads them
  }
  if (hasL) {
    T* sl = (T*)S.SL + (long long)sepL * NN;
#pragma unroll
    for (int q = 0; q < SQ; ++q) {
      const int s = w + NW * q;
      if (s < NS) {
        int a, b;
        upper_tile(s, a, b);
        tile_store<T>(sl + 16 * b * NB + 16 * a, NB, Sacc[q]);
      }
    }
  }
  if (tid == 0 && fail_i < len) {
    const long long blk = start + fail_i;
    atomicMin(p.info + sys, (S.orig ? S.orig[blk] : (int)blk) + 1);
  }
  if (p.pivmin && tid == 0) atomicMin(p.pivmin + sys, __float_as_uint(piv_run));
}

}  // namespace bss
