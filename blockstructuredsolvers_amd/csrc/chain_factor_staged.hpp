// chain_factor_staged.hpp -- chain factorisation for block sizes n = 96 and 128.
//
// Same algorithm and outputs as chain_factor_kernel (chain_factor.hpp): per chain
//   Z_s = B_L^T,  for each block i:  A_i -= C_{i-1}^T C_{i-1};  Uinv_i = chol(A_i)^{-1};
//   [C_i | W_i] = Uinv_i^T [B_i | Z_i];  Z_{i+1} = -C_i^T W_i;  S_L += W_i^T W_i
// but at these sizes four n x n blocks no longer fit in the 160 KB of LDS (one FP64
// block at n = 128 is 128 KB).  Only the upper triangle of the block being factored
// stays in LDS, as packed 16x16 tiles; it is factored in place (potrf, diagonal tiles
// kept inverted) and then inverted in place (trtri), giving Uinv.  B, Z and the
// products C, W live in their HBM output slots and reach the MFMAs through two LDS
// staging buffers:
//   * products over the depth of their operands (A - C^T C and the Schur products)
//     stage one 16-row slice of every operand per step;
//   * the triangular solve stages two 16-column panels of B (or Z) per step.
// Every staged element is read from memory once per product and shared by the eight
// waves, where multiplying straight from global memory re-reads each panel once per
// output tile (several MB per block: more than L2 holds for the chains in flight).
// The next step's slice is loaded into registers while the current one is multiplied
// (one barrier per step).  The slices are accumulated in the order of one product over
// the whole depth (tile_mma with K = n), 16 rows at a time.
#pragma once

#include "chain_factor.hpp"

namespace bss {

template <typename T, int NB>
struct StagedTri {
  static constexpr int TT = NB / 16;
  static constexpr int LDT = 16 + LdsPad<T>::value;  // 16-byte aligned columns
  static constexpr int TS = 16 * LDT;                 // elements per stored tile
  static constexpr int NS = TT * (TT + 1) / 2;        // upper tiles
  // upper tile (r, c), r <= c, packed column by column
  __device__ static constexpr int idx(int r, int c) { return (c * (c + 1) / 2 + r) * TS; }
};

// (r, c) of the s-th upper tile in packed column order
__device__ inline void upper_rc(int s, int& r, int& c) {
  c = 0;
  while (s > c) {
    s -= c + 1;
    ++c;
  }
  r = s;
}

// Global writes of this workgroup visible to all of its waves: the barrier's
// workgroup-scope release/acquire orders the stores, and the workgroup's waves share one
// CU's vector L1 (a device-scope fence would write back the whole L2).
__device__ inline void global_handoff() { __syncthreads(); }

// Four consecutive elements of one column ("quad"), 16-byte aligned in global memory,
// held in registers on their way to an LDS tile.
template <typename Ts>
__device__ inline void quad_load(Ts (&v)[4], const BSS_GLOBAL Ts* p) {
  if constexpr (sizeof(Ts) == 8) {
    typedef double v2_t __attribute__((ext_vector_type(2)));
    const v2_t a = *(const BSS_GLOBAL v2_t*)p;
    const v2_t b = *(const BSS_GLOBAL v2_t*)(p + 2);
    v[0] = a[0];
    v[1] = a[1];
    v[2] = b[0];
    v[3] = b[1];
  } else {
    typedef float v4_t __attribute__((ext_vector_type(4)));
    const v4_t a = *(const BSS_GLOBAL v4_t*)p;
    v[0] = a[0];
    v[1] = a[1];
    v[2] = a[2];
    v[3] = a[3];
  }
}

// The quad, converted to T, into four consecutive LDS elements (16-byte aligned)
template <typename T, typename Ts>
__device__ inline void quad_put(T* dst, const Ts (&v)[4]) {
  if constexpr (sizeof(T) == 8) {
    typedef double v2_t __attribute__((ext_vector_type(2)));
    *(v2_t*)dst = v2_t{(double)v[0], (double)v[1]};
    *(v2_t*)(dst + 2) = v2_t{(double)v[2], (double)v[3]};
  } else {
    typedef float v4_t __attribute__((ext_vector_type(4)));
    *(v4_t*)dst = v4_t{(float)v[0], (float)v[1], (float)v[2], (float)v[3]};
  }
}

// D-layout tile of a column-major global buffer, converted to T / stored from T
template <typename T, typename Tin>
__device__ inline typename Mfma<T>::acc_t gtile_load(const BSS_GLOBAL Tin* p, int ld) {
  const int l = lane_id();
  const int lr = l & 15, lk = l >> 4;
  typename Mfma<T>::acc_t a;
#pragma unroll
  for (int v = 0; v < 4; ++v) a[v] = (T)p[lr * ld + Mfma<T>::row(lk, v)];
  return a;
}
template <typename T>
__device__ inline void gtile_store(BSS_GLOBAL T* p, int ld, const typename Mfma<T>::acc_t& a) {
  const int l = lane_id();
  const int lr = l & 15, lk = l >> 4;
#pragma unroll
  for (int v = 0; v < 4; ++v) p[lr * ld + Mfma<T>::row(lk, v)] = a[v];
}

// Runs `steps` steps of a staged product.  Step s's operands are loaded into registers
// (load(s, regs)) while step s - 1 is multiplied, written to LDS buffer s & 1
// (put(regs, buf)) and multiplied after one barrier (mul(s, buf)).  When a wave passes
// the barrier of step s, every wave has finished mul(s - 1), so buffer (s + 1) & 1 is
// free for the next put.  The closing barrier leaves both buffers free.
// With one buffer (NBUF = 1) a second barrier per step waits for every mul(s - 1)
// before the buffer is overwritten.
template <typename Regs, int NBUF, typename Load, typename Put, typename Mul>
__device__ __attribute__((always_inline)) inline void staged_steps(int steps, Load&& load,
                                                                   Put&& put, Mul&& mul) {
  Regs r;
  if (steps > 0) load(0, r);
  auto step = [&](int s) {
    const int buf = NBUF == 1 ? 0 : (s & 1);
    if (NBUF == 1 && s > 0) __syncthreads();
    put(r, buf);
    __syncthreads();
    if (s + 1 < steps) load(s + 1, r);
    mul(s, buf);
  };
  if constexpr (NBUF == 1) {
    // the single-buffered variants run at 128 registers a wave: an unrolled step loop
    // spills inside the steps (28.9 against 31.2 ms for the level-0 FP32 factorisation
    // of 16 systems at n = 128)
#pragma unroll 1
    for (int s = 0; s < steps; ++s) step(s);
  } else {
    for (int s = 0; s < steps; ++s) step(s);
  }
  __syncthreads();
}

// Staging buffers and workgroups per CU.  FP32 factors single-buffer (66 KB of LDS at
// n = 128) so that two workgroups share a CU and one's potrf and trtri overlap the
// other's products, within 128 registers a wave (level-0 factorisation of 16 systems
// at n = 128: 28.9 ms against 37.2 ms with one double-buffered workgroup per CU).  FP64
// factors double-buffer with one workgroup per CU: at n = 96 two workgroups would fit
// in LDS, but the 128-register cap spills (41.1 ms against 34.0 ms).
template <typename T, int NB>
struct StagedCfg {
  static constexpr int NBUF = sizeof(T) == 4 ? 1 : 2;
  static constexpr int WG_PER_CU = NBUF == 2 ? 1 : 2;
};

// Quads per thread of one staged operand: a 16-row slice (4 quads per column) of an
// n-column block, or a pair of 16-column panels (n / 4 quads per column)
template <int NB, int NT>
struct StageShape {
  static constexpr int SLICE_Q = 4 * NB;
  static constexpr int PANEL_Q = 8 * NB;
  static constexpr int SLICE_QPT = (SLICE_Q + NT - 1) / NT;
  static constexpr int PANEL_QPT = (PANEL_Q + NT - 1) / NT;
};

template <typename Ts, int Q>
struct QuadRegs {
  Ts v[Q][4];
};

template <typename Tin, typename T, int NB, int NW>
__global__ void __launch_bounds__(64 * NW, (StagedCfg<T, NB>::WG_PER_CU * NW / 4))
    chain_factor_staged_kernel(FactorLevelArgs p) {
  typedef StagedTri<T, NB> Tri;
  constexpr int TT = Tri::TT, NS = Tri::NS, LDT = Tri::LDT, TS = Tri::TS;
  constexpr int NT = 64 * NW;
  constexpr long long NN = (long long)NB * NB;
  constexpr int SQ = (NS + NW - 1) / NW;            // upper-tile tasks per wave
  constexpr int ZQ = (TT * TT + NW - 1) / NW;       // full-tile tasks per wave
  constexpr int XQ = (2 * TT + NW - 1) / NW;        // solve tasks per wave and panel pair
  constexpr int OP = TT * TS;                       // one staged operand (TT tiles)
  typedef StageShape<NB, NT> Sh;
  static_assert(NW >= TT - 1, "the in-place inverse gives one tile row per wave");
  static_assert(TT % 2 == 0, "the triangular solve stages panels in pairs");
  typedef typename Mfma<T>::acc_t acc_t;

  __shared__ __attribute__((aligned(16))) T tri[NS * TS];
  constexpr int NBUF = StagedCfg<T, NB>::NBUF;
  __shared__ __attribute__((aligned(16))) T stg[NBUF * 2 * OP];  // buffers x 2 operands
  __shared__ int fail;

  const ChainRec cd = p.chains[blockIdx.x];
  const int start = cd.start, len = cd.len, sepL = cd.sepL, sepR = cd.sepR;
  if (len <= 0) return;
  const LevelSys S = p.sys[cd.slot];
  const int sys = S.sys;
  const bool hasL = sepL >= 0, hasR = sepR >= 0;
  const BSS_GLOBAL Tin* A = gptr((const Tin*)S.A);
  const BSS_GLOBAL Tin* B = gptr((const Tin*)S.B);
  BSS_GLOBAL T* U = gptr((T*)S.U);
  BSS_GLOBAL T* C = gptr((T*)S.C);
  BSS_GLOBAL T* W = gptr((T*)S.W);
  const int w = wave_id();
  const int tid = threadIdx.x;
  // thread index of the staging quads, re-derived for every block (opaque_v) so that
  // the per-thread staging addresses are computed where they are used instead of being
  // hoisted out of the block loop and held in registers for the whole chain
  int tq = tid;

  // k-slice staging: quad qi = (column qi / 4, rows 4 (qi % 4) ..) of rows 16 kt .. of
  // an n x n column-major block, into tile (column / 16) of an operand area
  auto slice_load = [&](auto& v, const auto* M, int kt) {
#pragma unroll
    for (int j = 0; j < Sh::SLICE_QPT; ++j) {
      const int qi = tq + NT * j;
      if (Sh::SLICE_Q % NT == 0 || qi < Sh::SLICE_Q)
        quad_load(v[j], M + (qi >> 2) * NB + 16 * kt + 4 * (qi & 3));
    }
  };
  auto slice_put = [&](T* area, const auto& v) {
#pragma unroll
    for (int j = 0; j < Sh::SLICE_QPT; ++j) {
      const int qi = tq + NT * j;
      const int c = qi >> 2;
      if (Sh::SLICE_Q % NT == 0 || qi < Sh::SLICE_Q)
        quad_put(area + (c >> 4) * TS + (c & 15) * LDT + 4 * (qi & 3), v[j]);
    }
  };
  // panel-pair staging: quad qi = (column c0 + qi / (n/4), rows 4 (qi % (n/4)) ..), into
  // tile (row / 16) of panel (column - c0) / 16: panel pp's tile t at area + (pp TT + t) TS
  auto panel_load = [&](auto& v, const auto* M, int c0) {
#pragma unroll
    for (int j = 0; j < Sh::PANEL_QPT; ++j) {
      const int qi = tq + NT * j;
      if (Sh::PANEL_Q % NT == 0 || qi < Sh::PANEL_Q)
        quad_load(v[j], M + (c0 + qi / (NB / 4)) * NB + 4 * (qi % (NB / 4)));
    }
  };
  auto panel_put = [&](T* area, const auto& v) {
#pragma unroll
    for (int j = 0; j < Sh::PANEL_QPT; ++j) {
      const int qi = tq + NT * j;
      const int cc = qi / (NB / 4), m = qi % (NB / 4);
      if (Sh::PANEL_Q % NT == 0 || qi < Sh::PANEL_Q)
        quad_put(area + ((cc >> 4) * TT + (m >> 2)) * TS + (cc & 15) * LDT + 4 * (m & 3), v[j]);
    }
  };

  acc_t Sacc[SQ];
#pragma unroll
  for (int q = 0; q < SQ; ++q) Sacc[q] = Mfma<T>::zero();
  if (tid == 0) fail = 0;
  int fail_i = len;      // first failed block of the chain (thread 0)
  float piv_run = 1.0f;  // smallest pivot ratio of the chain (wave 0)

  if (hasL) {  // Z_1 = B_L^T into the first block's panel slot
    const BSS_GLOBAL Tin* bl = B + (long long)(start - 1) * NN;
    BSS_GLOBAL T* z = W + (long long)start * NN;
    for (int e = tid; e < NN; e += NT) {
      const int row = e % NB, col = e / NB;
      z[e] = (T)bl[row * NB + col];
    }
  }

  for (int i = 0; i < len; ++i) {
    tq = opaque_v(tid);
    const long long blk = start + i;
    const bool last = (i == len - 1);
    const bool hasC = !last || hasR;
    BSS_GLOBAL T* Cb = C + blk * NN;
    BSS_GLOBAL T* Wb = W + blk * NN;

    // 1. upper tiles of A_blk - C_{blk-1}^T C_{blk-1} into the LDS triangle
    {
      acc_t acc[SQ];
#pragma unroll
      for (int q = 0; q < SQ; ++q) {
        const int s = w + NW * q;
        if (s < NS) {
          int a, b;
          upper_rc(s, a, b);
          acc[q] = gtile_load<T>(A + blk * NN + 16 * b * NB + 16 * a, NB);
        }
      }
      if (i > 0) {
        const BSS_GLOBAL T* Cp = C + (blk - 1) * NN;
        staged_steps<QuadRegs<T, Sh::SLICE_QPT>, NBUF>(
            TT, [&](int kt, auto& r) { slice_load(r.v, Cp, kt); },
            [&](const auto& r, int buf) { slice_put(stg + buf * 2 * OP, r.v); },
            [&](int kt, int buf) {
              const T* X = stg + buf * 2 * OP;
#pragma unroll
              for (int q = 0; q < SQ; ++q) {
                const int s = w + NW * q;
                if (s < NS) {
                  int a, b;
                  upper_rc(s, a, b);
                  tile_mma<T, true, false, true, 16>(acc[q], X + a * TS, LDT, X + b * TS, LDT);
                }
              }
            });
      }
#pragma unroll
      for (int q = 0; q < SQ; ++q) {
        const int s = w + NW * q;
        if (s < NS) {
          int a, b;
          upper_rc(s, a, b);
          tile_store<T>(tri + Tri::idx(a, b), LDT, acc[q]);
        }
      }
    }
    __syncthreads();  // triangle complete
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
      upper_rc(task, ii, jj);
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
    // Uinv tiles to U
    for (int s = w; s < NS; s += NW) {
      int a, b;
      upper_rc(s, a, b);
      gtile_store<T>(U + blk * NN + 16 * b * NB + 16 * a, NB, tile_load<T>(tri + Tri::idx(a, b), LDT));
    }
    // 4. [C | W] = Uinv^T [B | Z], two 16-column panels per step: output tile (r, c)
    //    = sum_{t <= r} Uinv(t, r)^T B(t, c).  Task k of a step is panel k / TT, row
    //    k % TT (the second panel's rows reversed, so that each wave's two tasks cost
    //    about the same).  C goes to its slot; W overwrites Z in place (a panel is staged
    //    before its outputs are stored, and the next pair is other columns).
    auto solve_mul = [&](BSS_GLOBAL T* out, int pair, int buf) {
      const T* X = stg + buf * 2 * OP;
#pragma unroll
      for (int q = 0; q < XQ; ++q) {
        const int k = w + NW * q;
        if (k < 2 * TT) {
          const int pp = k / TT, rr = k % TT;
          const int r = pp == 0 ? rr : TT - 1 - rr;
          acc_t acc = Mfma<T>::zero();
          for (int t = 0; t <= r; ++t)
            tile_mma<T, true, false, false, 16>(acc, tri + Tri::idx(t, r), LDT,
                                                X + (pp * TT + t) * TS, LDT);
          gtile_store<T>(out + 16 * (2 * pair + pp) * NB + 16 * r, NB, acc);
        }
      }
    };
    if (hasC) {
      const BSS_GLOBAL Tin* Bb = B + blk * NN;
      staged_steps<QuadRegs<Tin, Sh::PANEL_QPT>, NBUF>(
          TT / 2, [&](int pair, auto& r) { panel_load(r.v, Bb, 32 * pair); },
          [&](const auto& r, int buf) { panel_put(stg + buf * 2 * OP, r.v); },
          [&](int pair, int buf) { solve_mul(Cb, pair, buf); });
    }
    if (hasL) {
      staged_steps<QuadRegs<T, Sh::PANEL_QPT>, NBUF>(
          TT / 2, [&](int pair, auto& r) { panel_load(r.v, Wb, 32 * pair); },
          [&](const auto& r, int buf) { panel_put(stg + buf * 2 * OP, r.v); },
          [&](int pair, int buf) { solve_mul(Wb, pair, buf); });
    }
    global_handoff();  // C and W visible to every wave
    // 5. Schur products over 16-row slices of C (operand area 0) and W (area 1)
    if (!last && hasL) {  // S_L += W^T W;  Z_{i+1} = -C^T W into the next panel slot
      acc_t Z[ZQ];
#pragma unroll
      for (int q = 0; q < ZQ; ++q) Z[q] = Mfma<T>::zero();
      staged_steps<QuadRegs<T, 2 * Sh::SLICE_QPT>, NBUF>(
          TT,
          [&](int kt, auto& r) {
            slice_load(*(T(*)[Sh::SLICE_QPT][4])r.v, Cb, kt);
            slice_load(*(T(*)[Sh::SLICE_QPT][4])(r.v + Sh::SLICE_QPT), Wb, kt);
          },
          [&](const auto& r, int buf) {
            slice_put(stg + buf * 2 * OP, *(const T(*)[Sh::SLICE_QPT][4])r.v);
            slice_put(stg + buf * 2 * OP + OP, *(const T(*)[Sh::SLICE_QPT][4])(r.v + Sh::SLICE_QPT));
          },
          [&](int kt, int buf) {
            const T* X = stg + buf * 2 * OP;
            const T* Y = X + OP;
#pragma unroll
            for (int q = 0; q < SQ; ++q) {
              const int s = w + NW * q;
              if (s < NS) {
                int a, b;
                upper_rc(s, a, b);
                tile_mma<T, true, false, false, 16>(Sacc[q], Y + a * TS, LDT, Y + b * TS, LDT);
              }
            }
#pragma unroll
            for (int q = 0; q < ZQ; ++q) {
              const int task = w + NW * q;
              if (task < TT * TT) {
                const int a = task % TT, b = task / TT;
                tile_mma<T, true, false, true, 16>(Z[q], X + a * TS, LDT, Y + b * TS, LDT);
              }
            }
          });
      BSS_GLOBAL T* z = W + (blk + 1) * NN;
#pragma unroll
      for (int q = 0; q < ZQ; ++q) {
        const int task = w + NW * q;
        if (task < TT * TT) {
          const int a = task % TT, b = task / TT;
          gtile_store<T>(z + 16 * b * NB + 16 * a, NB, Z[q]);
        }
      }
    } else if (last && (hasL || hasR)) {
      // S_L += W^T W (hasL); S_R = C^T C and B' = -W^T C (hasR)
      acc_t R[SQ], Bp[ZQ];
#pragma unroll
      for (int q = 0; q < SQ; ++q) R[q] = Mfma<T>::zero();
#pragma unroll
      for (int q = 0; q < ZQ; ++q) Bp[q] = Mfma<T>::zero();
      staged_steps<QuadRegs<T, 2 * Sh::SLICE_QPT>, NBUF>(
          TT,
          [&](int kt, auto& r) {
            if (hasR) slice_load(*(T(*)[Sh::SLICE_QPT][4])r.v, Cb, kt);
            if (hasL) slice_load(*(T(*)[Sh::SLICE_QPT][4])(r.v + Sh::SLICE_QPT), Wb, kt);
          },
          [&](const auto& r, int buf) {
            if (hasR) slice_put(stg + buf * 2 * OP, *(const T(*)[Sh::SLICE_QPT][4])r.v);
            if (hasL)
              slice_put(stg + buf * 2 * OP + OP, *(const T(*)[Sh::SLICE_QPT][4])(r.v + Sh::SLICE_QPT));
          },
          [&](int kt, int buf) {
            const T* X = stg + buf * 2 * OP;
            const T* Y = X + OP;
#pragma unroll
            for (int q = 0; q < SQ; ++q) {
              const int s = w + NW * q;
              if (s < NS) {
                int a, b;
                upper_rc(s, a, b);
                if (hasL)
                  tile_mma<T, true, false, false, 16>(Sacc[q], Y + a * TS, LDT, Y + b * TS, LDT);
                if (hasR)
                  tile_mma<T, true, false, false, 16>(R[q], X + a * TS, LDT, X + b * TS, LDT);
              }
            }
            if (hasR && hasL) {
#pragma unroll
              for (int q = 0; q < ZQ; ++q) {
                const int task = w + NW * q;
                if (task < TT * TT) {
                  const int a = task % TT, b = task / TT;
                  tile_mma<T, true, false, true, 16>(Bp[q], Y + a * TS, LDT, X + b * TS, LDT);
                }
              }
            }
          });
      if (hasR) {
        BSS_GLOBAL T* sr = gptr((T*)S.SR) + (long long)sepR * NN;
#pragma unroll
        for (int q = 0; q < SQ; ++q) {
          const int s = w + NW * q;
          if (s < NS) {
            int a, b;
            upper_rc(s, a, b);
            gtile_store<T>(sr + 16 * b * NB + 16 * a, NB, R[q]);
          }
        }
        if (hasL) {
          BSS_GLOBAL T* bn = gptr((T*)S.Bn) + (long long)sepL * NN;
#pragma unroll
          for (int q = 0; q < ZQ; ++q) {
            const int task = w + NW * q;
            if (task < TT * TT) {
              const int a = task % TT, b = task / TT;
              gtile_store<T>(bn + 16 * b * NB + 16 * a, NB, Bp[q]);
            }
          }
        }
      }
    }
    global_handoff();  // Z_{i+1} and C_blk visible before the next block reads them
  }
  if (hasL) {
    BSS_GLOBAL T* sl = gptr((T*)S.SL) + (long long)sepL * NN;
#pragma unroll
    for (int q = 0; q < SQ; ++q) {
      const int s = w + NW * q;
      if (s < NS) {
        int a, b;
        upper_rc(s, a, b);
        gtile_store<T>(sl + 16 * b * NB + 16 * a, NB, Sacc[q]);
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
