// mixed.hip -- FP64 residual / norm kernels and the refinement loop (see mixed.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>

#include "common.hpp"
#include "mixed.hpp"
#include "pinned_pool.hpp"

namespace bss {
namespace {

thread_local std::string g_mixed_err;
// per workgroup partials: [0,16) max|r_j|, [16,32) max|x_j|, [32,48) max|d_j|,
// [48] max row sum of |M| (when requested)
constexpr int kParts = 64;
constexpr int kMnormSlot = 48;

size_t al(size_t b) { return (b + 255) & ~size_t(255); }

// NaN-propagating max (fmax would drop a NaN residual and report convergence)
__device__ inline double nmax(double a, double b) { return (b > a || b != b) ? b : a; }

// Residual pass over one work item (a block range [b0, b1) of one system):
//   x_new = x_old + dx (FP64; written for the workgroup's own blocks),
//   r = d - (B_{i-1}^T x_{i-1} + A_i x_i + B_i x_{i+1}) in FP64 on MFMA, stored as FP32
//   for the next correction solve, plus per-column max |r|, |x|, |d|.
// With `mnorm`, the row sums of |M| are accumulated from the operand values the
// products already hold, so ||M||_inf costs no extra pass over A and B.
// One wave per 16-row tile; a rolling three-block window of x_new lives in LDS and
// each block's A/B operands are issued before the window update to hide HBM latency.
// Every coupling block is read from HBM once: at block row i, B_i serves both
// B_i x_{i+1} (row i) and B_i^T x_i (row i+1, carried to the next iteration in
// registers); the second read of B_i follows the first immediately and hits L2.
// V1 (one right-hand side): the products are FP64 FMAs on the same operand registers
// instead of MFMA.  A 16x16x4 FP64 MFMA issues every 64 cycles on this GPU, the same
// FLOP rate as the vector FMAs, but with one column it would compute 16 columns to keep
// one; each lane then sums its 16 k-slots and the four lanes of a row combine.
// TS (V1 up to n = 64, no ||M|| pass): B_i^T x_i is not read again from L2 as a
// transposed operand; the products B_i[r][k] x_i[r] of the registers that serve B_i
// x_{i+1} go to an LDS tile, and the next row sums its columns.
// RQ (V1 from n = 96 on, where that tile would not fit in LDS, and at n = 64 with the
// ||M|| pass): the same products are summed over the 16 rows of each wave in registers
// (red16) and only the per-wave sums go to LDS, n values per wave; the next row adds the
// waves' sums.  The ||M|| pass sums |B_i| the same way.  The transposed re-read of B_i
// it replaces ran this kernel at 3.7 TB/s at n = 128 (its strided 32-byte reads),
// against 5.6 without it; at n = 64 the ||M|| residual went from 3.56 to 3.42 ms for
// 32 systems, while without ||M|| the LDS tile stays as fast (3.29 against 3.32 ms).
template <int KC>
struct Red16 {
  // halving steps (masks 8, 4, ..) while the slot count stays even; the rest exchange
  // whole values
  static constexpr int halvings() {
    int v = KC, h = 0;
    while (h < 4 && v % 2 == 0) {
      v /= 2;
      ++h;
    }
    return h;
  }
  static constexpr int H = halvings();
  static constexpr int CNT = KC >> H;  // sums each lane holds at the end
};

// Sums KC values of every lane over the 16 lanes of its row group (lanes l ^ 1, 2, 4,
// 8).  Halving step with mask m: the lane whose mask bit is clear keeps the first half of
// its slots, the other the second half, and each adds the partner's copy of the half it
// keeps; then whole exchanges for the masks left.  On return v[0 .. CNT) of lane lr are
// the sums of local slots red16_first(lr) .. + CNT (the same in lanes that differ only in
// the exchanged mask bits).
template <int KC>
__device__ inline void red16(double (&v)[KC], int lr) {
  constexpr int H = Red16<KC>::H;
  int V = KC;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const int m = 8 >> i;
    const bool up = (lr & m) != 0;
    const int half = V / 2;
#pragma unroll
    for (int j = 0; j < KC / 2; ++j) {
      if (j < half) {
        const double send = up ? v[j] : v[half + j];
        const double recv = __shfl_xor(send, m);
        v[j] = (up ? v[half + j] : v[j]) + recv;
      }
    }
    V = half;
  }
#pragma unroll
  for (int i = H; i < 4; ++i) {
    const int m = 8 >> i;
#pragma unroll
    for (int j = 0; j < Red16<KC>::CNT; ++j) v[j] += __shfl_xor(v[j], m);
  }
}
template <int KC>
__device__ inline int red16_first(int lr) {
  int f = 0, V = KC;
#pragma unroll
  for (int i = 0; i < Red16<KC>::H; ++i) {
    V /= 2;
    if (lr & (8 >> i)) f += V;
  }
  return f;
}

// MN: the ||M|| pass may run (taken from mnorm_arg for several columns; one-column
// kernels are launched with MN equal to it, which keeps branches out of their k loop).
template <int NB, bool V1, bool TS, bool MN>
__global__ void __launch_bounds__(4 * NB)
    residual_kernel(const ResItem* __restrict__ items, const ResSys* __restrict__ systems,
                    double* __restrict__ part, int nrhs, int xo_k, int dx_k, int xn_k, int ro_k,
                    int mnorm_arg) {
  static_assert(!TS || (V1 && NB <= 64 && !MN), "transpose-sum path: one column, n <= 64");
  constexpr int TT = NB / 16, NT = 64 * TT, LDV = NB + 2, KQ = NB / 4;
  constexpr long long NN = (long long)NB * NB;
  typedef Mfma<double>::acc_t acc_t;
  __shared__ double X[3][16 * LDV];
  __shared__ double red[4][4 * TT][16];
  constexpr int LDP = NB + 1;
  __shared__ double P[TS ? NB * LDP : 1];  // TS: B_i[r][k] x_i[r] at P[r * LDP + k]
  constexpr bool RQ = V1 && !TS && NB >= 64;
  // RQ: per-wave sums over rows r of B_i[r][k] x_i[r] (Q[0]) and |B_i[r][k]| (Q[1])
  __shared__ double Q[RQ ? 2 : 1][RQ ? TT * NB : 1];
  const int mnorm = !MN ? 0 : V1 ? 1 : mnorm_arg;
  const ResItem it = items[blockIdx.x];
  const ResSys& S = systems[it.sys];
  const int N = S.N;
  const int tid = threadIdx.x;
  const int w = wave_id(), l = lane_id(), lr = l & 15, lk = l >> 4;
  const long long ND = (long long)NB * nrhs;
  const auto As = gptr((const double*)S.A);
  const auto Bs = gptr((const double*)S.B);
  const auto ds = gptr((const double*)S.d);
  const auto xo = gptr((const double*)(xo_k >= 0 ? S.x[xo_k] : nullptr));
  const auto dxs = gptr((const float*)(dx_k >= 0 ? S.v[dx_k] : nullptr));
  const auto xn = gptr(S.x[xn_k]);
  const auto ro = gptr(S.v[ro_k]);
  for (int e = tid; e < 3 * 16 * LDV; e += NT) (&X[0][0])[e] = 0.0;
  double rmax = 0.0, xmax = 0.0, dmax = 0.0, mmax = 0.0;
  const int b0 = it.b0, b1 = it.b1;

  // x_new of block j (zero outside the system), elements e = tid + u*NT: loaded into
  // registers by stage_load, then written to its window slot (and, for the
  // workgroup's own blocks, to xnew) by stage_store.  The wait counters are in order,
  // so the x loads are issued before the block's operand loads, and the wait for them
  // leaves the operands in flight; every lane loads (clamped indices), as a predicated
  // load would make every later wait conservative.  x_old and dx are always given
  // (the first pass of a solve is init_pass_kernel).
  constexpr int SE = (NB * 16 + NT - 1) / NT;  // up to 16 right-hand sides
  auto stage_load = [&](int j, double (&v)[SE]) {
    const bool inside = j >= 0 && j < N;
    const long long jc = j < 0 ? 0 : (j >= N ? N - 1 : j);
#pragma unroll
    for (int u = 0; u < SE; ++u) {
      const int e = tid + u * NT;
      const long long o = jc * ND + (e < (int)ND ? e : 0);
      const double x = xo[o] + (double)dxs[o];
      v[u] = inside ? x : 0.0;
    }
  };
  auto stage_store = [&](int j, const double (&v)[SE]) {
    double* slot = X[((j % 3) + 3) % 3];
    const bool own = j >= b0 && j < b1;
#pragma unroll
    for (int u = 0; u < SE; ++u) {
      const int e = tid + u * NT;
      if (e < (int)ND) {
        slot[(e / NB) * LDV + (e % NB)] = v[u];
        if (own) xn[(long long)j * ND + e] = v[u];
      }
    }
  };
  auto stage = [&](int j) {
    double v[SE];
    stage_load(j, v);
    stage_store(j, v);
  };
  // k index of operand slot q: 4 consecutive k per lane (contiguous loads along k)
  auto kof = [&](int q) { return (q >> 2) * 16 + 4 * lk + (q & 3); };

  // sum over the four lanes (lk) that share row 16w + lr
  auto row_sum = [&](double v) { return row4_sum(v); };
  // carry: B_{i-1}^T x_{i-1} for the current row (and its |.| row sums for ||M||_inf);
  // V1 keeps it as one value per lane (row 16w + lr)
  acc_t carry = Mfma<double>::zero();
  [[maybe_unused]] double carry1 = 0.0;
  double rs_carry = 0.0;
  __syncthreads();  // window zeroed
  {
    double aBt[KQ];
    const bool has = b0 > 0;
    if (has) {
#pragma unroll
      for (int q = 0; q < KQ; ++q)
        aBt[q] = Bs[(b0 - 1) * NN + (long long)(16 * w + lr) * NB + kof(q)];
    }
    stage(b0 - 1);
    stage(b0);
    __syncthreads();
    if (has) {
      const double* Xm = X[(((b0 - 1) % 3) + 3) % 3];
      if constexpr (V1) {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < KQ; ++q) t = fma(aBt[q], Xm[kof(q)], t);
        carry1 = row_sum(t);
      }
#pragma unroll
      for (int q = 0; q < KQ; ++q) {
        if constexpr (!V1) carry = Mfma<double>::mma(aBt[q], Xm[lr * LDV + kof(q)], carry);
        if (mnorm) rs_carry += fabs(aBt[q]);
      }
      if constexpr (RQ) {
        if (mnorm) rs_carry = row4_sum(rs_carry);  // RQ keeps whole row sums
      }
    }
  }

  // the k range in NCH chunks of KC operand slots (two chunks from n = 64 on with MFMA),
  // which bounds the operand registers and measured faster at n = 64; V1 up to n = 64
  // holds fewer registers and takes one chunk (at n = 128 one chunk would spill)
  constexpr int KC = ((V1 && NB <= 64) || KQ <= 8 || KQ % 8 != 0) ? KQ : KQ / 2;
  constexpr int NCH = KQ / KC;
  static_assert(KC * NCH == KQ && KC % 4 == 0, "operand chunks");
  // PF: A of the next block row is loaded one row ahead, after this row's other loads,
  // so that it is still in flight while this row waits for them (V1 at n <= 64, where
  // the registers allow it; the last row of an item re-reads its own block from L2)
  constexpr bool PF = V1 && NB <= 64;
  double pA[KQ];
  if constexpr (PF) {
#pragma unroll
    for (int q = 0; q < KQ; ++q) pA[q] = As[b0 * NN + (long long)kof(q) * NB + 16 * w + lr];
  }
  // RQ runs in two chunks from n = 96 on (one at n = 64).  EO (explicit order): the
  // second chunk's loads are issued before the first chunk's work; left to itself the scheduler issues them after it,
  // and a work-group (one per CU at n = 128) then has half a block row in flight.
  // At n = 96 the registers this takes cost the second work-group per CU, and the
  // scheduler's own order measured faster.
  constexpr bool EO = RQ && NB >= 128;
  static_assert(!EO || NCH == 2, "explicit order: two chunks");
  for (int blk = b0; blk < b1; ++blk) {
    // operands of this block row, issued first; B_blk twice (row blk, and transposed
    // for row blk + 1 when that row is ours) unless TS or RQ carries B^T x
    const bool next = blk + 1 < N;
    const bool carry_on = next && blk + 1 < b1;
    double xs[SE];
    stage_load(blk + 1, xs);
    double aA[KC], aB[KC], aBn[KC];
    double cA[EO ? KC : 1], cB[EO ? KC : 1];  // EO: chunk 0
    // B_blk exists only below the last block; otherwise its loads read A_blk (always
    // there) and are zeroed, so no load is conditional
    const auto Bb = next ? Bs + blk * NN : As + blk * NN;
    auto load_ops = [&](int ch, double(&uA)[KC], double(&uB)[KC]) {
#pragma unroll
      for (int qq = 0; qq < KC; ++qq) {
        const int k = kof(ch * KC + qq);
        if constexpr (PF) uA[qq] = pA[ch * KC + qq];
        else uA[qq] = As[blk * NN + (long long)k * NB + 16 * w + lr];
        const double b = Bb[(long long)k * NB + 16 * w + lr];
        uB[qq] = next ? b : 0.0;
      }
    };
    auto load_chunk = [&](int ch) {
      load_ops(ch, aA, aB);
      if constexpr (!TS && !RQ) {
#pragma unroll
        for (int qq = 0; qq < KC; ++qq) {
          const double b = Bb[(long long)(16 * w + lr) * NB + kof(ch * KC + qq)];
          aBn[qq] = carry_on ? b : 0.0;
        }
      }
    };
    // d_blk, like x, ahead of the operands
    acc_t acc;
    [[maybe_unused]] double d1 = 0.0;
    if constexpr (V1) {
      d1 = ds[(long long)blk * ND + 16 * w + lr];
    } else {
      const int dc = lr < nrhs ? lr : 0;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * w + Mfma<double>::row(lk, v);
        const double x = ds[(long long)blk * ND + dc * NB + row];
        acc[v] = (lr < nrhs) ? x : 0.0;
      }
    }
    if constexpr (!EO) load_chunk(0);
    else load_ops(0, cA, cB);
    if constexpr (PF) {
      const int j = blk + 1 < b1 ? blk + 1 : blk;
#pragma unroll
      for (int q = 0; q < KQ; ++q) pA[q] = As[j * NN + (long long)kof(q) * NB + 16 * w + lr];
    }
    if constexpr (V1) {
      dmax = nmax(dmax, fabs(d1));
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) dmax = nmax(dmax, fabs(acc[v]));
    }
    __syncthreads();  // previous block's window reads are done
    if constexpr (TS) {
      // this row's carry: column 16w + lr of the previous row's product tile (written
      // after the previous row's second barrier, rewritten only after this row's)
      // (the four lanes of a row each sum NB / 4 of the tile's rows)
      if (blk > b0) {
        double t = 0.0;
#pragma unroll
        for (int r = 0; r < NB / 4; ++r) t += P[((NB / 4) * lk + r) * LDP + 16 * w + lr];
        carry1 = row_sum(t);
      }
    }
    if constexpr (RQ) {
      // this row's carry (and |B| row sum): the previous row's per-wave sums (written
      // after the previous row's second barrier, rewritten only after this row's)
      if (blk > b0) {
        double t = 0.0, u = 0.0;
#pragma unroll
        for (int v = 0; v < TT; ++v) {
          t += Q[0][v * NB + 16 * w + lr];
          if (mnorm) u += Q[1][v * NB + 16 * w + lr];
        }
        carry1 = t;
        if (mnorm) rs_carry = u;
      }
    }
    stage_store(blk + 1, xs);
    __syncthreads();
    const double* X0 = X[blk % 3];
    const double* Xp = X[(blk + 1) % 3];
    if constexpr (V1) {
      xmax = nmax(xmax, fabs(X0[16 * w + lr]));
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v)
        xmax = nmax(xmax, fabs(X0[lr * LDV + 16 * w + Mfma<double>::row(lk, v)]));
    }
    {
      if constexpr (!V1) {
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] -= carry[v];
      }
      acc_t cn = Mfma<double>::zero();
      // V1 partial sums: A x and B x in two chains, B^T x (the next row's carry)
      [[maybe_unused]] double sa = 0.0, sb = 0.0, sn = 0.0;
      double rs = RQ ? 0.0 : rs_carry, rsn = 0.0;
      // RQ: the next row's carry, B[r][k] x[r] summed over the wave's 16 rows (and
      // |B[r][k]| for ||M||), to Q[.][w NB + k], for slots q0 .. q0 + G of chunk ch as
      // their FMAs end (a whole chunk at once spills at n = 128).  No branch (a branch
      // splits the unrolled loop and its loads then spill): every lane stores, the lanes
      // that share a sum hold the same bits (its adds are commutative); rows without a
      // next row in this item store sums nobody reads.
      constexpr int G = KC % 8 == 0 ? 8 : 4;
      auto rq_sums = [&](int ch, int q0, const double(&uB)[KC]) {
        constexpr int CNT = Red16<G>::CNT;
        const int f = red16_first<G>(lr);
        const double xr = X0[16 * w + lr];
        double t[G];
#pragma unroll
        for (int j = 0; j < G; ++j) t[j] = uB[q0 + j] * xr;
        red16<G>(t, lr);
#pragma unroll
        for (int c = 0; c < CNT; ++c) Q[0][w * NB + kof(ch * KC + q0 + f + c)] = t[c];
        if (mnorm) {
#pragma unroll
          for (int j = 0; j < G; ++j) t[j] = fabs(uB[q0 + j]);
          red16<G>(t, lr);
#pragma unroll
          for (int c = 0; c < CNT; ++c) Q[1][w * NB + kof(ch * KC + q0 + f + c)] = t[c];
        }
      };
      if constexpr (EO) {
        auto chunk_work = [&](int ch, const double(&uA)[KC], const double(&uB)[KC]) {
#pragma unroll
          for (int qq = 0; qq < KC; ++qq) {
            const int k = kof(ch * KC + qq);
            sa = fma(uA[qq], X0[k], sa);
            sb = fma(uB[qq], Xp[k], sb);
            if (mnorm) rs += fabs(uA[qq]) + fabs(uB[qq]);
            if (qq % G == G - 1) rq_sums(ch, qq + 1 - G, uB);
          }
        };
        // chunk 1's loads, then chunk 0's work, kept in that order by the barrier
        double nA[KC], nB[KC];
        load_ops(1, nA, nB);
        __builtin_amdgcn_sched_barrier(0);
        chunk_work(0, cA, cB);
        chunk_work(1, nA, nB);
      } else {
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
          if (ch > 0) load_chunk(ch);
#pragma unroll
          for (int qq = 0; qq < KC; ++qq) {
            const int k = kof(ch * KC + qq);
            if constexpr (V1) {
              const double x0 = X0[k];
              sa = fma(aA[qq], x0, sa);
              sb = fma(aB[qq], Xp[k], sb);
              if constexpr (!TS && !RQ) sn = fma(aBn[qq], x0, sn);
            } else {
              acc = Mfma<double>::mma(-aA[qq], X0[lr * LDV + k], acc);
              acc = Mfma<double>::mma(-aB[qq], Xp[lr * LDV + k], acc);
              cn = Mfma<double>::mma(aBn[qq], X0[lr * LDV + k], cn);
            }
            if (mnorm) {
              rs += fabs(aA[qq]) + fabs(aB[qq]);
              if constexpr (!RQ) rsn += fabs(aBn[qq]);
            }
            if constexpr (RQ) {
              if (qq % G == G - 1) rq_sums(ch, qq + 1 - G, aB);
            }
          }
        }
      }
      if constexpr (TS) {
        d1 = d1 - carry1 - row_sum(sa + sb);
        if (carry_on) {
          const double xr = X0[16 * w + lr];
#pragma unroll
          for (int q = 0; q < KQ; ++q) P[(16 * w + lr) * LDP + kof(q)] = aB[q] * xr;
        }
      } else if constexpr (V1 && RQ) {
        d1 = d1 - carry1 - row_sum(sa + sb);  // the next carry comes through Q
      } else if constexpr (V1) {
        d1 = d1 - carry1 - row_sum(sa + sb);
        carry1 = row_sum(sn);
      } else {
        carry = cn;
      }
      if constexpr (!RQ) rs_carry = rsn;
      if (mnorm) {
        // the four lanes sharing lr hold the four k-slices of row 16w + lr
        rs = row4_sum(rs);
        if constexpr (RQ) rs += rs_carry;  // the |B_{i-1}| column sum, already whole
        mmax = nmax(mmax, rs);
      }
    }
    if constexpr (V1) {
      if (lk == 0) ro[(long long)blk * ND + 16 * w + lr] = (float)d1;
      rmax = nmax(rmax, fabs(d1));
    } else if (lr < nrhs) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        ro[(long long)blk * ND + lr * NB + 16 * w + Mfma<double>::row(lk, v)] = (float)acc[v];
        rmax = nmax(rmax, fabs(acc[v]));
      }
    }
  }
  if constexpr (V1) {
    // lanes hold rows of column 0: fold the wave into column slot 0
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      rmax = nmax(rmax, __shfl_xor(rmax, o));
      xmax = nmax(xmax, __shfl_xor(xmax, o));
      dmax = nmax(dmax, __shfl_xor(dmax, o));
    }
    if (lr != 0) rmax = xmax = dmax = 0.0;
  }
  red[0][4 * w + lk][lr] = rmax;
  red[1][4 * w + lk][lr] = xmax;
  red[2][4 * w + lk][lr] = dmax;
  red[3][4 * w + lk][lr] = mmax;
  __syncthreads();
  if (tid < kParts) {
    const int q = tid / 16, c = tid % 16;
    double m = 0.0;
    if (q < 3) {
      for (int k = 0; k < 4 * TT; ++k) m = nmax(m, red[q][k][c]);
    } else if (c == 0) {
      for (int k = 0; k < 4 * TT; ++k)
        for (int cc = 0; cc < 16; ++cc) m = nmax(m, red[3][k][cc]);
    }
    if (q < 3 || c == 0) part[(long long)blockIdx.x * kParts + (q < 3 ? tid : kMnormSlot)] = m;
  }
}

// Combine the partial maxima: berr per (system, column).  `init` records ||d_j||;
// `mnorm` first reduces ||M||_inf from the same pass.  One kRedThreads workgroup per
// system: each thread folds every kRedThreads-th partial (eight loads in flight), then
// LDS finishes the reduction.
constexpr int kRedThreads = 1024;

__global__ void __launch_bounds__(kRedThreads)
    berr_kernel(const ResSys* __restrict__ systems, const double* __restrict__ part,
                double* __restrict__ mnorm, double* __restrict__ dnorm,
                double* __restrict__ berr, int nrhs, int init, int want_mnorm) {
  static_assert(kRedThreads % kParts == 0, "a thread keeps one partial slot");
  __shared__ double red[kRedThreads];
  const int sys = blockIdx.x, tid = threadIdx.x;
  const int g0 = systems[sys].item0, g1 = systems[sys].item1;
  // the system's partials are contiguous (item-major, kParts slots each): thread t
  // reads every kRedThreads-th value, all of slot t % kParts, coalesced
  const auto base = gptr(part) + (long long)g0 * kParts;
  const long long total = (long long)(g1 - g0) * kParts;
  double m = 0.0;
  long long i = tid;
  // the partials were just written by other CUs: issue eight loads before using any
  for (; i + 7 * kRedThreads < total; i += 8 * kRedThreads) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = base[i + u * kRedThreads];
#pragma unroll
    for (int u = 0; u < 8; ++u) m = nmax(m, v[u]);
  }
  for (; i < total; i += kRedThreads) m = nmax(m, base[i]);
  red[tid] = m;
  __syncthreads();
  if (tid < kParts) {
    for (int k = 1; k < kRedThreads / kParts; ++k) m = nmax(m, red[tid + k * kParts]);
    red[tid] = m;
  }
  __syncthreads();
  if (tid >= 16) return;
  const double mn = want_mnorm ? red[kMnormSlot] : mnorm[sys];
  if (want_mnorm && tid == 0) mnorm[sys] = mn;
  const int c = tid;
  if (c >= nrhs) return;
  const double r = red[c], x = red[16 + c], dd = red[32 + c];
  if (init) dnorm[sys * 16 + c] = dd;
  const double den = mn * x + (init ? dd : dnorm[sys * 16 + c]);
  berr[sys * nrhs + c] = den > 0.0 ? r / den : r;
}

// blocks per residual workgroup: at least 32, and at most 256 workgroups per system
// (at N = 8192, two systems: two workgroups per CU, measured 3% faster than four with
// 16 blocks each; 24 or 64 blocks were 30% slower)
int blocks_per_item(int N) { return std::max(32, (N + 255) / 256); }

// The first pass of a solve (x = 0, so r = d): d into the FP32 vector, zeros into the
// FP64 iterate, and the item's per-column max |d_j| (as both the residual and the
// right-hand-side maxima; the iterate maximum is 0).  Elementwise, with no block
// window: one workgroup per residual work item, kInitThreads threads.
constexpr int kInitThreads = 256;

__global__ void __launch_bounds__(kInitThreads)
    init_pass_kernel(const ResItem* __restrict__ items, const ResSys* __restrict__ systems,
                     double* __restrict__ part, int n, int nrhs, int xn_k, int ro_k,
                     const int* __restrict__ skip) {
  __shared__ double red[kInitThreads / 64];
  if (skip && *skip == 0) return;  // the whole workgroup leaves (uniform)
  const ResItem it = items[blockIdx.x];
  const ResSys& S = systems[it.sys];
  const long long ND = (long long)n * nrhs;
  const auto ds = gptr(S.d);
  const auto xn = gptr(S.x[xn_k]);
  const auto ro = gptr(S.v[ro_k]);
  const int tid = threadIdx.x, w = wave_id(), l = lane_id();
  const int per = kInitThreads / n;  // blocks per sweep of the workgroup (n <= 128)
  double* out = part + (long long)blockIdx.x * kParts;
  for (int c = 0; c < nrhs; ++c) {
    double m = 0.0;
    for (long long j = it.b0 + tid / n; tid < per * n && j < it.b1; j += per) {
      const long long e = j * ND + (long long)c * n + tid % n;
      const double v = ds[e];
      ro[e] = (float)v;
      xn[e] = 0.0;
      m = nmax(m, fabs(v));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = nmax(m, __shfl_xor(m, o));
    if (l == 0) red[w] = m;
    __syncthreads();
    if (tid == 0) {
      double t = red[0];
      for (int k = 1; k < kInitThreads / 64; ++k) t = nmax(t, red[k]);
      out[c] = t;       // max |r_c| (r = d)
      out[16 + c] = 0;  // max |x_c|
      out[32 + c] = t;  // max |d_c|
    }
    __syncthreads();
  }
  if (tid == 0) out[kMnormSlot] = 0.0;
}

template <int NB>
hipError_t launch_residual(const MixedState& m, bool init, int xo_k, int dx_k, int xn_k,
                           int ro_k, int want_mnorm, hipStream_t s) {
  if (init)
    hipLaunchKernelGGL(init_pass_kernel, dim3(m.nitems), dim3(kInitThreads), 0, s, m.items,
                       m.sys_dev, m.part, NB, m.nrhs, xn_k, ro_k, m.init_skip);
  else if (m.nrhs == 1 && !want_mnorm)
    hipLaunchKernelGGL((residual_kernel<NB, true, NB <= 64, false>), dim3(m.nitems),
                       dim3(4 * NB), 0, s, m.items, m.sys_dev, m.part, m.nrhs, xo_k, dx_k, xn_k,
                       ro_k, 0);
  else if (m.nrhs == 1)
    hipLaunchKernelGGL((residual_kernel<NB, true, false, true>), dim3(m.nitems), dim3(4 * NB),
                       0, s, m.items, m.sys_dev, m.part, m.nrhs, xo_k, dx_k, xn_k, ro_k, 1);
  else
    hipLaunchKernelGGL((residual_kernel<NB, false, false, true>), dim3(m.nitems), dim3(4 * NB),
                       0, s, m.items, m.sys_dev, m.part, m.nrhs, xo_k, dx_k, xn_k, ro_k,
                       want_mnorm);
  return hipGetLastError();
}

// profiler kinds (enum bss_kernel_kind in bss.h)
constexpr int kKindResidual = 6, kKindResidualInit = 7, kKindBerr = 8;

// One refinement pass: the residual kernel, then the per-system backward error.
hipError_t residual_prof(const MixedState& m, const MixedSolveHooks& h, bool init, int xo_k,
                         int dx_k, int xn_k, int ro_k, int want_mnorm, hipStream_t s) {
  hipError_t e;
  {
    void* scope = h.prof_begin ? h.prof_begin(h.ctx, init ? kKindResidualInit : kKindResidual, s)
                               : nullptr;
    switch (m.n) {
      case 16: e = launch_residual<16>(m, init, xo_k, dx_k, xn_k, ro_k, want_mnorm, s); break;
      case 32: e = launch_residual<32>(m, init, xo_k, dx_k, xn_k, ro_k, want_mnorm, s); break;
      case 48: e = launch_residual<48>(m, init, xo_k, dx_k, xn_k, ro_k, want_mnorm, s); break;
      case 64: e = launch_residual<64>(m, init, xo_k, dx_k, xn_k, ro_k, want_mnorm, s); break;
      case 96: e = launch_residual<96>(m, init, xo_k, dx_k, xn_k, ro_k, want_mnorm, s); break;
      case 128: e = launch_residual<128>(m, init, xo_k, dx_k, xn_k, ro_k, want_mnorm, s); break;
      default: e = hipErrorInvalidValue;
    }
    if (scope) h.prof_end(scope);
  }
  if (e != hipSuccess) return e;
  void* scope = h.prof_begin ? h.prof_begin(h.ctx, kKindBerr, s) : nullptr;
  hipLaunchKernelGGL(berr_kernel, dim3(m.nsys), dim3(kRedThreads), 0, s, m.sys_dev, m.part,
                     m.mnorm, m.dnorm, m.berr, m.nrhs, init ? 1 : 0, want_mnorm);
  e = hipGetLastError();
  if (scope) h.prof_end(scope);
  return e;
}

}  // namespace

size_t MixedState::bytes(const std::vector<int>& Ns, int n, int nrhs) {
  size_t b = 0, items = 0;
  const size_t S = Ns.size();
  size_t nd_all = 0;
  for (int N : Ns) {
    nd_all += (size_t)N * n * nrhs;
    items += (N + blocks_per_item(N) - 1) / blocks_per_item(N);
  }
  b += 2 * al(nd_all * 8) + 2 * al(nd_all * 4);
  b += al(items * sizeof(ResItem)) + al(S * sizeof(ResSys)) + al(items * kParts * 8);
  b += al(S * 8) + al(S * 16 * 8) + al(S * 16 * 8);
  if (factor_fuses_forward(n, true, nrhs)) b += al(nd_all * 8) + al(sizeof(int));
  return b;
}

int MixedState::attach(char* base, const std::vector<int>& Ns_, int n_, int nrhs_) {
  Ns = Ns_;
  n = n_;
  nrhs = nrhs_;
  nsys = (int)Ns.size();
  size_t o = 0;
  auto take = [&](size_t bytes) {
    char* p = base + o;
    o += al(bytes);
    return p;
  };
  sys_host.assign(nsys, ResSys{});
  std::vector<ResItem> it;
  // each vector set packs the systems like the caller's d (system s at the sum of the
  // earlier systems' N * n * nrhs), so the final iterate leaves in one copy
  size_t nd_all = 0;
  for (int N : Ns) nd_all += (size_t)N * n * nrhs;
  double* x0 = (double*)take(nd_all * 8);
  double* x1 = (double*)take(nd_all * 8);
  float* v0 = (float*)take(nd_all * 4);
  float* v1 = (float*)take(nd_all * 4);
  size_t od = 0;
  for (int s = 0; s < nsys; ++s) {
    const int N = Ns[s];
    const size_t nd = (size_t)N * n * nrhs;
    ResSys& r = sys_host[s];
    r.x[0] = x0 + od;
    r.x[1] = x1 + od;
    r.v[0] = v0 + od;
    r.v[1] = v1 + od;
    od += nd;
    r.N = N;
    r.item0 = (int)it.size();
    const int bpw = blocks_per_item(N);
    for (int b0 = 0; b0 < N; b0 += bpw) it.push_back({s, b0, std::min(N, b0 + bpw), 0});
    r.item1 = (int)it.size();
  }
  nitems = (int)it.size();
  items = (ResItem*)take(it.size() * sizeof(ResItem));
  sys_dev = (ResSys*)take(nsys * sizeof(ResSys));
  part = (double*)take((size_t)nitems * kParts * 8);
  mnorm = (double*)take(nsys * 8);
  dnorm = (double*)take((size_t)nsys * 16 * 8);
  berr = (double*)take((size_t)nsys * 16 * 8);
  dsnap = nullptr;
  dflag = nullptr;
  if (factor_fuses_forward(n, true, nrhs)) {
    dsnap = (double*)take(nd_all * 8);
    dflag = (int*)take(sizeof(int));
  }
  if (hipMemcpy(items, it.data(), it.size() * sizeof(ResItem), hipMemcpyHostToDevice) !=
      hipSuccess) {
    g_mixed_err = "work-item upload failed";
    return -1;
  }
  if (!host_berr &&
      !(host_berr = (double*)pinned_take((size_t)nsys * 16 * sizeof(double)))) {
    g_mixed_err = "pinned host allocation failed";
    return -1;
  }
  if (!berr_ready && hipEventCreateWithFlags(&berr_ready, hipEventDisableTiming) != hipSuccess) {
    berr_ready = nullptr;
    g_mixed_err = "event creation failed";
    return -1;
  }
  last_berr.assign((size_t)nsys * nrhs, 0.0);
  return 0;
}

int MixedState::bind(const double* A, const double* B, const double* d) {
  size_t oa = 0, ob = 0, od = 0;
  const size_t nn = (size_t)n * n, ndv = (size_t)n * nrhs;
  for (int s = 0; s < nsys; ++s) {
    const int N = Ns[s];
    sys_host[s].A = A + oa;
    sys_host[s].B = B ? B + ob : nullptr;
    sys_host[s].d = d + od;
    oa += (size_t)N * nn;
    ob += (size_t)std::max(N - 1, 0) * nn;
    od += (size_t)N * ndv;
  }
  if (hipMemcpy(sys_dev, sys_host.data(), nsys * sizeof(ResSys), hipMemcpyHostToDevice) !=
      hipSuccess) {
    g_mixed_err = "system-table upload failed";
    return -1;
  }
  return 0;
}

void MixedState::release() {
  pinned_give(host_berr);
  host_berr = nullptr;
  if (berr_ready) (void)hipEventDestroy(berr_ready);
  berr_ready = nullptr;
}

hipError_t wait_event_spin(hipEvent_t ev) {
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    const hipError_t q = hipEventQuery(ev);
    if (q != hipErrorNotReady) return q;
    if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2))
      return hipEventSynchronize(ev);
  }
}

void mixed_after_factor(MixedState& m) { m.mnorm_stale = true; }

// below this tolerance the backward error after the first FP32 pass is not read
constexpr double kFirstCheckTol = 1e-9;

namespace {
__global__ void __launch_bounds__(256) snapshot_kernel(const ResSys* __restrict__ systems,
                                                       double* __restrict__ snap, long long n,
                                                       int* __restrict__ flag) {
  const double* __restrict__ d = systems[0].d;  // system 0 starts the packed d
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    snap[i] = d[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) *flag = 0;
}
}  // namespace

int mixed_snapshot_d(MixedState& m, hipStream_t s) {
  if (!m.dsnap || !m.dflag) {
    g_mixed_err = "no d snapshot in this workspace";
    return -1;
  }
  long long nd = 0;
  for (int N : m.Ns) nd += (long long)N * m.n * m.nrhs;
  const int grid = (int)std::min<long long>((nd + 255) / 256, 4096);
  hipLaunchKernelGGL(snapshot_kernel, dim3(grid), dim3(256), 0, s, m.sys_dev, m.dsnap, nd,
                     m.dflag);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    g_mixed_err = hipGetErrorString(e);
    return -1;
  }
  return 0;
}

int mixed_first_pass(MixedState& m, const MixedSolveHooks& h, hipStream_t s) {
  // x = 0, r = d (as FP32 into v[0]), ||d||.  With init_skip the kernel may leave at
  // once; the backward-error kernel after it still runs, on the partial norms the
  // skipped pass wrote the last time (nothing else writes them in between)
  const hipError_t e = residual_prof(m, h, true, -1, -1, 0, 0, 0, s);
  m.init_skip = nullptr;
  if (e != hipSuccess) {
    g_mixed_err = hipGetErrorString(e);
    return -1;
  }
  return 0;
}

namespace {
// copy the packed iterate set k into the caller's d
int write_iterate(MixedState& m, double* d, int k, hipStream_t s) {
  size_t nd_all = 0;
  for (int sy = 0; sy < m.nsys; ++sy) nd_all += (size_t)m.Ns[sy] * m.n * m.nrhs;
  const hipError_t e = hipMemcpyAsync(d, m.sys_host[0].x[k], nd_all * 8,
                                      hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) {
    g_mixed_err = hipGetErrorString(e);
    return -1;
  }
  return 0;
}

// berr to the host (one small copy, polled), into last_berr; *worst = max (NaN wins)
int read_berr(MixedState& m, hipStream_t s, double* worst) {
  hipError_t e = hipMemcpyAsync(m.host_berr, m.berr, sizeof(double) * m.nsys * m.nrhs,
                                hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipEventRecord(m.berr_ready, s);
  if (e == hipSuccess) e = wait_event_spin(m.berr_ready);
  if (e != hipSuccess) {
    g_mixed_err = hipGetErrorString(e);
    return -1;
  }
  double w = 0.0;
  for (int k = 0; k < m.nsys * m.nrhs; ++k) {
    const double b = m.host_berr[k];
    m.last_berr[k] = b;
    w = (b != b || w != w) ? NAN : std::max(w, b);
  }
  *worst = w;
  return 0;
}
}  // namespace

int mixed_solve(MixedState& m, const MixedSolveHooks& h, double* d, double tol, int max_iters,
                int* iters, bool* converged, bool write_unconverged, hipStream_t s,
                bool first_pass_done) {
  int cur = 0, xa = 0, it = 0;
  hipError_t e;
  *converged = false;
  *iters = 0;
  if (!first_pass_done && mixed_first_pass(m, h, s) != 0) return -1;
  bool prepared = h.prepare == nullptr;
  double prev = HUGE_VAL;  // backward error at the previous check
  while (true) {
    if (h.solve32(h.ctx, cur, s) != 0) {
      g_mixed_err = "FP32 solve failed";
      return -1;
    }
    if (!prepared) {
      // the first solve is queued before the caller's check; a changed solve path
      // redoes it (d is untouched until the end, so the x = 0 pass rebuilds r = d)
      prepared = true;
      const int pr = h.prepare(h.ctx, s);
      if (pr == 1) {
        if (mixed_first_pass(m, h, s) != 0) return -1;
        if (h.solve32(h.ctx, cur, s) != 0) {
          g_mixed_err = "FP32 solve failed";
          return -1;
        }
      } else if (pr == 2) {
        // an FP32 pivot failed: every iterate would be NaN
        for (double& b : m.last_berr) b = NAN;
        return 0;
      } else if (pr != 0) {
        g_mixed_err = "solve preparation failed";
        return -1;
      }
    }
    const int want_mnorm = m.mnorm_stale ? 1 : 0;
    e = residual_prof(m, h, false, xa, cur, 1 - xa, 1 - cur, want_mnorm, s);
    m.mnorm_stale = false;
    if (e != hipSuccess) {
      g_mixed_err = hipGetErrorString(e);
      return -1;
    }
    xa = 1 - xa;
    cur = 1 - cur;
    ++it;
    // One FP32 solve of d cannot reach a tolerance far below FP32's precision, so the
    // first pass is not checked: the host queues the next correction while the device
    // is still busy instead of waiting for the backward error (an exactly solvable
    // system then takes one more, harmless, correction).
    if (it == 1 && tol < kFirstCheckTol && max_iters > 1) continue;
    double worst;
    if (read_berr(m, s, &worst) != 0) return -1;
    if (worst <= tol) {
      *converged = true;
      break;
    }
    // stop short: NaN (never compares), the iteration cap, or stagnation -- refinement
    // with a contraction factor above 1/2 is not going to reach the target soon, and
    // above 1 it diverges
    if (worst != worst || it >= max_iters || !(worst <= 0.5 * prev)) break;
    prev = worst;
  }
  *iters = it;
  // the final iterate into the caller's d (the iterate set is packed like d)
  if ((*converged || write_unconverged) && write_iterate(m, d, xa, s) != 0) return -1;
  return 0;
}

int mixed_measure(MixedState& m, const MixedSolveHooks& h, hipStream_t s, double* worst) {
  size_t nd_all = 0;
  for (int sy = 0; sy < m.nsys; ++sy) nd_all += (size_t)m.Ns[sy] * m.n * m.nrhs;
  // a zero FP32 correction in set 0, then x_1 = x_0 + 0 and r = d - M x_0
  hipError_t e = hipMemsetAsync(m.sys_host[0].v[0], 0, nd_all * 4, s);
  if (e == hipSuccess) e = residual_prof(m, h, false, 0, 0, 1, 1, m.mnorm_stale ? 1 : 0, s);
  m.mnorm_stale = false;
  if (e != hipSuccess) {
    g_mixed_err = hipGetErrorString(e);
    return -1;
  }
  return read_berr(m, s, worst);
}

const char* mixed_last_error() { return g_mixed_err.c_str(); }

}  // namespace bss
