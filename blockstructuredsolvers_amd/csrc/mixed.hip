// mixed.hip -- FP64 residual / norm kernels and the refinement loop (see mixed.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "common.hpp"
#include "mixed.hpp"

namespace bss {
namespace {

thread_local std::string g_mixed_err;
constexpr int kMaxWG = 1024;  // residual workgroups per system (partial-max slots)
// per workgroup partials: [0,16) max|r_j|, [16,32) max|x_j|, [32,48) max|d_j|,
// [48] max row sum of |M| (when requested)
constexpr int kParts = 64;
constexpr int kMnormSlot = 48;

size_t al(size_t b) { return (b + 255) & ~size_t(255); }

// NaN-propagating max (fmax would drop a NaN residual and report convergence)
__device__ inline double nmax(double a, double b) { return (b > a || b != b) ? b : a; }

// Residual pass over the blocks [wg*bpw, (wg+1)*bpw) of one system:
//   x_new = x_old + dx (FP64; written for the workgroup's own blocks),
//   r = d - (B_{i-1}^T x_{i-1} + A_i x_i + B_i x_{i+1}) in FP64 on MFMA, stored as FP32
//   for the next correction solve, plus per-column max |r|, |x|, |d|.
// INIT: x = 0, so r = d with no products (the first pass of a solve).
// With `mnorm`, the row sums of |M| are accumulated from the operand values the
// products already hold, so ||M||_inf costs no extra pass over A and B.
// One wave per 16-row tile; a rolling three-block window of x_new lives in LDS and
// each block's A/B operands are issued before the window update to hide HBM latency.
template <int NB, bool INIT>
__global__ void __launch_bounds__(4 * NB)
    residual_kernel(const double* __restrict__ A, const double* __restrict__ B,
                    const double* __restrict__ d, const double* __restrict__ xold,
                    const float* __restrict__ dx, double* __restrict__ xnew,
                    float* __restrict__ rout, double* __restrict__ part, int N, int nrhs,
                    int bpw, int nwg, int mnorm) {
  constexpr int TT = NB / 16, NT = 64 * TT, LDV = NB + 2, KQ = NB / 4;
  constexpr long long NN = (long long)NB * NB;
  typedef Mfma<double>::acc_t acc_t;
  __shared__ double X[3][16 * LDV];
  __shared__ double red[4][4 * TT][16];
  const int sys = blockIdx.y, wg = blockIdx.x, tid = threadIdx.x;
  const int w = wave_id(), l = lane_id(), lr = l & 15, lk = l >> 4;
  const long long ND = (long long)NB * nrhs;
  const double* As = A + sys * (long long)N * NN;
  const double* Bs = B + sys * (long long)(N > 0 ? N - 1 : 0) * NN;
  const double* ds = d + sys * (long long)N * ND;
  const double* xo = xold ? xold + sys * (long long)N * ND : nullptr;
  const float* dxs = dx ? dx + sys * (long long)N * ND : nullptr;
  double* xn = xnew + sys * (long long)N * ND;
  float* ro = rout + sys * (long long)N * ND;
  for (int e = tid; e < 3 * 16 * LDV; e += NT) (&X[0][0])[e] = 0.0;
  double rmax = 0.0, xmax = 0.0, dmax = 0.0, mmax = 0.0;
  const int b0 = wg * bpw, b1 = min(N, b0 + bpw);

  // x_new of block j into its window slot (zero outside the system); own blocks are
  // also written to xnew
  auto stage = [&](int j) {
    double* slot = X[((j % 3) + 3) % 3];
    for (int e = tid; e < (int)ND; e += NT) {
      double v = 0.0;
      if (!INIT && j >= 0 && j < N) {
        if (xo) v = xo[(long long)j * ND + e];
        if (dxs) v += (double)dxs[(long long)j * ND + e];
      }
      slot[(e / NB) * LDV + (e % NB)] = v;
      if (j >= b0 && j < b1) xn[(long long)j * ND + e] = v;
    }
  };

  for (int blk = b0; blk < b1; ++blk) {
    // operands of this block row, issued first
    double aA[KQ], aBt[KQ], aB[KQ];
    if (!INIT) {
#pragma unroll
      for (int q = 0; q < KQ; ++q) {
        const int k = (q >> 2) * 16 + 4 * lk + (q & 3);
        aA[q] = As[blk * NN + (long long)k * NB + 16 * w + lr];
        aBt[q] = blk > 0 ? Bs[(blk - 1) * NN + (long long)(16 * w + lr) * NB + k] : 0.0;
        aB[q] = blk + 1 < N ? Bs[blk * NN + (long long)k * NB + 16 * w + lr] : 0.0;
      }
    }
    acc_t acc;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = 16 * w + Mfma<double>::row(lk, v);
      acc[v] = (lr < nrhs) ? ds[(long long)blk * ND + lr * NB + row] : 0.0;
      dmax = nmax(dmax, fabs(acc[v]));
    }
    __syncthreads();  // previous block's window reads are done
    if (blk == b0) {
      stage(blk - 1);
      stage(blk);
    }
    stage(blk + 1);
    __syncthreads();
    const double* Xm = X[(((blk - 1) % 3) + 3) % 3];
    const double* X0 = X[blk % 3];
    const double* Xp = X[(blk + 1) % 3];
#pragma unroll
    for (int v = 0; v < 4; ++v)
      xmax = nmax(xmax, fabs(X0[lr * LDV + 16 * w + Mfma<double>::row(lk, v)]));
    if (!INIT) {
      double rs = 0.0;
#pragma unroll
      for (int q = 0; q < KQ; ++q) {
        const int k = (q >> 2) * 16 + 4 * lk + (q & 3);
        acc = Mfma<double>::mma(-aA[q], X0[lr * LDV + k], acc);
        acc = Mfma<double>::mma(-aBt[q], Xm[lr * LDV + k], acc);
        acc = Mfma<double>::mma(-aB[q], Xp[lr * LDV + k], acc);
        if (mnorm) rs += fabs(aA[q]) + fabs(aBt[q]) + fabs(aB[q]);
      }
      if (mnorm) {
        // the four lanes sharing lr hold the four k-slices of row 16w + lr
        rs += __shfl_xor(rs, 16);
        rs += __shfl_xor(rs, 32);
        mmax = nmax(mmax, rs);
      }
    }
    if (lr < nrhs) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        ro[(long long)blk * ND + lr * NB + 16 * w + Mfma<double>::row(lk, v)] = (float)acc[v];
        rmax = nmax(rmax, fabs(acc[v]));
      }
    }
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
    if (q < 3 || c == 0) part[((long long)sys * nwg + wg) * kParts + (q < 3 ? tid : kMnormSlot)] = m;
  }
}

// Combine the partial maxima: berr per (system, column).  `init` records ||d_j||;
// `mnorm` first reduces ||M||_inf from the same pass.  One 256-thread workgroup per
// system: each thread folds every 256th partial record, then wave shuffles and LDS
// finish the reduction.
constexpr int kRedThreads = 256;
constexpr int kSlots = 49;  // 3 x 16 column maxima + the |M| row-sum maximum

__global__ void __launch_bounds__(kRedThreads)
    berr_kernel(const double* __restrict__ part, double* __restrict__ mnorm,
                double* __restrict__ dnorm, double* __restrict__ berr, int nwg, int nrhs,
                int init, int want_mnorm) {
  __shared__ double red[kRedThreads / 64][kSlots];
  const int sys = blockIdx.x, tid = threadIdx.x;
  double m[kSlots];
#pragma unroll
  for (int k = 0; k < kSlots; ++k) m[k] = 0.0;
  for (int g = tid; g < nwg; g += kRedThreads) {
    const double* pp = part + ((long long)sys * nwg + g) * kParts;
#pragma unroll
    for (int k = 0; k < 48; ++k) m[k] = nmax(m[k], pp[k]);
    m[48] = nmax(m[48], pp[kMnormSlot]);
  }
#pragma unroll
  for (int k = 0; k < kSlots; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m[k] = nmax(m[k], __shfl_xor(m[k], off));
  }
  if ((tid & 63) == 0) {
#pragma unroll
    for (int k = 0; k < kSlots; ++k) red[tid >> 6][k] = m[k];
  }
  __syncthreads();
  if (tid >= 16) return;
  auto fold = [&](int k) {
    double v = 0.0;
    for (int w = 0; w < kRedThreads / 64; ++w) v = nmax(v, red[w][k]);
    return v;
  };
  const double mn = want_mnorm ? fold(48) : mnorm[sys];
  if (want_mnorm && tid == 0) mnorm[sys] = mn;
  const int c = tid;
  if (c >= nrhs) return;
  const double r = fold(c), x = fold(16 + c), dd = fold(32 + c);
  if (init) dnorm[sys * 16 + c] = dd;
  const double den = mn * x + (init ? dd : dnorm[sys * 16 + c]);
  berr[sys * nrhs + c] = den > 0.0 ? r / den : r;
}

void geometry(int N, int* nwg, int* bpw) {
  *nwg = std::max(1, std::min(kMaxWG, N));
  *bpw = (N + *nwg - 1) / *nwg;
  *nwg = (N + *bpw - 1) / *bpw;
}

template <int NB>
hipError_t launch_residual(const MixedState& m, const double* A, const double* B,
                           const double* d, const double* xold, const float* dx, double* xnew,
                           float* rout, bool init, int want_mnorm, hipStream_t s) {
  if (init)
    hipLaunchKernelGGL((residual_kernel<NB, true>), dim3(m.nwg, m.nsys), dim3(4 * NB), 0, s, A,
                       B, d, xold, dx, xnew, rout, m.part, m.N, m.nrhs, m.bpw, m.nwg, 0);
  else
    hipLaunchKernelGGL((residual_kernel<NB, false>), dim3(m.nwg, m.nsys), dim3(4 * NB), 0, s, A,
                       B, d, xold, dx, xnew, rout, m.part, m.N, m.nrhs, m.bpw, m.nwg, want_mnorm);
  return hipGetLastError();
}

hipError_t residual(const MixedState& m, const double* A, const double* B, const double* d,
                    const double* xold, const float* dx, double* xnew, float* rout, int init,
                    int want_mnorm, hipStream_t s) {
  hipError_t e;
  switch (m.n) {
    case 16: e = launch_residual<16>(m, A, B, d, xold, dx, xnew, rout, init, want_mnorm, s); break;
    case 32: e = launch_residual<32>(m, A, B, d, xold, dx, xnew, rout, init, want_mnorm, s); break;
    case 48: e = launch_residual<48>(m, A, B, d, xold, dx, xnew, rout, init, want_mnorm, s); break;
    case 64: e = launch_residual<64>(m, A, B, d, xold, dx, xnew, rout, init, want_mnorm, s); break;
    default: return hipErrorInvalidValue;
  }
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(berr_kernel, dim3(m.nsys), dim3(kRedThreads), 0, s, m.part, m.mnorm, m.dnorm,
                     m.berr, m.nwg, m.nrhs, init, want_mnorm);
  return hipGetLastError();
}

}  // namespace

size_t MixedState::bytes(int N, int n, int nrhs, int nsys) {
  int nwg, bpw;
  geometry(N, &nwg, &bpw);
  const size_t nd = (size_t)N * n * nrhs * nsys;
  return 2 * al(nd * 8) + al(nd * 4) + al((size_t)nsys * nwg * kParts * 8) + al(nsys * 8) +
         al((size_t)nsys * 16 * 8) + al((size_t)nsys * 16 * 8);
}

void MixedState::attach(char* base, int N_, int n_, int nrhs_, int nsys_) {
  N = N_;
  n = n_;
  nrhs = nrhs_;
  nsys = nsys_;
  geometry(N, &nwg, &bpw);
  const size_t nd = (size_t)N * n * nrhs * nsys;
  size_t o = 0;
  x[0] = (double*)(base + o);
  o += al(nd * 8);
  x[1] = (double*)(base + o);
  o += al(nd * 8);
  corr = (float*)(base + o);
  o += al(nd * 4);
  part = (double*)(base + o);
  o += al((size_t)nsys * nwg * kParts * 8);
  mnorm = (double*)(base + o);
  o += al(nsys * 8);
  dnorm = (double*)(base + o);
  o += al((size_t)nsys * 16 * 8);
  berr = (double*)(base + o);
  if (!host_berr && hipHostMalloc((void**)&host_berr, (size_t)nsys * 16 * sizeof(double)) != hipSuccess)
    host_berr = nullptr;
  last_berr.assign((size_t)nsys * nrhs, 0.0);
}

void MixedState::release() {
  if (host_berr) (void)hipHostFree(host_berr);
  host_berr = nullptr;
}

int mixed_after_factor(MixedState& m, const double* A, const double* B, int n, int nsys,
                       hipStream_t s) {
  (void)A;
  (void)B;
  (void)n;
  (void)nsys;
  (void)s;
  m.mnorm_stale = true;  // recomputed by the next solve's first residual pass
  return 0;
}

namespace {
constexpr int kKindResidual = 6;  // BSS_K_RESIDUAL in bss.h

hipError_t residual_prof(const MixedState& m, const MixedSolveHooks& h, const double* A,
                         const double* B, const double* d, const double* xold,
                         const float* dx, double* xnew, float* rout, int init, int want_mnorm,
                         hipStream_t s) {
  void* scope = h.prof_begin ? h.prof_begin(h.ctx, kKindResidual, s) : nullptr;
  hipError_t e = residual(m, A, B, d, xold, dx, xnew, rout, init, want_mnorm, s);
  if (scope) h.prof_end(scope);
  return e;
}
}  // namespace

int mixed_solve(MixedState& m, const MixedSolveHooks& h, const double* A, const double* B,
                double* d, float* v0, int n, int nrhs, int nsys, double tol, int max_iters,
                int* iters, hipStream_t s) {
  (void)n;
  (void)nrhs;
  float* v[2] = {v0, m.corr};
  int cur = 0, xa = 0, it = 0;
  hipError_t e = residual_prof(m, h, A, B, d, nullptr, nullptr, m.x[0], v[0], 1, 0, s);
  if (e != hipSuccess) {
    g_mixed_err = hipGetErrorString(e);
    return -1;
  }
  while (true) {
    if (h.solve32(h.ctx, v[cur], s) != 0) {
      g_mixed_err = "FP32 solve failed";
      return -1;
    }
    const int want_mnorm = m.mnorm_stale ? 1 : 0;
    e = residual_prof(m, h, A, B, d, m.x[xa], v[cur], m.x[1 - xa], v[1 - cur], 0, want_mnorm, s);
    m.mnorm_stale = false;
    if (e == hipSuccess)
      e = hipMemcpyAsync(m.host_berr, m.berr, sizeof(double) * nsys * m.nrhs,
                         hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      g_mixed_err = hipGetErrorString(e);
      return -1;
    }
    xa = 1 - xa;
    cur = 1 - cur;
    ++it;
    double worst = 0.0;
    bool nan = false;
    for (int k = 0; k < nsys * m.nrhs; ++k) {
      const double b = m.host_berr[k];
      m.last_berr[k] = b;
      if (b != b) nan = true;
      worst = std::max(worst, b);
    }
    if ((!nan && worst <= tol) || it >= max_iters) break;
  }
  e = hipMemcpyAsync(d, m.x[xa], sizeof(double) * (size_t)m.N * m.n * m.nrhs * nsys,
                     hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) {
    g_mixed_err = hipGetErrorString(e);
    return -1;
  }
  *iters = it;
  return 0;
}

const char* mixed_last_error() { return g_mixed_err.c_str(); }

}  // namespace bss
