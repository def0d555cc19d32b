// mixed.hip -- FP64 residual / norm kernels and the refinement loop (see mixed.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "common.hpp"
#include "mixed.hpp"

namespace bss {
namespace {

thread_local std::string g_mixed_err;
constexpr int kMaxWG = 1024;   // residual workgroups per system (partial-max slots)
constexpr int kParts = 3 * 16; // per workgroup: rmax, xmax, dmax for up to 16 columns

size_t al(size_t b) { return (b + 255) & ~size_t(255); }

// NaN-propagating max (fmax would drop a NaN residual and report convergence)
__device__ inline double nmax(double a, double b) { return (b > a || b != b) ? b : a; }

// r = d - M x_new with x_new = x_old + dx over the blocks [wg*bpw, (wg+1)*bpw).
// One wave per 16-row tile; the three block products use the same MFMA tile helper as
// the solver, with A_i and B_i read from HBM straight into operand registers.
template <int NB>
__global__ void __launch_bounds__(4 * NB)
    residual_kernel(const double* __restrict__ A, const double* __restrict__ B,
                    const double* __restrict__ d, const double* __restrict__ xold,
                    const float* __restrict__ dx, double* __restrict__ xnew,
                    float* __restrict__ rout, double* __restrict__ part, int N, int nrhs,
                    int bpw, int nwg) {
  constexpr int TT = NB / 16, NT = 64 * TT, LDV = NB + 2;
  constexpr long long NN = (long long)NB * NB;
  typedef Mfma<double>::acc_t acc_t;
  __shared__ double X[3][16 * LDV];
  __shared__ double red[3][4 * TT][16];
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
  double rmax = 0.0, xmax = 0.0, dmax = 0.0;
  const int b0 = wg * bpw, b1 = min(N, b0 + bpw);
  for (int blk = b0; blk < b1; ++blk) {
    __syncthreads();
    for (int e = tid; e < 3 * (int)ND; e += NT) {
      const int k = e / (int)ND, ee = e % (int)ND;
      const long long j = (long long)blk - 1 + k;
      double v = 0.0;
      if (j >= 0 && j < N) {
        if (xo) v = xo[j * ND + ee];
        if (dxs) v += (double)dxs[j * ND + ee];
      }
      X[k][(ee / NB) * LDV + (ee % NB)] = v;
      if (k == 1) xn[(long long)blk * ND + ee] = v;
    }
    __syncthreads();
    acc_t acc;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = 16 * w + Mfma<double>::row(lk, v);
      acc[v] = (lr < nrhs) ? ds[(long long)blk * ND + lr * NB + row] : 0.0;
      dmax = nmax(dmax, fabs(acc[v]));
      xmax = nmax(xmax, fabs(X[1][lr * LDV + row]));
    }
    tile_mma<double, false, false, true, NB>(acc, As + blk * NN + 16 * w, NB, X[1], LDV);
    if (blk > 0)
      tile_mma<double, true, false, true, NB>(acc, Bs + (blk - 1) * NN + 16 * w * NB, NB, X[0], LDV);
    if (blk + 1 < N)
      tile_mma<double, false, false, true, NB>(acc, Bs + blk * NN + 16 * w, NB, X[2], LDV);
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
  __syncthreads();
  if (tid < kParts) {
    const int q = tid / 16, c = tid % 16;
    double m = 0.0;
    for (int k = 0; k < 4 * TT; ++k) m = nmax(m, red[q][k][c]);
    part[((long long)sys * nwg + wg) * kParts + tid] = m;
  }
}

// combine partial maxima -> berr per (system, column); `init` records ||d||.
__global__ void berr_kernel(const double* __restrict__ part, const double* __restrict__ mnorm,
                            double* __restrict__ dnorm, double* __restrict__ berr, int nwg,
                            int nrhs, int init) {
  const int sys = blockIdx.x, c = threadIdx.x;
  if (c >= nrhs) return;
  double r = 0.0, x = 0.0, dd = 0.0;
  for (int g = 0; g < nwg; ++g) {
    const double* pp = part + ((long long)sys * nwg + g) * kParts;
    r = nmax(r, pp[c]);
    x = nmax(x, pp[16 + c]);
    dd = nmax(dd, pp[32 + c]);
  }
  if (init) dnorm[sys * 16 + c] = dd;
  const double den = mnorm[sys] * x + dnorm[sys * 16 + c];
  berr[sys * nrhs + c] = den > 0.0 ? r / den : r;
}

// ||M||_inf: max over block rows of sum_c |A_i| + |B_{i-1}^T| + |B_i|, per system.
template <int NB>
__global__ void __launch_bounds__(4 * NB) mnorm_kernel(const double* __restrict__ A,
                                                       const double* __restrict__ B,
                                                       double* __restrict__ part, int N,
                                                       int bpw, int nwg) {
  constexpr long long NN = (long long)NB * NB;
  __shared__ double rs[4][NB];
  const int sys = blockIdx.y, wg = blockIdx.x, tid = threadIdx.x;
  const int r = tid % NB, q = tid / NB;
  const double* As = A + sys * (long long)N * NN;
  const double* Bs = B + sys * (long long)(N > 0 ? N - 1 : 0) * NN;
  double best = 0.0;
  const int b0 = wg * bpw, b1 = min(N, b0 + bpw);
  for (int blk = b0; blk < b1; ++blk) {
    double s = 0.0;
    for (int c = q; c < NB; c += 4) {
      s += fabs(As[blk * NN + c * NB + r]);
      if (blk > 0) s += fabs(Bs[(blk - 1) * NN + r * NB + c]);
      if (blk + 1 < N) s += fabs(Bs[blk * NN + c * NB + r]);
    }
    __syncthreads();
    rs[q][r] = s;
    __syncthreads();
    if (q == 0) best = nmax(best, rs[0][r] + rs[1][r] + rs[2][r] + rs[3][r]);
  }
  __syncthreads();
  if (q == 0) rs[0][r] = best;
  __syncthreads();
  if (tid == 0) {
    double m = 0.0;
    for (int k = 0; k < NB; ++k) m = nmax(m, rs[0][k]);
    part[((long long)sys * nwg + wg) * kParts] = m;
  }
}

__global__ void mnorm_reduce_kernel(const double* __restrict__ part, double* __restrict__ mnorm,
                                    int nwg) {
  const int sys = blockIdx.x;
  if (threadIdx.x != 0) return;
  double m = 0.0;
  for (int g = 0; g < nwg; ++g) m = nmax(m, part[((long long)sys * nwg + g) * kParts]);
  mnorm[sys] = m;
}

void geometry(int N, int* nwg, int* bpw) {
  *nwg = std::max(1, std::min(kMaxWG, N));
  *bpw = (N + *nwg - 1) / *nwg;
  *nwg = (N + *bpw - 1) / *bpw;
}

template <int NB>
hipError_t launch_residual(const MixedState& m, const double* A, const double* B,
                           const double* d, const double* xold, const float* dx, double* xnew,
                           float* rout, hipStream_t s) {
  hipLaunchKernelGGL((residual_kernel<NB>), dim3(m.nwg, m.nsys), dim3(4 * NB), 0, s, A, B, d,
                     xold, dx, xnew, rout, m.part, m.N, m.nrhs, m.bpw, m.nwg);
  return hipGetLastError();
}

hipError_t residual(const MixedState& m, const double* A, const double* B, const double* d,
                    const double* xold, const float* dx, double* xnew, float* rout, int init,
                    hipStream_t s) {
  hipError_t e;
  switch (m.n) {
    case 16: e = launch_residual<16>(m, A, B, d, xold, dx, xnew, rout, s); break;
    case 32: e = launch_residual<32>(m, A, B, d, xold, dx, xnew, rout, s); break;
    case 48: e = launch_residual<48>(m, A, B, d, xold, dx, xnew, rout, s); break;
    case 64: e = launch_residual<64>(m, A, B, d, xold, dx, xnew, rout, s); break;
    default: return hipErrorInvalidValue;
  }
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(berr_kernel, dim3(m.nsys), dim3(64), 0, s, m.part, m.mnorm, m.dnorm, m.berr,
                     m.nwg, m.nrhs, init);
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
  switch (n) {
#define BSS_MN(NB)                                                                          \
  case NB:                                                                                  \
    hipLaunchKernelGGL((mnorm_kernel<NB>), dim3(m.nwg, nsys), dim3(4 * NB), 0, s, A, B,      \
                       m.part, m.N, m.bpw, m.nwg);                                          \
    break;
    BSS_MN(16) BSS_MN(32) BSS_MN(48) BSS_MN(64)
#undef BSS_MN
    default: g_mixed_err = "unsupported n"; return -1;
  }
  hipLaunchKernelGGL(mnorm_reduce_kernel, dim3(nsys), dim3(64), 0, s, m.part, m.mnorm, m.nwg);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    g_mixed_err = hipGetErrorString(e);
    return -1;
  }
  return 0;
}

namespace {
constexpr int kKindResidual = 6;  // BSS_K_RESIDUAL in bss.h

hipError_t residual_prof(const MixedState& m, const MixedSolveHooks& h, const double* A,
                         const double* B, const double* d, const double* xold,
                         const float* dx, double* xnew, float* rout, int init, hipStream_t s) {
  void* scope = h.prof_begin ? h.prof_begin(h.ctx, kKindResidual, s) : nullptr;
  hipError_t e = residual(m, A, B, d, xold, dx, xnew, rout, init, s);
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
  hipError_t e = residual_prof(m, h, A, B, d, nullptr, nullptr, m.x[0], v[0], 1, s);
  if (e != hipSuccess) {
    g_mixed_err = hipGetErrorString(e);
    return -1;
  }
  while (true) {
    if (h.solve32(h.ctx, v[cur], s) != 0) {
      g_mixed_err = "FP32 solve failed";
      return -1;
    }
    e = residual_prof(m, h, A, B, d, m.x[xa], v[cur], m.x[1 - xa], v[1 - cur], 0, s);
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
