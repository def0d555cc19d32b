// mixed.hpp -- FP64 iterative refinement around the FP32 factorisation.
//
// The factorisation (potrf, trsm and Schur updates) runs in FP32 on MFMA; FP64 is used
// only for the residual r = d - M x over the block-tridiagonal operator (read from the
// caller's FP64 A and B), the solution update x += dx and the norms of the normwise
// backward error  berr_j = ||r_j||_inf / (||M||_inf ||x_j||_inf + ||d_j||_inf).
// The reference has no mixed-precision path; its FP32 and FP64 solves are separate
// instantiations (src/backbone_rocm.jl:9-12, 57-60).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

namespace bss {

struct MixedState {
  int N = 0, n = 0, nrhs = 0, nsys = 0, nwg = 0, bpw = 0;
  double* x[2] = {nullptr, nullptr};  // FP64 iterates (ping-pong), N*n*nrhs per system
  float* corr = nullptr;              // second FP32 vector (ping-pong with the level-0 d)
  double* part = nullptr;             // per-workgroup partial maxima [sys][wg][3][16]
  double* mnorm = nullptr;            // ||M||_inf per system
  double* dnorm = nullptr;            // ||d_j||_inf per system and column
  double* berr = nullptr;             // backward error per system and column
  double* host_berr = nullptr;        // pinned mirror of berr
  bool mnorm_stale = true;            // ||M|| must be recomputed (new factorisation)
  std::vector<double> last_berr;

  static size_t bytes(int N, int n, int nrhs, int nsys);
  void attach(char* base, int N, int n, int nrhs, int nsys);
  void release();
  ~MixedState() { release(); }
};

struct MixedSolveHooks {
  void* ctx = nullptr;
  int (*solve32)(void* ctx, float* d0, hipStream_t s) = nullptr;
  // optional profiler brackets around the residual launches (kind = BSS_K_RESIDUAL)
  void* (*prof_begin)(void* ctx, int kind, hipStream_t s) = nullptr;
  void (*prof_end)(void* scope) = nullptr;
};

// Marks ||M||_inf stale after a factorisation; the next solve's first residual pass
// recomputes it from the operands it reads anyway.
int mixed_after_factor(MixedState& m, const double* A, const double* B, int n, int nsys,
                       hipStream_t s);

// Refinement loop: x_0 = 0; repeat { dx = M32^{-1} r ; x += dx ; r = d - M x } until
// max berr <= tol or max_iters solves.  On exit d holds x.  Synchronises the stream
// once per iteration to test convergence.
int mixed_solve(MixedState& m, const MixedSolveHooks& h, const double* A, const double* B,
                double* d, float* v0, int n, int nrhs, int nsys, double tol, int max_iters,
                int* iters, hipStream_t s);

const char* mixed_last_error();

}  // namespace bss
