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

// one system's level-0 data for the residual pass
struct ResSys {
  const double* A;
  const double* B;
  const double* d;
  double* x[2];  // FP64 iterates (ping-pong)
  float* v[2];   // FP32 right-hand side / correction vectors (ping-pong)
  int N;
  int item0, item1;  // residual work items of this system
  int pad;
};

struct ResItem {
  int sys, b0, b1, pad;
};

struct MixedState {
  int n = 0, nrhs = 0, nsys = 0, nitems = 0;
  std::vector<int> Ns;
  std::vector<ResSys> sys_host;
  ResSys* sys_dev = nullptr;
  ResItem* items = nullptr;
  double* part = nullptr;   // per work item partial maxima
  double* mnorm = nullptr;  // ||M||_inf per system
  double* dnorm = nullptr;  // ||d_j||_inf per system and column
  double* berr = nullptr;   // backward error per system and column
  double* host_berr = nullptr;  // pinned mirror of berr
  hipEvent_t berr_ready = nullptr;  // recorded after each copy of berr to host_berr
  bool mnorm_stale = true;      // ||M|| must be recomputed (new factorisation)
  // the next x = 0 pass only: when set and *init_skip == 0 it does nothing (the same
  // pass already ran on the same d); mixed_first_pass clears it
  const int* init_skip = nullptr;
  // one right-hand side with a fused level-0 sweep (factor_fuses_forward): a copy of d
  // as the factorisation found it, and a flag set when d differs from it (a plan's
  // BSS_OPT_FACTOR_SWEEP); null otherwise
  double* dsnap = nullptr;
  int* dflag = nullptr;
  std::vector<double> last_berr;

  // device bytes for systems of sizes Ns (N_use per system)
  static size_t bytes(const std::vector<int>& Ns, int n, int nrhs);
  // carve the workspace at `base`; builds and uploads the work-item table (0 = ok)
  int attach(char* base, const std::vector<int>& Ns, int n, int nrhs);
  // (re)bind the caller's FP64 A, B, d (packed systems); uploads the system table
  int bind(const double* A, const double* B, const double* d);
  float* v_of(int sys, int k) const { return sys_host[sys].v[k]; }
  void release();
  ~MixedState() { release(); }
};

// Wait for `ev` by polling it for up to about two milliseconds, then by blocking.  The
// refinement loop waits for a few bytes once per iteration, with the device idle until
// the host has queued the next iteration: a blocking wait adds its wake-up time to
// that idle gap.
hipError_t wait_event_spin(hipEvent_t ev);

struct MixedSolveHooks {
  void* ctx = nullptr;
  // FP32 recursive solve in place in the level-0 vectors of ping-pong set k
  int (*solve32)(void* ctx, int k, hipStream_t s) = nullptr;
  // optional, called once right after the first solve32 is enqueued, so that the first
  // solve is queued behind earlier work while the caller waits on it: returns 0 when
  // that solve stands, 1 when the caller has changed the solve path (the x = 0 pass and
  // the first solve are then redone from d), 2 when refinement cannot succeed (an FP32
  // pivot failed: the loop stops, unconverged, without writing d), anything else on
  // failure
  int (*prepare)(void* ctx, hipStream_t s) = nullptr;
  // optional profiler brackets around the residual and backward-error launches
  void* (*prof_begin)(void* ctx, int kind, hipStream_t s) = nullptr;
  void (*prof_end)(void* scope) = nullptr;
};

// Marks ||M||_inf stale after a factorisation; the next solve's first residual pass
// recomputes it from the operands it reads anyway.
void mixed_after_factor(MixedState& m);

// Refinement loop: x_0 = 0; repeat { dx = M32^{-1} r ; x += dx ; r = d - M x } until
// max berr <= tol (*converged = true), or it stops short: a NaN backward error, a
// backward error that did not halve since the previous check (stagnation: the FP32
// factors cannot refine this system further), or max_iters solves.  d receives the
// last iterate when converged, or when `write_unconverged` is set.  Synchronises the
// stream once per checked iteration.
int mixed_solve(MixedState& m, const MixedSolveHooks& h, double* d, double tol, int max_iters,
                int* iters, bool* converged, bool write_unconverged, hipStream_t s,
                bool first_pass_done = false);

// Backward error of the FP64 solution held in iterate set 0 (the FP64 fallback's
// answer): one residual pass with a zero correction, which copies it into set 1 and
// measures berr per system and column into last_berr.  Returns the largest (NaN if
// any is NaN) through *worst.  The answer to hand out is then in set 1.
int mixed_measure(MixedState& m, const MixedSolveHooks& h, hipStream_t s, double* worst);

// Device address of iterate set k (FP64, packed like the caller's d).
inline double* mixed_iterate(const MixedState& m, int k) { return m.sys_host[0].x[k]; }

// The refinement loop's first (x = 0) residual pass on its own: r = d as FP32 into the
// level-0 vectors of set 0 and ||d||.  It reads only the caller's FP64 d, A and B, so it
// may run beside the factorisation; mixed_solve(..., first_pass_done = true) then
// starts at the first FP32 solve.
int mixed_first_pass(MixedState& m, const MixedSolveHooks& h, hipStream_t s);

// dsnap = the caller's packed d and *dflag = 0, in one kernel that reads d through the
// device system table (so a captured graph stays valid across rebinding).  Needs dsnap.
int mixed_snapshot_d(MixedState& m, hipStream_t s);

const char* mixed_last_error();

}  // namespace bss
