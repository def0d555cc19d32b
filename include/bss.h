/*
 * bss.h -- C ABI of libbss, the MI355X (gfx950) recursive Schur-complement
 * block-tridiagonal Cholesky solver.
 *
 * The solver replaces the reference's batched GPU path (src/gpu_batched.jl) and its
 * ROCm backbone (src/backbone_rocm.jl).  Host code (Julia through ccall, Python
 * through ctypes, C/C++ directly) binds these entry points; all arguments are plain
 * integers, doubles and device pointers, and no rocBLAS/rocSOLVER is involved.
 *
 * System:   M = tridiag(B_{i-1}^T, A_i, B_i), i = 1..N, blocks n x n, SPD.
 * Layout:   the reference tensors (src/gpu_seq.jl:29-43 create_matrix_list):
 *           every block column-major with leading dimension n, blocks contiguous,
 *           A : n*n*N, B : n*n*(N-1), d : n*nrhs*N elements per system; a batch of
 *           `nsys` same-shape systems stores system s at offset s*(that size).
 * Precision: dtype BSS_F64 / BSS_F32 for plain solves in that precision; mode
 *           BSS_MODE_MIXED (dtype must be BSS_F64) factors in FP32 on MFMA and
 *           refines in FP64 to a normwise backward error target.
 * Status:   every entry point returns 0 on success, a negative BSS_E* code on an
 *           argument or HIP error (message via bss_last_error), and never throws.
 *           Numerical failure (a non-positive pivot) is not an error status: it is
 *           reported per system by bss_info, like LAPACK's info.  A mixed-mode solve
 *           that ends above its backward-error target returns the positive
 *           BSS_NOT_CONVERGED instead of BSS_OK.
 * Threads:  calls on one plan must be serialised by the caller; different plans
 *           may be used from different threads.  Work is enqueued on the stream
 *           passed in (NULL = the default stream) and the calls return before it
 *           completes, as the reference's calls do (test/test_rocm.jl:97,101).
 */
#ifndef BSS_H_
#define BSS_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bss_plan* bss_plan_t;
typedef struct bss_batch* bss_batch_t;

enum bss_dtype { BSS_F32 = 0, BSS_F64 = 1 };
enum bss_mode { BSS_MODE_PLAIN = 0, BSS_MODE_MIXED = 1 };
/* BSS_SCHED_REFERENCE reproduces the reference's level schedule exactly
 * (find_sequence_upper, src/gpu_batched.jl:112-162): separators at both ends,
 * chains of m_i >= 3 blocks, and N padded up to N_use with identity blocks, which
 * the caller allocates (bss_plan_get(BSS_Q_N_USE)).  BSS_SCHED_AUTO uses this
 * library's schedule: no padding, chain lengths chosen per level (tuned table or
 * explicit bss_set_option(BSS_OPT_CHAIN_LEN*)). */
enum bss_schedule { BSS_SCHED_AUTO = 0, BSS_SCHED_REFERENCE = 1 };

enum bss_status {
  BSS_OK = 0,
  BSS_E_ARG = -1,      /* invalid argument */
  BSS_E_HIP = -2,      /* HIP runtime error */
  BSS_E_NOMEM = -3,    /* device allocation failed */
  BSS_E_STATE = -4,    /* call out of order (e.g. solve before factorize) */
  BSS_E_UNSUPPORTED = -5,
  /* positive: numerical outcome, not an error.  Mixed mode: the last solve did not
   * reach the backward-error target, neither by FP32-factor refinement nor by the FP64
   * fallback (or with the fallback off); d holds the last iterate and
   * bss_last_backward_error its backward error (NaN for a non-SPD system) */
  BSS_NOT_CONVERGED = 1
};

/* Plan queries (bss_plan_get). */
enum bss_query {
  BSS_Q_N_USE = 0,          /* blocks per system the caller's A/B/d must hold (system 0
                               of a group plan) */
  BSS_Q_LEVELS = 1,         /* recursion levels incl. the single-chain base */
  BSS_Q_WORKSPACE_BYTES = 2,
  BSS_Q_LEVEL_N = 16,       /* + level: blocks at that level (system 0) */
  BSS_Q_LEVEL_CHAINS = 48,  /* + level: chains of all systems (one workgroup each) */
  BSS_Q_LEVEL_MAXLEN = 80,  /* + level: longest chain */
  BSS_Q_LAST_ITERS = 112,   /* refinement iterations used by the last mixed solve */
  BSS_Q_REFACTORED = 113,   /* separator levels re-factored in FP64 (mixed mode) */
  BSS_Q_FALLBACK = 114      /* mixed mode: 1 when the solves since the last factorize
                               run through the plain FP64 fallback (see BSS_OPT_FALLBACK) */
};

/* Options (bss_set_option).  Schedule and storage options (CHAIN_LEN*, BASE_MAX,
 * PRESERVE_INPUT) must be set before the first query, bind or factorize. */
enum bss_option {
  BSS_OPT_TOL = 0,          /* mixed: target normwise backward error (default 1e-12) */
  BSS_OPT_MAX_ITERS = 1,    /* mixed: max refinement iterations (default 10) */
  BSS_OPT_PIVOT_TOL = 2,    /* mixed: threshold on U_jj^2 / A_jj of the FP32 factorisation
                               of the separator levels (default 1e-5).  When a system
                               falls below it, or an FP32 separator pivot fails, the
                               separator levels are re-factored in FP64 at the next
                               bss_solve or bss_info (which synchronise) */
  BSS_OPT_CHAIN_LEN = 3,    /* auto schedule: chain length for every level */
  BSS_OPT_BASE_MAX = 4,     /* auto schedule: largest single-chain base level */
  BSS_OPT_PRESERVE_INPUT = 5, /* plain mode: 1 (default) = keep A and B unchanged, the
                               factors go to the plan's workspace (twice the A + B
                               bytes of workspace); 0 = overwrite A and B with the
                               factors, as the reference does.  With 0, bss_info still
                               flags every system that failed, but the index it gives
                               is a failing block of the solver's own elimination order,
                               not necessarily the smallest one LAPACK order gives (the
                               input the sequential re-check needs is gone) */
  BSS_OPT_FACTOR_WAVES = 6,  /* waves per factor workgroup at n = 64: 4 (default) or 8 */
  BSS_OPT_GRAPH = 7,         /* 1 = capture each launch sequence (factorize; solve; in
                               mixed mode each FP32 correction solve) into a hipGraph on
                               first use and replay it afterwards (default 0).  Ignored
                               while the profiler is on. */
  BSS_OPT_PROFILE_KINDS = 8, /* profiler: bit k set = bracket launches of kernel kind k
                               (enum bss_kernel_kind; default all kinds).  Every
                               bracketed launch costs an event pair on the stream */
  BSS_OPT_PROFILE_MAX_LEVEL = 9, /* profiler: bracket launches at levels <= value only
                               (default: every level; the residual and backward-error
                               kernels count as level 0) */
  BSS_OPT_FALLBACK = 10,     /* mixed: 1 (default) = when refinement stops short of the
                               target (a NaN backward error, a backward error that no
                               longer halves per iteration, max_iters reached, or an FP32
                               level-0 pivot failure), factor all systems once in plain
                               FP64 (input preserved, workspace allocated on first use)
                               and solve this and every later right-hand side of the
                               factorisation with it, its backward error measured in
                               FP64; 0 = return BSS_NOT_CONVERGED with the last iterate */
  BSS_OPT_FACTOR_SWEEP = 11, /* mixed, one right-hand side, n <= 96: 1 (default) =
                               bss_factorize (and bss_batch_factorize) also runs the first
                               solve's level-0 forward sweep on the d it finds (inside the
                               level-0 factor kernel, as bss_factorize_solve does) and keeps
                               a device copy of d in the plan's workspace (n N_use 8 bytes
                               per system).  The next bss_solve compares d with that copy
                               on the device and skips its x = 0 pass and level-0 forward
                               sweep if every bit is unchanged, else runs them; the answer is
                               bitwise that of option value 0 either way.  No counterpart in
                               the reference (its factorize! never reads d). */
  BSS_OPT_CHAIN_LEN_LEVEL = 32 /* + level k (< 16): chain length of level k only
                                  (auto schedule); overrides BSS_OPT_CHAIN_LEN there */
};

/* Kernel kinds for the profiler (bss_profile_read). */
enum bss_kernel_kind {
  BSS_K_FACTOR = 0,        /* fused chain factorisation, one launch per level */
  BSS_K_ASSEMBLE_DIAG = 1, /* reduced diagonal assembly */
  BSS_K_FORWARD = 2,       /* forward substitution sweep */
  BSS_K_ASSEMBLE_RHS = 3,  /* reduced right-hand side assembly */
  BSS_K_BACKWARD = 4,      /* backward substitution sweep */
  BSS_K_SCATTER = 5,       /* separator solution scatter */
  BSS_K_RESIDUAL = 6,      /* FP64 solution update + residual (mixed mode) */
  BSS_K_RESIDUAL_INIT = 7, /* first refinement pass: x = 0, r = d (mixed mode) */
  BSS_K_BERR = 8,          /* per-system backward-error reduction (mixed mode) */
  BSS_K_COUNT = 9
};

/* Create a plan for `nsys` independent systems of N blocks of size n with `nrhs`
 * right-hand-side columns.  Replaces initialize_batched (src/gpu_batched.jl:165-182):
 * the schedule search and the per-level workspace allocation happen here (device
 * memory is allocated on the current HIP device); the caller owns A, B, d.
 * Supported n: 16, 32, 48, 64, 96, 128 (FP32 and FP64), 1 <= nrhs <= 16. */
int bss_plan_create(int N, int n, int nrhs, int nsys, int dtype, int mode, int schedule,
                    bss_plan_t* out);

/* Create a plan for `nsys` independent systems of different lengths: system s has
 * Ns[s] blocks of size n (an uneven batch).  Every system gets its own level
 * schedule, the same one it would get alone (so its arithmetic, and its result bit
 * for bit, do not depend on its co-residents), and each recursion level runs all
 * systems that have it in one launch.  The caller's A, B and d hold the systems
 * packed one after the other: A = sum_s N_use[s] blocks, B = sum_s (N_use[s]-1)
 * blocks, d = sum_s N_use[s] vector blocks (N_use from bss_plan_system_sizes).
 * Generalises initialize_batched (src/gpu_batched.jl:165-182), whose batch is
 * uniform (one N for all systems). */
int bss_plan_create_group(int nsys, const int* Ns, int n, int nrhs, int dtype, int mode,
                          int schedule, bss_plan_t* out);

/* Blocks each system occupies in the packed buffers (N_use[s]; equals Ns[s] except
 * under BSS_SCHED_REFERENCE, which pads).  Writes nsys values when n_use is not NULL
 * and returns nsys (or a negative status).  No counterpart in the reference (its
 * batch is uniform). */
int bss_plan_system_sizes(bss_plan_t plan, int* n_use);

/* Set a numeric option (see enum bss_option). */
int bss_set_option(bss_plan_t plan, int option, double value);

/* Query plan properties (see enum bss_query). */
int bss_plan_get(bss_plan_t plan, int query, int64_t* value);

/* Give the plan caller-owned device memory for its workspace instead of the
 * hipMalloc that bss_plan_bind would make: at least BSS_Q_WORKSPACE_BYTES bytes,
 * 256-byte aligned, valid until bss_plan_destroy (which then frees nothing), with no
 * device work pending on it when bss_plan_bind initialises it.  Call before
 * bss_plan_bind.  Lets a caller that recycles memory (a request loop on a
 * caching allocator) avoid a device allocation and free per plan.  No counterpart in
 * the reference (create_data allocates per level, src/gpu_batched.jl:73-110). */
int bss_plan_set_workspace(bss_plan_t plan, void* ws, size_t bytes);

/* Bind the caller's device buffers (reference tensors A_tensor, B_tensor, d_tensor
 * of BlockTriDiagData_batched, src/gpu_batched.jl:14-27).  A and B are read by
 * bss_factorize and left unchanged (the mixed mode's FP64 residual and bss_info's
 * sequential re-check read them again); in plain mode with BSS_OPT_PRESERVE_INPUT = 0
 * they are overwritten with the factors, as the reference overwrites them.  d holds
 * the right-hand sides and receives the solution. */
int bss_plan_bind(bss_plan_t plan, void* A, void* B, void* d);

/* Factorize all systems.  Replaces factorize!(::BlockTriDiagData_batched)
 * (src/gpu_batched.jl:184-227) and, through it, cholesky_factorize_batched!,
 * cholesky_solve_batched! (2n RHS) and compute_schur_complement!
 * (src/backbone_rocm.jl:62-138).  `stream` is a hipStream_t (NULL = default). */
int bss_factorize(bss_plan_t plan, void* stream);

/* Solve in place in d with the current factors; may be repeated with new d.
 * Replaces solve!(::BlockTriDiagData_batched) (src/gpu_batched.jl:229-259) and
 * compute_schur_RHS!, update_boundary!, cholesky_solve_batched! with nd = 1
 * (src/backbone_rocm.jl:84-170).  In mixed mode this runs FP32 solves inside FP64
 * iterative refinement and synchronises the stream once per iteration; it returns
 * BSS_OK when every system and column reached the target (by refinement, or by the
 * FP64 fallback of BSS_OPT_FALLBACK), else BSS_NOT_CONVERGED. */
int bss_solve(bss_plan_t plan, void* stream);

/* bss_factorize followed by bss_solve, with the same results and the same state
 * afterwards (later bss_solve calls reuse the factors), for a caller whose d is
 * final before the factorisation starts.  Levels 0 and 1 are factored on `stream`;
 * then the first solve's forward sweeps (and, in mixed mode, its x = 0 residual pass)
 * run on a plan-owned stream: those of levels 0 and 1 beside the factorisation of
 * level 2, and each later level's as soon as that level is factored.  The side stream
 * is joined back into `stream` before the backward sweeps.  Returns what bss_solve
 * returns.  Replaces the reference's back-to-back
 * factorize!(data); solve!(data) (src/gpu_batched.jl:184-259, as called in
 * test/test_rocm.jl:97-101); the reference has no combined call.  With BSS_OPT_GRAPH on,
 * or a single level, it is exactly bss_factorize then bss_solve. */
int bss_factorize_solve(bss_plan_t plan, void* stream);

/* Factorize and solve one system in plain mode (dtype BSS_F64 or BSS_F32; A and B left
 * unchanged, solution in place in d) without allocating: no hipMalloc/hipFree and no
 * plan.  The same kernels, schedule (auto, chain length 3, base 8, 4 factor waves) and
 * results, bit for bit, as bss_factorize + bss_solve on a plan created with those
 * defaults.  Workspace and device tables come from a 64 MiB device array of the library
 * (one per GPU, held until the process ends); systems whose workspace exceeds it (FP64:
 * about 550 blocks at n = 64, 2,200 at n = 32; BSS_Q_WORKSPACE_BYTES of a plan tells)
 * get BSS_E_ARG.  Set-up copies are ordered on `stream` (from pageable host memory,
 * so the call waits for the stream's earlier work).  Every call on one GPU shares that
 * array: calls that overlap on one GPU -- on different streams or threads -- race.  No
 * per-system info (use a plan and bss_info for the failing block).  No counterpart in
 * the reference: for callers that cannot allocate device memory (tight serving loops,
 * code that forbids allocation). */
int bss_factorize_solve_noalloc(int N, int n, int nrhs, int dtype, const void* A,
                                const void* B, void* d, void* stream);

/* Build the block-tridiagonal normal equations of a linear state-space smoother on the
 * device: for the model x_{i+1} = F x_i + w (w ~ N(0, Q)), y_i = H x_i + v (v ~ N(0, R)),
 * A_i = Q^{-1} + F^T Q^{-1} F + H^T R^{-1} H (all N blocks) and B_i = -Q^{-1} F (N-1
 * blocks), written in the reference layout (column-major n x n blocks, contiguous)
 * into A and B, as dtype, on `stream`.  F (n x n), H (m x n), Qinv (n x n) and Rinv
 * (m x m) are column-major FP64 device arrays.  The reference builds these blocks on
 * the host (benchmark/benchmark_kf.jl:79-154, with A_1 and A_N special-cased); here a
 * request carries only the model and its blocks never cross PCIe. */
int bss_build_smoother(int N, int n, int m, const double* F, const double* H,
                       const double* Qinv, const double* Rinv, void* A, void* B, int dtype,
                       void* stream);

/* Per-system numerical status after the last factorize (synchronises the stream):
 * info[s] = 0 if system s is positive definite, else the 1-based index of the block
 * at which the sequential block Cholesky of src/backbone_cpu.jl:5-15 (blocks 1, 2, ...
 * in order, LAPACK potrf on each Schur-updated diagonal block) meets its first
 * non-positive pivot -- the LAPACK-order answer, whatever the recursive elimination
 * order met first.  A system flagged by the recursive factorisation is re-factored
 * once, as a single chain in that order, from the preserved input (in FP64 for mixed
 * plans, whose info therefore describes the FP64 matrix); see BSS_OPT_PRESERVE_INPUT
 * for the one case without preserved input.  The reference passes an `info` buffer to
 * potrf and never reads it (src/backbone_rocm.jl:64-80); this exposes it. */
int bss_info(bss_plan_t plan, int* info_host);

/* Per-system, per-column backward error of the last mixed-mode solve
 * (nsys*nrhs doubles, system-major), as measured on the device in FP64. */
int bss_last_backward_error(bss_plan_t plan, double* berr_host);

/* Wait for all work enqueued by this plan (hipStreamSynchronize on its stream). */
int bss_sync(bss_plan_t plan);

/* Free the plan and its workspace (not the caller's A, B, d). */
int bss_plan_destroy(bss_plan_t plan);

/* Copy the calling thread's last error message into buf (always NUL-terminated);
 * returns the full message length. */
int bss_last_error(char* buf, size_t len);

/* ---- Uneven batches across block sizes (no counterpart in the reference, whose batch
 * is uniform: initialize_batched, src/gpu_batched.jl:165-182) ----
 * `nsys` independent systems, system s with Ns[s] blocks of size ns[s] (any mix of the
 * supported n), one nrhs, dtype, mode and schedule.  The systems are grouped by n into
 * one group plan each (bss_plan_create_group: every system keeps the schedule it gets
 * alone, so its result is bitwise the one it gets alone in plain mode).  With
 * BSS_OPT_GRAPH on (the default for batches) the whole factorisation -- every level of
 * every group, the groups as parallel branches -- is one hipGraph launch per call, and
 * so is a plain-mode solve; a mixed-mode solve runs each group's refinement loop on its
 * own host thread and stream.  Options (bss_batch_set_option) apply to every group. */
int bss_batch_create(int nsys, const int* Ns, const int* ns, int nrhs, int dtype, int mode,
                     int schedule, bss_batch_t* out);
int bss_batch_set_option(bss_batch_t batch, int option, double value);
/* Layout of the caller's packed buffers: element offsets of each system's A, B and d
 * (nsys values each, NULL to skip) and totals[3] = elements of A, B, d.  Groups are
 * packed in ascending n, each group's systems in input order, each system N_use blocks
 * in the reference layout.  Returns nsys. */
int bss_batch_layout(bss_batch_t batch, int64_t* offA, int64_t* offB, int64_t* offd,
                     int64_t* totals);
/* bss_plan_get over the groups: the sum for BSS_Q_WORKSPACE_BYTES and BSS_Q_FALLBACK
 * (groups on the fallback), the largest value for the other queries. */
int bss_batch_get(bss_batch_t batch, int query, int64_t* value);
int bss_batch_bind(bss_batch_t batch, void* A, void* B, void* d);
int bss_batch_factorize(bss_batch_t batch, void* stream);
/* Returns BSS_OK, or BSS_NOT_CONVERGED when a mixed-mode group did not reach its target. */
int bss_batch_solve(bss_batch_t batch, void* stream);
/* Per-system info (input order), as bss_info. */
int bss_batch_info(bss_batch_t batch, int* info_host);
/* Mixed mode: per-system, per-column backward error (input order, system-major). */
int bss_batch_last_backward_error(bss_batch_t batch, double* berr_host);
int bss_batch_sync(bss_batch_t batch);
int bss_batch_destroy(bss_batch_t batch);

/* ---- Profiling (no counterpart in the reference, whose benchmarks time whole
 * factorize!/solve! calls with AMDGPU.@elapsed, benchmark/benchmark_gpu.jl:143-150) ----
 * While enabled, every kernel launch of the plan is bracketed by HIP events on the
 * launch stream.  Enabling clears previous records. */
int bss_profile_enable(bss_plan_t plan, int on);

/* Sum of kernel durations (ms) and launch count recorded since the last enable, for
 * one kernel kind (enum bss_kernel_kind) at one level (-1 = all levels).
 * Synchronises the plan's stream. */
int bss_profile_read(bss_plan_t plan, int kind, int level, int* count, double* total_ms);

/* ---- Schedule introspection (host only, no device work; used by tests) ---- */

/* The reference's find_sequence_upper (src/gpu_batched.jl:155-162): writes N_use and
 * up to `cap` (N_i, m_i, P_i) triples, top level first, into levels[3*k..3*k+2];
 * returns the number of levels or a negative status. */
int bss_reference_schedule(int N, int* n_use, int* levels, int cap);

/* Chain layout of a plan level: for chain c, chains[4c..4c+3] = {start block, length,
 * has-left-separator, has-right-separator}; returns the number of chains. */
int bss_plan_level_chains(bss_plan_t plan, int level, int* chains, int cap);

#ifdef __cplusplus
}
#endif

#endif /* BSS_H_ */
