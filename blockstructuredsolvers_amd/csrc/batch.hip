// batch.hip -- uneven batches of independent systems with different N and different
// block sizes n, run as one unit: bss_batch_*.
//
// The reference's batch is uniform (initialize_batched, src/gpu_batched.jl:165-182:
// one N, one n).  Here the systems are grouped by n into one group plan each (every
// system keeps the level schedule it gets alone, and each level of a group is one
// launch over all its systems); the groups run side by side on their own streams.  In
// graph mode (the default) the whole batch's factorisation -- every level of every
// group, the groups as parallel branches -- is captured once into one hipGraph and
// replayed with one launch per call, and so is the plain-mode solve.  Mixed-mode
// solves run each group's refinement loop on its own host thread and stream, so the
// groups' iterations and host checks overlap.
//
// One fused kernel across block sizes is not used on purpose: every n wants its own
// workgroup size, register budget and LDS footprint (n = 16 runs one wave and ~9 KB of
// LDS per chain, n = 64 FP64 eight waves and 135 KB), and a launch shaped for the
// largest n would cut the small-n occupancy several-fold.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bss.h"

namespace bss {
int api_fail(int code, const char* msg);
int plan_factor_prepare(bss_plan* p, void* stream);
int plan_factor_kernels(bss_plan* p, hipStream_t c);
int plan_factor_finish(bss_plan* p, void* stream);
int plan_solve_kernels(bss_plan* p, hipStream_t c);
}  // namespace bss

struct bss_batch {
  int nsys = 0, nrhs = 0, dtype = 0, mode = 0, schedule = 0;
  std::vector<int> Ns, ns;
  std::vector<int> sizes;                 // distinct block sizes, ascending (one group each)
  std::vector<std::vector<int>> members;  // systems of each group, in input order
  std::vector<bss_plan_t> plans;          // one group plan per block size
  std::vector<int64_t> offA, offB, offd;  // per system: element offsets in the packed buffers
  std::vector<int64_t> gA, gB, gd;        // per group: offset of its first system
  int64_t totA = 0, totB = 0, totd = 0;
  std::vector<hipStream_t> streams;       // one per group
  std::vector<hipEvent_t> ev_join;        // group -> caller
  hipEvent_t ev_fork = nullptr;           // caller -> groups
  hipStream_t cap = nullptr;              // graph capture happens here
  hipGraphExec_t g_factor = nullptr, g_solve = nullptr;
  bool use_graph = true, bound = false, factored = false;
  hipStream_t last_stream = nullptr;
  void drop_graphs() {
    for (hipGraphExec_t* g : {&g_factor, &g_solve}) {
      if (*g) (void)hipGraphExecDestroy(*g);
      *g = nullptr;
    }
  }
  ~bss_batch() {
    drop_graphs();
    for (bss_plan_t p : plans) bss_plan_destroy(p);
    for (hipStream_t s : streams) (void)hipStreamDestroy(s);
    for (hipEvent_t e : ev_join) (void)hipEventDestroy(e);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (cap) (void)hipStreamDestroy(cap);
  }
};

namespace {

int fail(int code, const char* msg) { return bss::api_fail(code, msg); }

#define BATCH_HIP(expr)                                               \
  do {                                                                \
    const hipError_t e_ = (expr);                                     \
    if (e_ != hipSuccess) return fail(BSS_E_HIP, hipGetErrorString(e_)); \
  } while (0)

size_t elem_size(const bss_batch* b) { return b->dtype == BSS_F64 ? 8 : 4; }

int check(const bss_batch* b) { return b ? BSS_OK : fail(BSS_E_ARG, "null batch"); }

// Streams and events of the fork/join (created on first use, default priority).
int ensure_streams(bss_batch* b) {
  if (b->ev_fork) return BSS_OK;
  BATCH_HIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
  for (size_t g = 0; g < b->plans.size(); ++g) {
    hipStream_t s;
    hipEvent_t e;
    BATCH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    b->streams.push_back(s);
    BATCH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->ev_join.push_back(e);
  }
  BATCH_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
  return BSS_OK;
}

// Runs body(g, stream_g) for every group between a fork from `s` and a join back into
// `s`: group work is ordered after the caller's earlier work on `s`, and the caller's
// later work after every group.
template <typename F>
int fork_join(bss_batch* b, hipStream_t s, F body) {
  BATCH_HIP(hipEventRecord(b->ev_fork, s));
  int rc = BSS_OK;
  for (size_t g = 0; g < b->plans.size(); ++g) {
    BATCH_HIP(hipStreamWaitEvent(b->streams[g], b->ev_fork, 0));
    if (rc == BSS_OK) rc = body((int)g, b->streams[g]);
    // the join is recorded even after a failure, so `s` never runs ahead of group work
    BATCH_HIP(hipEventRecord(b->ev_join[g], b->streams[g]));
    BATCH_HIP(hipStreamWaitEvent(s, b->ev_join[g], 0));
  }
  return rc;
}

// body enqueued directly on `s` (graph off), or captured once from a fork/join on the
// capture stream into one graph and replayed on `s` with one launch.
template <typename F>
int run_batch_graph(bss_batch* b, hipGraphExec_t* ex, hipStream_t s, F body) {
  if (!b->use_graph) return fork_join(b, s, body);
  if (!*ex) {
    BATCH_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    const int rc = fork_join(b, b->cap, body);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(b->cap, &graph);
    if (rc != BSS_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (ec != hipSuccess) return fail(BSS_E_HIP, hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(ex, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      *ex = nullptr;
      return fail(BSS_E_HIP, hipGetErrorString(ei));
    }
  }
  BATCH_HIP(hipGraphLaunch(*ex, s));
  return BSS_OK;
}

}  // namespace

extern "C" {

int bss_batch_create(int nsys, const int* Ns, const int* ns, int nrhs, int dtype, int mode,
                     int schedule, bss_batch_t* out) {
  if (!out) return fail(BSS_E_ARG, "out is null");
  *out = nullptr;
  if (nsys < 1 || !Ns || !ns) return fail(BSS_E_ARG, "nsys must be >= 1 with Ns and ns given");
  bss_batch* b = new (std::nothrow) bss_batch();
  if (!b) return fail(BSS_E_NOMEM, "host allocation failed");
  b->nsys = nsys;
  b->nrhs = nrhs;
  b->dtype = dtype;
  b->mode = mode;
  b->schedule = schedule;
  b->Ns.assign(Ns, Ns + nsys);
  b->ns.assign(ns, ns + nsys);
  std::map<int, std::vector<int>> by_n;
  for (int s = 0; s < nsys; ++s) by_n[ns[s]].push_back(s);
  for (auto& kv : by_n) {
    std::vector<int> gN;
    for (int s : kv.second) gN.push_back(Ns[s]);
    bss_plan_t p = nullptr;
    const int rc = bss_plan_create_group((int)gN.size(), gN.data(), kv.first, nrhs, dtype, mode,
                                         schedule, &p);
    if (rc != BSS_OK) {
      delete b;
      return rc;
    }
    b->sizes.push_back(kv.first);
    b->members.push_back(kv.second);
    b->plans.push_back(p);
  }
  *out = b;
  return BSS_OK;
}

// Layout of the packed buffers (computed on first use, after every option): groups in
// ascending n, each group's systems in input order, every system N_use blocks.
static int batch_layout(bss_batch* b) {
  if (!b->offA.empty()) return BSS_OK;
  b->offA.assign(b->nsys, 0);
  b->offB.assign(b->nsys, 0);
  b->offd.assign(b->nsys, 0);
  int64_t a = 0, bb = 0, d = 0;
  for (size_t g = 0; g < b->plans.size(); ++g) {
    const int n = b->sizes[g];
    std::vector<int> nu(b->members[g].size());
    const int rc = bss_plan_system_sizes(b->plans[g], nu.data());
    if (rc < 0) return rc;
    b->gA.push_back(a);
    b->gB.push_back(bb);
    b->gd.push_back(d);
    for (size_t k = 0; k < nu.size(); ++k) {
      const int s = b->members[g][k];
      b->offA[s] = a;
      b->offB[s] = bb;
      b->offd[s] = d;
      a += (int64_t)nu[k] * n * n;
      bb += (int64_t)std::max(nu[k] - 1, 0) * n * n;
      d += (int64_t)nu[k] * n * b->nrhs;
    }
  }
  b->totA = a;
  b->totB = bb;
  b->totd = d;
  return BSS_OK;
}

int bss_batch_set_option(bss_batch_t b, int option, double value) {
  if (int e = check(b)) return e;
  if (option == BSS_OPT_GRAPH) {
    b->use_graph = value != 0.0;
    b->drop_graphs();
    return BSS_OK;
  }
  for (bss_plan_t p : b->plans)
    if (int e = bss_set_option(p, option, value)) return e;
  b->drop_graphs();
  return BSS_OK;
}

int bss_batch_layout(bss_batch_t b, int64_t* offA, int64_t* offB, int64_t* offd,
                     int64_t* totals) {
  if (int e = check(b)) return e;
  if (int e = batch_layout(b)) return e;
  for (int s = 0; s < b->nsys; ++s) {
    if (offA) offA[s] = b->offA[s];
    if (offB) offB[s] = b->offB[s];
    if (offd) offd[s] = b->offd[s];
  }
  if (totals) {
    totals[0] = b->totA;
    totals[1] = b->totB;
    totals[2] = b->totd;
  }
  return b->nsys;
}

int bss_batch_get(bss_batch_t b, int query, int64_t* value) {
  if (int e = check(b)) return e;
  if (!value) return fail(BSS_E_ARG, "value is null");
  int64_t acc = 0;
  for (bss_plan_t p : b->plans) {
    int64_t v = 0;
    if (int e = bss_plan_get(p, query, &v)) return e;
    if (query == BSS_Q_WORKSPACE_BYTES || query == BSS_Q_FALLBACK)
      acc += v;  // total workspace; number of groups on the fallback
    else
      acc = std::max(acc, v);  // levels, iterations, refactored levels: the largest
  }
  *value = acc;
  return BSS_OK;
}

int bss_batch_bind(bss_batch_t b, void* A, void* B, void* d) {
  if (int e = check(b)) return e;
  if (int e = batch_layout(b)) return e;
  if (!A || !d || (b->totB > 0 && !B)) return fail(BSS_E_ARG, "null buffer");
  const size_t es = elem_size(b);
  for (size_t g = 0; g < b->plans.size(); ++g) {
    void* Bg = B ? (void*)((char*)B + b->gB[g] * es) : nullptr;
    if (int e = bss_plan_bind(b->plans[g], (char*)A + b->gA[g] * es, Bg,
                              (char*)d + b->gd[g] * es))
      return e;
  }
  if (int e = ensure_streams(b)) return e;
  b->bound = true;
  b->factored = false;
  return BSS_OK;
}

int bss_batch_factorize(bss_batch_t b, void* stream) {
  if (int e = check(b)) return e;
  if (!b->bound) return fail(BSS_E_STATE, "bss_batch_bind must precede bss_batch_factorize");
  hipStream_t s = (hipStream_t)stream;
  b->last_stream = s;
  for (size_t g = 0; g < b->plans.size(); ++g)
    if (int e = bss::plan_factor_prepare(b->plans[g], b->streams[g])) return e;
  int rc = run_batch_graph(b, &b->g_factor, s, [b](int g, hipStream_t sg) {
    return bss::plan_factor_kernels(b->plans[g], sg);
  });
  // each group's bookkeeping (mixed: pivot copies) after the whole factorisation
  for (size_t g = 0; g < b->plans.size() && rc == BSS_OK; ++g)
    rc = bss::plan_factor_finish(b->plans[g], s);
  // later per-group calls (info, solves) order themselves on the groups' streams
  if (rc == BSS_OK) {
    BATCH_HIP(hipEventRecord(b->ev_fork, s));
    for (size_t g = 0; g < b->plans.size(); ++g)
      BATCH_HIP(hipStreamWaitEvent(b->streams[g], b->ev_fork, 0));
  }
  b->factored = rc == BSS_OK;
  return rc;
}

int bss_batch_solve(bss_batch_t b, void* stream) {
  if (int e = check(b)) return e;
  if (!b->factored) return fail(BSS_E_STATE, "bss_batch_factorize must precede bss_batch_solve");
  hipStream_t s = (hipStream_t)stream;
  b->last_stream = s;
  if (b->mode != BSS_MODE_MIXED)
    return run_batch_graph(b, &b->g_solve, s, [b](int g, hipStream_t sg) {
      return bss::plan_solve_kernels(b->plans[g], sg);
    });
  // mixed: each group's refinement loop on its own host thread and stream
  BATCH_HIP(hipEventRecord(b->ev_fork, s));
  for (size_t g = 0; g < b->plans.size(); ++g)
    BATCH_HIP(hipStreamWaitEvent(b->streams[g], b->ev_fork, 0));
  const size_t G = b->plans.size();
  std::vector<int> rcs(G, BSS_OK);
  std::vector<std::string> msgs(G);  // error messages are per thread
  auto run = [b, &rcs, &msgs](size_t g) {
    rcs[g] = bss_solve(b->plans[g], b->streams[g]);
    if (rcs[g] < 0) {
      char buf[512];
      bss_last_error(buf, sizeof(buf));
      msgs[g] = buf;
    }
  };
  std::vector<std::thread> th;
  for (size_t g = 1; g < G; ++g) th.emplace_back(run, g);
  run(0);
  for (auto& t : th) t.join();
  for (size_t g = 0; g < G; ++g) {
    BATCH_HIP(hipEventRecord(b->ev_join[g], b->streams[g]));
    BATCH_HIP(hipStreamWaitEvent(s, b->ev_join[g], 0));
  }
  int rc = BSS_OK;
  for (size_t g = 0; g < G; ++g) {
    if (rcs[g] < 0) return fail(rcs[g], msgs[g].c_str());  // an error wins
    rc = std::max(rc, rcs[g]);
  }
  return rc;
}

int bss_batch_info(bss_batch_t b, int* info_host) {
  if (int e = check(b)) return e;
  if (!info_host) return fail(BSS_E_ARG, "info_host is null");
  if (!b->factored) return fail(BSS_E_STATE, "no factorization yet");
  for (size_t g = 0; g < b->plans.size(); ++g) {
    std::vector<int> v(b->members[g].size());
    if (int e = bss_info(b->plans[g], v.data())) return e;
    for (size_t k = 0; k < v.size(); ++k) info_host[b->members[g][k]] = v[k];
  }
  return BSS_OK;
}

int bss_batch_last_backward_error(bss_batch_t b, double* berr_host) {
  if (int e = check(b)) return e;
  if (!berr_host) return fail(BSS_E_ARG, "berr_host is null");
  for (size_t g = 0; g < b->plans.size(); ++g) {
    std::vector<double> v(b->members[g].size() * b->nrhs);
    if (int e = bss_last_backward_error(b->plans[g], v.data())) return e;
    for (size_t k = 0; k < b->members[g].size(); ++k)
      for (int c = 0; c < b->nrhs; ++c)
        berr_host[(size_t)b->members[g][k] * b->nrhs + c] = v[k * b->nrhs + c];
  }
  return BSS_OK;
}

int bss_batch_sync(bss_batch_t b) {
  if (int e = check(b)) return e;
  BATCH_HIP(hipStreamSynchronize(b->last_stream));
  return BSS_OK;
}

int bss_batch_destroy(bss_batch_t b) {
  if (!b) return BSS_OK;
  if (b->bound) (void)hipStreamSynchronize(b->last_stream);
  delete b;
  return BSS_OK;
}

}  // extern "C"
