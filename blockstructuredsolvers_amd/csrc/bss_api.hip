// bss_api.hip -- C ABI of libbss: plans (groups of systems), per-level workspace and
// tables, the level loop.
//
// Replaces the reference's host side of the batched path: initialize_batched /
// create_data (src/gpu_batched.jl:73-182) become bss_plan_create (schedule + one
// workspace allocation), factorize!/solve! (src/gpu_batched.jl:184-259) become
// bss_factorize/bss_solve, which enqueue one launch per level and kernel kind on the
// caller's stream instead of ~31 library calls and pointer-array slicing per level.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/bss.h"
#include "launch.hpp"
#include "plan_tables.hpp"
#include "mixed.hpp"
#include "schedule.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(BSS_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),     \
                  __FILE__, __LINE__);                                                  \
  } while (0)


constexpr int kInfoNone = 0x7f7f7f7f;  // memset byte 0x7f: "no failure" sentinel
constexpr size_t kNone = ~size_t(0);   // buffer not allocated

// one system at one level: its chain layout and where its buffers live
struct SysLevel {
  bss::LevelShape shape;
  int P = 0;
  size_t oA = kNone, oB = kNone, od = kNone, oW = kNone, oSR = kNone, oSL = kNone,
         orR = kNone, orL = kNone, oU = kNone, oC = kNone, oOrig = kNone;
};

struct SysPlan {
  int N = 0, N_use = 0;
  std::vector<SysLevel> lv;
};

// one global level: the chains and separators of every system that has this level
struct GLevel {
  std::vector<bss::ChainRec> chains;
  std::vector<bss::SepRec> seps;
  std::vector<int> slot_sys;  // slot -> system
  bss::ChainRec* dchains = nullptr;
  bss::SepRec* dseps = nullptr;
  bss::LevelSys* dtab[2] = {nullptr, nullptr};  // [1] differs at level 0 only (d vector)
  int maxlen = 0;
};

}  // namespace

struct bss_plan {
  int n = 0, nrhs = 0, nsys = 0, dtype = 0, mode = 0, schedule = 0;
  std::vector<int> Ns;  // user sizes
  int device = 0;
  std::vector<SysPlan> sp;
  std::vector<GLevel> gl;
  char* ws = nullptr;
  size_t ws_bytes = 0;
  void *A = nullptr, *B = nullptr, *d = nullptr;  // caller buffers (packed systems)
  int* dinfo = nullptr;
  bool factored = false;
  bool laid_out = false;
  bool built = false;
  double tol = 1e-12, pivot_tol = 1e-5;
  int max_iters = 10;
  std::vector<int> chain_len = std::vector<int>(16, 0);  // per-level overrides
  int chain_len_default = 3;
  int base_max = 8;
  int factor_waves = 4;
  bool preserve = false;  // plain mode: factors to workspace, A/B untouched
  // mixed-precision state
  bss::MixedState mx;
  size_t mixed_off = kNone;
  int last_iters = 0, refactored = 0;
  hipStream_t last_stream = nullptr;
  // profiler: event pairs around kernel launches while enabled
  struct ProfRec {
    int kind, level;
    hipEvent_t a, b;
  };
  bool prof = false;
  std::vector<ProfRec> recs;
  std::vector<hipEvent_t> pool;
  size_t pool_next = 0;
  // table offsets
  struct TabOff {
    size_t chains = kNone, seps = kNone, tab[2] = {kNone, kNone};
  };
  std::vector<TabOff> toff;
  size_t info_off = kNone;
  ~bss_plan() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
};

namespace {

size_t fsize(const bss_plan* p) {  // factor element size
  return (p->dtype == BSS_F64 && p->mode == BSS_MODE_PLAIN) ? 8 : 4;
}
size_t isize(const bss_plan* p) { return p->dtype == BSS_F64 ? 8 : 4; }  // caller's type

// factors of level 0 live in the caller's A/B (reference semantics) unless the plan
// is mixed or asked to preserve its input
bool inplace0(const bss_plan* p) { return p->mode == BSS_MODE_PLAIN && !p->preserve; }

hipEvent_t prof_event(bss_plan* p) {
  if (p->pool_next == p->pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    p->pool.push_back(e);
  }
  return p->pool[p->pool_next++];
}

// Brackets one kernel launch with events when profiling is on.
struct ProfScope {
  bss_plan* p;
  hipStream_t s;
  long idx = -1;
  ProfScope(bss_plan* p_, int kind, int level, hipStream_t s_) : p(p_), s(s_) {
    if (!p->prof) return;
    hipEvent_t a = prof_event(p), b = prof_event(p);
    if (!a || !b || hipEventRecord(a, s) != hipSuccess) return;
    p->recs.push_back({kind, level, a, b});
    idx = (long)p->recs.size() - 1;
  }
  ~ProfScope() {
    if (idx >= 0) (void)hipEventRecord(p->recs[idx].b, s);
  }
};

struct Bump {
  size_t off = 0;
  size_t take(size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~size_t(255);
    return o;
  }
};

std::vector<bss::LevelShape> schedule_of(const bss_plan* p, int N, int* n_use) {
  if (p->schedule == BSS_SCHED_REFERENCE) return bss::reference_schedule_levels(N, n_use);
  *n_use = N;
  return bss::auto_schedule(N, p->chain_len, p->chain_len_default, std::max(1, p->base_max));
}

// Schedules and workspace layout (host only).
int layout(bss_plan* p) {
  if (p->laid_out) return BSS_OK;
  const size_t nn = (size_t)p->n * p->n, nd = (size_t)p->n * p->nrhs;
  const size_t fs = fsize(p);
  const bool mixed = p->mode == BSS_MODE_MIXED;
  p->sp.assign(p->nsys, SysPlan());
  size_t L = 0;
  for (int s = 0; s < p->nsys; ++s) {
    SysPlan& S = p->sp[s];
    S.N = p->Ns[s];
    auto shapes = schedule_of(p, S.N, &S.N_use);
    S.lv.resize(shapes.size());
    for (size_t l = 0; l < shapes.size(); ++l) {
      S.lv[l].shape = shapes[l];
      S.lv[l].P = (int)shapes[l].seppos.size();
    }
    L = std::max(L, shapes.size());
  }
  Bump bump;
  for (int s = 0; s < p->nsys; ++s) {
    SysPlan& S = p->sp[s];
    for (size_t l = 0; l < S.lv.size(); ++l) {
      SysLevel& V = S.lv[l];
      const size_t NL = (size_t)V.shape.N;
      if (l > 0) {
        V.oA = bump.take(NL * nn * fs);
        V.oB = bump.take(std::max<size_t>(NL - 1, 1) * nn * fs);
        V.od = bump.take(NL * nd * fs);
      } else if (mixed || p->preserve) {
        V.oU = bump.take(NL * nn * fs);
        V.oC = bump.take(std::max<size_t>(NL - 1, 1) * nn * fs);
      }
      bool anyL = false;
      for (auto& c : V.shape.chains) anyL = anyL || c.sepL >= 0;
      if (anyL) V.oW = bump.take(NL * nn * fs);
      if (V.P > 0) {
        V.oSR = bump.take(V.P * nn * fs);
        V.oSL = bump.take(V.P * nn * fs);
        V.orR = bump.take(V.P * nd * fs);
        V.orL = bump.take(V.P * nd * fs);
      }
      V.oOrig = bump.take(sizeof(int) * std::max<size_t>(NL, 1));
    }
  }
  // global levels
  p->gl.assign(L, GLevel());
  for (size_t l = 0; l < L; ++l) {
    GLevel& G = p->gl[l];
    for (int s = 0; s < p->nsys; ++s) {
      SysPlan& S = p->sp[s];
      if (l >= S.lv.size()) continue;
      const int slot = (int)G.slot_sys.size();
      G.slot_sys.push_back(s);
      const SysLevel& V = S.lv[l];
      for (auto& c : V.shape.chains) {
        G.chains.push_back({c.start, c.len, c.sepL, c.sepR, slot, 0, 0, 0});
        G.maxlen = std::max(G.maxlen, c.len);
      }
      for (int k = 0; k < V.P; ++k) G.seps.push_back({slot, k, V.shape.seppos[k], 0});
    }
  }
  p->toff.assign(L, bss_plan::TabOff());
  for (size_t l = 0; l < L; ++l) {
    GLevel& G = p->gl[l];
    p->toff[l].chains = bump.take(sizeof(bss::ChainRec) * std::max<size_t>(G.chains.size(), 1));
    p->toff[l].seps = bump.take(sizeof(bss::SepRec) * std::max<size_t>(G.seps.size(), 1));
    for (int k = 0; k < (l == 0 ? 2 : 1); ++k)
      p->toff[l].tab[k] = bump.take(sizeof(bss::LevelSys) * G.slot_sys.size());
  }
  p->info_off = bump.take(sizeof(int) * p->nsys);
  if (mixed) {
    std::vector<int> nu(p->nsys);
    for (int s = 0; s < p->nsys; ++s) nu[s] = p->sp[s].N_use;
    p->mixed_off = bump.take(bss::MixedState::bytes(nu, p->n, p->nrhs));
  }
  p->ws_bytes = bump.off;
  p->laid_out = true;
  return BSS_OK;
}

// Device pointers of one system at one level; `k` selects the level-0 vector set.
bss::LevelSys level_sys(const bss_plan* p, int s, size_t l, int k) {
  const SysPlan& S = p->sp[s];
  const SysLevel& V = S.lv[l];
  auto at = [&](size_t o) -> void* { return o != kNone ? (void*)(p->ws + o) : nullptr; };
  bss::LevelSys t{};
  const size_t nn = (size_t)p->n * p->n, nd = (size_t)p->n * p->nrhs;
  if (l == 0) {
    size_t oa = 0, ob = 0, od = 0;  // packed caller buffers
    for (int q = 0; q < s; ++q) {
      oa += (size_t)p->sp[q].N_use * nn;
      ob += (size_t)std::max(p->sp[q].N_use - 1, 0) * nn;
      od += (size_t)p->sp[q].N_use * nd;
    }
    const size_t is = isize(p);
    t.A = p->A ? (const char*)p->A + oa * is : nullptr;
    t.B = p->B ? (const char*)p->B + ob * is : nullptr;
    if (inplace0(p)) {
      t.U = (void*)t.A;
      t.C = (void*)t.B;
    } else {
      t.U = at(V.oU);
      t.C = at(V.oC);
    }
    if (p->mode == BSS_MODE_MIXED)
      t.d = p->mx.sys_host.empty() ? nullptr : (void*)p->mx.v_of(s, k);
    else
      t.d = p->d ? (char*)p->d + od * is : nullptr;
  } else {
    t.A = at(V.oA);
    t.B = at(V.oB);
    t.U = (void*)t.A;
    t.C = (void*)t.B;
    t.d = at(V.od);
  }
  t.W = at(V.oW);
  t.SR = at(V.oSR);
  t.SL = at(V.oSL);
  t.rR = at(V.orR);
  t.rL = at(V.orL);
  if (l + 1 < S.lv.size()) {
    const SysLevel& Nx = S.lv[l + 1];
    t.An = at(Nx.oA);
    t.Bn = at(Nx.oB);
    t.dn = at(Nx.od);
  }
  t.orig = (const int*)at(V.oOrig);
  t.sys = s;
  t.N = V.shape.N;
  return t;
}

int upload_tables(bss_plan* p, size_t l) {
  GLevel& G = p->gl[l];
  for (int k = 0; k < (l == 0 ? 2 : 1); ++k) {
    std::vector<bss::LevelSys> tab;
    for (int s : G.slot_sys) tab.push_back(level_sys(p, s, l, k));
    HIP_TRY(hipMemcpy(G.dtab[k], tab.data(), tab.size() * sizeof(bss::LevelSys),
                      hipMemcpyHostToDevice));
  }
  return BSS_OK;
}

// Layout + workspace allocation + static tables (device work).
int build(bss_plan* p) {
  if (p->built) return BSS_OK;
  if (int e = layout(p)) return e;
  hipError_t e = hipMalloc((void**)&p->ws, p->ws_bytes);
  if (e != hipSuccess) {
    p->ws = nullptr;
    return fail(BSS_E_NOMEM, "hipMalloc(%zu bytes) failed: %s", p->ws_bytes, hipGetErrorString(e));
  }
  HIP_TRY(hipMemset(p->ws, 0, p->ws_bytes));
  p->dinfo = (int*)(p->ws + p->info_off);
  if (p->mode == BSS_MODE_MIXED) {
    std::vector<int> nu(p->nsys);
    for (int s = 0; s < p->nsys; ++s) nu[s] = p->sp[s].N_use;
    if (p->mx.attach(p->ws + p->mixed_off, nu, p->n, p->nrhs) != 0)
      return fail(BSS_E_HIP, "mixed workspace: %s", bss::mixed_last_error());
  }
  for (size_t l = 0; l < p->gl.size(); ++l) {
    GLevel& G = p->gl[l];
    G.dchains = (bss::ChainRec*)(p->ws + p->toff[l].chains);
    G.dseps = (bss::SepRec*)(p->ws + p->toff[l].seps);
    for (int k = 0; k < 2; ++k)
      G.dtab[k] = (bss::LevelSys*)(p->ws + p->toff[l].tab[l == 0 ? k : 0]);
    if (!G.chains.empty())
      HIP_TRY(hipMemcpy(G.dchains, G.chains.data(), G.chains.size() * sizeof(bss::ChainRec),
                        hipMemcpyHostToDevice));
    if (!G.seps.empty())
      HIP_TRY(hipMemcpy(G.dseps, G.seps.data(), G.seps.size() * sizeof(bss::SepRec),
                        hipMemcpyHostToDevice));
  }
  // original (level-0) block index of every block, per system
  for (int s = 0; s < p->nsys; ++s) {
    SysPlan& S = p->sp[s];
    std::vector<int> orig(S.lv[0].shape.N);
    for (size_t i = 0; i < orig.size(); ++i) orig[i] = (int)i;
    for (size_t l = 0; l < S.lv.size(); ++l) {
      SysLevel& V = S.lv[l];
      HIP_TRY(hipMemcpy(p->ws + V.oOrig, orig.data(), orig.size() * sizeof(int),
                        hipMemcpyHostToDevice));
      std::vector<int> next(V.P);
      for (int k = 0; k < V.P; ++k) next[k] = orig[V.shape.seppos[k]];
      orig.swap(next);
    }
  }
  p->built = true;
  for (size_t l = 1; l < p->gl.size(); ++l)
    if (int e2 = upload_tables(p, l)) return e2;
  return BSS_OK;
}

template <typename Tin, typename T>
int factor_levels(bss_plan* p, hipStream_t s) {
  const int nb = p->n;
  HIP_TRY(hipMemsetAsync(p->dinfo, 0x7f, sizeof(int) * p->nsys, s));
  for (size_t l = 0; l < p->gl.size(); ++l) {
    GLevel& G = p->gl[l];
    const bss::FactorLevelArgs a{G.dchains, G.dtab[0], p->dinfo};
    hipError_t e;
    {
      ProfScope ps(p, BSS_K_FACTOR, (int)l, s);
      if (l == 0)
        e = bss::launch_factor<Tin, T>(nb, p->factor_waves, a, (int)G.chains.size(), s);
      else
        e = bss::launch_factor<T, T>(nb, p->factor_waves, a, (int)G.chains.size(), s);
    }
    if (e != hipSuccess) return fail(BSS_E_HIP, "factor level %zu launch: %s", l, hipGetErrorString(e));
    if (!G.seps.empty()) {
      ProfScope ps(p, BSS_K_ASSEMBLE_DIAG, (int)l, s);
      if (l == 0)
        e = bss::launch_assemble_diag<Tin, T>(G.dseps, G.dtab[0], nb, (int)G.seps.size(), s);
      else
        e = bss::launch_assemble_diag<T, T>(G.dseps, G.dtab[0], nb, (int)G.seps.size(), s);
      if (e != hipSuccess) return fail(BSS_E_HIP, "assemble level %zu: %s", l, hipGetErrorString(e));
    }
  }
  return BSS_OK;
}

// Solve in place in the level-0 vectors of set k (plain: the caller's d; mixed: the
// FP32 ping-pong vector k).
template <typename T>
int solve_levels(bss_plan* p, int k, hipStream_t s) {
  const int nb = p->n;
  const int nd = p->n * p->nrhs;
  const size_t L = p->gl.size();
  for (size_t l = 0; l < L; ++l) {
    GLevel& G = p->gl[l];
    const bss::SolveLevelArgs a{G.dchains, G.dtab[l == 0 ? k : 0], p->nrhs};
    hipError_t e;
    {
      ProfScope ps(p, BSS_K_FORWARD, (int)l, s);
      e = bss::launch_forward<T>(nb, a, (int)G.chains.size(), s);
    }
    if (e != hipSuccess) return fail(BSS_E_HIP, "forward level %zu: %s", l, hipGetErrorString(e));
    if (!G.seps.empty()) {
      ProfScope ps(p, BSS_K_ASSEMBLE_RHS, (int)l, s);
      e = bss::launch_assemble_rhs<T>(G.dseps, a.sys, nd, (int)G.seps.size(), s);
      if (e != hipSuccess) return fail(BSS_E_HIP, "rhs assemble %zu: %s", l, hipGetErrorString(e));
    }
  }
  for (size_t li = L; li-- > 0;) {
    GLevel& G = p->gl[li];
    const bss::SolveLevelArgs a{G.dchains, G.dtab[li == 0 ? k : 0], p->nrhs};
    hipError_t e;
    {
      ProfScope ps(p, BSS_K_BACKWARD, (int)li, s);
      e = bss::launch_backward<T>(nb, a, (int)G.chains.size(), s);
    }
    if (e != hipSuccess) return fail(BSS_E_HIP, "backward level %zu: %s", li, hipGetErrorString(e));
    if (!G.seps.empty()) {
      ProfScope ps(p, BSS_K_SCATTER, (int)li, s);
      e = bss::launch_scatter_sep<T>(G.dseps, a.sys, nd, (int)G.seps.size(), s);
      if (e != hipSuccess) return fail(BSS_E_HIP, "scatter %zu: %s", li, hipGetErrorString(e));
    }
  }
  return BSS_OK;
}

int check_plan(bss_plan* p) {
  if (!p) return fail(BSS_E_ARG, "null plan");
  return BSS_OK;
}

int create(int nsys, const int* Ns, int n, int nrhs, int dtype, int mode, int schedule,
           bss_plan_t* out) {
  if (!out) return fail(BSS_E_ARG, "out is null");
  *out = nullptr;
  if (nsys < 1 || nsys > (1 << 20)) return fail(BSS_E_ARG, "nsys must be in [1, 2^20] (got %d)", nsys);
  if (!Ns) return fail(BSS_E_ARG, "Ns is null");
  for (int s = 0; s < nsys; ++s)
    if (Ns[s] < 1) return fail(BSS_E_ARG, "N must be >= 1 (system %d has %d)", s, Ns[s]);
  if (n != 16 && n != 32 && n != 48 && n != 64)
    return fail(BSS_E_UNSUPPORTED, "block size n=%d not supported (16, 32, 48, 64)", n);
  if (nrhs < 1 || nrhs > 16) return fail(BSS_E_UNSUPPORTED, "nrhs=%d not in [1, 16]", nrhs);
  if (dtype != BSS_F32 && dtype != BSS_F64) return fail(BSS_E_ARG, "bad dtype %d", dtype);
  if (mode != BSS_MODE_PLAIN && mode != BSS_MODE_MIXED) return fail(BSS_E_ARG, "bad mode %d", mode);
  if (mode == BSS_MODE_MIXED && dtype != BSS_F64)
    return fail(BSS_E_ARG, "mixed mode needs FP64 input (dtype BSS_F64)");
  if (schedule != BSS_SCHED_AUTO && schedule != BSS_SCHED_REFERENCE)
    return fail(BSS_E_ARG, "bad schedule %d", schedule);
  bss_plan* p = new (std::nothrow) bss_plan();
  if (!p) return fail(BSS_E_NOMEM, "host allocation failed");
  p->Ns.assign(Ns, Ns + nsys);
  p->n = n;
  p->nrhs = nrhs;
  p->nsys = nsys;
  p->dtype = dtype;
  p->mode = mode;
  p->schedule = schedule;
  if (hipGetDevice(&p->device) != hipSuccess) p->device = 0;
  *out = p;
  return BSS_OK;
}

}  // namespace

extern "C" {

int bss_last_error(char* buf, size_t len) {
  if (buf && len) {
    size_t k = std::min(len - 1, g_last_error.size());
    memcpy(buf, g_last_error.data(), k);
    buf[k] = 0;
  }
  return (int)g_last_error.size();
}

int bss_plan_create(int N, int n, int nrhs, int nsys, int dtype, int mode, int schedule,
                    bss_plan_t* out) {
  if (nsys < 1 || nsys > (1 << 20)) return fail(BSS_E_ARG, "nsys must be in [1, 2^20] (got %d)", nsys);
  if (N < 1) return fail(BSS_E_ARG, "N must be >= 1 (got %d)", N);
  std::vector<int> Ns(nsys, N);
  return create(nsys, Ns.data(), n, nrhs, dtype, mode, schedule, out);
}

int bss_plan_create_group(int nsys, const int* Ns, int n, int nrhs, int dtype, int mode,
                          int schedule, bss_plan_t* out) {
  return create(nsys, Ns, n, nrhs, dtype, mode, schedule, out);
}

int bss_plan_system_sizes(bss_plan_t p, int* n_use) {
  if (int e = check_plan(p)) return e;
  if (int e = layout(p)) return e;
  if (n_use)
    for (int s = 0; s < p->nsys; ++s) n_use[s] = p->sp[s].N_use;
  return p->nsys;
}

int bss_set_option(bss_plan_t p, int option, double value) {
  if (int e = check_plan(p)) return e;
  const bool sched_opt = option == BSS_OPT_CHAIN_LEN || option == BSS_OPT_BASE_MAX ||
                         option == BSS_OPT_PRESERVE_INPUT ||
                         (option >= BSS_OPT_CHAIN_LEN_LEVEL && option < BSS_OPT_CHAIN_LEN_LEVEL + 16);
  if (sched_opt && p->laid_out)
    return fail(BSS_E_STATE, "schedule and storage options must precede queries, bind and factorize");
  switch (option) {
    case BSS_OPT_TOL:
      if (!(value > 0)) return fail(BSS_E_ARG, "tol must be > 0");
      p->tol = value;
      return BSS_OK;
    case BSS_OPT_MAX_ITERS:
      if (value < 1 || value > 100) return fail(BSS_E_ARG, "max_iters must be in [1, 100]");
      p->max_iters = (int)value;
      return BSS_OK;
    case BSS_OPT_PIVOT_TOL:
      if (!(value >= 0)) return fail(BSS_E_ARG, "pivot_tol must be >= 0");
      p->pivot_tol = value;
      return BSS_OK;
    case BSS_OPT_CHAIN_LEN:
      if (value < 1 || value > 1 << 20) return fail(BSS_E_ARG, "chain length must be >= 1");
      p->chain_len_default = (int)value;
      return BSS_OK;
    case BSS_OPT_FACTOR_WAVES:
      if (value != 4 && value != 8) return fail(BSS_E_ARG, "factor waves must be 4 or 8");
      p->factor_waves = (int)value;
      return BSS_OK;
    case BSS_OPT_BASE_MAX:
      if (value < 1 || value > 1 << 20) return fail(BSS_E_ARG, "base_max must be >= 1");
      p->base_max = (int)value;
      return BSS_OK;
    case BSS_OPT_PRESERVE_INPUT:
      p->preserve = value != 0.0;
      return BSS_OK;
    default:
      if (option >= BSS_OPT_CHAIN_LEN_LEVEL && option < BSS_OPT_CHAIN_LEN_LEVEL + 16) {
        if (value < 1 || value > 1 << 20) return fail(BSS_E_ARG, "chain length must be >= 1");
        p->chain_len[option - BSS_OPT_CHAIN_LEN_LEVEL] = (int)value;
        return BSS_OK;
      }
      return fail(BSS_E_ARG, "unknown option %d", option);
  }
}

int bss_plan_get(bss_plan_t p, int query, int64_t* value) {
  if (int e = check_plan(p)) return e;
  if (!value) return fail(BSS_E_ARG, "value is null");
  if (int e = layout(p)) return e;
  const int L = (int)p->gl.size();
  const SysPlan& S0 = p->sp[0];
  if (query == BSS_Q_N_USE) {
    *value = S0.N_use;
  } else if (query == BSS_Q_LEVELS) {
    *value = L;
  } else if (query == BSS_Q_WORKSPACE_BYTES) {
    *value = (int64_t)p->ws_bytes;
  } else if (query >= BSS_Q_LEVEL_N && query < BSS_Q_LEVEL_N + 32) {
    const int l = query - BSS_Q_LEVEL_N;
    if (l >= L) return fail(BSS_E_ARG, "level %d out of range", l);
    *value = l < (int)S0.lv.size() ? S0.lv[l].shape.N : 0;
  } else if (query >= BSS_Q_LEVEL_CHAINS && query < BSS_Q_LEVEL_CHAINS + 32) {
    const int l = query - BSS_Q_LEVEL_CHAINS;
    if (l >= L) return fail(BSS_E_ARG, "level %d out of range", l);
    *value = (int64_t)p->gl[l].chains.size();
  } else if (query >= BSS_Q_LEVEL_MAXLEN && query < BSS_Q_LEVEL_MAXLEN + 32) {
    const int l = query - BSS_Q_LEVEL_MAXLEN;
    if (l >= L) return fail(BSS_E_ARG, "level %d out of range", l);
    *value = p->gl[l].maxlen;
  } else if (query == BSS_Q_LAST_ITERS) {
    *value = p->last_iters;
  } else if (query == BSS_Q_REFACTORED) {
    *value = p->refactored;
  } else {
    return fail(BSS_E_ARG, "unknown query %d", query);
  }
  return BSS_OK;
}

int bss_plan_level_chains(bss_plan_t p, int level, int* chains, int cap) {
  if (int e = check_plan(p)) return e;
  if (int e = layout(p)) return e;
  const SysPlan& S0 = p->sp[0];
  if (level < 0 || level >= (int)S0.lv.size()) return fail(BSS_E_ARG, "level out of range");
  const auto& ch = S0.lv[level].shape.chains;
  for (int c = 0; c < (int)ch.size() && c < cap; ++c) {
    chains[4 * c + 0] = ch[c].start;
    chains[4 * c + 1] = ch[c].len;
    chains[4 * c + 2] = ch[c].sepL >= 0 ? 1 : 0;
    chains[4 * c + 3] = ch[c].sepR >= 0 ? 1 : 0;
  }
  return (int)ch.size();
}

int bss_plan_bind(bss_plan_t p, void* A, void* B, void* d) {
  if (int e = check_plan(p)) return e;
  if (int e = layout(p)) return e;
  bool needB = false;
  for (auto& S : p->sp) needB = needB || S.N_use > 1;
  if (!A || !d || (!B && needB)) return fail(BSS_E_ARG, "null buffer");
  if (int e = build(p)) return e;
  p->A = A;
  p->B = B;
  p->d = d;
  if (p->mode == BSS_MODE_MIXED && p->mx.bind((const double*)A, (const double*)B, (const double*)d) != 0)
    return fail(BSS_E_HIP, "mixed bind: %s", bss::mixed_last_error());
  if (int e = upload_tables(p, 0)) return e;
  p->factored = false;
  return BSS_OK;
}

int bss_factorize(bss_plan_t p, void* stream) {
  if (int e = check_plan(p)) return e;
  if (!p->A) return fail(BSS_E_STATE, "bss_plan_bind must precede bss_factorize");
  hipStream_t s = (hipStream_t)stream;
  p->last_stream = s;
  int rc;
  if (p->mode == BSS_MODE_MIXED) {
    rc = factor_levels<double, float>(p, s);
    if (rc == BSS_OK) bss::mixed_after_factor(p->mx);
  } else if (p->dtype == BSS_F64) {
    rc = factor_levels<double, double>(p, s);
  } else {
    rc = factor_levels<float, float>(p, s);
  }
  if (rc == BSS_OK) p->factored = true;
  return rc;
}

int bss_solve(bss_plan_t p, void* stream) {
  if (int e = check_plan(p)) return e;
  if (!p->factored) return fail(BSS_E_STATE, "bss_factorize must precede bss_solve");
  hipStream_t s = (hipStream_t)stream;
  p->last_stream = s;
  if (p->mode == BSS_MODE_MIXED) {
    bss::MixedSolveHooks hooks;
    hooks.ctx = p;
    hooks.solve32 = [](void* ctx, int k, hipStream_t st) -> int {
      return solve_levels<float>((bss_plan*)ctx, k, st);
    };
    hooks.prof_begin = [](void* ctx, int kind, hipStream_t st) -> void* {
      return new ProfScope((bss_plan*)ctx, kind, 0, st);
    };
    hooks.prof_end = [](void* scope) { delete (ProfScope*)scope; };
    int iters = 0;
    int rc = bss::mixed_solve(p->mx, hooks, (double*)p->d, p->tol, p->max_iters, &iters, s);
    p->last_iters = iters;
    if (rc != 0) return fail(BSS_E_HIP, "mixed solve failed: %s", bss::mixed_last_error());
    return BSS_OK;
  }
  if (p->dtype == BSS_F64) return solve_levels<double>(p, 0, s);
  return solve_levels<float>(p, 0, s);
}

int bss_info(bss_plan_t p, int* info_host) {
  if (int e = check_plan(p)) return e;
  if (!info_host) return fail(BSS_E_ARG, "info_host is null");
  if (!p->built || !p->factored) return fail(BSS_E_STATE, "no factorization yet");
  HIP_TRY(hipStreamSynchronize(p->last_stream));
  HIP_TRY(hipMemcpy(info_host, p->dinfo, sizeof(int) * p->nsys, hipMemcpyDeviceToHost));
  for (int s = 0; s < p->nsys; ++s)
    if (info_host[s] == kInfoNone) info_host[s] = 0;
  return BSS_OK;
}

int bss_last_backward_error(bss_plan_t p, double* berr_host) {
  if (int e = check_plan(p)) return e;
  if (p->mode != BSS_MODE_MIXED) return fail(BSS_E_STATE, "backward error is tracked in mixed mode");
  if (!berr_host) return fail(BSS_E_ARG, "berr_host is null");
  for (size_t i = 0; i < p->mx.last_berr.size(); ++i) berr_host[i] = p->mx.last_berr[i];
  return BSS_OK;
}

int bss_sync(bss_plan_t p) {
  if (int e = check_plan(p)) return e;
  HIP_TRY(hipStreamSynchronize(p->last_stream));
  return BSS_OK;
}

int bss_plan_destroy(bss_plan_t p) {
  if (!p) return BSS_OK;
  if (p->ws) {
    (void)hipStreamSynchronize(p->last_stream);
    (void)hipFree(p->ws);
  }
  delete p;
  return BSS_OK;
}

int bss_profile_enable(bss_plan_t p, int on) {
  if (int e = check_plan(p)) return e;
  if (on) {
    if (p->prof) HIP_TRY(hipStreamSynchronize(p->last_stream));
    p->recs.clear();
    p->pool_next = 0;
  }
  p->prof = on != 0;
  return BSS_OK;
}

int bss_profile_read(bss_plan_t p, int kind, int level, int* count, double* total_ms) {
  if (int e = check_plan(p)) return e;
  if (!count || !total_ms) return fail(BSS_E_ARG, "null output");
  HIP_TRY(hipStreamSynchronize(p->last_stream));
  int c = 0;
  double t = 0.0;
  for (const auto& r : p->recs) {
    if (r.kind != kind || (level >= 0 && r.level != level)) continue;
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
    ++c;
    t += ms;
  }
  *count = c;
  *total_ms = t;
  return BSS_OK;
}

int bss_reference_schedule(int N, int* n_use, int* levels, int cap) {
  if (N < 1 || !n_use) return fail(BSS_E_ARG, "bad arguments");
  std::vector<std::array<int, 3>> lv;
  *n_use = bss::reference_schedule(N, &lv);
  for (int k = 0; k < (int)lv.size() && k < cap; ++k)
    for (int j = 0; j < 3; ++j) levels[3 * k + j] = lv[k][j];
  return (int)lv.size();
}

}  // extern "C"
