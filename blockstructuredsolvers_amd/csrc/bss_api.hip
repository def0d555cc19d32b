// bss_api.hip -- C ABI of libbss: plans, per-level workspace, the level loop.
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

struct Level {
  bss::LevelShape shape;
  int P = 0, nch = 0, maxlen = 0;
  int4* dchains = nullptr;
  int* dseppos = nullptr;
  int* dorig = nullptr;
  // buffers; element type is the plan's input type at level 0 (A, B), factor type
  // elsewhere.  In plain mode U/C alias A/B.
  void *A = nullptr, *B = nullptr, *d = nullptr;
  void *U = nullptr, *C = nullptr, *W = nullptr;
  void *SR = nullptr, *SL = nullptr, *rR = nullptr, *rL = nullptr;
  // per-system element strides
  long long sA = 0, sB = 0, sD = 0, sW = 0, sS = 0, sR = 0;
};

}  // namespace

struct bss_plan {
  int N = 0, n = 0, nrhs = 0, nsys = 0, dtype = 0, mode = 0, schedule = 0;
  int N_use = 0;
  int device = 0;
  std::vector<Level> lv;
  char* ws = nullptr;
  size_t ws_bytes = 0;
  void *A = nullptr, *B = nullptr, *d = nullptr;  // caller buffers
  int* dinfo = nullptr;
  bool factored = false;
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
  ~bss_plan() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
};

namespace {

size_t fsize(const bss_plan* p) {  // factor element size
  return (p->dtype == BSS_F64 && p->mode == BSS_MODE_PLAIN) ? 8 : 4;
}

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

// Lay out the schedule and the workspace; with `alloc` also allocate the workspace
// and upload the chain tables (device work).  Layout alone is host-only.
int build(bss_plan* p, bool alloc = true) {
  if (p->built) return BSS_OK;
  std::vector<bss::LevelShape> shapes;
  if (p->schedule == BSS_SCHED_REFERENCE) {
    int n_use = 0;
    shapes = bss::reference_schedule_levels(p->N, &n_use);
    p->N_use = n_use;
  } else {
    shapes = bss::auto_schedule(p->N, p->chain_len, p->chain_len_default,
                                std::max(1, p->base_max));
    p->N_use = p->N;
  }
  const size_t nn = (size_t)p->n * p->n, nd = (size_t)p->n * p->nrhs;
  const size_t fs = fsize(p), S = (size_t)p->nsys;
  const bool mixed = p->mode == BSS_MODE_MIXED;
  Bump bump;
  constexpr size_t kNone = ~size_t(0);  // buffer not allocated at this level
  struct Off {
    size_t A = kNone, B = kNone, d = kNone, W = kNone, SR = kNone, SL = kNone, rR = kNone,
           rL = kNone, U = kNone, C = kNone, chains = kNone, seppos = kNone, orig = kNone;
  };
  std::vector<Off> off(shapes.size());
  p->lv.assign(shapes.size(), Level());
  for (size_t l = 0; l < shapes.size(); ++l) {
    Level& L = p->lv[l];
    L.shape = shapes[l];
    L.P = (int)L.shape.seppos.size();
    L.nch = (int)L.shape.chains.size();
    for (auto& c : L.shape.chains) L.maxlen = std::max(L.maxlen, c.len);
    const size_t NL = (size_t)L.shape.N;
    Off& o = off[l];
    o = Off{};
    L.sA = (long long)(NL * nn);
    L.sB = (long long)((NL > 0 ? NL - 1 : 0) * nn);
    L.sD = (long long)(NL * nd);
    L.sW = (long long)(NL * nn);
    L.sS = (long long)((size_t)std::max(L.P, 1) * nn);
    L.sR = (long long)((size_t)std::max(L.P, 1) * nd);
    if (l > 0) {
      o.A = bump.take(S * NL * nn * fs);
      o.B = bump.take(S * std::max<size_t>(NL - 1, 1) * nn * fs);
      o.d = bump.take(S * NL * nd * fs);
    } else if (mixed) {
      o.U = bump.take(S * NL * nn * fs);
      o.C = bump.take(S * std::max<size_t>(NL - 1, 1) * nn * fs);
      o.d = bump.take(S * NL * nd * fs);  // FP32 correction vector
    } else if (p->preserve) {
      o.U = bump.take(S * NL * nn * fs);
      o.C = bump.take(S * std::max<size_t>(NL - 1, 1) * nn * fs);
    }
    bool anyL = false;
    for (auto& c : L.shape.chains) anyL = anyL || c.sepL >= 0;
    if (anyL) o.W = bump.take(S * NL * nn * fs);
    if (!anyL) L.sW = 0;
    if (L.P > 0) {
      o.SR = bump.take(S * L.P * nn * fs);
      o.SL = bump.take(S * L.P * nn * fs);
      o.rR = bump.take(S * L.P * nd * fs);
      o.rL = bump.take(S * L.P * nd * fs);
    }
    o.chains = bump.take(sizeof(int4) * std::max(L.nch, 1));
    o.seppos = bump.take(sizeof(int) * std::max(L.P, 1));
    o.orig = bump.take(sizeof(int) * std::max<size_t>(NL, 1));
  }
  const size_t info_off = bump.take(sizeof(int) * S);
  size_t mixed_off = 0, mixed_bytes = 0;
  if (mixed) {
    mixed_bytes = bss::MixedState::bytes(p->N_use, p->n, p->nrhs, p->nsys);
    mixed_off = bump.take(mixed_bytes);
  }
  p->ws_bytes = bump.off;
  if (!alloc) return BSS_OK;
  hipError_t e = hipMalloc((void**)&p->ws, p->ws_bytes);
  if (e != hipSuccess) {
    p->ws = nullptr;
    return fail(BSS_E_NOMEM, "hipMalloc(%zu bytes) failed: %s", p->ws_bytes, hipGetErrorString(e));
  }
  HIP_TRY(hipMemset(p->ws, 0, p->ws_bytes));
  p->dinfo = (int*)(p->ws + info_off);
  if (mixed) p->mx.attach(p->ws + mixed_off, p->N_use, p->n, p->nrhs, p->nsys);

  // original (level-0) block index of every block, for info reporting
  std::vector<int> orig(p->lv[0].shape.N);
  for (size_t i = 0; i < orig.size(); ++i) orig[i] = (int)i;
  for (size_t l = 0; l < p->lv.size(); ++l) {
    Level& L = p->lv[l];
    Off& o = off[l];
    auto at = [&](size_t x) -> void* { return x != kNone ? (void*)(p->ws + x) : nullptr; };
    if (l > 0) {
      L.A = at(o.A);
      L.B = at(o.B);
      L.d = at(o.d);
      L.U = L.A;
      L.C = L.B;
    } else if (mixed || p->preserve) {
      L.U = at(o.U);
      L.C = at(o.C);
      L.d = at(o.d);
    }
    L.W = at(o.W);
    L.SR = at(o.SR);
    L.SL = at(o.SL);
    L.rR = at(o.rR);
    L.rL = at(o.rL);
    L.dchains = (int4*)(p->ws + o.chains);
    L.dseppos = (int*)(p->ws + o.seppos);
    L.dorig = (int*)(p->ws + o.orig);
    std::vector<int4> ch;
    for (auto& c : L.shape.chains) ch.push_back(make_int4(c.start, c.len, c.sepL, c.sepR));
    if (!ch.empty())
      HIP_TRY(hipMemcpy(L.dchains, ch.data(), ch.size() * sizeof(int4), hipMemcpyHostToDevice));
    if (L.P > 0)
      HIP_TRY(hipMemcpy(L.dseppos, L.shape.seppos.data(), L.P * sizeof(int),
                        hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(L.dorig, orig.data(), orig.size() * sizeof(int), hipMemcpyHostToDevice));
    std::vector<int> next(L.P);
    for (int k = 0; k < L.P; ++k) next[k] = orig[L.shape.seppos[k]];
    orig.swap(next);
  }
  p->built = true;
  return BSS_OK;
}

template <typename Tin, typename T>
int factor_levels(bss_plan* p, hipStream_t s) {
  const int nb = p->n;
  HIP_TRY(hipMemsetAsync(p->dinfo, 0x7f, sizeof(int) * p->nsys, s));
  for (size_t l = 0; l < p->lv.size(); ++l) {
    Level& L = p->lv[l];
    Level* Nx = (l + 1 < p->lv.size()) ? &p->lv[l + 1] : nullptr;
    bss::FactorLevelArgs<T> a{};
    a.A_in = (l == 0) ? p->A : L.A;
    a.B_in = (l == 0) ? p->B : L.B;
    a.U_out = (T*)((l == 0 && inplace0(p)) ? p->A : L.U);
    a.C_out = (T*)((l == 0 && inplace0(p)) ? p->B : L.C);
    a.W_out = (T*)L.W;
    a.SR = (T*)L.SR;
    a.SL = (T*)L.SL;
    a.Bnext = Nx ? (T*)Nx->B : nullptr;
    a.chains = L.dchains;
    a.orig = L.dorig;
    a.info = p->dinfo;
    a.sysA = L.sA;
    a.sysB = L.sB;
    a.sysU = L.sA;
    a.sysC = L.sB;
    a.sysW = L.sW;
    a.sysS = L.sS;
    a.sysBn = Nx ? Nx->sB : 0;
    hipError_t e;
    {
      ProfScope ps(p, BSS_K_FACTOR, (int)l, s);
      if (l == 0)
        e = bss::launch_factor<Tin, T>(nb, p->factor_waves, a, L.nch, p->nsys, s);
      else
        e = bss::launch_factor<T, T>(nb, p->factor_waves, a, L.nch, p->nsys, s);
    }
    if (e != hipSuccess) return fail(BSS_E_HIP, "factor level %zu launch: %s", l, hipGetErrorString(e));
    if (Nx) {
      ProfScope ps(p, BSS_K_ASSEMBLE_DIAG, (int)l, s);
      if (l == 0)
        e = bss::launch_assemble_diag<Tin, T>((const Tin*)p->A, (const T*)L.SR, (const T*)L.SL,
                                              L.dseppos, (T*)Nx->A, nb, L.P, p->nsys, L.sA,
                                              L.sS, Nx->sA, s);
      else
        e = bss::launch_assemble_diag<T, T>((const T*)L.A, (const T*)L.SR, (const T*)L.SL,
                                            L.dseppos, (T*)Nx->A, nb, L.P, p->nsys, L.sA, L.sS,
                                            Nx->sA, s);
      if (e != hipSuccess) return fail(BSS_E_HIP, "assemble level %zu: %s", l, hipGetErrorString(e));
    }
  }
  return BSS_OK;
}

// Solve in place in the level-0 vector `d0` (element type T) with the factors.
template <typename T>
int solve_levels(bss_plan* p, T* d0, hipStream_t s) {
  const int nb = p->n;
  const int nd = p->n * p->nrhs;
  const size_t L0 = p->lv.size();
  const bool inpl = inplace0(p);
  auto U0 = [&]() -> const T* { return (const T*)(inpl ? p->A : p->lv[0].U); };
  auto C0 = [&]() -> const T* { return (const T*)(inpl ? p->B : p->lv[0].C); };
  for (size_t l = 0; l < L0; ++l) {
    Level& L = p->lv[l];
    Level* Nx = (l + 1 < L0) ? &p->lv[l + 1] : nullptr;
    bss::SolveLevelArgs<T> a{};
    a.U = l == 0 ? U0() : (const T*)L.U;
    a.C = l == 0 ? C0() : (const T*)L.C;
    a.W = (const T*)L.W;
    a.d = l == 0 ? d0 : (T*)L.d;
    a.rR = (T*)L.rR;
    a.rL = (T*)L.rL;
    a.chains = L.dchains;
    a.nrhs = p->nrhs;
    a.sysU = L.sA;
    a.sysC = L.sB;
    a.sysW = L.sW;
    a.sysD = L.sD;
    a.sysR = L.sR;
    hipError_t e;
    {
      ProfScope ps(p, BSS_K_FORWARD, (int)l, s);
      e = bss::launch_forward<T>(nb, a, L.nch, p->nsys, s);
    }
    if (e != hipSuccess) return fail(BSS_E_HIP, "forward level %zu: %s", l, hipGetErrorString(e));
    if (Nx) {
      ProfScope ps(p, BSS_K_ASSEMBLE_RHS, (int)l, s);
      e = bss::launch_assemble_rhs<T>(a.d, a.rR, a.rL, L.dseppos, (T*)Nx->d, nd, L.P, p->nsys,
                                      L.sD, L.sR, Nx->sD, s);
      if (e != hipSuccess) return fail(BSS_E_HIP, "rhs assemble %zu: %s", l, hipGetErrorString(e));
    }
  }
  for (size_t li = L0; li-- > 0;) {
    Level& L = p->lv[li];
    Level* Nx = (li + 1 < L0) ? &p->lv[li + 1] : nullptr;
    bss::SolveLevelArgs<T> a{};
    a.U = li == 0 ? U0() : (const T*)L.U;
    a.C = li == 0 ? C0() : (const T*)L.C;
    a.W = (const T*)L.W;
    a.d = li == 0 ? d0 : (T*)L.d;
    a.xsep = Nx ? (const T*)Nx->d : nullptr;
    a.chains = L.dchains;
    a.nrhs = p->nrhs;
    a.sysU = L.sA;
    a.sysC = L.sB;
    a.sysW = L.sW;
    a.sysD = L.sD;
    a.sysX = Nx ? Nx->sD : 0;
    hipError_t e;
    {
      ProfScope ps(p, BSS_K_BACKWARD, (int)li, s);
      e = bss::launch_backward<T>(nb, a, L.nch, p->nsys, s);
    }
    if (e != hipSuccess) return fail(BSS_E_HIP, "backward level %zu: %s", li, hipGetErrorString(e));
    if (Nx) {
      ProfScope ps(p, BSS_K_SCATTER, (int)li, s);
      e = bss::launch_scatter_sep<T>((const T*)Nx->d, L.dseppos, a.d, nd, L.P, p->nsys, Nx->sD,
                                     L.sD, s);
      if (e != hipSuccess) return fail(BSS_E_HIP, "scatter %zu: %s", li, hipGetErrorString(e));
    }
  }
  return BSS_OK;
}

int check_plan(bss_plan* p) {
  if (!p) return fail(BSS_E_ARG, "null plan");
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
  if (!out) return fail(BSS_E_ARG, "out is null");
  *out = nullptr;
  if (N < 1) return fail(BSS_E_ARG, "N must be >= 1 (got %d)", N);
  if (n != 16 && n != 32 && n != 48 && n != 64)
    return fail(BSS_E_UNSUPPORTED, "block size n=%d not supported (16, 32, 48, 64)", n);
  if (nrhs < 1 || nrhs > 16) return fail(BSS_E_UNSUPPORTED, "nrhs=%d not in [1, 16]", nrhs);
  if (nsys < 1 || nsys > 65535) return fail(BSS_E_ARG, "nsys must be in [1, 65535] (got %d)", nsys);
  if (dtype != BSS_F32 && dtype != BSS_F64) return fail(BSS_E_ARG, "bad dtype %d", dtype);
  if (mode != BSS_MODE_PLAIN && mode != BSS_MODE_MIXED) return fail(BSS_E_ARG, "bad mode %d", mode);
  if (mode == BSS_MODE_MIXED && dtype != BSS_F64)
    return fail(BSS_E_ARG, "mixed mode needs FP64 input (dtype BSS_F64)");
  if (schedule != BSS_SCHED_AUTO && schedule != BSS_SCHED_REFERENCE)
    return fail(BSS_E_ARG, "bad schedule %d", schedule);
  bss_plan* p = new (std::nothrow) bss_plan();
  if (!p) return fail(BSS_E_NOMEM, "host allocation failed");
  p->N = N;
  p->n = n;
  p->nrhs = nrhs;
  p->nsys = nsys;
  p->dtype = dtype;
  p->mode = mode;
  p->schedule = schedule;
  if (hipGetDevice(&p->device) != hipSuccess) p->device = 0;
  if (schedule == BSS_SCHED_REFERENCE) {
    p->N_use = bss::reference_schedule(N, nullptr);
  } else {
    p->N_use = N;
  }
  *out = p;
  return BSS_OK;
}

int bss_set_option(bss_plan_t p, int option, double value) {
  if (int e = check_plan(p)) return e;
  if (p->built && (option == BSS_OPT_CHAIN_LEN || option == BSS_OPT_BASE_MAX))
    return fail(BSS_E_STATE, "schedule options must be set before the first bind/factorize");
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
      if (p->built) return fail(BSS_E_STATE, "preserve-input must be set before the first bind");
      p->preserve = value != 0.0;
      return BSS_OK;
    default:
      if (option >= BSS_OPT_CHAIN_LEN_LEVEL && option < BSS_OPT_CHAIN_LEN_LEVEL + 16) {
        if (p->built) return fail(BSS_E_STATE, "schedule options must precede the first bind");
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
  if (query == BSS_Q_N_USE) {
    *value = p->N_use;
    return BSS_OK;
  }
  if (int e = build(p, false)) return e;
  const int L = (int)p->lv.size();
  if (query == BSS_Q_LEVELS) {
    *value = L;
  } else if (query == BSS_Q_WORKSPACE_BYTES) {
    *value = (int64_t)p->ws_bytes;
  } else if (query >= BSS_Q_LEVEL_N && query < BSS_Q_LEVEL_N + 32) {
    const int l = query - BSS_Q_LEVEL_N;
    if (l >= L) return fail(BSS_E_ARG, "level %d out of range", l);
    *value = p->lv[l].shape.N;
  } else if (query >= BSS_Q_LEVEL_CHAINS && query < BSS_Q_LEVEL_CHAINS + 32) {
    const int l = query - BSS_Q_LEVEL_CHAINS;
    if (l >= L) return fail(BSS_E_ARG, "level %d out of range", l);
    *value = p->lv[l].nch;
  } else if (query >= BSS_Q_LEVEL_MAXLEN && query < BSS_Q_LEVEL_MAXLEN + 32) {
    const int l = query - BSS_Q_LEVEL_MAXLEN;
    if (l >= L) return fail(BSS_E_ARG, "level %d out of range", l);
    *value = p->lv[l].maxlen;
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
  if (int e = build(p, false)) return e;
  if (level < 0 || level >= (int)p->lv.size()) return fail(BSS_E_ARG, "level out of range");
  const auto& ch = p->lv[level].shape.chains;
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
  if (!A || !d || (!B && p->N_use > 1)) return fail(BSS_E_ARG, "null buffer");
  if (int e = build(p)) return e;
  p->A = A;
  p->B = B;
  p->d = d;
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
    if (rc == BSS_OK) {
      if (bss::mixed_after_factor(p->mx, (const double*)p->A, (const double*)p->B, p->n,
                                  p->nsys, s) != 0)
        rc = fail(BSS_E_HIP, "norm kernel: %s", bss::mixed_last_error());
    }
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
    hooks.solve32 = [](void* ctx, float* d0, hipStream_t st) -> int {
      return solve_levels<float>((bss_plan*)ctx, d0, st);
    };
    hooks.prof_begin = [](void* ctx, int kind, hipStream_t st) -> void* {
      return new ProfScope((bss_plan*)ctx, kind, 0, st);
    };
    hooks.prof_end = [](void* scope) { delete (ProfScope*)scope; };
    int iters = 0;
    int rc = bss::mixed_solve(p->mx, hooks, (const double*)p->A, (const double*)p->B,
                              (double*)p->d, (float*)p->lv[0].d, p->n, p->nrhs, p->nsys, p->tol,
                              p->max_iters, &iters, s);
    p->last_iters = iters;
    if (rc != 0) return fail(BSS_E_HIP, "mixed solve failed: %s", bss::mixed_last_error());
    return BSS_OK;
  }
  if (p->dtype == BSS_F64) return solve_levels<double>(p, (double*)p->d, s);
  return solve_levels<float>(p, (float*)p->d, s);
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
