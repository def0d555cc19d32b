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
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/bss.h"
#include "launch.hpp"
#include "plan_tables.hpp"
#include "mixed.hpp"
#include "pinned_pool.hpp"
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


constexpr int kInfoNone = 0x7f7f7f7f;  // "no failure" sentinel

// Resets the per-system status words (info, and the pivot monitor when `piv` is set) in
// one launch.  A kernel rather than hipMemsetAsync, so the factorisation can be captured
// into a hipGraph.
__global__ void reset_status_kernel(int* __restrict__ info, int ninfo, int vinfo,
                                    int* __restrict__ piv, int npiv, int vpiv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ninfo) info[i] = vinfo;
  if (piv && i < npiv) piv[i] = vpiv;
}
// *changed = 1 if any of the n 64-bit words of a and b differ (grid-stride; the flag
// was cleared before)
__global__ void __launch_bounds__(256) words_differ_kernel(const unsigned long long* __restrict__ a,
                                                           const unsigned long long* __restrict__ b,
                                                           long long n, int* __restrict__ changed) {
  bool diff = false;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    diff |= a[i] != b[i];
  if (diff) *changed = 1;
}
constexpr int kPivNone = 0x7f800000;  // float +inf bits: no pivot seen
constexpr size_t kNone = ~size_t(0);   // buffer not allocated

// one system at one level: its chain layout and where its buffers live
struct SysLevel {
  bss::LevelShape shape;
  int P = 0;
  size_t oA = kNone, oB = kNone, od = kNone, oW = kNone, oSR = kNone, oSL = kNone,
         orR = kNone, orL = kNone, oU = kNone, oC = kNone, oOrig = kNone, oPrev = kNone;
  // mixed mode, levels >= 1: FP64 copies used when the separator levels are
  // re-factored in FP64
  size_t wA = kNone, wB = kNone, wd = kNone, wW = kNone, wSR = kNone, wSL = kNone,
         wrR = kNone, wrL = kNone;
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
  // [k]: level-0 vector set k (mixed mode ping-pong); differs at levels 0 and 1 only
  bss::LevelSys* dtab[2] = {nullptr, nullptr};
  // mixed mode: level >= 1 -> FP64 tables; level 0 -> hand-off tables whose next-level
  // pointers (An, dn) name the FP64 copies of level 1
  bss::LevelSys* dtab64[2] = {nullptr, nullptr};
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
  bool ws_owned = true;  // false: caller's memory (bss_plan_set_workspace)
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
  bool preserve = true;  // plain mode: factors to workspace, A/B untouched
  // bss_info: per-system LAPACK-order info of the last factorisation, computed once
  std::vector<int> info_exact;
  bool info_valid = false;
  // mixed mode: plain FP64 plan over the same systems and buffers (input preserved),
  // used when refinement stops short (BSS_OPT_FALLBACK)
  bool fallback_on = true;
  bss_plan* fb = nullptr;
  bool fb_factored = false;  // fb holds the factors of the current A, B
  bool fb_active = false;    // solves of this factorisation go through fb
  bool fp32_failed = false;  // an FP32 level-0 pivot failed in the last factorisation
  // set-up traffic (workspace clear, table uploads): synchronous on the null stream, or
  // stream-ordered on setup_stream (the allocation-free entry point)
  bool setup_on_stream = false;
  hipStream_t setup_stream = nullptr;
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
  unsigned prof_kinds = ~0u;
  int prof_max_level = 1 << 30;
  std::vector<ProfRec> recs;
  std::vector<hipEvent_t> pool;
  size_t pool_next = 0;
  // table offsets
  struct TabOff {
    size_t chains = kNone, seps = kNone, tab[2] = {kNone, kNone}, tab64[2] = {kNone, kNone};
  };
  std::vector<TabOff> toff;
  size_t info_off = kNone;
  // host-written set-up ranges of the workspace (see layout)
  size_t static_begin = 0, static_end = 0, bound_begin = 0, bound_end = 0;
  // separator re-factorisation (mixed mode)
  size_t widen_off = kNone, narrow_off = kNone;
  unsigned* dpiv = nullptr;
  bss::ConvRec<float, double>* dwiden = nullptr;
  bss::ConvRec<double, float>* dnarrow = nullptr;
  int nconv = 0;
  bool pivots_pending = false;  // factorized, pivots not yet inspected
  unsigned* host_piv = nullptr;  // pinned: [0, nsys) FP32 level-0 info words, [nsys,
                                 // 2 nsys) FP32 separator info words, [2 nsys, 3 nsys)
                                 // pivot ratios, copied after each factorize
  hipEvent_t piv_event = nullptr;
  bool sep64 = false;           // separator levels currently factored in FP64
  hipGraphExec_t g_solve64[2] = {nullptr, nullptr};
  // hipGraph replay of the launch sequences (option BSS_OPT_GRAPH)
  bool use_graph = false;
  hipStream_t cap_stream = nullptr;  // capture happens here, replay on the caller's stream
  hipGraphExec_t g_factor = nullptr, g_solve[2] = {nullptr, nullptr};
  // bss_factorize_solve: the first solve's forward sweeps of levels [0, fwd_done) run
  // on side_stream while the later levels are factored; the next solve32 skips them
  hipStream_t side_stream = nullptr;
  hipEvent_t ev_l0_factored = nullptr, ev_l0_solved = nullptr;
  std::vector<hipEvent_t> ev_level;  // level l factored (levels beyond the first two)
  size_t fwd_done = 0;
  // BSS_OPT_FACTOR_SWEEP: bss_factorize ran the first solve's x = 0 pass and level-0
  // forward sweep on d (fsweep_pending) and copied d to mx.dsnap; the next bss_solve
  // sets *mx.dflag = 1 on the device if d differs from it, and its first pass and
  // level-0 forward sweep (fsweep_flag, consumed by the first solve32) run only then
  bool fsweep_on = true, fsweep_pending = false;
  const int* fsweep_flag = nullptr;
  void drop_graphs() {
    for (hipGraphExec_t* g : {&g_factor, &g_solve[0], &g_solve[1], &g_solve64[0], &g_solve64[1]}) {
      if (*g) (void)hipGraphExecDestroy(*g);
      *g = nullptr;
    }
  }
  ~bss_plan() {
    drop_graphs();
    if (piv_event) (void)hipEventDestroy(piv_event);
    if (host_piv) bss::pinned_give(host_piv);
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    if (ev_l0_factored) (void)hipEventDestroy(ev_l0_factored);
    if (ev_l0_solved) (void)hipEventDestroy(ev_l0_solved);
    for (hipEvent_t e : ev_level) (void)hipEventDestroy(e);
    if (side_stream) (void)hipStreamDestroy(side_stream);
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
};

namespace bss {
int plan_factor_prepare(bss_plan* p, void* stream);
int plan_factor_kernels(bss_plan* p, hipStream_t c);
int plan_factor_finish(bss_plan* p, void* stream);
int plan_solve_kernels(bss_plan* p, hipStream_t c);
}  // namespace bss

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
    if (!p->prof || !((p->prof_kinds >> kind) & 1u) || level > p->prof_max_level) return;
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
      if (mixed && l > 0) {
        // the FP32 factor keeps its C apart from B, so the reduced couplings survive
        // for an FP64 re-factorisation
        V.oC = bump.take(std::max<size_t>(NL - 1, 1) * nn * fs);
        V.wA = bump.take(NL * nn * 8);
        V.wB = bump.take(std::max<size_t>(NL - 1, 1) * nn * 8);
        V.wd = bump.take(NL * nd * 8);
        if (anyL) V.wW = bump.take(NL * nn * 8);
        if (V.P > 0) {
          V.wSR = bump.take(V.P * nn * 8);
          V.wSL = bump.take(V.P * nn * 8);
          V.wrR = bump.take(V.P * nd * 8);
          V.wrL = bump.take(V.P * nd * 8);
        }
      }
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
  // Everything the host writes sits in two contiguous ranges, each uploaded in one copy:
  // the static range (index maps, chain and separator records, the tables of levels
  // >= 2, conversion records), written once when the workspace is first set up, and the
  // bound range (the tables of levels 0 and 1, which name the caller's buffers),
  // rewritten at every bind.
  p->static_begin = bump.off;
  for (int s = 0; s < p->nsys; ++s) {
    SysPlan& S = p->sp[s];
    for (size_t l = 0; l < S.lv.size(); ++l) {
      SysLevel& V = S.lv[l];
      const size_t NL = (size_t)V.shape.N;
      V.oOrig = bump.take(sizeof(int) * std::max<size_t>(NL, 1));
      if (l > 0) V.oPrev = bump.take(sizeof(int) * std::max<size_t>(NL, 1));
    }
  }
  p->toff.assign(L, bss_plan::TabOff());
  auto take_tables = [&](size_t l) {
    const GLevel& G = p->gl[l];
    for (int k = 0; k < (l <= 1 ? 2 : 1); ++k) {
      p->toff[l].tab[k] = bump.take(sizeof(bss::LevelSys) * G.slot_sys.size());
      if (mixed) p->toff[l].tab64[k] = bump.take(sizeof(bss::LevelSys) * G.slot_sys.size());
    }
  };
  for (size_t l = 0; l < L; ++l) {
    GLevel& G = p->gl[l];
    p->toff[l].chains = bump.take(sizeof(bss::ChainRec) * std::max<size_t>(G.chains.size(), 1));
    p->toff[l].seps = bump.take(sizeof(bss::SepRec) * std::max<size_t>(G.seps.size(), 1));
    if (l >= 2) take_tables(l);
  }
  if (mixed) {
    p->widen_off = bump.take(sizeof(bss::ConvRec<float, double>) * p->nsys);
    p->narrow_off = bump.take(sizeof(bss::ConvRec<double, float>) * p->nsys);
  }
  p->static_end = p->bound_begin = bump.off;
  for (size_t l = 0; l < std::min<size_t>(L, 2); ++l) take_tables(l);
  p->bound_end = bump.off;
  // status words: [0, nsys) info of level 0 and FP64 levels, [nsys, 2 nsys) info of the
  // FP32 separator levels; mixed mode adds [2 nsys, 3 nsys) pivot monitor, so that the
  // host reads separator info and pivots in one copy
  p->info_off = bump.take(sizeof(int) * (mixed ? 3 : 2) * p->nsys);
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
// `wide` (mixed mode): levels >= 1 are the FP64 copies; at level 0 only the next-level
// pointers change (hand-off into FP64 level 1).
bss::LevelSys level_sys(const bss_plan* p, int s, size_t l, int k, bool wide = false) {
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
    t.A = at(wide ? V.wA : V.oA);
    t.B = at(wide ? V.wB : V.oB);
    t.U = (void*)t.A;
    t.C = (!wide && V.oC != kNone) ? at(V.oC) : (void*)t.B;
    t.d = at(wide ? V.wd : V.od);
  }
  const bool w = wide && l > 0;
  t.W = at(w ? V.wW : V.oW);
  t.SR = at(w ? V.wSR : V.oSR);
  t.SL = at(w ? V.wSL : V.oSL);
  t.rR = at(w ? V.wrR : V.orR);
  t.rL = at(w ? V.wrL : V.orL);
  if (l + 1 < S.lv.size()) {
    const SysLevel& Nx = S.lv[l + 1];
    t.An = at(wide ? Nx.wA : Nx.oA);
    t.Bn = at(wide ? Nx.wB : Nx.oB);
    t.dn = at(wide ? Nx.wd : Nx.od);
  }
  t.orig = (const int*)at(V.oOrig);
  t.sys = s;
  t.N = V.shape.N;
  // the sweeps do the level hand-off themselves (not in the FP64 separator tables,
  // whose top hand-off converts between types in separate kernels)
  t.fused = wide ? 0 : 1;
  if (!wide && l > 0) {
    const SysLevel& Pv = S.lv[l - 1];
    t.dp = level_sys(p, s, l - 1, k, false).d;
    t.rRp = at(Pv.orR);
    t.rLp = at(Pv.orL);
    t.pp = (const int*)at(V.oPrev);
  }
  return t;
}

// Host -> device copy of plan set-up data: synchronous, or stream-ordered on the plan's
// set-up stream.  A stream-ordered copy from pageable host memory has taken the source
// bytes when the call returns, so the host vectors may go out of scope right after.
hipError_t h2d(const bss_plan* p, void* dst, const void* src, size_t bytes) {
  if (!bytes) return hipSuccess;
  if (p->setup_on_stream)
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, p->setup_stream);
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
}

__global__ void zero_kernel(uint4* __restrict__ p, size_t n16) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Host image of one set-up range of the workspace: every piece is written at its
// workspace offset, and the range then goes to the device in one copy.
struct SetupImage {
  size_t base;
  std::vector<char> bytes;
  bool ok = true;
  SetupImage(size_t begin, size_t end) : base(begin), bytes(end - begin, 0) {}
  void put(size_t off, const void* src, size_t n) {
    if (!n) return;
    if (off < base || off + n > base + bytes.size()) {
      ok = false;
      return;
    }
    std::memcpy(bytes.data() + (off - base), src, n);
  }
  void put_at(const bss_plan* p, const void* dev, const void* src, size_t n) {
    put((size_t)((const char*)dev - p->ws), src, n);
  }
  int upload(const bss_plan* p) const {
    if (!ok) return fail(BSS_E_STATE, "set-up data outside its workspace range");
    HIP_TRY(h2d(p, p->ws + base, bytes.data(), bytes.size()));
    return BSS_OK;
  }
};

void fill_tables(const bss_plan* p, size_t l, SetupImage& img) {
  const GLevel& G = p->gl[l];
  for (int k = 0; k < (l <= 1 ? 2 : 1); ++k) {
    for (int wide = 0; wide < (p->mode == BSS_MODE_MIXED ? 2 : 1); ++wide) {
      if (wide && l == 1 && k == 1) continue;  // FP64 tables vary at level 0 only
      std::vector<bss::LevelSys> tab;
      for (int s : G.slot_sys) tab.push_back(level_sys(p, s, l, k, wide != 0));
      img.put_at(p, wide ? G.dtab64[k] : G.dtab[k], tab.data(),
                 tab.size() * sizeof(bss::LevelSys));
    }
  }
}

// The tables of levels 0 and 1, which name the caller's buffers, in one copy.
int upload_bound_tables(bss_plan* p) {
  SetupImage img(p->bound_begin, p->bound_end);
  for (size_t l = 0; l < std::min<size_t>(2, p->gl.size()); ++l) fill_tables(p, l, img);
  return img.upload(p);
}

// Layout + workspace allocation + static tables (device work).
int build(bss_plan* p) {
  if (p->built) return BSS_OK;
  if (int e = layout(p)) return e;
  if (!p->ws) {
    hipError_t e = hipMalloc((void**)&p->ws, p->ws_bytes);
    if (e != hipSuccess) {
      p->ws = nullptr;
      return fail(BSS_E_NOMEM, "hipMalloc(%zu bytes) failed: %s", p->ws_bytes, hipGetErrorString(e));
    }
    p->ws_owned = true;
  }
  if (p->setup_on_stream) {
    // stream-ordered behind the caller's earlier work on the same memory (a fill
    // kernel rather than a memset node, so a caller may capture it into a graph)
    const size_t n16 = p->ws_bytes / 16;  // the layout rounds every piece to 256 bytes
    const unsigned grid = (unsigned)std::min<size_t>((n16 + 255) / 256, 4096);
    hipLaunchKernelGGL(zero_kernel, dim3(std::max(grid, 1u)), dim3(256), 0, p->setup_stream,
                       (uint4*)p->ws, n16);
    HIP_TRY(hipGetLastError());
  } else {
    // set-up traffic goes through the null stream and is complete before this returns,
    // so kernels on any (also non-blocking) stream see it
    HIP_TRY(hipMemsetAsync(p->ws, 0, p->ws_bytes, nullptr));
    HIP_TRY(hipStreamSynchronize(nullptr));
  }
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
    for (int k = 0; k < 2; ++k) {
      G.dtab[k] = (bss::LevelSys*)(p->ws + p->toff[l].tab[l <= 1 ? k : 0]);
      if (p->mode == BSS_MODE_MIXED)
        G.dtab64[k] = (bss::LevelSys*)(p->ws + p->toff[l].tab64[l == 0 ? k : 0]);
    }
  }
  // the static set-up range, assembled on the host and uploaded in one copy
  SetupImage img(p->static_begin, p->static_end);
  for (size_t l = 0; l < p->gl.size(); ++l) {
    const GLevel& G = p->gl[l];
    img.put_at(p, G.dchains, G.chains.data(), G.chains.size() * sizeof(bss::ChainRec));
    img.put_at(p, G.dseps, G.seps.data(), G.seps.size() * sizeof(bss::SepRec));
  }
  // original (level-0) block index of every block, per system
  for (int s = 0; s < p->nsys; ++s) {
    SysPlan& S = p->sp[s];
    std::vector<int> orig(S.lv[0].shape.N);
    for (size_t i = 0; i < orig.size(); ++i) orig[i] = (int)i;
    for (size_t l = 0; l < S.lv.size(); ++l) {
      SysLevel& V = S.lv[l];
      img.put(V.oOrig, orig.data(), orig.size() * sizeof(int));
      std::vector<int> next(V.P);
      for (int k = 0; k < V.P; ++k) next[k] = orig[V.shape.seppos[k]];
      orig.swap(next);
      if (l + 1 < S.lv.size() && V.P > 0)  // next level's block q sits at seppos[q] here
        img.put(S.lv[l + 1].oPrev, V.shape.seppos.data(), V.P * sizeof(int));
    }
  }
  if (p->mode == BSS_MODE_MIXED) {
    // level-1 couplings FP32 -> FP64 and level-1 vectors FP64 -> FP32, per system
    p->dpiv = (unsigned*)(p->dinfo + 2 * p->nsys);
    p->dwiden = (bss::ConvRec<float, double>*)(p->ws + p->widen_off);
    p->dnarrow = (bss::ConvRec<double, float>*)(p->ws + p->narrow_off);
    std::vector<bss::ConvRec<float, double>> wr;
    std::vector<bss::ConvRec<double, float>> nr;
    const long long nn = (long long)p->n * p->n, nd = (long long)p->n * p->nrhs;
    for (int s = 0; s < p->nsys; ++s) {
      const SysPlan& S = p->sp[s];
      if (S.lv.size() < 2) continue;
      const SysLevel& V = S.lv[1];
      wr.push_back({(const float*)(p->ws + V.oB), (double*)(p->ws + V.wB),
                    (long long)std::max(V.shape.N - 1, 0) * nn});
      nr.push_back({(const double*)(p->ws + V.wd), (float*)(p->ws + V.od),
                    (long long)V.shape.N * nd});
    }
    p->nconv = (int)wr.size();
    img.put_at(p, p->dwiden, wr.data(), wr.size() * sizeof(wr[0]));
    img.put_at(p, p->dnarrow, nr.data(), nr.size() * sizeof(nr[0]));
    // kept from an earlier set-up that failed later on
    if (!p->host_piv) p->host_piv = (unsigned*)bss::pinned_take(sizeof(unsigned) * 3 * p->nsys);
    if (!p->host_piv) return fail(BSS_E_NOMEM, "pinned host allocation failed");
    if (!p->piv_event) HIP_TRY(hipEventCreateWithFlags(&p->piv_event, hipEventDisableTiming));
  }
  for (size_t l = 2; l < p->gl.size(); ++l) fill_tables(p, l, img);
  if (int e = img.upload(p)) return e;
  p->built = true;
  return BSS_OK;
}

// fwd0: level 0's kernel also runs that level's forward sweep of vector set 0 (one
// right-hand side, bss::factor_fuses_forward)
template <typename Tin, typename T>
int factor_levels(bss_plan* p, hipStream_t s, size_t l_begin = 0, size_t l_end = ~size_t(0),
                  bool fwd0 = false) {
  const int nb = p->n;
  const bool mixed = p->mode == BSS_MODE_MIXED;
  if (l_begin == 0) {
    hipLaunchKernelGGL(reset_status_kernel, dim3((2 * p->nsys + 255) / 256), dim3(256), 0, s,
                       p->dinfo, 2 * p->nsys, kInfoNone, mixed ? (int*)p->dpiv : nullptr,
                       p->nsys, kPivNone);
    HIP_TRY(hipGetLastError());
  }
  for (size_t l = l_begin; l < std::min(l_end, p->gl.size()); ++l) {
    GLevel& G = p->gl[l];
    // mixed mode: FP32 separator levels report into their own info words and feed the
    // pivot monitor that decides on an FP64 re-factorisation
    bss::FactorLevelArgs a{G.dchains, G.dtab[0], p->dinfo + ((mixed && l > 0) ? p->nsys : 0),
                           (mixed && l > 0) ? p->dpiv : nullptr};
    a.fwd = (fwd0 && l == 0) ? 1 : 0;
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
        e = bss::launch_assemble_diag<Tin, T, T>(G.dseps, G.dtab[0], nb, (int)G.seps.size(), s);
      else
        e = bss::launch_assemble_diag<T, T, T>(G.dseps, G.dtab[0], nb, (int)G.seps.size(), s);
      if (e != hipSuccess) return fail(BSS_E_HIP, "assemble level %zu: %s", l, hipGetErrorString(e));
    }
  }
  return BSS_OK;
}

// Mixed mode: factor the separator levels (>= 1) again in FP64 from the FP64 input
// separator blocks and the level-0 contributions, into the FP64 copies.
int refactor_separators64(bss_plan* p, hipStream_t s) {
  const int nb = p->n;
  GLevel& G0 = p->gl[0];
  hipError_t e = bss::launch_assemble_diag<double, float, double>(G0.dseps, G0.dtab64[0], nb,
                                                                 (int)G0.seps.size(), s);
  if (e == hipSuccess) e = bss::launch_convert<float, double>(p->dwiden, p->nconv, s);
  if (e != hipSuccess) return fail(BSS_E_HIP, "FP64 hand-off: %s", hipGetErrorString(e));
  for (size_t l = 1; l < p->gl.size(); ++l) {
    GLevel& G = p->gl[l];
    const bss::FactorLevelArgs a{G.dchains, G.dtab64[0], p->dinfo, nullptr};
    {
      ProfScope ps(p, BSS_K_FACTOR, (int)l, s);
      e = bss::launch_factor<double, double>(nb, p->factor_waves, a, (int)G.chains.size(), s);
    }
    if (e == hipSuccess && !G.seps.empty()) {
      ProfScope ps(p, BSS_K_ASSEMBLE_DIAG, (int)l, s);
      e = bss::launch_assemble_diag<double, double, double>(G.dseps, G.dtab64[0], nb,
                                                            (int)G.seps.size(), s);
    }
    if (e != hipSuccess) return fail(BSS_E_HIP, "FP64 separator level %zu: %s", l, hipGetErrorString(e));
  }
  return BSS_OK;
}

// Mixed mode, once per factorisation (synchronises s): inspect the FP32 separator
// pivots and re-factor the separator levels in FP64 when any system's smallest
// U_jj^2 / A_jj is below the pivot tolerance or an FP32 separator pivot failed.
int ensure_pivots(bss_plan* p, hipStream_t s) {
  if (p->mode != BSS_MODE_MIXED || !p->pivots_pending) return BSS_OK;
  p->pivots_pending = false;
  p->sep64 = false;
  p->refactored = 0;
  p->fp32_failed = false;
  HIP_TRY(bss::wait_event_spin(p->piv_event));
  for (int q = 0; q < p->nsys; ++q) p->fp32_failed |= (int)p->host_piv[q] != kInfoNone;
  // an FP32 level-0 failure cannot be refined away: the solve goes to the fallback
  if (p->fp32_failed || p->gl.size() < 2) return BSS_OK;
  bool need = false;
  for (int q = 0; q < p->nsys; ++q) {
    float r;
    memcpy(&r, &p->host_piv[2 * p->nsys + q], sizeof(r));
    if (!(r >= (float)p->pivot_tol) || (int)p->host_piv[p->nsys + q] != kInfoNone) need = true;
  }
  if (!need) return BSS_OK;
  if (int e = refactor_separators64(p, s)) return e;
  p->sep64 = true;
  p->refactored = (int)p->gl.size() - 1;
  return BSS_OK;
}

// Solve in place in the level-0 vectors of set k (plain: the caller's d; mixed: the
// FP32 ping-pong vector k).
template <typename T>
int solve_levels(bss_plan* p, int k, hipStream_t s, size_t l_begin = 0,
                 const int* skip0 = nullptr) {
  const int nb = p->n;
  const size_t L = p->gl.size();
  for (size_t l = l_begin; l < L; ++l) {
    GLevel& G = p->gl[l];
    bss::SolveLevelArgs a{G.dchains, G.dtab[k], p->nrhs};
    if (l == 0) a.skip = skip0;
    hipError_t e;
    {
      ProfScope ps(p, BSS_K_FORWARD, (int)l, s);
      e = bss::launch_forward<T>(nb, a, (int)G.chains.size(), s);
    }
    if (e != hipSuccess) return fail(BSS_E_HIP, "forward level %zu: %s", l, hipGetErrorString(e));
  }
  for (size_t li = L; li-- > 0;) {
    GLevel& G = p->gl[li];
    const bss::SolveLevelArgs a{G.dchains, G.dtab[k], p->nrhs};
    hipError_t e;
    {
      ProfScope ps(p, BSS_K_BACKWARD, (int)li, s);
      e = bss::launch_backward<T>(nb, a, (int)G.chains.size(), s);
    }
    if (e != hipSuccess) return fail(BSS_E_HIP, "backward level %zu: %s", li, hipGetErrorString(e));
  }
  return BSS_OK;
}

// Runs `enqueue` on stream s directly, or -- with graphs on and the profiler off --
// replays a graph captured from it on the plan's private capture stream (the caller's
// stream may be the legacy null stream, which cannot be captured).  Every kernel
// argument is a plan-owned address, so one capture stays valid across rebinding.
template <typename F>
int run_graph(bss_plan* p, hipGraphExec_t* ex, hipStream_t s, F enqueue) {
  if (!p->use_graph || p->prof) return enqueue(s);
  if (!*ex) {
    if (!p->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(p->cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(p->cap_stream, &g);
    if (rc != BSS_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (ec != hipSuccess) return fail(BSS_E_HIP, "graph capture: %s", hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      *ex = nullptr;
      return fail(BSS_E_HIP, "graph instantiate: %s", hipGetErrorString(ei));
    }
  }
  HIP_TRY(hipGraphLaunch(*ex, s));
  return BSS_OK;
}

// Mixed-mode FP32 correction solve after an FP64 separator re-factorisation: level 0
// in FP32 (vector set k), separator levels in FP64, FP32 <-> FP64 at the hand-off.
int solve_levels_sep64(bss_plan* p, int k, hipStream_t s) {
  const int nb = p->n;
  const int nd = p->n * p->nrhs;
  const size_t L = p->gl.size();
  GLevel& G0 = p->gl[0];
  const bss::SolveLevelArgs a0{G0.dchains, G0.dtab[k], p->nrhs};
  hipError_t e;
  {
    ProfScope ps(p, BSS_K_FORWARD, 0, s);
    e = bss::launch_forward<float>(nb, a0, (int)G0.chains.size(), s);
  }
  if (e == hipSuccess) {
    ProfScope ps(p, BSS_K_ASSEMBLE_RHS, 0, s);
    e = bss::launch_assemble_rhs<float, double>(G0.dseps, G0.dtab64[k], nd, (int)G0.seps.size(), s);
  }
  if (e != hipSuccess) return fail(BSS_E_HIP, "FP32 level of the FP64-separator solve: %s", hipGetErrorString(e));
  for (size_t l = 1; l < L; ++l) {
    GLevel& G = p->gl[l];
    const bss::SolveLevelArgs a{G.dchains, G.dtab64[0], p->nrhs};
    {
      ProfScope ps(p, BSS_K_FORWARD, (int)l, s);
      e = bss::launch_forward<double>(nb, a, (int)G.chains.size(), s);
    }
    if (e == hipSuccess && !G.seps.empty()) {
      ProfScope ps(p, BSS_K_ASSEMBLE_RHS, (int)l, s);
      e = bss::launch_assemble_rhs<double, double>(G.dseps, a.sys, nd, (int)G.seps.size(), s);
    }
    if (e != hipSuccess) return fail(BSS_E_HIP, "FP64 forward level %zu: %s", l, hipGetErrorString(e));
  }
  for (size_t l = L - 1; l >= 1; --l) {
    GLevel& G = p->gl[l];
    const bss::SolveLevelArgs a{G.dchains, G.dtab64[0], p->nrhs};
    {
      ProfScope ps(p, BSS_K_BACKWARD, (int)l, s);
      e = bss::launch_backward<double>(nb, a, (int)G.chains.size(), s);
    }
    if (e == hipSuccess && !G.seps.empty()) {
      ProfScope ps(p, BSS_K_SCATTER, (int)l, s);
      e = bss::launch_scatter_sep<double, double>(G.dseps, a.sys, nd, (int)G.seps.size(), s);
    }
    if (e != hipSuccess) return fail(BSS_E_HIP, "FP64 backward level %zu: %s", l, hipGetErrorString(e));
  }
  e = bss::launch_convert<double, float>(p->dnarrow, p->nconv, s);
  if (e == hipSuccess) {
    ProfScope ps(p, BSS_K_BACKWARD, 0, s);
    e = bss::launch_backward<float>(nb, a0, (int)G0.chains.size(), s);
  }
  if (e != hipSuccess) return fail(BSS_E_HIP, "FP32 back-substitution: %s", hipGetErrorString(e));
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
  if (n != 16 && n != 32 && n != 48 && n != 64 && n != 96 && n != 128)
    return fail(BSS_E_UNSUPPORTED, "block size n=%d not supported (16, 32, 48, 64, 96, 128)", n);
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

// Refinement-loop hooks over this plan.  The first solve32 after bss_factorize_solve
// (vector set 0, fwd_done set) skips the forward sweeps that already ran (level 0's
// inside its factor kernel when fused, the others on the side stream), and runs
// without a graph.  That first solve32 always runs on the FP32
// separators: the pivot check (prepare) follows it, and an FP64 separator
// re-factorisation there makes the loop redo it from d on the FP64 separators.
bss::MixedSolveHooks mixed_hooks(bss_plan* p) {
  bss::MixedSolveHooks hooks;
  hooks.ctx = p;
  // the first solve32 runs on the FP32 separators before the pivot check; an FP64
  // re-factorisation asks for it to be redone, a failed FP32 level-0 pivot for the
  // refinement to stop
  hooks.prepare = [](void* ctx, hipStream_t st) -> int {
    bss_plan* q = (bss_plan*)ctx;
    const bool was64 = q->sep64;
    const int rc = ensure_pivots(q, st);
    if (rc != BSS_OK) return rc;
    if (q->fp32_failed) return 2;
    return q->sep64 && !was64 ? 1 : 0;
  };
  hooks.solve32 = [](void* ctx, int k, hipStream_t st) -> int {
    bss_plan* q = (bss_plan*)ctx;
    const size_t skip = k == 0 ? q->fwd_done : 0;
    const int* fsweep = k == 0 ? q->fsweep_flag : nullptr;
    q->fwd_done = 0;
    q->fsweep_flag = nullptr;
    if (q->sep64)
      return run_graph(q, &q->g_solve64[k], st,
                       [q, k](hipStream_t c) { return solve_levels_sep64(q, k, c); });
    if (skip) return solve_levels<float>(q, k, st, skip);
    if (fsweep) return solve_levels<float>(q, k, st, 0, fsweep);
    return run_graph(q, &q->g_solve[k], st, [q, k](hipStream_t c) { return solve_levels<float>(q, k, c); });
  };
  hooks.prof_begin = [](void* ctx, int kind, hipStream_t st) -> void* {
    return new ProfScope((bss_plan*)ctx, kind, 0, st);
  };
  hooks.prof_end = [](void* scope) { delete (ProfScope*)scope; };
  return hooks;
}

size_t packed_vector_elems(const bss_plan* p) {
  size_t nd = 0;
  for (const SysPlan& S : p->sp) nd += (size_t)S.N_use * p->n * p->nrhs;
  return nd;
}

// Mixed mode, refinement stopped short: solve with a plain FP64 factorisation of the
// same systems (created and bound on first use, factored once per factorisation).  It
// solves in iterate set 0 (d copied there first, so the caller's d is the right-hand
// side of the backward-error pass), the answer's backward error is measured in FP64 by
// the refinement's own residual pass, and the answer goes to d.
int fallback_solve(bss_plan* p, hipStream_t s) {
  bss::MixedState& m = p->mx;
  double* x0 = bss::mixed_iterate(m, 0);
  if (!p->fb) {
    bss_plan* q = nullptr;
    if (int e = create(p->nsys, p->Ns.data(), p->n, p->nrhs, BSS_F64, BSS_MODE_PLAIN,
                       p->schedule, &q))
      return e;
    q->chain_len = p->chain_len;
    q->chain_len_default = p->chain_len_default;
    q->base_max = p->base_max;
    q->factor_waves = p->factor_waves;
    q->preserve = true;
    if (int e = bss_plan_bind(q, p->A, p->B, x0)) {
      bss_plan_destroy(q);
      return e;
    }
    p->fb = q;
    p->fb_factored = false;
  }
  const size_t nd = packed_vector_elems(p);
  HIP_TRY(hipMemcpyAsync(x0, p->d, nd * 8, hipMemcpyDeviceToDevice, s));
  if (!p->fb_factored) {
    if (int e = bss_factorize(p->fb, s)) return e;
    p->fb_factored = true;
  }
  if (int e = bss_solve(p->fb, s)) return e;
  double worst = 0.0;
  if (bss::mixed_measure(m, mixed_hooks(p), s, &worst) != 0)
    return fail(BSS_E_HIP, "FP64 fallback backward error: %s", bss::mixed_last_error());
  HIP_TRY(hipMemcpyAsync(p->d, bss::mixed_iterate(m, 1), nd * 8, hipMemcpyDeviceToDevice, s));
  return worst <= p->tol ? BSS_OK : BSS_NOT_CONVERGED;
}

int run_mixed_solve(bss_plan* p, hipStream_t s, bool first_pass_done) {
  if (p->fb_active) {
    p->last_iters = 0;
    p->fwd_done = 0;
    return fallback_solve(p, s);
  }
  const bss::MixedSolveHooks hooks = mixed_hooks(p);
  int iters = 0;
  bool converged = false;
  const int rc = bss::mixed_solve(p->mx, hooks, (double*)p->d, p->tol, p->max_iters, &iters,
                                  &converged, !p->fallback_on, s, first_pass_done);
  p->last_iters = iters;
  if (rc != 0) return fail(BSS_E_HIP, "mixed solve failed: %s", bss::mixed_last_error());
  if (converged) return BSS_OK;
  if (!p->fallback_on) return BSS_NOT_CONVERGED;
  p->fb_active = true;
  return fallback_solve(p, s);
}

// LAPACK-order info of system s: its blocks factored again as one chain from block 0
// (the sequential order of src/backbone_cpu.jl:5-15) from the preserved input, in FP64
// for a mixed plan, into a temporary plan's workspace; the chain reports its first
// failed block.  Only systems the recursive factorisation flagged come here.
int sequential_info(bss_plan* p, int s, hipStream_t st, int* out) {
  const int NU = p->sp[s].N_use;
  const size_t nn = (size_t)p->n * p->n, nd = (size_t)p->n * p->nrhs;
  size_t oa = 0, ob = 0, od = 0;
  for (int q = 0; q < s; ++q) {
    oa += (size_t)p->sp[q].N_use * nn;
    ob += (size_t)std::max(p->sp[q].N_use - 1, 0) * nn;
    od += (size_t)p->sp[q].N_use * nd;
  }
  const size_t is = isize(p);
  const int dt = p->mode == BSS_MODE_MIXED ? BSS_F64 : p->dtype;
  bss_plan* q = nullptr;
  if (int e = create(1, &NU, p->n, 1, dt, BSS_MODE_PLAIN, BSS_SCHED_AUTO, &q)) return e;
  q->base_max = NU;  // one level: a single chain over all blocks
  q->preserve = true;
  // the chain's factorisation reads A and B only; d is bound because a plan needs one
  int rc = bss_plan_bind(q, (char*)p->A + oa * is,
                         NU > 1 ? (void*)((char*)p->B + ob * is) : nullptr,
                         (char*)p->d + od * is);
  if (rc == BSS_OK) rc = bss_factorize(q, st);
  int v = kInfoNone;
  if (rc == BSS_OK) {
    const hipError_t e1 = hipStreamSynchronize(st);
    const hipError_t e2 = e1 == hipSuccess ? hipMemcpy(&v, q->dinfo, sizeof(int),
                                                       hipMemcpyDeviceToHost)
                                           : e1;
    if (e2 != hipSuccess)
      rc = fail(BSS_E_HIP, "sequential re-check: %s", hipGetErrorString(e2));
  }
  bss_plan_destroy(q);
  *out = v == kInfoNone ? 0 : v;
  return rc;
}

}  // namespace

namespace bss {
int api_fail(int code, const char* msg) { return fail(code, "%s", msg); }

// The pieces of bss_factorize / bss_solve, for callers inside the library that enqueue
// several plans' kernels into one graph (batch.hip): host-side state before the
// kernels, the kernel sequence itself (no host copies, so it can be captured), and the
// host-visible bookkeeping after it.
int plan_factor_prepare(bss_plan* p, void* stream) {
  if (int e = check_plan(p)) return e;
  if (!p->A) return fail(BSS_E_STATE, "bss_plan_bind must precede bss_factorize");
  p->last_stream = (hipStream_t)stream;
  p->info_valid = false;
  p->fb_active = p->fb_factored = false;
  p->fsweep_pending = false;
  return BSS_OK;
}

bool plan_factor_sweeps(const bss_plan* p) {
  return p->fsweep_on && p->mode == BSS_MODE_MIXED && p->mx.dsnap &&
         factor_fuses_forward(p->n, true, p->nrhs);
}

int plan_factor_kernels(bss_plan* p, hipStream_t c) {
  if (plan_factor_sweeps(p)) {
    // BSS_OPT_FACTOR_SWEEP: d kept in dsnap, the first correction's x = 0 pass, and the
    // level-0 factor kernel running its level-0 forward sweep (kernels only: capturable)
    if (mixed_snapshot_d(p->mx, c) != 0 || mixed_first_pass(p->mx, mixed_hooks(p), c) != 0)
      return fail(BSS_E_HIP, "factor-time first pass: %s", mixed_last_error());
    return factor_levels<double, float>(p, c, 0, ~size_t(0), true);
  }
  if (p->mode == BSS_MODE_MIXED) return factor_levels<double, float>(p, c);
  if (p->dtype == BSS_F64) return factor_levels<double, double>(p, c);
  return factor_levels<float, float>(p, c);
}

int plan_factor_finish(bss_plan* p, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (p->mode == BSS_MODE_MIXED) {
    // FP32 info words and pivot ratios to pinned memory; read at the first solve
    HIP_TRY(hipMemcpyAsync(p->host_piv, p->dinfo, sizeof(int) * 3 * p->nsys,
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipEventRecord(p->piv_event, s));
    bss::mixed_after_factor(p->mx);
    p->pivots_pending = true;
    p->sep64 = false;
  }
  p->factored = true;
  // plan_factor_kernels ran the first correction's level-0 sweep: the next solve checks d
  p->fsweep_pending = plan_factor_sweeps(p);
  return BSS_OK;
}

// plain mode: the solve's kernel sequence
int plan_solve_kernels(bss_plan* p, hipStream_t c) {
  if (p->dtype == BSS_F64) return solve_levels<double>(p, 0, c);
  return solve_levels<float>(p, 0, c);
}
}  // namespace bss

// Workspace of bss_factorize_solve_noalloc: a file-scope device array, one copy on every
// GPU the library runs on, held until the process ends.  In host code its name is not a
// device address; each device's copy is located once with hipGetSymbolAddress and only
// that address is used (kernel arguments, copies).
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch` (its address looked up on first use there).
int scratch_of_current_device(char** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(BSS_E_UNSUPPORTED, "device %d", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(BSS_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_scratch_ptr[dev] = (char*)ptr;
  }
  *out = g_scratch_ptr[dev];
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
      p->drop_graphs();
      return BSS_OK;
    case BSS_OPT_GRAPH:
      p->use_graph = value != 0.0;
      if (!p->use_graph) p->drop_graphs();
      return BSS_OK;
    case BSS_OPT_BASE_MAX:
      if (value < 1 || value > 1 << 20) return fail(BSS_E_ARG, "base_max must be >= 1");
      p->base_max = (int)value;
      return BSS_OK;
    case BSS_OPT_PRESERVE_INPUT:
      p->preserve = value != 0.0;
      return BSS_OK;
    case BSS_OPT_PROFILE_KINDS:
      if (value < 0 || value >= (double)(1u << BSS_K_COUNT) || value != (double)(unsigned)value)
        return fail(BSS_E_ARG, "profile kinds: a bit mask below 2^%d", BSS_K_COUNT);
      p->prof_kinds = (unsigned)value;
      return BSS_OK;
    case BSS_OPT_PROFILE_MAX_LEVEL:
      if (value < 0) return fail(BSS_E_ARG, "profile max level must be >= 0");
      p->prof_max_level = value > (double)(1 << 30) ? 1 << 30 : (int)value;
      return BSS_OK;
    case BSS_OPT_FALLBACK:
      p->fallback_on = value != 0.0;
      return BSS_OK;
    case BSS_OPT_FACTOR_SWEEP:
      p->fsweep_on = value != 0.0;
      if (!p->fsweep_on) p->fsweep_pending = false;
      p->drop_graphs();  // the factorisation's launch sequence changes
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
  } else if (query == BSS_Q_FALLBACK) {
    *value = p->fb_active ? 1 : 0;
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
  if (p->fb) {  // the FP64 fallback plan names the previous buffers
    bss_plan_destroy(p->fb);
    p->fb = nullptr;
  }
  p->fb_active = p->fb_factored = false;
  p->info_valid = false;
  if (p->mode == BSS_MODE_MIXED && p->mx.bind((const double*)A, (const double*)B, (const double*)d) != 0)
    return fail(BSS_E_HIP, "mixed bind: %s", bss::mixed_last_error());
  // levels 0 and 1 name the caller's buffers (level 1 reads level 0's vectors)
  if (int e = upload_bound_tables(p)) return e;
  HIP_TRY(hipStreamSynchronize(nullptr));
  p->factored = false;
  return BSS_OK;
}

int bss_factorize(bss_plan_t p, void* stream) {
  if (int e = bss::plan_factor_prepare(p, stream)) return e;
  hipStream_t s = (hipStream_t)stream;
  int rc = run_graph(p, &p->g_factor, s,
                     [p](hipStream_t c) { return bss::plan_factor_kernels(p, c); });
  if (rc != BSS_OK) return rc;
  return bss::plan_factor_finish(p, stream);
}

int bss_solve(bss_plan_t p, void* stream) {
  if (int e = check_plan(p)) return e;
  if (!p->factored) return fail(BSS_E_STATE, "bss_factorize must precede bss_solve");
  hipStream_t s = (hipStream_t)stream;
  p->last_stream = s;
  p->fwd_done = 0;
  if (p->mode == BSS_MODE_MIXED) {
    const bool fsweep = p->fsweep_pending && !p->fb_active;
    p->fsweep_pending = false;
    if (fsweep) {
      // d unchanged since bss_factorize, bit for bit, or the x = 0 pass and the level-0
      // forward sweep run again (the flag is read on the device: no host wait)
      const long long nd = (long long)packed_vector_elems(p);
      const int grid = (int)std::min<long long>((nd + 255) / 256, 4096);
      hipLaunchKernelGGL(words_differ_kernel, dim3(grid), dim3(256), 0, s,
                         (const unsigned long long*)p->d,
                         (const unsigned long long*)p->mx.dsnap, nd, p->mx.dflag);
      HIP_TRY(hipGetLastError());
      p->mx.init_skip = p->mx.dflag;
      p->fsweep_flag = p->mx.dflag;
    }
    const int rc = run_mixed_solve(p, s, false);
    p->mx.init_skip = nullptr;
    p->fsweep_flag = nullptr;
    return rc;
  }
  return run_graph(p, &p->g_solve[0], s,
                   [p](hipStream_t c) { return bss::plan_solve_kernels(p, c); });
}

int bss_factorize_solve(bss_plan_t p, void* stream) {
  if (int e = check_plan(p)) return e;
  if (!p->A) return fail(BSS_E_STATE, "bss_plan_bind must precede bss_factorize_solve");
  if (p->use_graph || p->gl.size() < 2) {
    if (int e = bss_factorize(p, stream)) return e;
    return bss_solve(p, stream);
  }
  hipStream_t s = (hipStream_t)stream;
  p->last_stream = s;
  p->factored = false;
  p->fwd_done = 0;
  p->fsweep_pending = false;
  p->info_valid = false;
  p->fb_active = p->fb_factored = false;
  if (!p->side_stream) {
    HIP_TRY(hipStreamCreateWithFlags(&p->side_stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&p->ev_l0_factored, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&p->ev_l0_solved, hipEventDisableTiming));
  }
  const size_t L = p->gl.size();
  while (p->ev_level.size() < L) {
    hipEvent_t ev;
    HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    p->ev_level.push_back(ev);
  }
  const bool mixed = p->mode == BSS_MODE_MIXED;
  const bool f64 = !mixed && p->dtype == BSS_F64;
  // fuse: the level-0 factor kernel runs level 0's forward sweep itself, on the factors
  // while they are in LDS (one right-hand side; the sweep kernel's arithmetic, so the
  // same bits), which saves the sweep's pass over the level-0 factors
  const bool fuse = bss::factor_fuses_forward(p->n, !f64, p->nrhs);
  auto factor = [&](size_t lb, size_t le) {
    if (mixed) return factor_levels<double, float>(p, s, lb, le, fuse);
    if (f64) return factor_levels<double, double>(p, s, lb, le, fuse);
    return factor_levels<float, float>(p, s, lb, le, fuse);
  };
  // mixed: the x = 0 residual pass writes r = d in FP32 (the first solve's right-hand
  // side), before the level-0 kernel reads it when that kernel runs the sweep
  const bss::MixedSolveHooks hooks = mixed_hooks(p);
  if (mixed && fuse && bss::mixed_first_pass(p->mx, hooks, s) != 0)
    return fail(BSS_E_HIP, "first residual pass: %s", bss::mixed_last_error());
  // levels 0 and 1 with their hand-offs on the caller's stream (ordered after the
  // caller's writes of A, B and d); the side stream starts when they are done.  Level 1
  // still fills the device, so the side stream's sweeps (levels 0 -- unless fused -- and
  // 1, then each later level once it is factored) only have idle compute units to use
  // beside the later, latency-bound levels.
  const size_t split = 2;
  if (int e = factor(0, split)) return e;
  hipStream_t s2 = p->side_stream;
  HIP_TRY(hipEventRecord(p->ev_l0_factored, s));
  HIP_TRY(hipStreamWaitEvent(s2, p->ev_l0_factored, 0));
  // side stream: the first solve's forward sweeps of the factored levels beside the
  // later levels' factorisation
  if (mixed && !fuse && bss::mixed_first_pass(p->mx, hooks, s2) != 0)
    return fail(BSS_E_HIP, "first residual pass: %s", bss::mixed_last_error());
  hipError_t e = hipSuccess;
  for (size_t l = fuse ? 1 : 0; l < split && e == hipSuccess; ++l) {
    GLevel& G = p->gl[l];
    const bss::SolveLevelArgs a{G.dchains, G.dtab[0], p->nrhs};
    ProfScope ps(p, BSS_K_FORWARD, (int)l, s2);
    e = f64 ? bss::launch_forward<double>(p->n, a, (int)G.chains.size(), s2)
            : bss::launch_forward<float>(p->n, a, (int)G.chains.size(), s2);
  }
  // each later level: factored on the caller's stream, then its forward sweep on the
  // side stream while the next level is factored
  int rc = BSS_OK;
  for (size_t l = split; l < L && rc == BSS_OK && e == hipSuccess; ++l) {
    rc = factor(l, l + 1);
    if (rc != BSS_OK) break;
    e = hipEventRecord(p->ev_level[l], s);
    if (e == hipSuccess) e = hipStreamWaitEvent(s2, p->ev_level[l], 0);
    if (e != hipSuccess) break;
    GLevel& G = p->gl[l];
    const bss::SolveLevelArgs a{G.dchains, G.dtab[0], p->nrhs};
    ProfScope ps(p, BSS_K_FORWARD, (int)l, s2);
    e = f64 ? bss::launch_forward<double>(p->n, a, (int)G.chains.size(), s2)
            : bss::launch_forward<float>(p->n, a, (int)G.chains.size(), s2);
  }
  // mixed: FP32 info words and pivot ratios to pinned memory, as bss_factorize does,
  // before the join (the host's pivot check waits for the factorisation only)
  hipError_t ep = hipSuccess;
  if (mixed && rc == BSS_OK) {
    ep = hipMemcpyAsync(p->host_piv, p->dinfo, sizeof(int) * 3 * p->nsys,
                        hipMemcpyDeviceToHost, s);
    if (ep == hipSuccess) ep = hipEventRecord(p->piv_event, s);
  }
  // the join is recorded and awaited even after a failure, so that the caller's stream
  // never runs ahead of side-stream work
  const hipError_t ej = hipEventRecord(p->ev_l0_solved, s2);
  if (ej == hipSuccess) HIP_TRY(hipStreamWaitEvent(s, p->ev_l0_solved, 0));
  if (e != hipSuccess) return fail(BSS_E_HIP, "side-stream forward sweep: %s", hipGetErrorString(e));
  if (ej != hipSuccess) return fail(BSS_E_HIP, "side-stream join: %s", hipGetErrorString(ej));
  if (rc != BSS_OK) return rc;
  if (ep != hipSuccess) return fail(BSS_E_HIP, "pivot copy: %s", hipGetErrorString(ep));
  if (mixed) {
    bss::mixed_after_factor(p->mx);
    p->pivots_pending = true;
    p->sep64 = false;
  }
  p->factored = true;
  p->fwd_done = L;
  int rs;
  if (mixed)
    rs = run_mixed_solve(p, s, true);
  else
    rs = f64 ? solve_levels<double>(p, 0, s, L) : solve_levels<float>(p, 0, s, L);
  p->fwd_done = 0;
  return rs;
}

int bss_factorize_solve_noalloc(int N, int n, int nrhs, int dtype, const void* A, const void* B,
                                void* d, void* stream) {
  if (N < 1) return fail(BSS_E_ARG, "N must be >= 1 (got %d)", N);
  if (!A || !d || (N > 1 && !B)) return fail(BSS_E_ARG, "null buffer");
  char* base = nullptr;
  if (int e = scratch_of_current_device(&base)) return e;
  hipStream_t s = (hipStream_t)stream;
  // a plan on the stack: schedule and layout on the host, every device byte in scratch
  bss_plan P;
  P.Ns.assign(1, N);
  P.n = n;
  P.nrhs = nrhs;
  P.nsys = 1;
  P.dtype = dtype;
  P.mode = BSS_MODE_PLAIN;
  P.schedule = BSS_SCHED_AUTO;
  if (n != 16 && n != 32 && n != 48 && n != 64 && n != 96 && n != 128)
    return fail(BSS_E_UNSUPPORTED, "block size n=%d not supported (16, 32, 48, 64, 96, 128)", n);
  if (nrhs < 1 || nrhs > 16) return fail(BSS_E_UNSUPPORTED, "nrhs=%d not in [1, 16]", nrhs);
  if (dtype != BSS_F32 && dtype != BSS_F64) return fail(BSS_E_ARG, "bad dtype %d", dtype);
  if (hipGetDevice(&P.device) != hipSuccess) P.device = 0;
  if (int e = layout(&P)) return e;
  if (P.ws_bytes > kScratchBytes)
    return fail(BSS_E_ARG,
                "N=%d, n=%d, nrhs=%d needs %zu bytes of workspace; the allocation-free entry "
                "point has %zu (use a plan)", N, n, nrhs, P.ws_bytes, kScratchBytes);
  P.ws = base;
  P.ws_owned = false;
  P.setup_on_stream = true;
  P.setup_stream = s;
  P.A = (void*)A;
  P.B = (void*)B;
  P.d = d;
  P.last_stream = s;
  if (int e = build(&P)) return e;  // clears scratch, uploads the static tables
  if (int e = upload_bound_tables(&P)) return e;
  int rc = dtype == BSS_F64 ? factor_levels<double, double>(&P, s)
                            : factor_levels<float, float>(&P, s);
  if (rc == BSS_OK)
    rc = dtype == BSS_F64 ? solve_levels<double>(&P, 0, s) : solve_levels<float>(&P, 0, s);
  return rc;
}

int bss_info(bss_plan_t p, int* info_host) {
  if (int e = check_plan(p)) return e;
  if (!info_host) return fail(BSS_E_ARG, "info_host is null");
  if (!p->built || !p->factored) return fail(BSS_E_STATE, "no factorization yet");
  if (!p->info_valid) {
    if (int e = ensure_pivots(p, p->last_stream)) return e;
    HIP_TRY(hipStreamSynchronize(p->last_stream));
    std::vector<int> info(2 * p->nsys);
    HIP_TRY(hipMemcpy(info.data(), p->dinfo, sizeof(int) * 2 * p->nsys, hipMemcpyDeviceToHost));
    p->info_exact.assign(p->nsys, 0);
    for (int s = 0; s < p->nsys; ++s) {
      // FP32 separator failures count unless the separator levels were re-factored in
      // FP64 (whose failures land in the first word, with level 0's)
      int v = info[s];
      if (p->mode == BSS_MODE_MIXED && !p->sep64) v = std::min(v, info[p->nsys + s]);
      if (v == kInfoNone) continue;
      // the recursive order meets a failure at some block of a non-SPD system; the
      // LAPACK-order index needs the sequential re-check (preserved input only)
      if (inplace0(p))
        p->info_exact[s] = v;
      else if (int e = sequential_info(p, s, p->last_stream, &p->info_exact[s]))
        return e;
    }
    p->info_valid = true;
  }
  for (int s = 0; s < p->nsys; ++s) info_host[s] = p->info_exact[s];
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

int bss_plan_set_workspace(bss_plan_t p, void* ws, size_t bytes) {
  if (int e = check_plan(p)) return e;
  if (p->built) return fail(BSS_E_STATE, "the workspace must be set before bind");
  if (int e = layout(p)) return e;
  if (!ws) return fail(BSS_E_ARG, "workspace is null");
  if (((uintptr_t)ws & 255) != 0) return fail(BSS_E_ARG, "workspace must be 256-byte aligned");
  if (bytes < p->ws_bytes)
    return fail(BSS_E_ARG, "workspace of %zu bytes is smaller than the %zu the plan needs", bytes,
                p->ws_bytes);
  p->ws = (char*)ws;
  p->ws_owned = false;
  return BSS_OK;
}

int bss_plan_destroy(bss_plan_t p) {
  if (!p) return BSS_OK;
  if (p->fb) bss_plan_destroy(p->fb);
  if (p->ws) {
    (void)hipStreamSynchronize(p->last_stream);
    if (p->ws_owned) (void)hipFree(p->ws);
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
