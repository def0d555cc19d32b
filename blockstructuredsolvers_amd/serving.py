"""Request loop: continuous batching of independent block-tridiagonal systems on one
GPU, admitted by HBM footprint, with factors kept resident for repeated solves.

The reference has no serving layer: its callers factorize and solve one data object at
a time (benchmark/benchmark_kf.jl:227-331 times one smoother system per call,
test/test_rocm.jl:95-107 one system per test).  This module is the request loop
around the same plans:

* ``Request`` -- one system (A, B in pinned host memory, reference block layout) with
  1..8 right-hand sides to solve against one factorisation, and an arrival time.
* ``Server.serve`` releases requests at their arrival times into a FIFO.  The head of
  the FIFO is admitted as soon as (a) a worker is free and (b) its device footprint --
  plan workspace + A, B, d, as the plan reports it before any allocation -- fits in the
  HBM budget beside the requests in flight.  A coalesced group is admitted by the sum of
  its systems' footprints at the group's width, an upper bound of its group plan's.  There is no batch boundary: a request
  that arrives while others run starts as soon as a slot and memory free up.
* Requests are coalesced: when a worker is free, the FIFO head is admitted together
  with every other waiting request of the same block size n and column class (up to
  ``max_group`` requests and the HBM budget), and the group runs as one group plan
  (``bss_plan_create_group``), so its systems share every launch.  A request's columns
  are its right-hand sides side by side (``width`` = count x nrhs).  Requests of one
  column go to one-column plans (the single-column sweeps); wider requests go to plans
  as wide as the widest request of the group, at most 16 columns, and a narrower
  request fills its spare columns with copies of its first one.
* Each worker thread owns one HIP stream and runs its group end to end: blocks
  uploaded (or built on the device from the model), one ``factorize`` of all systems,
  then one ``solve`` of every request's columns at once (rounds of 16 columns for
  wider requests), and the downloads.  The factors stay resident between rounds.  A
  group with a single round uploads its right-hand sides before the factorisation,
  which in mixed mode with one column then runs the first correction's level-0 sweep
  (BSS_OPT_FACTOR_SWEEP); a solve's device time then excludes that sweep.
  Work of different workers overlaps on the device (ctypes releases the GIL during
  library calls).  Plans take their workspace from torch's caching allocator, so
  admission and release make no device allocation or free.
* Every system runs the level schedule ``SERVE_SCHEDULE`` (long level-0 chains) unless
  the server is given one; admission footprints are queried on that schedule.
* Latency of a request = host time its last solution is available - arrival time; a
  solve's latency is the device time of the solve that produced it.

A system's arithmetic in a group plan is the arithmetic it gets alone (its own level
schedule, separate launch slots, no cross-system reductions), and the multi-column
sweeps compute every column on its own, whatever the width, so a served answer equals
the answer of the same request solved alone (``solve_alone``; bitwise for plain plans,
within the refinement tolerance in mixed mode, where a group refines until all its
systems meet the target).
"""
from __future__ import annotations

import ctypes
import functools
import queue
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _lib as L
from .solver import (_dtype_code, build_smoother, factorize, initialize_batched,
                     initialize_batched_group, solve)


@dataclass
class Request:
    """One system and its right-hand sides.  The blocks come either as dense arrays (A,
    B) or as a state-space model (``model``, a ``workloads.SmootherModel``) whose blocks
    the device builds (``bss_build_smoother``), so they never cross PCIe."""
    rid: int
    arrival: float                      # seconds after the start of serve()
    rhs: List[np.ndarray]               # each (n, nrhs, N)
    A: Optional[np.ndarray] = None      # (n, n, N) column-major blocks
    B: Optional[np.ndarray] = None      # (n, n, N-1)
    model: Optional[object] = None

    @property
    def N(self) -> int:
        return self.rhs[0].shape[2]

    @property
    def n(self) -> int:
        return self.rhs[0].shape[0]

    @property
    def nrhs(self) -> int:
        return self.rhs[0].shape[1]

    @property
    def width(self) -> int:
        """Columns to solve: the right-hand sides side by side."""
        return len(self.rhs) * self.nrhs

    def dense(self):
        """(A, B) as host arrays (from the model when the request carries one)."""
        if self.A is not None:
            return self.A, self.B
        Ai, Bi = self.model.blocks()
        A = np.asfortranarray(np.broadcast_to(Ai[:, :, None], (self.n, self.n, self.N)))
        B = np.asfortranarray(np.broadcast_to(Bi[:, :, None], (self.n, self.n, max(self.N - 1, 0))))
        return A, B


@dataclass
class Result:
    rid: int
    x: List[np.ndarray]                 # one (n, nrhs, N) solution per right-hand side
                                        # (views of a pinned host buffer)
    arrival: float
    admitted: float
    finished: float
    solve_ms: List[float] = field(default_factory=list)  # per right-hand side: device
                                                         # time of the solve that made it
    info: int = 0

    @property
    def latency_ms(self) -> float:
        return (self.finished - self.arrival) * 1e3


MAX_COLS = 16  # columns one solve takes (the plans' nrhs limit)


def solve_width(width: int) -> int:
    """Plan width for a request of `width` columns: 1 for a single column (the
    single-column sweeps), otherwise the columns up to 16 per solve (the multi-column
    sweeps)."""
    return 1 if width == 1 else min(MAX_COLS, width)


def group_width(group: List["Request"]) -> int:
    """Plan width of a coalesced group (all one-column requests, or all wider)."""
    return solve_width(max(r.width for r in group))


def _storage_order(blocks: np.ndarray) -> torch.Tensor:
    """(r, c, i) column-major blocks -> pinned (i, c, r) C-order tensor, which is the
    plan tensors' storage order (one contiguous DMA per upload)."""
    return torch.from_numpy(np.ascontiguousarray(np.transpose(blocks, (2, 1, 0)))).pin_memory()


# The request loop's level schedule for every system (unless the server is given schedule
# options): level-0 chains of 31 blocks, 3 below, base 4.  On the served smoother stream
# (n in {16, 32, 64}, N ~ U[128, 4096], all requests present at once) it raised the
# served rate from 1,753 to 2,449 (plain) and 1,831 to 2,385 (mixed) systems/s against
# the library default of chains of 3 and base 8.  ``solve_alone`` uses it too, so a
# served answer still equals the request solved alone.
SERVE_SCHEDULE = {"chain_len_levels": [31] + [3] * 15, "base_max": 4}
_SCHEDULE_KEYS = ("chain_len", "chain_len_levels", "base_max")


def _with_schedule(plan_options: dict) -> dict:
    """`plan_options`, with SERVE_SCHEDULE unless they name a schedule themselves."""
    if any(plan_options.get(k) is not None for k in _SCHEDULE_KEYS):
        return dict(plan_options)
    return {**plan_options, **SERVE_SCHEDULE}


def _schedule_key(schedule: Optional[dict]) -> tuple:
    schedule = schedule or {}
    levels = schedule.get("chain_len_levels")
    return (schedule.get("chain_len"), tuple(levels) if levels is not None else None,
            schedule.get("base_max"))


def _apply_schedule(plan, key: tuple) -> None:
    chain_len, levels, base_max = key
    if chain_len is not None:
        L.call("bss_set_option", plan, L.BSS_OPT_CHAIN_LEN, float(chain_len))
    if base_max is not None:
        L.call("bss_set_option", plan, L.BSS_OPT_BASE_MAX, float(base_max))
    for k, m in enumerate(list(levels or [])[:16]):
        L.call("bss_set_option", plan, L.BSS_OPT_CHAIN_LEN_LEVEL + k, float(m))


@functools.lru_cache(maxsize=4096)
def _workspace_bytes(N: int, n: int, nrhs: int, dcode: int, mcode: int,
                     sched: tuple = (None, None, None)) -> int:
    p = ctypes.c_void_p()
    L.call("bss_plan_create", N, n, nrhs, 1, dcode, mcode, L.BSS_SCHED_AUTO, ctypes.byref(p))
    try:
        _apply_schedule(p, sched)
        v = ctypes.c_int64()
        L.call("bss_plan_get", p, L.BSS_Q_WORKSPACE_BYTES, ctypes.byref(v))
        return int(v.value)
    finally:
        L.lib().bss_plan_destroy(p)


def request_footprint(N: int, n: int, nrhs: int, dtype=torch.float64, mode: str = "plain",
                      schedule: Optional[dict] = None) -> int:
    """Device bytes a request holds while in flight: the plan's workspace (queried from a
    host-only plan on `schedule`, a dict of ``chain_len``/``chain_len_levels``/
    ``base_max``, or the library default; nothing allocated) plus its A, B and d."""
    mcode = L.BSS_MODE_MIXED if mode == "mixed" else L.BSS_MODE_PLAIN
    ws = _workspace_bytes(N, n, nrhs, _dtype_code(dtype), mcode, _schedule_key(schedule))
    es = 8 if dtype == torch.float64 else 4
    return ws + 256 + es * n * (2 * N * n + N * nrhs)


def _storage_view(t: torch.Tensor) -> torch.Tensor:
    """The (blocks, cols, rows) C-order storage behind a per-system (rows, cols, blocks)
    view of a plan tensor."""
    return t.permute(2, 1, 0)


def _model_host(model) -> torch.Tensor:
    """F, H, Q^-1, R^-1 of a state-space model, each column-major, packed one after the
    other into one pinned FP64 buffer (one host-to-device copy per request)."""
    parts = [np.asarray(M, dtype=np.float64).ravel(order="F")
             for M in (model.F, model.H, model.Qinv, model.Rinv)]
    return torch.from_numpy(np.concatenate(parts)).pin_memory()


def _build_smoother_into(data, A: torch.Tensor, B: torch.Tensor, req: Request,
                         host: Optional[torch.Tensor] = None) -> None:
    """bss_build_smoother into one system's A, B views of a group plan, on the current
    stream.  `host` is the model packed by ``_model_host`` (packed here when absent);
    its device copy stays referenced by `data` until the plan is closed."""
    n, m = req.n, req.model.H.shape[0]
    dev = data.device
    if host is None:
        host = _model_host(req.model)
    buf = host.to(dev, non_blocking=True)
    base, es = buf.data_ptr(), buf.element_size()
    oF, oH = 0, n * n
    oQ = oH + m * n
    oR = oQ + n * n
    code = L.BSS_F64 if data.dtype == torch.float64 else L.BSS_F32
    L.call("bss_build_smoother", req.N, n, m, base + oF * es, base + oH * es, base + oQ * es,
           base + oR * es, A.data_ptr(), B.data_ptr() if req.N > 1 else None, code,
           torch.cuda.current_stream(dev).cuda_stream)
    keep = getattr(data, "_model_tensors", [])
    keep.append(buf)
    data._model_tensors = keep


def group_footprint(Ns: List[int], n: int, nrhs: int, dtype=torch.float64,
                    mode: str = "plain", schedule: Optional[dict] = None) -> int:
    """Device bytes a coalesced group holds while in flight: its group plan's workspace
    (host-only query on `schedule`, as for ``request_footprint``) plus the systems' A, B
    and d."""
    mcode = L.BSS_MODE_MIXED if mode == "mixed" else L.BSS_MODE_PLAIN
    p = ctypes.c_void_p()
    arr = (ctypes.c_int * len(Ns))(*Ns)
    L.call("bss_plan_create_group", len(Ns), arr, n, nrhs, _dtype_code(dtype), mcode,
           L.BSS_SCHED_AUTO, ctypes.byref(p))
    try:
        _apply_schedule(p, _schedule_key(schedule))
        v = ctypes.c_int64()
        L.call("bss_plan_get", p, L.BSS_Q_WORKSPACE_BYTES, ctypes.byref(v))
        ws = int(v.value)
    finally:
        L.lib().bss_plan_destroy(p)
    es = 8 if dtype == torch.float64 else 4
    return ws + 256 + sum(es * n * (2 * N * n + N * nrhs) for N in Ns)


class _Staged:
    """A request's host data in pinned memory, prepared before it is served: its dense
    blocks, or its packed model when the device builds the blocks; its right-hand sides;
    and one download buffer per right-hand side."""

    def __init__(self, req: Request, dtype):
        npt = np.float64 if dtype == torch.float64 else np.float32
        self.A = self.B = None
        self.model = _model_host(req.model) if req.model is not None else None
        if req.model is None:
            self.A = _storage_order(req.A.astype(npt))
            self.B = _storage_order(req.B.astype(npt))
        # all columns side by side: (n, width, N) blocks in (N, width, n) storage order
        self.d = _storage_order(np.concatenate([d.astype(npt) for d in req.rhs], axis=1))
        self.out = torch.empty(self.d.shape, dtype=self.d.dtype, pin_memory=True)


class Server:
    """Continuous-batching request loop over libbss plans on one device.

    ``mode``/``dtype`` apply to every request; ``hbm_budget`` (bytes) bounds the summed
    footprint of the requests in flight (default: 80% of the device's free memory);
    ``workers`` is the number of groups in flight at most (one HIP stream each) and
    ``max_group`` the most requests coalesced into one group plan (1: no coalescing).
    ``plan_options`` go to every group plan; without schedule options the plans run
    ``SERVE_SCHEDULE``."""

    def __init__(self, mode: str = "plain", dtype=torch.float64, device="cuda",
                 hbm_budget: Optional[int] = None, workers: int = 8, max_group: int = 64,
                 **plan_options):
        if not torch.cuda.is_available():
            raise RuntimeError("Server needs a HIP device; none is visible")
        self.mode = mode
        self.dtype = dtype
        self.device = torch.device(device)
        self.workers = workers
        self.max_group = max(1, int(max_group))
        self.plan_options = _with_schedule(plan_options)
        self.schedule = {k: self.plan_options.get(k) for k in _SCHEDULE_KEYS}
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.8 * free)
        self.hbm_budget = int(hbm_budget)
        self.stats = {"max_inflight_bytes": 0, "max_inflight": 0, "groups": 0,
                      "max_group": 0}

    # -- admission ---------------------------------------------------------------------
    def footprint(self, r: Request) -> int:
        """Device bytes request `r` holds when it runs alone."""
        return request_footprint(r.N, r.n, solve_width(r.width), self.dtype, self.mode,
                                 self.schedule)

    def _group_fp(self, group: List[Request], width: Optional[int] = None) -> int:
        """Admission bound of a group: the sum of its systems' single-system footprints
        at the group's plan width.  A group plan lays out each system's pieces as a
        single-system plan does and rounds the pieces it shares once, so its workspace
        (``group_footprint``) never exceeds this sum, and the bound needs no plan layout
        (cached per N and width)."""
        P = group_width(group) if width is None else width
        return sum(request_footprint(r.N, r.n, P, self.dtype, self.mode, self.schedule)
                   for r in group)

    @staticmethod
    def _same_class(a: Request, b: Request) -> bool:
        return a.n == b.n and (a.width > 1) == (b.width > 1)

    # -- one coalesced group on one worker stream -----------------------------------
    def _run(self, reqs: List[Request], staged: List[_Staged], stream: torch.cuda.Stream,
             t0: float, admitted: float) -> List[Result]:
        n, P = reqs[0].n, group_width(reqs)
        with torch.cuda.stream(stream):
            data = initialize_batched_group([r.N for r in reqs], n, self.dtype, self.device,
                                            nrhs=P, mode=self.mode, torch_workspace=True,
                                            **self.plan_options)
            rounds = max((r.width + P - 1) // P for r in reqs)

            def put_d(j):
                # column round j of every request that has one into d (spare columns:
                # copies of the first); returns the requests in the round
                c0 = j * P
                live = [k for k, r in enumerate(reqs) if c0 < r.width]
                for k in live:
                    c = min(P, reqs[k].width - c0)
                    dst = _storage_view(data.d_tensors[k])      # (N, P, n)
                    dst[:, :c].copy_(staged[k].d[:, c0:c0 + c], non_blocking=True)
                    if c < P:
                        dst[:, c:].copy_(dst[:, :1].expand(-1, P - c, -1))
                return live

            # one round: d goes in before the factorisation, whose level-0 kernel then
            # also runs the first correction's level-0 sweep (mixed, one column); with
            # several rounds d is written after it and that sweep would be thrown away
            if rounds > 1:
                L.call("bss_set_option", data._plan, L.BSS_OPT_FACTOR_SWEEP, 0.0)
            try:
                for k, (r, st) in enumerate(zip(reqs, staged)):
                    A, B = data.A_tensors[k], data.B_tensors[k]
                    if r.model is not None:
                        _build_smoother_into(data, A, B, r, st.model)
                    else:
                        _storage_view(A).copy_(st.A, non_blocking=True)
                        if r.N > 1:
                            _storage_view(B).copy_(st.B, non_blocking=True)
                first = put_d(0) if rounds == 1 else None
                factorize(data)
                col_ms = [[] for _ in reqs]     # per column: its solve's device time
                finished = [0.0] * len(reqs)
                for j in range(rounds):
                    c0 = j * P
                    live = first if first is not None else put_d(j)
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    solve(data)
                    e1.record(stream)
                    for k in live:
                        c = min(P, reqs[k].width - c0)
                        staged[k].out[:, c0:c0 + c].copy_(
                            _storage_view(data.d_tensors[k])[:, :c], non_blocking=True)
                    stream.synchronize()
                    ms = e0.elapsed_time(e1)
                    now = time.perf_counter() - t0
                    for k in live:
                        col_ms[k] += [ms] * min(P, reqs[k].width - c0)
                        finished[k] = now
                info = data.info()
            finally:
                data.close()
        out = []
        for k, r in enumerate(reqs):
            w = r.nrhs
            # right-hand side q: columns [q w, (q + 1) w), as an (n, nrhs, N) view
            xs = [np.transpose(staged[k].out[:, q * w:(q + 1) * w].numpy(), (2, 1, 0))
                  for q in range(len(r.rhs))]
            ms = [col_ms[k][q * w] for q in range(len(r.rhs))]
            out.append(Result(r.rid, xs, r.arrival, admitted, finished[k], ms, int(info[k])))
        return out

    # -- the loop ----------------------------------------------------------------------
    def serve(self, requests: List[Request], realtime: bool = True) -> List[Result]:
        """Serve `requests` (arrival times relative to the call) and return one Result
        per request, in request order.  With ``realtime=False`` every request is
        available at once (a closed-loop throughput run)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            fp = self.footprint(r)
            if fp > self.hbm_budget:
                raise ValueError(f"request {r.rid} needs {fp} bytes, more than the HBM "
                                 f"budget of {self.hbm_budget}")
        staged = {r.rid: _Staged(r, self.dtype) for r in reqs}
        streams = [torch.cuda.Stream(self.device) for _ in range(self.workers)]
        free_streams = list(range(self.workers))
        done: "queue.Queue[Tuple[int, int, object]]" = queue.Queue()
        results: Dict[int, Result] = {}
        used = 0
        inflight: Dict[int, int] = {}   # rid of a group's head -> the group's footprint
        groups_inflight: Dict[int, List[Request]] = {}
        fifo: List[Request] = []
        nxt = 0
        t0 = time.perf_counter()

        def worker(group: List[Request], si: int, admitted: float):
            try:
                res = self._run(group, [staged[r.rid] for r in group], streams[si], t0,
                                admitted)
                done.put((group, si, res))
            except BaseException as exc:  # reported to the loop, re-raised there
                done.put((group, si, exc))

        threads = []
        while len(results) < len(reqs):
            now = time.perf_counter() - t0
            while nxt < len(reqs) and (not realtime or reqs[nxt].arrival <= now):
                fifo.append(reqs[nxt])
                nxt += 1
            # admit in arrival order while a worker and memory are free: the FIFO head,
            # coalesced with the waiting requests of the same shape that still fit
            while fifo and free_streams:
                head = fifo[0]
                if used + self.footprint(head) > self.hbm_budget:
                    break
                group = [head]
                P = group_width(group)
                fp = self._group_fp(group, P)
                for r in fifo[1:]:
                    if len(group) >= self.max_group:
                        break
                    if not self._same_class(r, head):
                        continue
                    P2 = max(P, solve_width(r.width))
                    fp2 = (fp + request_footprint(r.N, r.n, P, self.dtype, self.mode,
                                                  self.schedule)
                           if P2 == P else self._group_fp(group + [r], P2))
                    if used + fp2 <= self.hbm_budget:
                        group.append(r)
                        P, fp = P2, fp2
                ids = {r.rid for r in group}
                fifo = [r for r in fifo if r.rid not in ids]
                used += fp
                inflight[head.rid] = fp
                self.stats["max_inflight_bytes"] = max(self.stats["max_inflight_bytes"], used)
                inflight_reqs = sum(len(g) for g in groups_inflight.values()) + len(group)
                groups_inflight[head.rid] = group
                self.stats["max_inflight"] = max(self.stats["max_inflight"], inflight_reqs)
                self.stats["groups"] += 1
                self.stats["max_group"] = max(self.stats["max_group"], len(group))
                si = free_streams.pop()
                th = threading.Thread(target=worker,
                                      args=(group, si, time.perf_counter() - t0), daemon=True)
                th.start()
                threads.append(th)
            # completions free their worker and memory
            timeout = 0.0005
            if nxt < len(reqs) and realtime:
                timeout = max(0.0, min(timeout, reqs[nxt].arrival - now))
            try:
                group, si, res = done.get(timeout=timeout if (fifo or nxt < len(reqs)) else 1.0)
            except queue.Empty:
                continue
            while True:
                if isinstance(res, BaseException):
                    for th in threads:
                        th.join()
                    raise res
                for r, out in zip(group, res):
                    results[r.rid] = out
                used -= inflight.pop(group[0].rid)
                groups_inflight.pop(group[0].rid)
                free_streams.append(si)
                try:
                    group, si, res = done.get_nowait()
                except queue.Empty:
                    break
        for th in threads:
            th.join()
        return [results[r.rid] for r in requests]


def group_plan_rate(server: Server, requests: List[Request], repeats: int = 3) -> float:
    """Systems/s of `requests` all present at once and fully coalesced: one group plan
    per block size and column class, each run by the server's group code on its own
    stream and thread (staging, blocks, factorisation, solves and transfers as served).
    The rate the request loop approaches at saturation; best of `repeats` runs."""
    classes: Dict[Tuple[int, bool], List[Request]] = {}
    for r in requests:
        classes.setdefault((r.n, r.width > 1), []).append(r)
    staged = {r.rid: _Staged(r, server.dtype) for r in requests}
    groups = list(classes.values())
    streams = [torch.cuda.Stream(server.device) for _ in groups]
    best = 0.0
    for _ in range(repeats):
        torch.cuda.synchronize(server.device)
        t0 = time.perf_counter()
        errs: List[BaseException] = []

        def run(g, s):
            try:
                server._run(g, [staged[r.rid] for r in g], s, t0, 0.0)
            except BaseException as exc:  # re-raised below
                errs.append(exc)

        th = [threading.Thread(target=run, args=(g, s)) for g, s in zip(groups, streams)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        if errs:
            raise errs[0]
        torch.cuda.synchronize(server.device)
        best = max(best, len(requests) / (time.perf_counter() - t0))
    return best


def summarize(results: List[Result]) -> dict:
    """p50/p99 request latency, throughput and solve-latency percentiles."""
    lat = np.array([r.latency_ms for r in results])
    solves = np.array([t for r in results for t in r.solve_ms])
    span = max(r.finished for r in results) - min(r.arrival for r in results)
    return {
        "requests": len(results),
        "systems_per_s": len(results) / span if span > 0 else float("inf"),
        "solves_per_s": len(solves) / span if span > 0 else float("inf"),
        "latency_ms_p50": float(np.percentile(lat, 50)),
        "latency_ms_p99": float(np.percentile(lat, 99)),
        "solve_ms_p50": float(np.percentile(solves, 50)),
        "solve_ms_p99": float(np.percentile(solves, 99)),
        "queue_ms_p50": float(np.percentile([(r.admitted - r.arrival) * 1e3 for r in results], 50)),
    }


def smoother_requests(count: int, rate: float, n_choices=(16, 32, 64), N_range=(128, 4096),
                      rhs_range=(1, 8), nrhs: int = 1, seed: int = 0,
                      device_blocks: bool = False) -> List[Request]:
    """Served-stream workload: Poisson arrivals at `rate` per second of state-space
    smoother systems (workloads.smoother_model), state size n uniform over
    `n_choices`, horizon N ~ U[N_range], 1..8 right-hand sides per factorisation.
    With ``device_blocks`` a request carries the model and the device builds its
    blocks; otherwise it carries the dense blocks."""
    from .workloads import smoother_model
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(count):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        n = int(rng.choice(n_choices))
        N = int(rng.integers(N_range[0], N_range[1] + 1))
        reps = int(rng.integers(rhs_range[0], rhs_range[1] + 1))
        model = smoother_model(n, np.random.default_rng(int(rng.integers(1 << 31))))
        rhs = [np.asfortranarray(rng.uniform(size=(n, nrhs, N))) for _ in range(reps)]
        req = Request(k, t, rhs, model=model)
        if not device_blocks:
            req.A, req.B = req.dense()
            req.model = None
        out.append(req)
    return out


def solve_alone(req: Request, mode: str = "plain", dtype=torch.float64, device="cuda",
                **plan_options) -> List[np.ndarray]:
    """The same request on its own plan, on the default stream, nothing else in
    flight: the reference point for served answers.  The columns are solved as the
    server solves them, ``solve_width(req.width)`` at a time (spare columns of the
    last round filled with copies of its first), on the server's default schedule
    (``SERVE_SCHEDULE``) unless `plan_options` name one."""
    W = req.width
    P = solve_width(W)
    data = initialize_batched(req.N, req.n, dtype, device, nrhs=P, mode=mode,
                              **_with_schedule(plan_options))
    L.call("bss_set_option", data._plan, L.BSS_OPT_FACTOR_SWEEP, 0.0)  # d comes later
    try:
        npt = np.float64 if dtype == torch.float64 else np.float32
        if req.model is not None:
            build_smoother(data, req.model)
        else:
            data.A_tensor.copy_(torch.from_numpy(req.A.astype(npt)))
            if req.N > 1:
                data.B_tensor.copy_(torch.from_numpy(req.B.astype(npt)))
        factorize(data)
        cols = np.concatenate([d.astype(npt) for d in req.rhs], axis=1)  # (n, W, N)
        x = np.empty_like(cols)
        for c0 in range(0, W, P):
            c = min(P, W - c0)
            blk = np.empty((req.n, P, req.N), dtype=npt, order="F")
            blk[:, :c] = cols[:, c0:c0 + c]
            blk[:, c:] = cols[:, c0:c0 + 1]
            data.d_tensor.copy_(torch.from_numpy(blk))
            solve(data)
            torch.cuda.synchronize()
            x[:, c0:c0 + c] = data.d_tensor.cpu().numpy()[:, :c]
        w = req.nrhs
        return [np.asfortranarray(x[:, q * w:(q + 1) * w]) for q in range(len(req.rhs))]
    finally:
        data.close()
