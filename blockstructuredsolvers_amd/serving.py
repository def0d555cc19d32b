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
  HBM budget beside the requests in flight.  There is no batch boundary: a request
  that arrives while others run starts as soon as a slot and memory free up.
* Requests are coalesced: when a worker is free, the FIFO head is admitted together
  with every other waiting request of the same block size n and right-hand-side width
  (up to ``max_group`` requests and the HBM budget), and the group runs as one group
  plan (``bss_plan_create_group``), so its systems share every launch.
* Each worker thread owns one HIP stream and runs its group end to end: blocks
  uploaded (or built on the device from the model), one ``factorize`` of all systems,
  then solve rounds: round k solves the k-th right-hand side of every request that has
  one (one ``solve`` over the group) and downloads those solutions.  The factors stay
  resident between the rounds.  Work of different workers overlaps on the device
  (ctypes releases the GIL during library calls).  Plans take their workspace from
  torch's caching allocator, so admission and release make no device allocation or
  free.
* Latency of a request = host time its last solution is available - arrival time; a
  solve's latency is its round's device time.

A system's arithmetic in a group plan is the arithmetic it gets alone (its own level
schedule, separate launch slots, no cross-system reductions), so a served answer equals
the answer of the same request solved alone (bitwise for plain plans; within the
refinement tolerance in mixed mode, where a group refines until all its systems meet
the target).
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
    arrival: float
    admitted: float
    finished: float
    solve_ms: List[float] = field(default_factory=list)  # per right-hand side, device
    info: int = 0

    @property
    def latency_ms(self) -> float:
        return (self.finished - self.arrival) * 1e3


def _storage_order(blocks: np.ndarray) -> torch.Tensor:
    """(r, c, i) column-major blocks -> pinned (i, c, r) C-order tensor, which is the
    plan tensors' storage order (one contiguous DMA per upload)."""
    return torch.from_numpy(np.ascontiguousarray(np.transpose(blocks, (2, 1, 0)))).pin_memory()


@functools.lru_cache(maxsize=4096)
def _workspace_bytes(N: int, n: int, nrhs: int, dcode: int, mcode: int) -> int:
    p = ctypes.c_void_p()
    L.call("bss_plan_create", N, n, nrhs, 1, dcode, mcode, L.BSS_SCHED_AUTO, ctypes.byref(p))
    try:
        v = ctypes.c_int64()
        L.call("bss_plan_get", p, L.BSS_Q_WORKSPACE_BYTES, ctypes.byref(v))
        return int(v.value)
    finally:
        L.lib().bss_plan_destroy(p)


def request_footprint(N: int, n: int, nrhs: int, dtype=torch.float64, mode: str = "plain") -> int:
    """Device bytes a request holds while in flight: the plan's workspace (queried from a
    host-only plan, nothing allocated) plus its A, B and d tensors."""
    mcode = L.BSS_MODE_MIXED if mode == "mixed" else L.BSS_MODE_PLAIN
    ws = _workspace_bytes(N, n, nrhs, _dtype_code(dtype), mcode)
    es = 8 if dtype == torch.float64 else 4
    return ws + 256 + es * n * (2 * N * n + N * nrhs)


def _storage_view(t: torch.Tensor) -> torch.Tensor:
    """The (blocks, cols, rows) C-order storage behind a per-system (rows, cols, blocks)
    view of a plan tensor."""
    return t.permute(2, 1, 0)


def _build_smoother_into(data, A: torch.Tensor, B: torch.Tensor, req: Request) -> None:
    """bss_build_smoother into one system's A, B views of a group plan, on the current
    stream; the model's device copies stay referenced by `data` until it is closed."""
    n = req.n
    dev = data.device
    F, H, Qi, Ri = [torch.from_numpy(np.ascontiguousarray(np.asarray(M, dtype=np.float64).T)).to(dev)
                    for M in (req.model.F, req.model.H, req.model.Qinv, req.model.Rinv)]
    code = L.BSS_F64 if data.dtype == torch.float64 else L.BSS_F32
    L.call("bss_build_smoother", req.N, n, req.model.H.shape[0], F.data_ptr(), H.data_ptr(),
           Qi.data_ptr(), Ri.data_ptr(), A.data_ptr(), B.data_ptr() if req.N > 1 else None,
           code, torch.cuda.current_stream(dev).cuda_stream)
    keep = getattr(data, "_model_tensors", [])
    keep.append((F, H, Qi, Ri))
    data._model_tensors = keep


def group_footprint(Ns: List[int], n: int, nrhs: int, dtype=torch.float64,
                    mode: str = "plain") -> int:
    """Device bytes a coalesced group holds while in flight: its group plan's workspace
    (host-only query) plus the systems' A, B and d."""
    mcode = L.BSS_MODE_MIXED if mode == "mixed" else L.BSS_MODE_PLAIN
    p = ctypes.c_void_p()
    arr = (ctypes.c_int * len(Ns))(*Ns)
    L.call("bss_plan_create_group", len(Ns), arr, n, nrhs, _dtype_code(dtype), mcode,
           L.BSS_SCHED_AUTO, ctypes.byref(p))
    try:
        v = ctypes.c_int64()
        L.call("bss_plan_get", p, L.BSS_Q_WORKSPACE_BYTES, ctypes.byref(v))
        ws = int(v.value)
    finally:
        L.lib().bss_plan_destroy(p)
    es = 8 if dtype == torch.float64 else 4
    return ws + 256 + sum(es * n * (2 * N * n + N * nrhs) for N in Ns)


class _Staged:
    """A request's host data in pinned memory, prepared before it is served (no A, B
    when the device builds the blocks from the request's model)."""

    def __init__(self, req: Request, dtype):
        npt = np.float64 if dtype == torch.float64 else np.float32
        self.A = self.B = None
        if req.model is None:
            self.A = _storage_order(req.A.astype(npt))
            self.B = _storage_order(req.B.astype(npt))
        self.d = [_storage_order(d.astype(npt)) for d in req.rhs]
        self.out = [torch.empty_like(d).pin_memory() for d in self.d]


class Server:
    """Continuous-batching request loop over libbss plans on one device.

    ``mode``/``dtype`` apply to every request; ``hbm_budget`` (bytes) bounds the summed
    footprint of the requests in flight (default: 80% of the device's free memory);
    ``workers`` is the number of groups in flight at most (one HIP stream each) and
    ``max_group`` the most requests coalesced into one group plan (1: no coalescing)."""

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
        self.plan_options = plan_options
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.8 * free)
        self.hbm_budget = int(hbm_budget)
        self.stats = {"max_inflight_bytes": 0, "max_inflight": 0, "groups": 0,
                      "max_group": 0}

    # -- admission ---------------------------------------------------------------------
    def footprint(self, N: int, n: int, nrhs: int) -> int:
        return request_footprint(N, n, nrhs, self.dtype, self.mode)

    def _group_fp(self, group: List[Request]) -> int:
        if len(group) == 1:
            r = group[0]
            return self.footprint(r.N, r.n, r.nrhs)
        return group_footprint([r.N for r in group], group[0].n, group[0].nrhs, self.dtype,
                               self.mode)

    # -- one coalesced group on one worker stream -----------------------------------
    def _run(self, reqs: List[Request], staged: List[_Staged], stream: torch.cuda.Stream,
             t0: float, admitted: float) -> List[Result]:
        n, nrhs = reqs[0].n, reqs[0].nrhs
        with torch.cuda.stream(stream):
            data = initialize_batched_group([r.N for r in reqs], n, self.dtype, self.device,
                                            nrhs=nrhs, mode=self.mode, torch_workspace=True,
                                            **self.plan_options)
            try:
                for k, (r, st) in enumerate(zip(reqs, staged)):
                    A, B = data.A_tensors[k], data.B_tensors[k]
                    if r.model is not None:
                        _build_smoother_into(data, A, B, r)
                    else:
                        _storage_view(A).copy_(st.A, non_blocking=True)
                        if r.N > 1:
                            _storage_view(B).copy_(st.B, non_blocking=True)
                factorize(data)
                xs = [[] for _ in reqs]
                solve_ms = [[] for _ in reqs]
                finished = [0.0] * len(reqs)
                rounds = max(len(r.rhs) for r in reqs)
                for j in range(rounds):
                    live = [k for k, r in enumerate(reqs) if j < len(r.rhs)]
                    for k in live:
                        _storage_view(data.d_tensors[k]).copy_(staged[k].d[j], non_blocking=True)
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    solve(data)
                    e1.record(stream)
                    for k in live:
                        staged[k].out[j].copy_(_storage_view(data.d_tensors[k]), non_blocking=True)
                    stream.synchronize()
                    ms = e0.elapsed_time(e1)
                    now = time.perf_counter() - t0
                    for k in live:
                        solve_ms[k].append(ms)
                        xs[k].append(np.transpose(staged[k].out[j].numpy(), (2, 1, 0)).copy())
                        finished[k] = now
                info = data.info()
            finally:
                data.close()
        return [Result(r.rid, xs[k], r.arrival, admitted, finished[k], solve_ms[k], int(info[k]))
                for k, r in enumerate(reqs)]

    # -- the loop ----------------------------------------------------------------------
    def serve(self, requests: List[Request], realtime: bool = True) -> List[Result]:
        """Serve `requests` (arrival times relative to the call) and return one Result
        per request, in request order.  With ``realtime=False`` every request is
        available at once (a closed-loop throughput run)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            fp = self.footprint(r.N, r.n, r.nrhs)
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
                if used + self.footprint(head.N, head.n, head.nrhs) > self.hbm_budget:
                    break
                group = [head]
                for r in fifo[1:]:
                    if len(group) >= self.max_group:
                        break
                    if r.n == head.n and r.nrhs == head.nrhs and used + self._group_fp(
                            group + [r]) <= self.hbm_budget:
                        group.append(r)
                fp = self._group_fp(group)
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
    flight: the reference point for served answers."""
    data = initialize_batched(req.N, req.n, dtype, device, nrhs=req.nrhs, mode=mode,
                              **plan_options)
    try:
        npt = np.float64 if dtype == torch.float64 else np.float32
        if req.model is not None:
            build_smoother(data, req.model)
        else:
            data.A_tensor.copy_(torch.from_numpy(req.A.astype(npt)))
            if req.N > 1:
                data.B_tensor.copy_(torch.from_numpy(req.B.astype(npt)))
        factorize(data)
        xs = []
        for d in req.rhs:
            data.d_tensor.copy_(torch.from_numpy(d.astype(npt)))
            solve(data)
            torch.cuda.synchronize()
            xs.append(data.d_tensor.cpu().numpy().copy())
        return xs
    finally:
        data.close()
