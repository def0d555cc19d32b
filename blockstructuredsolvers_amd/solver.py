"""Host mirror of the reference's batched GPU API on top of libbss.

Reference interface (src/BlockStructuredSolvers.jl:9-11, src/gpu_batched.jl):

    data = initialize_batched(N, n, T, M)      # gpu_batched.jl:165-182
    copyto!(data.A_tensor, A); copyto!(data.B_tensor, B); copyto!(data.d_tensor, d)
    factorize!(data)                           # gpu_batched.jl:184-227
    solve!(data)                               # gpu_batched.jl:229-259
    x = data.d_list[1:N]

Python has no ``!`` in names: ``factorize`` / ``solve`` (aliases ``factorize_`` /
``solve_``) play ``factorize!`` / ``solve!``.  The device array type ``M`` of the
reference is a torch device here (``"cuda"`` = the HIP device); torch only owns the
memory and supplies the stream, every computation is a libbss call.  Tensors keep the
reference layout: ``A_tensor`` is (n, n, N_use) with column-major blocks, ``d_tensor``
is (n, nrhs, N_use); for a batch of ``nsys`` systems a trailing system axis is added.

Like the reference, ``factorize``/``solve`` are asynchronous on the current stream
and a non-SPD input is not an exception: ``data.info()`` reports it per system.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional

import numpy as np
import torch

from . import _lib as L


def _dtype_code(T) -> int:
    if T in (torch.float64, np.float64, float, "float64", "Float64"):
        return L.BSS_F64
    if T in (torch.float32, np.float32, "float32", "Float32"):
        return L.BSS_F32
    raise TypeError(f"unsupported element type {T!r} (Float32 or Float64)")


def _torch_dtype(code: int):
    return torch.float64 if code == L.BSS_F64 else torch.float32


def _stream_handle(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _set_options(plan, chain_len, base_max, chain_len_levels, factor_waves, preserve_input,
                 tol, max_iters, graph=False, pivot_tol=None, fallback=None) -> None:
    if chain_len is not None:
        L.call("bss_set_option", plan, L.BSS_OPT_CHAIN_LEN, float(chain_len))
    if base_max is not None:
        L.call("bss_set_option", plan, L.BSS_OPT_BASE_MAX, float(base_max))
    if chain_len_levels is not None:
        for k, m in enumerate(list(chain_len_levels)[:16]):
            L.call("bss_set_option", plan, L.BSS_OPT_CHAIN_LEN_LEVEL + k, float(m))
    if factor_waves is not None:
        L.call("bss_set_option", plan, L.BSS_OPT_FACTOR_WAVES, float(factor_waves))
    if preserve_input is not None:
        L.call("bss_set_option", plan, L.BSS_OPT_PRESERVE_INPUT, 1.0 if preserve_input else 0.0)
    if tol is not None:
        L.call("bss_set_option", plan, L.BSS_OPT_TOL, float(tol))
    if max_iters is not None:
        L.call("bss_set_option", plan, L.BSS_OPT_MAX_ITERS, float(max_iters))
    if graph:
        L.call("bss_set_option", plan, L.BSS_OPT_GRAPH, 1.0)
    if pivot_tol is not None:
        L.call("bss_set_option", plan, L.BSS_OPT_PIVOT_TOL, float(pivot_tol))
    if fallback is not None:
        L.call("bss_set_option", plan, L.BSS_OPT_FALLBACK, 1.0 if fallback else 0.0)


def _torch_workspace(plan, device) -> torch.Tensor:
    """Workspace for the plan from torch's caching allocator (bss_plan_set_workspace):
    no hipMalloc/hipFree per plan, the memory returns to the allocator's pool of the
    current stream when the data object is closed."""
    v = ctypes.c_int64()
    L.call("bss_plan_get", plan, L.BSS_Q_WORKSPACE_BYTES, ctypes.byref(v))
    ws = torch.empty(int(v.value) + 256, dtype=torch.uint8, device=device)
    off = (-ws.data_ptr()) % 256
    L.call("bss_plan_set_workspace", plan, ws.data_ptr() + off, int(v.value))
    # bind initialises the workspace outside the stream order: a block recycled by the
    # allocator must have no pending work left
    torch.cuda.current_stream(device).synchronize()
    return ws


class BlockTriDiagData_batched:
    """Counterpart of ``BlockTriDiagData_batched`` (src/gpu_batched.jl:1-71): the
    user-visible tensors plus an opaque libbss plan holding the schedule and the
    per-level workspace (the reference's LHS/M/F/G/b buffers and NextData chain)."""

    def __init__(self, N: int, n: int, T, device, nrhs: int, nsys: int, mode: str,
                 schedule: str, chain_len: Optional[int], base_max: Optional[int],
                 preserve_input: bool = True, tol: Optional[float] = None,
                 max_iters: Optional[int] = None,
                 chain_len_levels: Optional[List[int]] = None,
                 factor_waves: Optional[int] = None, graph: bool = False,
                 pivot_tol: Optional[float] = None, torch_workspace: bool = False,
                 fallback: Optional[bool] = None):
        self.n_user = N
        self.n = n
        self.nrhs = nrhs
        self.nsys = nsys
        self.mode = mode
        self.schedule = schedule
        self.device = torch.device(device)
        code = _dtype_code(T)
        mode_code = {"plain": L.BSS_MODE_PLAIN, "mixed": L.BSS_MODE_MIXED}[mode]
        sched_code = {"auto": L.BSS_SCHED_AUTO, "reference": L.BSS_SCHED_REFERENCE}[schedule]
        plan = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            L.call("bss_plan_create", N, n, nrhs, nsys, code, mode_code, sched_code,
                   ctypes.byref(plan))
        self._plan = plan
        _set_options(plan, chain_len, base_max, chain_len_levels, factor_waves, preserve_input,
                     tol, max_iters, graph, pivot_tol, fallback)
        self.N = self._get(L.BSS_Q_N_USE)
        self.dtype = _torch_dtype(code)
        NU = self.N
        # column-major blocks: storage (S, N, n, n) C-order, exposed as (n, n, N[, S])
        a = torch.zeros((nsys, NU, n, n), dtype=self.dtype, device=self.device)
        b = torch.zeros((nsys, max(NU - 1, 0), n, n), dtype=self.dtype, device=self.device)
        dd = torch.zeros((nsys, NU, nrhs, n), dtype=self.dtype, device=self.device)
        self._a, self._b, self._d = a, b, dd
        perm = (3, 2, 1, 0)
        if nsys == 1:
            self.A_tensor = a[0].permute(2, 1, 0)
            self.B_tensor = b[0].permute(2, 1, 0)
            self.d_tensor = dd[0].permute(2, 1, 0)
        else:
            self.A_tensor = a.permute(*perm)
            self.B_tensor = b.permute(*perm)
            self.d_tensor = dd.permute(*perm)
        # identity padding of the blocks the reference pads (gpu_batched.jl:177-179:
        # blocks N..N_use, 1-based, i.e. from the user's last block on)
        if NU > N:
            eye = torch.eye(n, dtype=self.dtype, device=self.device)
            a[:, N - 1:] = eye
        self._ws = _torch_workspace(plan, self.device) if torch_workspace else None
        with torch.cuda.device(self.device):
            L.call("bss_plan_bind", plan, a.data_ptr(), b.data_ptr() if b.numel() else None,
                   dd.data_ptr())
        self.levels = self._get(L.BSS_Q_LEVELS)
        self.P = self._get(L.BSS_Q_LEVEL_N + 1) if self.levels > 1 else 0
        self.nchains = [self._get(L.BSS_Q_LEVEL_CHAINS + k) for k in range(self.levels)]

    # -- reference-style per-block views (gpu_seq.jl:29-43 create_matrix_list) --------
    @property
    def A_list(self) -> List[torch.Tensor]:
        return [self.A_tensor[:, :, i] for i in range(self.N)]

    @property
    def B_list(self) -> List[torch.Tensor]:
        return [self.B_tensor[:, :, i] for i in range(self.N - 1)]

    @property
    def d_list(self) -> List[torch.Tensor]:
        return [self.d_tensor[:, :, i] for i in range(self.N)]

    def _get(self, q: int) -> int:
        v = ctypes.c_int64()
        L.call("bss_plan_get", self._plan, q, ctypes.byref(v))
        return int(v.value)

    def level_sizes(self) -> List[int]:
        return [self._get(L.BSS_Q_LEVEL_N + k) for k in range(self.levels)]

    def info(self) -> np.ndarray:
        """Per-system potrf status (0 ok, else 1-based first failing block)."""
        out = (ctypes.c_int * self.nsys)()
        with torch.cuda.device(self.device):
            L.call("bss_info", self._plan, out)
        return np.array(out[:], dtype=np.int64)

    def backward_error(self) -> np.ndarray:
        """Mixed mode: device-measured normwise backward error per (system, column)."""
        out = (ctypes.c_double * (self.nsys * self.nrhs))()
        L.call("bss_last_backward_error", self._plan, out)
        return np.array(out[:]).reshape(self.nsys, self.nrhs)

    @property
    def iterations(self) -> int:
        return self._get(L.BSS_Q_LAST_ITERS)

    @property
    def fallback(self) -> bool:
        """Mixed mode: the solves since the last factorize went through the plain
        FP64 fallback (refinement on the FP32 factors stopped short of ``tol``)."""
        return bool(self._get(L.BSS_Q_FALLBACK))

    @property
    def converged(self) -> bool:
        """Mixed mode: the last solve met ``tol`` for every system and column (the
        status the solve call returned was BSS_OK, not BSS_NOT_CONVERGED)."""
        return getattr(self, "last_status", L.BSS_OK) == L.BSS_OK

    @property
    def refactored(self) -> int:
        """Mixed mode: separator levels re-factored in FP64 after the last factorize
        (known once a solve or info() has inspected the FP32 pivots)."""
        return self._get(L.BSS_Q_REFACTORED)

    def synchronize(self) -> None:
        with torch.cuda.device(self.device):
            L.call("bss_sync", self._plan)

    def close(self) -> None:
        if getattr(self, "_plan", None) is not None and self._plan.value:
            L.lib().bss_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()
        self._ws = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BlockTriDiagData_group(BlockTriDiagData_batched):
    """An uneven batch: ``nsys`` independent systems of one block size n whose block
    counts differ (system s has ``Ns[s]`` blocks).  The reference's batch is uniform
    (src/gpu_batched.jl:165-182); here every system keeps its own level schedule (the
    one it gets alone) and each level runs all systems that have it in one launch.

    Per-system views replace the single tensors: ``A_tensors[s]`` is (n, n, N_use[s]),
    ``B_tensors[s]`` (n, n, N_use[s]-1), ``d_tensors[s]`` (n, nrhs, N_use[s]), all
    column-major blocks packed one system after the other in three device buffers."""

    def __init__(self, Ns: List[int], n: int, T, device, nrhs: int, mode: str,
                 schedule: str, chain_len: Optional[int], base_max: Optional[int],
                 preserve_input: bool = True, tol: Optional[float] = None,
                 max_iters: Optional[int] = None,
                 chain_len_levels: Optional[List[int]] = None,
                 factor_waves: Optional[int] = None, graph: bool = False,
                 pivot_tol: Optional[float] = None, torch_workspace: bool = False,
                 fallback: Optional[bool] = None):
        Ns = [int(v) for v in Ns]
        self.Ns_user = Ns
        self.n = n
        self.nrhs = nrhs
        self.nsys = len(Ns)
        self.mode = mode
        self.schedule = schedule
        self.device = torch.device(device)
        code = _dtype_code(T)
        mode_code = {"plain": L.BSS_MODE_PLAIN, "mixed": L.BSS_MODE_MIXED}[mode]
        sched_code = {"auto": L.BSS_SCHED_AUTO, "reference": L.BSS_SCHED_REFERENCE}[schedule]
        plan = ctypes.c_void_p()
        arr = (ctypes.c_int * len(Ns))(*Ns)
        with torch.cuda.device(self.device):
            L.call("bss_plan_create_group", len(Ns), arr, n, nrhs, code, mode_code,
                   sched_code, ctypes.byref(plan))
        self._plan = plan
        _set_options(plan, chain_len, base_max, chain_len_levels, factor_waves, preserve_input,
                     tol, max_iters, graph, pivot_tol, fallback)
        nu = (ctypes.c_int * len(Ns))()
        L.call("bss_plan_system_sizes", plan, nu)
        self.N_use = [int(v) for v in nu]
        self.dtype = _torch_dtype(code)
        totA = sum(self.N_use)
        totB = sum(max(v - 1, 0) for v in self.N_use)
        a = torch.zeros((totA, n, n), dtype=self.dtype, device=self.device)
        b = torch.zeros((max(totB, 1), n, n), dtype=self.dtype, device=self.device)
        dd = torch.zeros((totA, nrhs, n), dtype=self.dtype, device=self.device)
        self._a, self._b, self._d = a, b, dd
        self.A_tensors, self.B_tensors, self.d_tensors = [], [], []
        oa = ob = 0
        for N, NU in zip(Ns, self.N_use):
            nb = max(NU - 1, 0)
            self.A_tensors.append(a[oa:oa + NU].permute(2, 1, 0))
            self.B_tensors.append(b[ob:ob + nb].permute(2, 1, 0))
            self.d_tensors.append(dd[oa:oa + NU].permute(2, 1, 0))
            if NU > N:  # identity padding (gpu_batched.jl:177-179)
                a[oa + N - 1:oa + NU] = torch.eye(n, dtype=self.dtype, device=self.device)
            oa += NU
            ob += nb
        self._ws = _torch_workspace(plan, self.device) if torch_workspace else None
        with torch.cuda.device(self.device):
            L.call("bss_plan_bind", plan, a.data_ptr(), b.data_ptr(), dd.data_ptr())
        self.levels = self._get(L.BSS_Q_LEVELS)
        self.nchains = [self._get(L.BSS_Q_LEVEL_CHAINS + k) for k in range(self.levels)]
        self.N = self.N_use[0]


def initialize_batched(N: int, n: int, T=torch.float64, M="cuda", *, nrhs: int = 1,
                       nsys: int = 1, mode: str = "plain", schedule: str = "auto",
                       chain_len: Optional[int] = None,
                       base_max: Optional[int] = None, preserve_input: bool = True,
                       tol: Optional[float] = None,
                       max_iters: Optional[int] = None,
                       chain_len_levels: Optional[List[int]] = None,
                       factor_waves: Optional[int] = None,
                       tuned: bool = True, graph: bool = False,
                       pivot_tol: Optional[float] = None,
                       torch_workspace: bool = False,
                       fallback: Optional[bool] = None) -> BlockTriDiagData_batched:
    """``initialize_batched(N, n, T, M)`` (src/gpu_batched.jl:165-182).

    ``M`` names the device (``"cuda"``/``"cuda:k"`` or a torch.device).  Extensions
    over the reference: ``nrhs`` right-hand-side columns (the reference hard-codes 1,
    gpu_batched.jl:257), ``nsys`` same-shape systems in one plan, ``mode="mixed"``
    (FP32 factor + FP64 refinement, T must be Float64) and ``schedule="reference"``
    (the reference's level schedule and identity padding to N_use) or ``"auto"``;
    ``preserve_input`` (default) keeps A and B unchanged in plain mode (factors go to
    the plan's workspace; ``False`` overwrites them as the reference does, and
    ``info()`` then gives a failing block that need not be the LAPACK-order first one);
    ``tol``/``max_iters`` control mixed-mode refinement and ``fallback`` (default on)
    its plain FP64 fallback when refinement stops short; ``chain_len``,
    ``chain_len_levels``, ``base_max`` and ``factor_waves`` fix the auto schedule,
    otherwise (``tuned=True``) it comes from the autotuner cache when one exists;
    ``graph=True`` replays each launch sequence from a captured hipGraph;
    ``pivot_tol`` (mixed) is the FP32 separator pivot threshold U_jj^2 / A_jj below
    which the separator levels are re-factored in FP64; ``torch_workspace`` takes the
    plan's workspace from torch's caching allocator instead of a hipMalloc."""
    if not torch.cuda.is_available():
        raise RuntimeError("initialize_batched needs a HIP device; none is visible")
    device = torch.device(M) if not isinstance(M, torch.device) else M
    if device.type != "cuda":
        raise ValueError(f"M must name a GPU device, got {M!r}")
    if (tuned and schedule == "auto" and chain_len is None and base_max is None
            and chain_len_levels is None and factor_waves is None):
        # schedule from the autotuner cache (keyed by N, n, type, mode and the largest
        # tuned batch-size bucket <= nsys)
        from .autotune import lookup
        cfg = lookup(N, n, T, mode, nsys=nsys)
        if cfg:
            chain_len_levels = cfg.get("chain_len_levels")
            base_max = cfg.get("base_max")
            factor_waves = cfg.get("factor_waves")
    return BlockTriDiagData_batched(N, n, T, device, nrhs, nsys, mode, schedule, chain_len,
                                    base_max, preserve_input, tol, max_iters,
                                    chain_len_levels, factor_waves, graph, pivot_tol,
                                    torch_workspace, fallback)


class BlockTriDiagData_seq(BlockTriDiagData_batched):
    """Counterpart of ``BlockTriDiagData_seq`` (src/gpu_seq.jl:1-27): one system
    eliminated block after block, with no recursion -- the plan's schedule is a single
    chain of N blocks (one workgroup runs the whole potrf / trsm / Schur-update
    sequence of ``cholesky_factorize!`` and the substitutions of ``cholesky_solve!``,
    src/backbone_rocm.jl:14-53).  Besides the tensors and per-block lists it exposes the
    reference's flat buffers (``A_vec``: all blocks as one (N*n*n, 1) column) and
    per-block device address arrays (``A_ptrs``, int64, ``device_batch``'s
    counterpart, src/backbone_rocm.jl:4-7)."""

    def __init__(self, N: int, n: int, T, device, nrhs: int):
        super().__init__(N, n, T, device, nrhs, 1, "plain", "auto", None, max(N, 1))
        if self.levels != 1:
            raise RuntimeError(f"sequential plan has {self.levels} levels, expected 1")
        es = self._a.element_size()

        def ptrs(t: torch.Tensor, count: int, block: int) -> torch.Tensor:
            base = t.data_ptr()
            return torch.tensor([base + i * block * es for i in range(count)],
                                dtype=torch.int64, device=self.device)

        self.A_vec = self._a.view(-1, 1)
        self.B_vec = self._b.view(-1, 1)
        self.d_vec = self._d.view(-1, 1)
        self.A_ptrs = ptrs(self._a, N, n * n)
        self.B_ptrs = ptrs(self._b, max(N - 1, 0), n * n)
        self.d_ptrs = ptrs(self._d, N, n * nrhs)


def initialize_seq(N: int, n: int, T=torch.float64, M="cuda", *,
                   nrhs: int = 1) -> BlockTriDiagData_seq:
    """``initialize_seq(N, n, T, M)`` (src/gpu_seq.jl:45-59): the sequential solver's
    data (zeroed A, B, d in the reference layout).  ``factorize``/``solve`` accept it
    like the batched data.  Extension: ``nrhs`` columns (the reference solves one)."""
    if not torch.cuda.is_available():
        raise RuntimeError("initialize_seq needs a HIP device; none is visible")
    device = torch.device(M) if not isinstance(M, torch.device) else M
    if device.type != "cuda":
        raise ValueError(f"M must name a GPU device, got {M!r}")
    return BlockTriDiagData_seq(N, n, T, device, nrhs)


def initialize_batched_group(Ns: List[int], n: int, T=torch.float64, M="cuda", *,
                             nrhs: int = 1, mode: str = "plain", schedule: str = "auto",
                             chain_len: Optional[int] = None,
                             base_max: Optional[int] = None, preserve_input: bool = True,
                             tol: Optional[float] = None, max_iters: Optional[int] = None,
                             chain_len_levels: Optional[List[int]] = None,
                             factor_waves: Optional[int] = None,
                             tuned: bool = True, graph: bool = False,
                             pivot_tol: Optional[float] = None,
                             torch_workspace: bool = False,
                             fallback: Optional[bool] = None) -> BlockTriDiagData_group:
    """Uneven batch: like ``initialize_batched`` but system s has ``Ns[s]`` blocks.
    Schedule options apply to every system.  With ``tuned=True`` the autotuner cache is
    used only when it names one configuration for all the sizes in ``Ns``, so each
    system keeps the schedule it would get alone."""
    if not torch.cuda.is_available():
        raise RuntimeError("initialize_batched_group needs a HIP device; none is visible")
    if len(Ns) < 1:
        raise ValueError("Ns must name at least one system")
    device = torch.device(M) if not isinstance(M, torch.device) else M
    if device.type != "cuda":
        raise ValueError(f"M must name a GPU device, got {M!r}")
    if (tuned and schedule == "auto" and chain_len is None and base_max is None
            and chain_len_levels is None and factor_waves is None):
        from .autotune import lookup
        cfgs = [lookup(N, n, T, mode, nsys=len(Ns)) for N in Ns]
        if cfgs[0] and all(c == cfgs[0] for c in cfgs):
            chain_len_levels = cfgs[0].get("chain_len_levels")
            base_max = cfgs[0].get("base_max")
            factor_waves = cfgs[0].get("factor_waves")
    return BlockTriDiagData_group(list(Ns), n, T, device, nrhs, mode, schedule, chain_len,
                                  base_max, preserve_input, tol, max_iters, chain_len_levels,
                                  factor_waves, graph, pivot_tol, torch_workspace, fallback)


class BlockTriDiagData_batch:
    """An uneven batch across block sizes: system s has ``Ns[s]`` blocks of size
    ``ns[s]``.  The reference's batch is uniform (src/gpu_batched.jl:165-182); here the
    systems are grouped by n into one group plan each, and by default the whole batch's
    factorisation (and a plain-mode solve) is one hipGraph launch with the groups as
    parallel branches (``bss_batch_*``).  Per-system views ``A_tensors[s]`` (n, n,
    N_use), ``B_tensors[s]``, ``d_tensors[s]`` (n, nrhs, N_use), in input order, over
    three packed device buffers."""

    def __init__(self, Ns: List[int], ns: List[int], T, device, nrhs: int, mode: str,
                 schedule: str, graph: bool, tol: Optional[float],
                 max_iters: Optional[int], preserve_input: bool, fallback: Optional[bool],
                 chain_len_levels: Optional[List[int]] = None,
                 base_max: Optional[int] = None):
        if len(Ns) != len(ns) or not Ns:
            raise ValueError("Ns and ns must name the same, non-zero number of systems")
        self.Ns_user = [int(v) for v in Ns]
        self.ns = [int(v) for v in ns]
        self.nsys = len(Ns)
        self.nrhs = nrhs
        self.mode = mode
        self.device = torch.device(device)
        code = _dtype_code(T)
        mode_code = {"plain": L.BSS_MODE_PLAIN, "mixed": L.BSS_MODE_MIXED}[mode]
        sched_code = {"auto": L.BSS_SCHED_AUTO, "reference": L.BSS_SCHED_REFERENCE}[schedule]
        b = ctypes.c_void_p()
        arrN = (ctypes.c_int * self.nsys)(*self.Ns_user)
        arrn = (ctypes.c_int * self.nsys)(*self.ns)
        with torch.cuda.device(self.device):
            L.call("bss_batch_create", self.nsys, arrN, arrn, nrhs, code, mode_code, sched_code,
                   ctypes.byref(b))
        self._batch = b
        opts = [(L.BSS_OPT_GRAPH, 1.0 if graph else 0.0),
                (L.BSS_OPT_PRESERVE_INPUT, 1.0 if preserve_input else 0.0)]
        if tol is not None:
            opts.append((L.BSS_OPT_TOL, float(tol)))
        if max_iters is not None:
            opts.append((L.BSS_OPT_MAX_ITERS, float(max_iters)))
        if fallback is not None:
            opts.append((L.BSS_OPT_FALLBACK, 1.0 if fallback else 0.0))
        if base_max is not None:
            opts.append((L.BSS_OPT_BASE_MAX, float(base_max)))
        for k, m in enumerate(list(chain_len_levels or [])[:16]):
            opts.append((L.BSS_OPT_CHAIN_LEN_LEVEL + k, float(m)))
        for opt, v in opts:
            L.call("bss_batch_set_option", b, opt, v)
        I64 = ctypes.c_int64
        oa, ob, od = (I64 * self.nsys)(), (I64 * self.nsys)(), (I64 * self.nsys)()
        tot = (I64 * 3)()
        L.call("bss_batch_layout", b, oa, ob, od, tot)
        self.dtype = _torch_dtype(code)
        a = torch.zeros(max(int(tot[0]), 1), dtype=self.dtype, device=self.device)
        bb = torch.zeros(max(int(tot[1]), 1), dtype=self.dtype, device=self.device)
        dd = torch.zeros(max(int(tot[2]), 1), dtype=self.dtype, device=self.device)
        self._a, self._b, self._d = a, bb, dd
        self.A_tensors, self.B_tensors, self.d_tensors, self.N_use = [], [], [], []
        # N_use per system: the d blocks are packed without gaps, so a system's d range
        # ends where the next one (in packed order) starts
        order = sorted(range(self.nsys), key=lambda s: int(od[s]))
        ends = {}
        for i, s in enumerate(order):
            ends[s] = int(od[order[i + 1]]) if i + 1 < len(order) else int(tot[2])
        for s in range(self.nsys):
            n = self.ns[s]
            NU = (ends[s] - int(od[s])) // (n * nrhs)
            self.N_use.append(NU)
            A = a[int(oa[s]):int(oa[s]) + NU * n * n].view(NU, n, n).permute(2, 1, 0)
            nb = max(NU - 1, 0)
            B = bb[int(ob[s]):int(ob[s]) + nb * n * n].view(nb, n, n).permute(2, 1, 0)
            d = dd[int(od[s]):int(od[s]) + NU * nrhs * n].view(NU, nrhs, n).permute(2, 1, 0)
            self.A_tensors.append(A)
            self.B_tensors.append(B)
            self.d_tensors.append(d)
            if NU > self.Ns_user[s]:  # identity padding (gpu_batched.jl:177-179)
                A[:, :, self.Ns_user[s] - 1:] = torch.eye(n, dtype=self.dtype,
                                                          device=self.device)[:, :, None]
        with torch.cuda.device(self.device):
            L.call("bss_batch_bind", b, a.data_ptr(), bb.data_ptr(), dd.data_ptr())

    def _get(self, q: int) -> int:
        v = ctypes.c_int64()
        L.call("bss_batch_get", self._batch, q, ctypes.byref(v))
        return int(v.value)

    def info(self) -> np.ndarray:
        out = (ctypes.c_int * self.nsys)()
        with torch.cuda.device(self.device):
            L.call("bss_batch_info", self._batch, out)
        return np.array(out[:], dtype=np.int64)

    def backward_error(self) -> np.ndarray:
        out = (ctypes.c_double * (self.nsys * self.nrhs))()
        L.call("bss_batch_last_backward_error", self._batch, out)
        return np.array(out[:]).reshape(self.nsys, self.nrhs)

    @property
    def iterations(self) -> int:
        """Mixed mode: the most refinement iterations any group used in the last solve."""
        return self._get(L.BSS_Q_LAST_ITERS)

    @property
    def converged(self) -> bool:
        return getattr(self, "last_status", L.BSS_OK) == L.BSS_OK

    def synchronize(self) -> None:
        with torch.cuda.device(self.device):
            L.call("bss_batch_sync", self._batch)

    def close(self) -> None:
        if getattr(self, "_batch", None) is not None and self._batch.value:
            L.lib().bss_batch_destroy(self._batch)
            self._batch = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def initialize_batch(Ns: List[int], ns: List[int], T=torch.float64, M="cuda", *,
                     nrhs: int = 1, mode: str = "plain", schedule: str = "auto",
                     graph: bool = True, tol: Optional[float] = None,
                     max_iters: Optional[int] = None, preserve_input: bool = True,
                     fallback: Optional[bool] = None,
                     chain_len_levels: Optional[List[int]] = None,
                     base_max: Optional[int] = None) -> BlockTriDiagData_batch:
    """Uneven batch across block sizes (``bss_batch_create``): system s has ``Ns[s]``
    blocks of size ``ns[s]``.  ``factorize``/``solve`` take the returned data like the
    single-plan data; ``graph`` (default on) runs each call as one hipGraph launch.
    ``chain_len_levels``/``base_max`` fix the auto schedule of every system (as for
    ``initialize_batched``)."""
    if not torch.cuda.is_available():
        raise RuntimeError("initialize_batch needs a HIP device; none is visible")
    device = torch.device(M) if not isinstance(M, torch.device) else M
    if device.type != "cuda":
        raise ValueError(f"M must name a GPU device, got {M!r}")
    return BlockTriDiagData_batch(list(Ns), list(ns), T, device, nrhs, mode, schedule, graph,
                                  tol, max_iters, preserve_input, fallback, chain_len_levels,
                                  base_max)


def factorize(data) -> None:
    """``factorize!(data)`` (src/gpu_batched.jl:184-227), asynchronous."""
    with torch.cuda.device(data.device):
        if isinstance(data, BlockTriDiagData_batch):
            L.call("bss_batch_factorize", data._batch, _stream_handle(data.device))
        else:
            L.call("bss_factorize", data._plan, _stream_handle(data.device))


def solve(data: BlockTriDiagData_batched) -> int:
    """``solve!(data)`` (src/gpu_batched.jl:229-259): solution in place in d_tensor.
    Returns the status: 0, or (mixed mode) ``BSS_NOT_CONVERGED`` = 1 when some system
    ended above the backward-error target even after the FP64 fallback; like the
    reference, nothing is raised for a numerical outcome (``data.converged``)."""
    with torch.cuda.device(data.device):
        if isinstance(data, BlockTriDiagData_batch):
            rc = L.call("bss_batch_solve", data._batch, _stream_handle(data.device))
        else:
            rc = L.call("bss_solve", data._plan, _stream_handle(data.device))
    data.last_status = rc
    return rc


def factorize_solve(data: BlockTriDiagData_batched) -> int:
    """``factorize!(data); solve!(data)`` (src/gpu_batched.jl:184-259) in one call, for
    a d that is final before the factorisation: with FP32 factors and one right-hand
    side the level-0 factor kernel also runs the first solve's level-0 forward sweep;
    the other forward sweeps overlap the separator levels' factorisation on a second
    stream.  Same results (bitwise), state and status as ``factorize`` then ``solve``;
    asynchronous."""
    with torch.cuda.device(data.device):
        rc = L.call("bss_factorize_solve", data._plan, _stream_handle(data.device))
    data.last_status = rc
    return rc


factorize_ = factorize
solve_ = solve


def build_smoother(data: BlockTriDiagData_batched, model) -> None:
    """Fill ``data.A_tensor``/``data.B_tensor`` (a single-system plan) with the
    smoother normal equations of ``model`` (``workloads.SmootherModel``), built on the
    device from F, H, Q^{-1}, R^{-1} (``bss_build_smoother``), asynchronously on the
    current stream.  The blocks never cross PCIe."""
    if data.nsys != 1 or isinstance(data, BlockTriDiagData_group):
        raise ValueError("build_smoother fills a single-system plan")
    n = data.n
    if model.F.shape != (n, n):
        raise ValueError(f"model state size {model.F.shape} does not match n={n}")
    dev = data.device
    # column-major device copies: the row-major storage of M^T is M column by column
    F, H, Qi, Ri = [torch.from_numpy(np.ascontiguousarray(np.asarray(M, dtype=np.float64).T)).to(dev)
                    for M in (model.F, model.H, model.Qinv, model.Rinv)]
    code = L.BSS_F64 if data.dtype == torch.float64 else L.BSS_F32
    with torch.cuda.device(dev):
        L.call("bss_build_smoother", data.n_user, n, model.H.shape[0], F.data_ptr(),
               H.data_ptr(), Qi.data_ptr(), Ri.data_ptr(), data._a.data_ptr(),
               data._b.data_ptr() if data._b.numel() else None, code, _stream_handle(dev))
    data._model_tensors = (F, H, Qi, Ri)  # alive until the stream has used them


def reference_schedule(N: int):
    """The reference's find_sequence_upper (src/gpu_batched.jl:155-162) as computed by
    libbss: (N_use, [(N_i, m_i, P_i) top level first])."""
    n_use = ctypes.c_int()
    buf = (ctypes.c_int * (3 * 64))()
    k = L.call("bss_reference_schedule", N, ctypes.byref(n_use), buf, 64)
    return n_use.value, [tuple(buf[3 * i:3 * i + 3]) for i in range(k)]
