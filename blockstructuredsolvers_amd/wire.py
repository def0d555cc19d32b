"""On-wire request and result frames for the request loop, and the JSON result records.

The frames carry the reference's layout contract unchanged: every matrix is a packed
column-major (Fortran order) n1 x n2 x N block tensor, block i at offset i*n1*n2
(``create_matrix_list``, src/gpu_seq.jl:29-43), so a frame's payload can be copied
into a plan's A/B/d tensors without reordering.  All integers and floats are
little-endian.

Request frame (48-byte header, then the payload):

    offset size field
    0      4    magic b"BSSQ"
    4      2    version (1)
    6      2    flags: bit 0 = the blocks come as a state-space model (F, H, Q^-1,
                R^-1) that the device expands (bss_build_smoother), not as dense A, B
    8      8    request id (u64)
    16     8    arrival time in seconds (f64)
    24     4    N, blocks (u32, >= 1)
    28     4    n, block size (u32, >= 1)
    32     4    nrhs, columns per right-hand side (u32, >= 1)
    36     4    k, right-hand sides solved with one factorisation (u32, >= 1)
    40     4    m, observation size (u32; 0 without a model)
    44     1    dtype of A, B and d: 0 = f64, 1 = f32 (the model is always f64)
    45     3    reserved, zero
    48     ...  dense: A (n, n, N), B (n, n, N-1); model: F (n, n), H (m, n),
                Qinv (n, n), Rinv (m, m); then k right-hand sides d (n, nrhs, N)

Result frame (40-byte header, then k solutions (n, nrhs, N) in f64):

    0 magic b"BSSR" | 4 version u16 | 6 flags u16 (0) | 8 request id u64 |
    16 info i32 (0, or the 1-based first block that is not positive definite) |
    20 k u32 | 24 N u32 | 28 n u32 | 32 nrhs u32 | 36 dtype u8 (0) | 37 reserved

The JSON records (``result_record``, ``benchmark_record``) keep the fields the
reference's GPU benchmark writes per (N, n, solver) -- init, factorize, solve and total
milliseconds as mean and standard deviation, and the residual
(benchmark/benchmark_gpu.jl:227-239) -- plus the per-request serving fields.
"""
from __future__ import annotations

import struct
from typing import Iterable, List, Sequence, Tuple

import numpy as np

from .serving import Request, Result

REQ_MAGIC = b"BSSQ"
RES_MAGIC = b"BSSR"
VERSION = 1
FLAG_MODEL = 1

_REQ = struct.Struct("<4sHHQdIIIIIB3x")
_RES = struct.Struct("<4sHHQiIIIIB3x")
assert _REQ.size == 48 and _RES.size == 40

_DTYPES = {0: np.dtype("<f8"), 1: np.dtype("<f4")}
_CODES = {np.dtype("<f8"): 0, np.dtype("<f4"): 1}

RESULT_SCHEMA = "bss.result/1"
BENCHMARK_SCHEMA = "bss.benchmark/1"


def _put(arr: np.ndarray, shape: Tuple[int, ...], dt: np.dtype, name: str) -> bytes:
    a = np.asarray(arr)
    if a.shape != shape:
        raise ValueError(f"{name}: shape {a.shape}, expected {shape}")
    return np.asfortranarray(a, dtype=dt).tobytes(order="F")


class _Reader:
    def __init__(self, buf: bytes, off: int):
        self.buf, self.off = memoryview(buf), off

    def take(self, shape: Tuple[int, ...], dt: np.dtype, name: str) -> np.ndarray:
        count = int(np.prod(shape))
        nbytes = count * dt.itemsize
        if self.off + nbytes > len(self.buf):
            raise ValueError(f"frame truncated in {name}: needs {nbytes} bytes at offset "
                             f"{self.off}, {len(self.buf) - self.off} left")
        a = np.frombuffer(self.buf, dtype=dt, count=count, offset=self.off)
        self.off += nbytes
        # native-endian, writable, column-major copy
        return np.asfortranarray(a.reshape(shape, order="F").astype(dt.newbyteorder("="), copy=True))

    def done(self) -> None:
        if self.off != len(self.buf):
            raise ValueError(f"{len(self.buf) - self.off} trailing bytes after the payload")


def encode_request(req: Request, dtype=np.float64) -> bytes:
    """One request frame.  ``dtype`` (float64 or float32) is the payload type of A, B
    and the right-hand sides."""
    dt = np.dtype(dtype).newbyteorder("<")
    if dt not in _CODES:
        raise ValueError(f"unsupported payload dtype {dtype}")
    if not req.rhs:
        raise ValueError("a request needs at least one right-hand side")
    N, n, nrhs, k = req.N, req.n, req.nrhs, len(req.rhs)
    model = req.model is not None
    m = req.model.H.shape[0] if model else 0
    head = _REQ.pack(REQ_MAGIC, VERSION, FLAG_MODEL if model else 0, req.rid,
                     float(req.arrival), N, n, nrhs, k, m, _CODES[dt])
    parts = [head]
    f8 = np.dtype("<f8")
    if model:
        mo = req.model
        parts += [_put(mo.F, (n, n), f8, "F"), _put(mo.H, (m, n), f8, "H"),
                  _put(mo.Qinv, (n, n), f8, "Qinv"), _put(mo.Rinv, (m, m), f8, "Rinv")]
    else:
        B = req.B if req.B is not None or N > 1 else np.zeros((n, n, 0))
        if req.A is None or B is None:
            raise ValueError("a request needs dense A and B or a model")
        parts += [_put(req.A, (n, n, N), dt, "A"), _put(B, (n, n, N - 1), dt, "B")]
    for j, d in enumerate(req.rhs):
        parts.append(_put(d, (n, nrhs, N), dt, f"d[{j}]"))
    return b"".join(parts)


def decode_request(buf: bytes) -> Request:
    """Inverse of ``encode_request``; raises ValueError on a malformed frame."""
    if len(buf) < _REQ.size:
        raise ValueError(f"frame of {len(buf)} bytes is shorter than the {_REQ.size}-byte header")
    magic, ver, flags, rid, arrival, N, n, nrhs, k, m, code = _REQ.unpack_from(buf, 0)
    if magic != REQ_MAGIC:
        raise ValueError(f"bad magic {magic!r}, expected {REQ_MAGIC!r}")
    if ver != VERSION:
        raise ValueError(f"unsupported frame version {ver}")
    if flags & ~FLAG_MODEL:
        raise ValueError(f"unknown flags 0x{flags:x}")
    if code not in _DTYPES:
        raise ValueError(f"unknown dtype code {code}")
    if min(N, n, nrhs, k) < 1:
        raise ValueError(f"N, n, nrhs and k must be >= 1 (got {N}, {n}, {nrhs}, {k})")
    model = bool(flags & FLAG_MODEL)
    if model != (m > 0):
        raise ValueError("the observation size m must be > 0 exactly when a model is sent")
    dt = _DTYPES[code]
    r = _Reader(buf, _REQ.size)
    A = B = mo = None
    if model:
        from .workloads import SmootherModel
        f8 = np.dtype("<f8")
        mo = SmootherModel(r.take((n, n), f8, "F"), r.take((m, n), f8, "H"),
                           r.take((n, n), f8, "Qinv"), r.take((m, m), f8, "Rinv"))
    else:
        A = r.take((n, n, N), dt, "A")
        B = r.take((n, n, N - 1), dt, "B")
    rhs = [r.take((n, nrhs, N), dt, f"d[{j}]") for j in range(k)]
    r.done()
    return Request(rid=rid, arrival=arrival, rhs=rhs, A=A, B=B, model=mo)


def encode_result(res: Result) -> bytes:
    """One result frame: the solutions in f64."""
    if not res.x:
        raise ValueError("a result carries at least one solution")
    n, nrhs, N = res.x[0].shape
    head = _RES.pack(RES_MAGIC, VERSION, 0, res.rid, int(res.info), len(res.x), N, n, nrhs, 0)
    f8 = np.dtype("<f8")
    return head + b"".join(_put(x, (n, nrhs, N), f8, f"x[{j}]") for j, x in enumerate(res.x))


def decode_result(buf: bytes):
    """(rid, info, [x_j]) from a result frame; raises ValueError on a malformed frame."""
    if len(buf) < _RES.size:
        raise ValueError(f"frame of {len(buf)} bytes is shorter than the {_RES.size}-byte header")
    magic, ver, flags, rid, info, k, N, n, nrhs, code = _RES.unpack_from(buf, 0)
    if magic != RES_MAGIC:
        raise ValueError(f"bad magic {magic!r}, expected {RES_MAGIC!r}")
    if ver != VERSION or flags != 0 or code != 0:
        raise ValueError(f"unsupported result frame (version {ver}, flags {flags}, dtype {code})")
    if min(N, n, nrhs, k) < 1:
        raise ValueError(f"N, n, nrhs and k must be >= 1 (got {N}, {n}, {nrhs}, {k})")
    r = _Reader(buf, _RES.size)
    xs = [r.take((n, nrhs, N), np.dtype("<f8"), f"x[{j}]") for j in range(k)]
    r.done()
    return rid, info, xs


def result_record(res: Result) -> dict:
    """JSON-serialisable record of one served request (no solution payload)."""
    n, nrhs, N = res.x[0].shape
    return {"schema": RESULT_SCHEMA, "rid": int(res.rid), "info": int(res.info),
            "N": int(N), "n": int(n), "nrhs": int(nrhs), "rhs": len(res.x),
            "arrival_s": float(res.arrival), "admitted_s": float(res.admitted),
            "finished_s": float(res.finished), "latency_ms": float(res.latency_ms),
            "solve_ms": [float(t) for t in res.solve_ms]}


def _mean_std(v: Sequence[float]) -> Tuple[float, float]:
    a = np.asarray(v, dtype=np.float64)
    # the reference's std (Statistics.std) is the corrected sample deviation
    return float(a.mean()), float(a.std(ddof=1)) if a.size > 1 else 0.0


def benchmark_record(N: int, n: int, solver: str, timings: Iterable[Tuple[float, float, float]],
                     residuals: Iterable[float]) -> dict:
    """The per-(N, n, solver) summary of benchmark/benchmark_gpu.jl:227-239 as JSON:
    ``timings`` are (init, factorize, solve) seconds per iteration, total = factorize +
    solve; times in ms as mean and standard deviation."""
    t = np.asarray(list(timings), dtype=np.float64).reshape(-1, 3)
    res = list(residuals)
    out = {"schema": BENCHMARK_SCHEMA, "N": int(N), "n": int(n), "total_size": int(N * n),
           "solver": solver, "iterations": int(t.shape[0])}
    for name, col in (("init", t[:, 0]), ("factorize", t[:, 1]), ("solve", t[:, 2]),
                      ("total", t[:, 1] + t[:, 2])):
        mu, sd = _mean_std(col * 1e3)
        out[f"{name}_ms"] = {"mean": mu, "std": sd}
    mu, sd = _mean_std(res)
    out["residual"] = {"mean": mu, "std": sd}
    return out


def serve_frames(server, frames: Sequence[bytes], realtime: bool = True):
    """Decode request frames, run them through ``server`` (a ``serving.Server``) and
    return (result frames, JSON records), both in request order."""
    reqs: List[Request] = [decode_request(f) for f in frames]
    if len({q.rid for q in reqs}) != len(reqs):
        raise ValueError("request ids must be unique within one serve_frames call")
    results = server.serve(reqs, realtime=realtime)
    by_id = {r.rid: r for r in results}
    ordered = [by_id[q.rid] for q in reqs]
    return [encode_result(r) for r in ordered], [result_record(r) for r in ordered]
