"""CPU oracle for the block-tridiagonal Cholesky path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import
this module, and only as the checker (or the timed CPU baseline).  The product
package never imports it.

Two restatements of the reference live here:

* ``factorize`` / ``solve`` -- ctypes bindings of ``bss_oracle.c``, the C
  restatement of the sequential block Cholesky of ``src/cpu.jl`` +
  ``src/backbone_cpu.jl:5-33`` (the FP64 block-Cholesky path of the reference).
* ``ref_build_sequence`` / ``ref_find_sequence_upper`` -- the reference's integer
  level-schedule search, ``src/gpu_batched.jl:112-162``, and
  ``ref_batched_factorize`` / ``ref_batched_solve`` -- a small numpy restatement of
  the reference's recursive batched Schur algorithm (``src/gpu_batched.jl:184-259``
  with the batched backbone ``src/backbone_rocm.jl:62-170``), used to show that the
  recursive algorithm and the sequential oracle agree and to pin the schedule.

Array convention (same as the reference's tensors, ``src/gpu_seq.jl:29-43``):
Fortran-ordered numpy arrays A (n, n, N), B (n, n, N-1), d (n, nrhs, N).
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run `make -C oracle`)")
        lib = ctypes.CDLL(path)
        dp = ctypes.c_void_p
        for sfx in ("f64", "f32"):
            getattr(lib, f"oracle_factorize_{sfx}").argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
            getattr(lib, f"oracle_factorize_{sfx}").restype = ctypes.c_int
            getattr(lib, f"oracle_solve_{sfx}").argtypes = [ctypes.c_int] * 3 + [dp, dp, dp]
            getattr(lib, f"oracle_solve_{sfx}").restype = None
            getattr(lib, f"oracle_matvec_{sfx}").argtypes = [ctypes.c_int] * 3 + [dp] * 4
            getattr(lib, f"oracle_matvec_{sfx}").restype = None
        lib.oracle_backward_error_f64.argtypes = [ctypes.c_int] * 3 + [dp] * 6
        lib.oracle_backward_error_f64.restype = ctypes.c_double
        _LIB = lib
    return _LIB


def _sfx(a: np.ndarray) -> str:
    if a.dtype == np.float64:
        return "f64"
    if a.dtype == np.float32:
        return "f32"
    raise TypeError(f"unsupported dtype {a.dtype}")


def _check(a: np.ndarray, shape, dtype):
    if a.shape != tuple(shape) or a.dtype != dtype or not a.flags.f_contiguous:
        raise ValueError(f"expected Fortran-ordered {dtype} array of shape {shape}, got "
                         f"{a.dtype} {a.shape} f_contiguous={a.flags.f_contiguous}")


def factorize(A: np.ndarray, B: np.ndarray) -> int:
    """In-place sequential block Cholesky (backbone_cpu.jl:5-15). Returns info (0 = ok,
    k = first block, 1-based, whose potrf failed)."""
    n, _, N = A.shape
    _check(A, (n, n, N), A.dtype)
    _check(B, (n, n, max(N - 1, 0)), A.dtype)
    f = getattr(_lib(), f"oracle_factorize_{_sfx(A)}")
    return int(f(N, n, A.ctypes.data, B.ctypes.data if B.size else None))


def solve(A: np.ndarray, B: np.ndarray, d: np.ndarray) -> None:
    """In-place forward/back substitution with the factors (backbone_cpu.jl:17-33)."""
    n, _, N = A.shape
    nrhs = d.shape[1]
    _check(d, (n, nrhs, N), A.dtype)
    f = getattr(_lib(), f"oracle_solve_{_sfx(A)}")
    f(N, n, nrhs, A.ctypes.data, B.ctypes.data if B.size else None, d.ctypes.data)


def matvec(A: np.ndarray, B: np.ndarray, x: np.ndarray) -> np.ndarray:
    """y = M x for the block-tridiagonal operator (test/utils.jl:22-27)."""
    n, _, N = A.shape
    y = np.zeros_like(x, order="F")
    f = getattr(_lib(), f"oracle_matvec_{_sfx(A)}")
    f(N, n, x.shape[1], A.ctypes.data, B.ctypes.data if B.size else None, x.ctypes.data,
      y.ctypes.data)
    return y


def backward_error(A: np.ndarray, B: np.ndarray, x: np.ndarray, d: np.ndarray) -> np.ndarray:
    """Per-column normwise backward error ||d-Mx||_inf / (||M||_inf ||x||_inf + ||d||_inf),
    evaluated in FP64 on FP64 copies of the inputs."""
    A64 = np.asfortranarray(A, dtype=np.float64)
    B64 = np.asfortranarray(B, dtype=np.float64)
    x64 = np.asfortranarray(x, dtype=np.float64)
    d64 = np.asfortranarray(d, dtype=np.float64)
    n, nrhs, N = x64.shape
    work = np.zeros_like(x64, order="F")
    berr = np.zeros(nrhs, dtype=np.float64)
    _lib().oracle_backward_error_f64(N, n, nrhs, A64.ctypes.data,
                                     B64.ctypes.data if B64.size else None, x64.ctypes.data,
                                     d64.ctypes.data, work.ctypes.data, berr.ctypes.data)
    return berr


def solve_system(A: np.ndarray, B: np.ndarray, d: np.ndarray) -> Tuple[np.ndarray, int]:
    """factorize! then solve! on copies (the reference's cpu.jl flow); returns (x, info)."""
    Af, Bf, x = (np.array(a, order="F", copy=True) for a in (A, B, d))
    info = factorize(Af, Bf)
    solve(Af, Bf, x)
    return x, info


def bss_norm(blocks: np.ndarray) -> float:
    """``_bss_norm`` of test/test_cpu.jl:1-7: 2-norm of the concatenated blocks."""
    return float(np.sqrt(np.sum(np.asarray(blocks, dtype=np.float64) ** 2)))


# ---------------------------------------------------------------------------------------
# Reference level schedule, src/gpu_batched.jl:112-162 (integer logic only).
# ---------------------------------------------------------------------------------------
def ref_build_sequence(start: int) -> Tuple[List[int], List[int], List[int]]:
    """Restates ``_build_sequence`` (gpu_batched.jl:112-143): walk down from N, at each
    level taking the smallest m >= 3 with (N + m) % (m + 1) == 0 and P >= 3, preferring
    m < P and otherwise the first acceptable m; lists are ordered bottom -> top."""
    P_list: List[int] = []
    N_list: List[int] = []
    m_list: List[int] = []
    ncur = start
    while True:
        best = None
        backup = None
        max_m = (ncur - 3) // 2
        for m in range(3, max_m + 1):
            num, den = ncur + m, m + 1
            if num % den == 0:
                pprev = num // den
                if pprev >= 3:
                    if m < pprev:
                        best = (m, pprev)
                        break
                    if backup is None:
                        backup = (m, pprev)
        if best is None and backup is not None:
            best = backup
        if best is None:
            break
        P_list.insert(0, best[1])
        N_list.insert(0, ncur)
        m_list.insert(0, best[0])
        ncur = best[1]
    return P_list, N_list, m_list


def ref_find_sequence_upper(N: int) -> Tuple[int, List[int], List[int], List[int]]:
    """Restates ``find_sequence_upper`` (gpu_batched.jl:155-162).  The reference returns
    only the three lists; N_use (the padded size its docstring promises) is
    ``N_list[-1]``, the top level, and is returned here explicitly."""
    n_use = N
    while True:
        P_list, N_list, m_list = ref_build_sequence(n_use)
        if P_list:
            return n_use, P_list, N_list, m_list
        n_use += 1


# ---------------------------------------------------------------------------------------
# Reference recursive batched algorithm, numpy restatement (small sizes only).
# ---------------------------------------------------------------------------------------
def _chol_u(a):
    return np.linalg.cholesky(a).T  # upper U with A = U^T U


def _seq_factorize(A, B):
    """Sequential block Cholesky on lists of blocks (backbone_rocm.jl:14-26 semantics)."""
    N = len(A)
    A[0] = _chol_u(A[0])
    for i in range(1, N):
        B[i - 1] = np.linalg.solve(A[i - 1].T, B[i - 1])
        A[i] = _chol_u(A[i] - B[i - 1].T @ B[i - 1])


def _seq_solve(A, B, d):
    """backbone_rocm.jl:28-53."""
    N = len(A)
    d[0] = np.linalg.solve(A[0].T, d[0])
    for i in range(1, N):
        d[i] = np.linalg.solve(A[i].T, d[i] - B[i - 1].T @ d[i - 1])
    d[N - 1] = np.linalg.solve(A[N - 1], d[N - 1])
    for i in range(N - 2, -1, -1):
        d[i] = np.linalg.solve(A[i], d[i] - B[i] @ d[i + 1])


class RefBatchedLevel:
    """One level of ``BlockTriDiagData_batched`` (gpu_batched.jl:1-71): separators at
    0-based blocks k(m+1), k = 0..P-1; P-1 chains of m interior blocks."""

    def __init__(self, N, m, P, n):
        self.N, self.m, self.P, self.n = N, m, P, n
        self.F = None
        self.M1 = self.M2 = None
        self.next = None


def ref_batched_factorize(A, B, levels):
    """Restates factorize!(::BlockTriDiagData_batched) (gpu_batched.jl:184-227) on Python
    lists of blocks: gather separators and couplings, factor the P-1 interior chains
    (backbone_rocm.jl:62-82), F = T^{-1} G with 2n right-hand sides
    (backbone_rocm.jl:84-120), Schur complement into the reduced tridiagonal system
    (backbone_rocm.jl:122-138), then recurse; the last level is sequential."""
    lev = levels[0]
    if not isinstance(lev, RefBatchedLevel):
        _seq_factorize(A, B)
        return
    N, m, P, n = lev.N, lev.m, lev.P, lev.n
    s = [k * (m + 1) for k in range(P)]
    lev.M1 = [B[s[k]].copy() for k in range(P - 1)]
    lev.M2 = [B[s[k] + m].copy() for k in range(P - 1)]
    LHS_A = [A[s[k]].copy() for k in range(P)]
    LHS_B = [np.zeros((n, n)) for _ in range(P - 1)]
    G = [np.zeros((n, 2 * n)) for _ in range(N)]
    for k in range(P - 1):
        G[s[k] + 1][:, :n] = lev.M1[k].T
        G[s[k] + m][:, n:] = lev.M2[k]
    F = [g.copy() for g in G]
    for k in range(P - 1):
        idx = list(range(s[k] + 1, s[k] + m + 1))
        Ac = [A[i] for i in idx]
        Bc = [B[i] for i in idx[:-1]]
        _seq_factorize(Ac, Bc)
        for j, i in enumerate(idx):
            A[i] = Ac[j]
        for j, i in enumerate(idx[:-1]):
            B[i] = Bc[j]
        Fc = [F[i] for i in idx]
        _seq_solve(Ac, Bc, Fc)
        for j, i in enumerate(idx):
            F[i] = Fc[j]
        M2n = sum(G[i].T @ F[i] for i in idx)
        LHS_A[k] = LHS_A[k] - M2n[:n, :n]
        LHS_A[k + 1] = LHS_A[k + 1] - M2n[n:, n:]
        LHS_B[k] = LHS_B[k] - M2n[:n, n:]
    lev.F = F
    lev.next_A, lev.next_B = LHS_A, LHS_B
    ref_batched_factorize(LHS_A, LHS_B, levels[1:])
    lev.next_A, lev.next_B = LHS_A, LHS_B


def ref_batched_solve(A, B, d, levels):
    """Restates solve!(::BlockTriDiagData_batched) (gpu_batched.jl:229-259): reduced RHS
    (backbone_rocm.jl:140-154), recursive separator solve, boundary update
    (backbone_rocm.jl:156-170), interior chain solves."""
    lev = levels[0]
    if not isinstance(lev, RefBatchedLevel):
        _seq_solve(A, B, d)
        return
    N, m, P, n = lev.N, lev.m, lev.P, lev.n
    s = [k * (m + 1) for k in range(P)]
    for k in range(P - 1):
        idx = range(s[k] + 1, s[k] + m + 1)
        b = -sum(lev.F[i].T @ d[i] for i in idx)
        d[s[k]] = d[s[k]] + b[:n]
        d[s[k + 1]] = d[s[k + 1]] + b[n:]
    dn = [d[s[k]].copy() for k in range(P)]
    ref_batched_solve(lev.next_A, lev.next_B, dn, levels[1:])
    for k in range(P):
        d[s[k]] = dn[k]
    for k in range(P - 1):
        d[s[k] + 1] = d[s[k] + 1] - lev.M1[k].T @ d[s[k]]
        d[s[k] + m] = d[s[k] + m] - lev.M2[k] @ d[s[k + 1]]
        idx = list(range(s[k] + 1, s[k] + m + 1))
        Ac = [A[i] for i in idx]
        Bc = [B[i] for i in idx[:-1]]
        dc = [d[i] for i in idx]
        _seq_solve(Ac, Bc, dc)
        for j, i in enumerate(idx):
            d[i] = dc[j]


def ref_batched_solve_system(A, B, d):
    """initialize_batched (gpu_batched.jl:165-182: schedule + identity padding to N_use)
    + factorize! + solve! on copies; returns the leading N blocks of the solution.
    Sizes are restricted to what a pure-numpy loop finishes in seconds."""
    n, _, N = A.shape
    nrhs = d.shape[1]
    n_use, P_list, N_list, m_list = ref_find_sequence_upper(N)
    Al = [np.eye(n) for _ in range(n_use)]
    Bl = [np.zeros((n, n)) for _ in range(n_use - 1)]
    dl = [np.zeros((n, nrhs)) for _ in range(n_use)]
    for i in range(N):
        Al[i] = np.array(A[:, :, i], dtype=np.float64)
        dl[i] = np.array(d[:, :, i], dtype=np.float64)
    for i in range(N - 1):
        Bl[i] = np.array(B[:, :, i], dtype=np.float64)
    # levels top -> bottom, then the sequential base
    levels = [RefBatchedLevel(Ni, mi, Pi, n)
              for Ni, mi, Pi in reversed(list(zip(N_list, m_list, P_list)))]
    levels.append("seq")
    ref_batched_factorize(Al, Bl, levels)
    ref_batched_solve(Al, Bl, dl, levels)
    x = np.zeros((n, nrhs, N), order="F")
    for i in range(N):
        x[:, :, i] = dl[i]
    return x
