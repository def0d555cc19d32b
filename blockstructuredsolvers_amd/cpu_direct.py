"""CPU direct-solver baselines on the assembled matrix (context for the GPU numbers, not
a parity check).

The reference's benchmark/benchmark_cpu.jl:21-99 factors the assembled sparse matrix
M = tridiag(B^T, A, B) (N n x N n) with HSL MA57, SuiteSparse CHOLMOD and
LDLFactorizations, and reports init / factorize / solve times and the residual
||M x - d||_2 (compute_residual_cpu, benchmark_cpu.jl:103-119).  None of those libraries
is in this image; scipy is, so the same measurement runs two direct solvers it has:

* LAPACK banded Cholesky (``dpbtrf`` + ``dpbtrs`` through scipy.linalg.lapack): M is
  banded with half-bandwidth 2n - 1, so this is the classic dense-band direct method
  (its ``info`` is the first non-positive pivot column, in the same sequential order as
  the block Cholesky of backbone_cpu.jl:5-15);
* SuperLU (``scipy.sparse.linalg.splu``), a general sparse LU, with the natural
  ordering and symmetric mode (the band is already the best ordering for this matrix).

    python -m blockstructuredsolvers_amd.cpu_direct --N 8192 --n 64
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from typing import Dict

import numpy as np

from .workloads import BlockSystem, block_matvec, spd_system


def assemble_banded(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """Upper band storage of M for LAPACK ?pbtrf (u = 2n - 1 superdiagonals):
    ab[u + r - j, j] = M[r, j] for j - u <= r <= j."""
    n, _, N = A.shape
    u = 2 * n - 1
    ab = np.zeros((u + 1, N * n), dtype=np.float64)
    base = np.arange(N) * n
    a, c = np.triu_indices(n)  # diagonal blocks, a <= c
    ab[(u - c + a)[:, None], base[None, :] + c[:, None]] = A[a, c, :]
    if N > 1:  # B_{i-1} sits in rows (i-1) n + a of block column i
        a2, c2 = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        a2, c2 = a2.ravel(), c2.ravel()
        ab[(u - n + a2 - c2)[:, None], base[None, 1:] + c2[:, None]] = B[a2, c2, :]
    return ab


def assemble_csc(A: np.ndarray, B: np.ndarray):
    """M as a scipy CSC matrix (both triangles)."""
    import scipy.sparse as sp
    n, _, N = A.shape
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    off = (np.arange(N) * n)[:, None]
    rows = [(off + ii).ravel()]
    cols = [(off + jj).ravel()]
    vals = [np.moveaxis(A, 2, 0).reshape(N, -1).ravel()]
    if N > 1:
        offB = off[:-1]
        Bv = np.moveaxis(B, 2, 0).reshape(N - 1, -1).ravel()
        rows += [(offB + ii).ravel(), (offB + n + jj).ravel()]
        cols += [(offB + n + jj).ravel(), (offB + ii).ravel()]
        vals += [Bv, Bv]
    return sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(N * n, N * n))


def _flat(d: np.ndarray) -> np.ndarray:
    """(n, nrhs, N) column-major blocks -> (N n, nrhs)."""
    n, nrhs, N = d.shape
    return np.moveaxis(d, 2, 0).reshape(N * n, nrhs, order="C")


def _blocks(x: np.ndarray, n: int) -> np.ndarray:
    N = x.shape[0] // n
    return np.asfortranarray(np.moveaxis(x.reshape(N, n, -1), 0, 2))


def banded_cholesky(sysm: BlockSystem) -> Dict:
    """dpbtrf + dpbtrs: times (s), solution in the block layout, LAPACK info (the
    1-based column of the first non-positive pivot, 0 = SPD)."""
    from scipy.linalg import lapack
    n = sysm.A.shape[0]
    t0 = time.perf_counter()
    ab = assemble_banded(sysm.A, sysm.B)
    t1 = time.perf_counter()
    cb, info = lapack.dpbtrf(ab, lower=0, overwrite_ab=1)
    t2 = time.perf_counter()
    x = None
    if info == 0:
        x, info2 = lapack.dpbtrs(cb, _flat(sysm.d), lower=0)
        if info2 != 0:
            raise RuntimeError(f"dpbtrs info {info2}")
    t3 = time.perf_counter()
    return {"init_s": t1 - t0, "factor_s": t2 - t1, "solve_s": t3 - t2, "info": int(info),
            "x": None if x is None else _blocks(x, n)}


def superlu(sysm: BlockSystem) -> Dict:
    """scipy's SuperLU with the natural ordering in symmetric mode."""
    import scipy.sparse.linalg as spl
    n = sysm.A.shape[0]
    t0 = time.perf_counter()
    M = assemble_csc(sysm.A, sysm.B)
    t1 = time.perf_counter()
    lu = spl.splu(M, permc_spec="NATURAL", options=dict(SymmetricMode=True))
    t2 = time.perf_counter()
    x = lu.solve(_flat(sysm.d))
    t3 = time.perf_counter()
    return {"init_s": t1 - t0, "factor_s": t2 - t1, "solve_s": t3 - t2,
            "fill_nnz": int(lu.L.nnz + lu.U.nnz), "x": _blocks(x, n)}


def residual_2norm(sysm: BlockSystem, x: np.ndarray) -> float:
    """||M x - d||_2 over all blocks and columns (benchmark_cpu.jl:103-119)."""
    return float(np.linalg.norm(block_matvec(sysm.A, sysm.B, x) - sysm.d))


# SuperLU (32-bit internal indexing) runs out of memory above about this many unknowns
SUPERLU_MAX_UNKNOWNS = 2048 * 64


def run(N: int, n: int, seed: int = 42, solvers=("banded_cholesky", "superlu")) -> Dict:
    """One seeded system (the benchmark generator) through each solver: the fields of
    benchmark_cpu.jl's report, in ms, plus the residual and the BLAS threads used.
    SuperLU gets the leading SUPERLU_MAX_UNKNOWNS // n blocks when the system is larger
    (its record names its own N)."""
    sysm = spd_system(N, n, nrhs=1, seed=seed)
    out = {"N": N, "n": n, "unknowns": N * n,
           "threads": os.environ.get("OMP_NUM_THREADS", "library default")}
    for name in solvers:
        s_run, N_run = sysm, N
        if name == "superlu" and N * n > SUPERLU_MAX_UNKNOWNS:
            N_run = SUPERLU_MAX_UNKNOWNS // n
            s_run = BlockSystem(np.asfortranarray(sysm.A[:, :, :N_run]),
                                np.asfortranarray(sysm.B[:, :, :N_run - 1]),
                                np.asfortranarray(sysm.d[:, :, :N_run]))
        r = {"banded_cholesky": banded_cholesky, "superlu": superlu}[name](s_run)
        rec = {k: r[k] * 1e3 for k in ("init_s", "factor_s", "solve_s")}
        rec = {k.replace("_s", "_ms"): v for k, v in rec.items()}
        rec["total_ms"] = rec["factor_ms"] + rec["solve_ms"]
        rec["N"] = N_run
        if r.get("x") is not None:
            rec["residual_2norm"] = residual_2norm(s_run, r["x"])
        for k in ("info", "fill_nnz"):
            if k in r:
                rec[k] = r[k]
        out[name] = rec
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=8192)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--solvers", default="banded_cholesky,superlu")
    a = ap.parse_args(argv)
    print(json.dumps(run(a.N, a.n, a.seed, tuple(a.solvers.split(",")))), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
