"""Seeded synthetic block-tridiagonal SPD systems.

The generators restate the reference's test/benchmark data (test/utils.jl:1-30,
benchmark/utils.jl:5-34): A_i = G_i G_i^T + alpha*n*I with G_i n x n i.i.d. N(0,1),
off-diagonal couplings B_i i.i.d. N(0,1) (scaled by ``bscale``), right-hand sides
d_i ~ U(0,1).  The matrix is M = tridiag(B_{i-1}^T, A_i, B_i).

All arrays use the reference tensor layout (src/gpu_seq.jl:29-43): Fortran-ordered
numpy arrays A (n, n, N), B (n, n, N-1), d (n, nrhs, N), i.e. every block is
column-major and blocks are contiguous.  ``numpy.random.default_rng(seed)`` (PCG64)
makes every system reproducible from its seed on any machine.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np


@dataclass
class BlockSystem:
    A: np.ndarray  # (n, n, N) Fortran order
    B: np.ndarray  # (n, n, N-1)
    d: np.ndarray  # (n, nrhs, N)
    x: Optional[np.ndarray] = None  # manufactured solution, when known

    @property
    def N(self) -> int:
        return self.A.shape[2]

    @property
    def n(self) -> int:
        return self.A.shape[0]

    @property
    def nrhs(self) -> int:
        return self.d.shape[1]


def _blocks_to_F(blocks_rc: np.ndarray) -> np.ndarray:
    """(N, n, m) array with [i, r, c] = X_i[r, c] -> Fortran (n, m, N)."""
    return np.asfortranarray(np.transpose(blocks_rc, (1, 2, 0)))


def block_matvec(A: np.ndarray, B: np.ndarray, x: np.ndarray) -> np.ndarray:
    """y = M x with numpy batched products (any dtype; used for manufactured RHS)."""
    At = np.transpose(A, (2, 0, 1))
    Bt = np.transpose(B, (2, 0, 1))
    xt = np.transpose(x, (2, 0, 1))
    y = At @ xt
    if Bt.shape[0]:
        y[:-1] += Bt @ xt[1:]
        y[1:] += np.transpose(Bt, (0, 2, 1)) @ xt[:-1]
    return _blocks_to_F(y)


def spd_system(N: int, n: int, nrhs: int = 1, seed: int = 42, alpha: float = 1.0,
               bscale: float = 1.0, dtype=np.float64, manufactured: bool = False
               ) -> BlockSystem:
    """Config-1/2/3 generator.  With ``manufactured`` the RHS is d = M x for
    x ~ U(0,1) (the reference tests' pattern, test/test_cpu.jl:27-36); otherwise
    d ~ U(0,1) directly (the benchmark configs)."""
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((N, n, n))
    A = G @ np.transpose(G, (0, 2, 1))
    A[:, np.arange(n), np.arange(n)] += alpha * n
    B = bscale * rng.standard_normal((max(N - 1, 0), n, n))
    if manufactured:
        x = _blocks_to_F(rng.uniform(size=(N, n, nrhs)))
        Af, Bf = _blocks_to_F(A), _blocks_to_F(B)
        d = block_matvec(Af, Bf, x)
        return BlockSystem(Af.astype(dtype, order="F"), Bf.astype(dtype, order="F"),
                           d.astype(dtype, order="F"), x.astype(dtype, order="F"))
    d = rng.uniform(size=(N, n, nrhs))
    return BlockSystem(_blocks_to_F(A).astype(dtype, order="F"),
                       _blocks_to_F(B).astype(dtype, order="F"),
                       _blocks_to_F(d).astype(dtype, order="F"))


def spd_bidiag_system(N: int, n: int, nrhs: int = 1, seed: int = 42, alpha: float = 1.0,
                      dtype=np.float64) -> BlockSystem:
    """SPD-by-construction variant of the mixed-precision config: M = R^T R + alpha*n*I
    with R block upper-bidiagonal (G_i on the diagonal, 0.5*H_i above), G_i, H_i
    i.i.d. N(0,1).  Blocks: A_i = G_i^T G_i + 0.25 H_{i-1}^T H_{i-1} + alpha*n*I,
    B_i = 0.5 G_i^T H_i.  Small alpha gives condition numbers ~1e2..1e6."""
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((N, n, n))
    H = rng.standard_normal((max(N - 1, 0), n, n))
    Gt = np.transpose(G, (0, 2, 1))
    A = Gt @ G
    if N > 1:
        A[1:] += 0.25 * np.transpose(H, (0, 2, 1)) @ H
    A[:, np.arange(n), np.arange(n)] += alpha * n
    B = 0.5 * Gt[:-1] @ H if N > 1 else np.zeros((0, n, n))
    d = rng.uniform(size=(N, n, nrhs))
    return BlockSystem(_blocks_to_F(A).astype(dtype, order="F"),
                       _blocks_to_F(B).astype(dtype, order="F"),
                       _blocks_to_F(d).astype(dtype, order="F"))


@dataclass
class SmootherModel:
    """Linear state-space model x_{i+1} = F x_i + w, y_i = H x_i + v with w ~ N(0, Q),
    v ~ N(0, R); kept as the inverses the normal equations use."""
    F: np.ndarray     # (n, n)
    H: np.ndarray     # (m, n)
    Qinv: np.ndarray  # (n, n)
    Rinv: np.ndarray  # (m, m)

    def blocks(self):
        """(A_i, B_i): A = Q^{-1} + F^T Q^{-1} F + H^T R^{-1} H, B = -Q^{-1} F."""
        A = self.Qinv + self.F.T @ self.Qinv @ self.F + self.H.T @ self.Rinv @ self.H
        return A, -self.Qinv @ self.F


def smoother_model(n: int, rng: np.random.Generator) -> SmootherModel:
    """The served-stream model: F = 0.95*Q_orth (orthogonal factor of a Gaussian with
    sign fix, benchmark/benchmark_kf.jl:80-82), H (n/2) x n N(0,1), Q = 0.1 I, R = I."""
    Z = rng.standard_normal((n, n))
    Qo, Rr = np.linalg.qr(Z)
    Qo = Qo * np.sign(np.diag(Rr))[None, :]
    H = rng.standard_normal((max(n // 2, 1), n))
    return SmootherModel(0.95 * Qo, H, 10.0 * np.eye(n), np.eye(H.shape[0]))


def smoother_system(N: int, n: int, nrhs: int = 1, seed: int = 42,
                    dtype=np.float64) -> BlockSystem:
    """State-space smoother normal equations (benchmark/benchmark_kf.jl:79-154 as the
    served-stream config states them): F = 0.95*Q_orth, H (n/2) x n N(0,1), Q = 0.1 I,
    R = I; A_i = Q^{-1} + F^T Q^{-1} F + H^T R^{-1} H, B_i = -Q^{-1} F.  The reference
    benchmark uses the transpose of B_i and special-cases A_1, A_N; this follows the
    config text."""
    rng = np.random.default_rng(seed)
    Ai, Bi = smoother_model(n, rng).blocks()
    A = np.broadcast_to(Ai, (N, n, n)).copy()
    B = np.broadcast_to(Bi, (max(N - 1, 0), n, n)).copy()
    d = rng.uniform(size=(N, n, nrhs))
    return BlockSystem(_blocks_to_F(A).astype(dtype, order="F"),
                       _blocks_to_F(B).astype(dtype, order="F"),
                       _blocks_to_F(d).astype(dtype, order="F"))
