"""Writes the golden fixtures under tests/golden/ (run from the repo root:
``python tests/golden/make_golden.py``).

The reference holds no golden vectors or fixtures for this path (its tests are
random with tolerance checks, test/test_cpu.jl:71,92; test/test_rocm.jl:118) and its
Julia code cannot run here (no Julia toolchain).  The fixtures are therefore produced
by this repo's FP64 oracle (oracle/bss_oracle.c, the restatement of
src/backbone_cpu.jl) and accepted only where scipy's dense Cholesky of the assembled
matrix agrees to 1e-13 relative.  Each fixture stores its inputs (small cases) or the
generator spec plus a SHA-256 of the generated inputs (config 1), and the solution.
"""
import hashlib
import os
import sys

import numpy as np
import scipy.linalg as sla

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from blockstructuredsolvers_amd import workloads  # noqa: E402
from oracle import oracle  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def dense(A, B):
    n, _, N = A.shape
    M = np.zeros((N * n, N * n))
    for i in range(N):
        M[i * n:(i + 1) * n, i * n:(i + 1) * n] = A[:, :, i]
        if i + 1 < N:
            M[i * n:(i + 1) * n, (i + 1) * n:(i + 2) * n] = B[:, :, i]
            M[(i + 1) * n:(i + 2) * n, i * n:(i + 1) * n] = B[:, :, i].T
    return M


def input_digest(s):
    h = hashlib.sha256()
    for a in (s.A, s.B, s.d):
        h.update(np.ascontiguousarray(np.asfortranarray(a)).tobytes(order="F"))
    return h.hexdigest()


def solve_checked(s):
    x, info = oracle.solve_system(s.A, s.B, s.d)
    assert info == 0
    n, nrhs, N = s.d.shape
    D = np.concatenate([s.d[:, :, i] for i in range(N)], axis=0)
    xd = sla.cho_solve(sla.cho_factor(dense(s.A, s.B)), D)
    X = np.concatenate([x[:, :, i] for i in range(N)], axis=0)
    rel = np.abs(X - xd).max() / np.abs(xd).max()
    assert rel <= 1e-13, rel
    return x, rel


def main():
    small = {}
    for name, (N, n, nrhs, seed, alpha) in {
        "N9_n16_r2": (9, 16, 2, 1, 1.0),
        "N50_n16_r1": (50, 16, 1, 2, 1.0),
        "N17_n32_r3": (17, 32, 3, 3, 1.0),
        "N2_n48_r1": (2, 48, 1, 4, 1.0),
    }.items():
        s = workloads.spd_system(N, n, nrhs=nrhs, seed=seed, alpha=alpha)
        x, rel = solve_checked(s)
        small[f"{name}_A"] = s.A
        small[f"{name}_B"] = s.B
        small[f"{name}_d"] = s.d
        small[f"{name}_x"] = x
        print(name, "oracle vs scipy dense", rel)
    np.savez_compressed(os.path.join(HERE, "small_systems.npz"), **small)

    s = workloads.spd_system(100, 32, nrhs=1, seed=42)
    x, rel = solve_checked(s)
    print("config1 oracle vs scipy dense", rel)
    np.savez_compressed(os.path.join(HERE, "config1_N100_n32.npz"), x=x,
                        spec=np.array([100, 32, 1, 42]), alpha=np.array(1.0),
                        digest=np.array(input_digest(s)))


if __name__ == "__main__":
    main()
