"""Golden fixtures (tests/golden/, written by make_golden.py): the oracle and the GPU
path reproduce the stored solutions, and the seeded generator still produces the
exact config-1 inputs the fixture was made from."""
import hashlib
import os

import numpy as np
import pytest

from blockstructuredsolvers_amd import workloads
from oracle import oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def small_cases():
    z = np.load(os.path.join(HERE, "small_systems.npz"))
    names = sorted({k.rsplit("_", 1)[0] for k in z.files})
    return [(nm, z[nm + "_A"], z[nm + "_B"], z[nm + "_d"], z[nm + "_x"]) for nm in names]


def config1():
    z = np.load(os.path.join(HERE, "config1_N100_n32.npz"))
    N, n, nrhs, seed = (int(v) for v in z["spec"])
    s = workloads.spd_system(N, n, nrhs=nrhs, seed=seed, alpha=float(z["alpha"]))
    return s, z["x"], str(z["digest"])


def digest(s):
    h = hashlib.sha256()
    for a in (s.A, s.B, s.d):
        h.update(np.asfortranarray(a).tobytes(order="F"))
    return h.hexdigest()


@pytest.mark.parametrize("case", small_cases(), ids=lambda c: c[0])
def test_oracle_reproduces_small_fixtures(case):
    _, A, B, d, x = case
    xo, info = oracle.solve_system(np.asfortranarray(A), np.asfortranarray(B),
                                   np.asfortranarray(d))
    assert info == 0
    assert np.abs(xo - x).max() <= 1e-13 * np.abs(x).max()


def test_generator_and_oracle_reproduce_config1():
    s, x, dig = config1()
    assert digest(s) == dig, "seeded generator output changed"
    xo, _ = oracle.solve_system(s.A, s.B, s.d)
    assert np.abs(xo - x).max() <= 1e-13 * np.abs(x).max()


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["auto", "reference"])
def test_gpu_reproduces_fixtures(schedule):
    import torch

    from blockstructuredsolvers_amd import factorize, initialize_batched, solve
    cases = [(nm, A, B, d, x) for nm, A, B, d, x in small_cases()]
    s, x1, _ = config1()
    cases.append(("config1", s.A, s.B, s.d, x1))
    for nm, A, B, d, x in cases:
        n, nrhs, N = d.shape
        data = initialize_batched(N, n, torch.float64, "cuda", nrhs=nrhs, schedule=schedule)
        data.A_tensor[:, :, :N].copy_(torch.from_numpy(np.asfortranarray(A)))
        data.B_tensor[:, :, :N - 1].copy_(torch.from_numpy(np.asfortranarray(B)))
        data.d_tensor[:, :, :N].copy_(torch.from_numpy(np.asfortranarray(d)))
        factorize(data)
        solve(data)
        torch.cuda.synchronize()
        got = data.d_tensor[:, :, :N].cpu().numpy()
        rel = np.abs(got - x).max() / np.abs(x).max()
        assert rel <= 1e-12, f"{nm}: {rel:.3e}"
