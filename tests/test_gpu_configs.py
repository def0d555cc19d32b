"""The benchmark configurations at full size against the FP64 oracle.

C2: N = 8192 blocks, n in {32, 64, 128}, 1 and 16 right-hand sides, plain FP64 (rel.
    difference <= 1e-12, backward error <= 1e-13) and plain FP32 (the reference test's
    bound ||x - x_ref|| <= 10 sqrt(eps32) ||x_ref||, test/test_rocm.jl:118), through
    the separate calls and through factorize_solve (bitwise equal).
C3: N = 8192, n = 64, mixed: the literal generator (A_i = G G^T + alpha n I, B_i =
    0.5 N(0,1)) at alpha in {1, 1e-2, 1e-4} -- the oracle's failing block where it is
    not SPD, else the 1e-12 backward error -- and the SPD-by-construction variant at the
    same alphas (backward error <= 1e-12, iterations reported).
C4: 512 systems, N_k log-uniform in [16, 4096], n_k uniform over the six block sizes,
    generated on the device, one group plan per n on its own stream, plain FP64 and
    mixed; two systems per n checked against the oracle.

The oracle runs in a thread beside the device work (its ctypes calls release the GIL).
"""
import concurrent.futures as cf

import numpy as np
import pytest
import torch

from blockstructuredsolvers_amd import (factorize, factorize_solve, initialize_batched, solve,
                                        workloads)
from blockstructuredsolvers_amd import _lib as L
from oracle import oracle

pytestmark = pytest.mark.gpu

_POOL = cf.ThreadPoolExecutor(max_workers=8)


def upload(data, sysm, T):
    data.A_tensor.copy_(torch.from_numpy(np.asfortranarray(sysm.A, dtype=T)))
    data.B_tensor.copy_(torch.from_numpy(np.asfortranarray(sysm.B, dtype=T)))
    data.d_tensor.copy_(torch.from_numpy(np.asfortranarray(sysm.d, dtype=T)))


def gpu_solve(sysm, T, combined=False, **kw):
    N, n, nrhs = sysm.A.shape[2], sysm.A.shape[0], sysm.d.shape[1]
    tt = torch.float64 if T == np.float64 else torch.float32
    data = initialize_batched(N, n, tt, "cuda", nrhs=nrhs, **kw)
    upload(data, sysm, T)
    rc = factorize_solve(data) if combined else (factorize(data), solve(data))[1]
    torch.cuda.synchronize()
    x = data.d_tensor.double().cpu().numpy()
    info = int(data.info()[0])
    out = dict(rc=rc, info=info, iterations=data.iterations if kw.get("mode") == "mixed" else 0)
    data.close()
    return x, out


@pytest.mark.parametrize("n,nrhs", [(32, 1), (32, 16), (64, 1), (64, 16), (128, 1), (128, 16)])
def test_c2_full_size(n, nrhs):
    N = 8192
    sysm = workloads.spd_system(N, n, nrhs=nrhs, seed=42)
    ref = _POOL.submit(oracle.solve_system, sysm.A, sysm.B, sysm.d)
    x64, o64 = gpu_solve(sysm, np.float64)
    xc, _ = gpu_solve(sysm, np.float64, combined=True)
    x32, o32 = gpu_solve(sysm, np.float32)
    xo, info = ref.result()
    assert info == 0 and o64["info"] == 0 and o32["info"] == 0
    assert np.array_equal(x64, xc)
    rel = np.abs(x64 - xo).max() / np.abs(xo).max()
    berr = oracle.backward_error(sysm.A, sysm.B, x64, sysm.d).max()
    err32 = oracle.bss_norm(x32 - xo) / oracle.bss_norm(xo)
    print(f"C2 N={N} n={n} nrhs={nrhs}: FP64 rel {rel:.2e} berr {berr:.2e}; FP32 rel "
          f"{err32:.2e}")
    assert rel <= 1e-12 and berr <= 1e-13
    assert err32 <= 10 * np.sqrt(np.finfo(np.float32).eps)


@pytest.mark.parametrize("alpha", [1.0, 1e-2, 1e-4])
def test_c3_literal_generator(alpha):
    """The config's literal generator: the oracle's info decides; where it is not SPD the
    device reports the same block (and no converged solve), else it meets 1e-12."""
    N, n = 8192, 64
    sysm = workloads.spd_system(N, n, seed=42, alpha=alpha, bscale=0.5)
    ref = _POOL.submit(oracle.solve_system, sysm.A, sysm.B, sysm.d)
    x, o = gpu_solve(sysm, np.float64, mode="mixed")
    xo, info = ref.result()
    print(f"C3 literal alpha={alpha:g}: oracle info {info}, device info {o['info']}, "
          f"status {o['rc']}, iterations {o['iterations']}")
    assert o["info"] == info
    if info == 0:
        assert o["rc"] == L.BSS_OK
        assert oracle.backward_error(sysm.A, sysm.B, x, sysm.d).max() <= 1e-12
    else:
        assert o["rc"] == L.BSS_NOT_CONVERGED


@pytest.mark.parametrize("alpha", [1.0, 1e-2, 1e-4])
def test_c3_spd_variant(alpha):
    N, n = 8192, 64
    sysm = workloads.spd_bidiag_system(N, n, seed=42, alpha=alpha)
    ref = _POOL.submit(oracle.solve_system, sysm.A, sysm.B, sysm.d)
    x, o = gpu_solve(sysm, np.float64, mode="mixed")
    xo, info = ref.result()
    berr = oracle.backward_error(sysm.A, sysm.B, x, sysm.d).max()
    rel = np.abs(x - xo).max() / np.abs(xo).max()
    print(f"C3 SPD variant alpha={alpha:g}: iterations {o['iterations']}, berr {berr:.2e}, "
          f"rel vs oracle {rel:.2e}, status {o['rc']}")
    assert info == 0 and o["info"] == 0 and o["rc"] == L.BSS_OK
    assert berr <= 1e-12
    assert 1 <= o["iterations"] <= 10


def test_c4_uneven_batch():
    """512 systems over all six block sizes in six group plans (one stream each), plain
    FP64 and mixed; per n the smallest system and one of median size (capped at 2048
    blocks to bound the oracle's time) are checked against the oracle."""
    from blockstructuredsolvers_amd import batch_bench
    Ns, ns = batch_bench.draw(512, 7)
    groups = [(n, [N for N, m in zip(Ns, ns) if m == n]) for n in batch_bench.SIZES]
    groups = [(n, g) for n, g in groups if g]
    assert len(groups) == 6 and sum(len(g) for _, g in groups) == 512
    pick = []
    for gi, (n, g) in enumerate(groups):
        order = np.argsort(g)
        small = int(order[0])
        mid = [int(s) for s in order if g[s] <= 2048][len([v for v in g if v <= 2048]) // 2]
        pick += [(gi, small), (gi, mid)] if mid != small else [(gi, small)]
    for mode in ("plain", "mixed"):
        plans, ms, samples = batch_bench.run(mode, groups, steps=1, warmup=0, seed=7, check=0,
                                             picks=pick)
        for data in plans:
            assert np.all(data.info() == 0)
        # inputs and device answers of the picked systems (d0: the right-hand side)
        jobs = []
        for gi, s, d0 in samples:
            data = plans[gi]
            A = np.asfortranarray(data.A_tensors[s].double().cpu().numpy())
            B = np.asfortranarray(data.B_tensors[s].double().cpu().numpy())
            x = data.d_tensors[s].double().cpu().numpy()
            d0 = np.asfortranarray(d0)
            jobs.append((groups[gi][0], groups[gi][1][s], A, B, x, d0,
                         _POOL.submit(oracle.solve_system, A, B, d0)))
        assert len(jobs) == len(pick)
        for n, N, A, B, x, d0, fut in jobs:
            xo, info = fut.result()
            assert info == 0
            berr = oracle.backward_error(A, B, x, d0).max()
            rel = np.abs(x - xo).max() / np.abs(xo).max()
            print(f"C4 {mode} n={n} N={N}: berr {berr:.2e} rel {rel:.2e}")
            if mode == "plain":
                assert rel <= 1e-12 and berr <= 1e-13
            else:
                assert berr <= 1e-12
        print(f"C4 {mode}: {ms:.1f} ms per step for 512 systems")
        for data in plans:
            data.close()
        del plans
        torch.cuda.empty_cache()
