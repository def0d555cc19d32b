"""Uneven batches across block sizes (bss_batch_*): the whole batch in one call, one
group plan per block size, by default one hipGraph launch per call.  Each system's
answer is bitwise the one it gets alone in plain mode (graph replay, new data each
time), mixed mode meets the 1e-12 backward error per system, and info maps back to the
input order."""
import numpy as np
import pytest
import torch

from blockstructuredsolvers_amd import (factorize, initialize_batch, initialize_batched, solve,
                                        workloads)
from blockstructuredsolvers_amd import _lib as L
from oracle import oracle

pytestmark = pytest.mark.gpu

NS = [16, 64, 32, 128, 16, 96, 48, 64]
NN = [40, 300, 7, 150, 1, 90, 500, 2]


def systems(seed, nrhs):
    return [workloads.spd_system(N, n, nrhs=nrhs, seed=seed + k)
            for k, (N, n) in enumerate(zip(NN, NS))]


def upload(data, syss):
    for s, sysm in enumerate(syss):
        N = sysm.A.shape[2]
        data.A_tensors[s][:, :, :N].copy_(torch.from_numpy(sysm.A))
        if N > 1:
            data.B_tensors[s][:, :, :N - 1].copy_(torch.from_numpy(sysm.B))
        data.d_tensors[s][:, :, :N].copy_(torch.from_numpy(sysm.d))


def alone(sysm, **kw):
    N, n, nrhs = sysm.A.shape[2], sysm.A.shape[0], sysm.d.shape[1]
    data = initialize_batched(N, n, torch.float64, "cuda", nrhs=nrhs, tuned=False, **kw)
    data.A_tensor[:, :, :N].copy_(torch.from_numpy(sysm.A))
    data.B_tensor[:, :, :N - 1].copy_(torch.from_numpy(sysm.B))
    data.d_tensor[:, :, :N].copy_(torch.from_numpy(sysm.d))
    factorize(data)
    solve(data)
    torch.cuda.synchronize()
    x = data.d_tensor[:, :, :N].cpu().numpy()
    data.close()
    return x


@pytest.mark.parametrize("graph,sched", [(True, {}), (False, {}),
                                         (True, {"chain_len_levels": [15] + [3] * 15,
                                                 "base_max": 4})])
def test_batch_plain_bitwise_alone(graph, sched):
    """With the default schedule, and with one fixed for every system (the uneven-batch
    benchmark's), each answer is bitwise that of the system alone on the same schedule."""
    data = initialize_batch(NN, NS, torch.float64, "cuda", nrhs=2, graph=graph, **sched)
    for rep in range(2):  # the second round replays the captured graphs on new data
        syss = systems(100 + 10 * rep, 2)
        upload(data, syss)
        factorize(data)
        assert solve(data) == L.BSS_OK
        torch.cuda.synchronize()
        assert np.all(data.info() == 0)
        for s, sysm in enumerate(syss):
            N = sysm.A.shape[2]
            x = data.d_tensors[s][:, :, :N].cpu().numpy()
            assert np.array_equal(x, alone(sysm, **sched)), (rep, s, NS[s], N)
            assert oracle.backward_error(sysm.A, sysm.B, x, sysm.d).max() <= 1e-13


def test_batch_mixed_and_info():
    data = initialize_batch(NN, NS, torch.float64, "cuda", nrhs=1, mode="mixed")
    syss = systems(300, 1)
    upload(data, syss)
    factorize(data)
    assert solve(data) == L.BSS_OK
    torch.cuda.synchronize()
    dev = data.backward_error()
    for s, sysm in enumerate(syss):
        N = sysm.A.shape[2]
        x = data.d_tensors[s][:, :, :N].cpu().numpy()
        b = oracle.backward_error(sysm.A, sysm.B, x, sysm.d).max()
        assert b <= 1e-12 and dev[s].max() <= 1e-12, (s, b)
    assert 2 <= data.iterations <= 10
    # a non-SPD block in system 6 (n = 48): its info, in input order, and the status
    bad = syss[6]
    A = bad.A.copy(order="F")
    A[:, :, 123] = -np.eye(48)
    syss[6] = workloads.BlockSystem(A, bad.B, bad.d)
    upload(data, syss)
    factorize(data)
    rc = solve(data)
    torch.cuda.synchronize()
    info = data.info()
    ref = oracle.factorize(A.copy(order="F"), bad.B.copy(order="F"))
    assert ref == 124
    assert info.tolist() == [0] * 6 + [ref] + [0]
    assert rc == L.BSS_NOT_CONVERGED
