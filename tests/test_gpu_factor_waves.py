"""At n = 64 the factorisation runs with four or eight waves per chain.  The eight-wave
kernel moves the work no later block waits for (the panel products, the Schur update's
off-diagonal tiles, the factor stores and the next blocks' loads) into the idle windows
of the next block's potrf; the four-wave kernel runs a block's steps in order.  Both
form the same products in the same order, so every output must agree bit for bit:
solutions, info (compared with the oracle too) and the mixed path's pivot decisions."""
import numpy as np
import pytest
import torch

from blockstructuredsolvers_amd import factorize, initialize_batched, solve, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


def as_f32(s):
    return workloads.BlockSystem(s.A.astype(np.float32, order="F"),
                                 s.B.astype(np.float32, order="F"),
                                 s.d.astype(np.float32, order="F"))


def run(systems, T, mode, waves, **kw):
    N, n, nrhs, S = systems[0].N, systems[0].n, systems[0].nrhs, len(systems)
    data = initialize_batched(N, n, T, "cuda", nrhs=nrhs, nsys=S, mode=mode, tuned=False,
                              factor_waves=waves, **kw)
    for s, sysm in enumerate(systems):
        sl = (Ellipsis, s) if S > 1 else (Ellipsis,)
        data.A_tensor[sl].copy_(torch.from_numpy(sysm.A))
        if N > 1:
            data.B_tensor[sl].copy_(torch.from_numpy(sysm.B))
        data.d_tensor[sl].copy_(torch.from_numpy(sysm.d))
    factorize(data)
    rc = solve(data)
    torch.cuda.synchronize()
    x = data.d_tensor.cpu().numpy().copy()
    info = np.atleast_1d(data.info()).copy()
    data.close()
    return x, info, rc


CASES = [  # (N, nsys, chain_len, nrhs): single blocks, one chain, short and long chains,
    # a partial last chain, and batches whose top level has more chains than CUs
    (1, 1, 3, 1), (2, 1, 3, 2), (3, 1, 1, 1), (7, 1, 2, 1), (64, 1, 63, 1), (300, 1, 15, 3),
    (1000, 2, 3, 1), (2000, 1, 3, 1), (700, 3, 7, 2),
]


@pytest.mark.parametrize("mode,T", [("plain", torch.float64), ("plain", torch.float32),
                                    ("mixed", torch.float64)])
@pytest.mark.parametrize("N,nsys,chain_len,nrhs", CASES)
def test_eight_waves_equal_four(mode, T, N, nsys, chain_len, nrhs):
    systems = [workloads.spd_system(N, 64, nrhs=nrhs, seed=900 + 7 * N + k) for k in range(nsys)]
    if T == torch.float32:
        systems = [as_f32(s) for s in systems]
    x4, i4, rc4 = run(systems, T, mode, 4, chain_len=chain_len)
    x8, i8, rc8 = run(systems, T, mode, 8, chain_len=chain_len)
    assert rc4 == rc8 == 0
    assert np.array_equal(i4, i8) and not i8.any()
    assert np.array_equal(x4, x8)
    if T == torch.float64:
        for s, sysm in enumerate(systems):
            xs = x8[..., s] if nsys > 1 else x8
            assert oracle.backward_error(sysm.A, sysm.B, xs, sysm.d).max() <= (
                1e-13 if mode == "plain" else 1e-12)


@pytest.mark.parametrize("mode", ["plain", "mixed"])
@pytest.mark.parametrize("where", ["first", "inside", "separator", "last", "two"])
def test_eight_waves_info(mode, where):
    """A non-SPD block at the start, inside a chain, on a separator, at the end, or two
    of them: the same LAPACK-order info from both kernels, equal to the oracle's."""
    N, chain_len = 600, 7
    sysm = workloads.spd_system(N, 64, nrhs=1, seed=77)
    A = sysm.A.copy(order="F")
    bad = {"first": [0], "inside": [123], "separator": [8 * (chain_len + 1) - 1],
           "last": [N - 1], "two": [451, 130]}[where]
    for b in bad:
        A[:, :, b] = -np.eye(64)
    sysm = workloads.BlockSystem(A, sysm.B, sysm.d)
    ref = oracle.factorize(A.copy(order="F"), sysm.B.copy(order="F"))
    assert ref == min(bad) + 1
    _, i4, _ = run([sysm], torch.float64, mode, 4, chain_len=chain_len)
    _, i8, _ = run([sysm], torch.float64, mode, 8, chain_len=chain_len)
    assert i4.tolist() == i8.tolist() == [ref]


def test_eight_waves_pivot_refactor():
    """Mixed mode with the FP32 separator pivots forced below threshold: both kernels
    take the FP64 separator path and give the same bits."""
    systems = [workloads.spd_system(900, 64, nrhs=1, seed=31 + k) for k in range(2)]
    x4, _, _ = run(systems, torch.float64, "mixed", 4, chain_len=3, pivot_tol=2.0)
    x8, _, _ = run(systems, torch.float64, "mixed", 8, chain_len=3, pivot_tol=2.0)
    assert np.array_equal(x4, x8)
