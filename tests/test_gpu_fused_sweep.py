"""bss_factorize_solve with FP32 factors and one right-hand side runs level 0's forward
sweep inside the level-0 factor kernel (the factors are still in LDS), with the sweep
kernel's arithmetic.  The solution, info and refinement history must be bitwise those
of factorize then solve, across block sizes, both n = 64 workgroup sizes, chain
lengths (single blocks, chains with and without separators, a partial last chain),
batches and groups."""
import numpy as np
import pytest
import torch

from blockstructuredsolvers_amd import (factorize, factorize_solve, initialize_batched,
                                        initialize_batched_group, solve, workloads)
from oracle import oracle

pytestmark = pytest.mark.gpu


def as_f32(s):
    return workloads.BlockSystem(s.A.astype(np.float32, order="F"),
                                 s.B.astype(np.float32, order="F"),
                                 s.d.astype(np.float32, order="F"))


def run(systems, T, mode, combined, **kw):
    N, n, S = systems[0].N, systems[0].n, len(systems)
    data = initialize_batched(N, n, T, "cuda", nrhs=1, nsys=S, mode=mode, tuned=False, **kw)
    for s, sysm in enumerate(systems):
        sl = (Ellipsis, s) if S > 1 else (Ellipsis,)
        data.A_tensor[sl].copy_(torch.from_numpy(sysm.A))
        if N > 1:
            data.B_tensor[sl].copy_(torch.from_numpy(sysm.B))
        data.d_tensor[sl].copy_(torch.from_numpy(sysm.d))
    if combined:
        rc = factorize_solve(data)
    else:
        factorize(data)
        rc = solve(data)
    torch.cuda.synchronize()
    out = (data.d_tensor.cpu().numpy().copy(), np.atleast_1d(data.info()).copy(), rc,
           data.iterations if mode == "mixed" else 0)
    data.close()
    return out


CASES = [  # (n, N, nsys, chain_len, waves)
    (16, 300, 1, 3, None), (32, 257, 2, 2, None), (48, 120, 1, 15, None),
    (64, 1, 1, 3, 8), (64, 5, 1, 1, 8), (64, 600, 1, 3, 8), (64, 600, 3, 7, 8),
    (64, 900, 1, 15, 4), (64, 1000, 2, 63, 8), (96, 200, 1, 3, None),
]


@pytest.mark.parametrize("mode,T", [("plain", torch.float32), ("mixed", torch.float64)])
@pytest.mark.parametrize("n,N,nsys,chain_len,waves", CASES)
def test_fused_sweep_bitwise(mode, T, n, N, nsys, chain_len, waves):
    systems = [workloads.spd_system(N, n, nrhs=1, seed=4000 + 13 * N + n + k) for k in range(nsys)]
    if T == torch.float32:
        systems = [as_f32(s) for s in systems]
    kw = dict(chain_len=chain_len)
    if waves is not None:
        kw["factor_waves"] = waves
    xs, is_, rs, its = run(systems, T, mode, False, **kw)
    xc, ic, rc, itc = run(systems, T, mode, True, **kw)
    assert np.array_equal(xs, xc)
    assert np.array_equal(is_, ic) and not ic.any()
    assert rs == rc == 0 and its == itc
    if mode == "mixed":
        for k, sysm in enumerate(systems):
            x = xc[..., k] if nsys > 1 else xc
            assert oracle.backward_error(sysm.A, sysm.B, x, sysm.d).max() <= 1e-12


def test_fused_sweep_group_and_info():
    """An uneven group, then a non-SPD block: the same bits and the same LAPACK-order
    info through the fused path."""
    systems = [workloads.spd_system(N, 64, nrhs=1, seed=4100 + N) for N in (3, 200, 777)]
    outs = []
    for combined in (False, True):
        data = initialize_batched_group([s.N for s in systems], 64, torch.float64, "cuda",
                                        nrhs=1, mode="mixed", tuned=False)
        for k, sysm in enumerate(systems):
            N = sysm.N
            data.A_tensors[k][:, :, :N].copy_(torch.from_numpy(sysm.A))
            if N > 1:
                data.B_tensors[k][:, :, :N - 1].copy_(torch.from_numpy(sysm.B))
            data.d_tensors[k][:, :, :N].copy_(torch.from_numpy(sysm.d))
        if combined:
            factorize_solve(data)
        else:
            factorize(data)
            solve(data)
        torch.cuda.synchronize()
        outs.append([data.d_tensors[k][:, :, :s.N].cpu().numpy().copy()
                     for k, s in enumerate(systems)] + [data.info().tolist()])
        data.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    assert outs[1][-1] == [0, 0, 0]
    sysm = workloads.spd_system(600, 64, nrhs=1, seed=4200)
    A = sysm.A.copy(order="F")
    A[:, :, 321] = -np.eye(64)
    bad = workloads.BlockSystem(A, sysm.B, sysm.d)
    ref = oracle.factorize(A.copy(order="F"), sysm.B.copy(order="F"))
    _, i_sep, _, _ = run([bad], torch.float64, "mixed", False, chain_len=3)
    _, i_cmb, _, _ = run([bad], torch.float64, "mixed", True, chain_len=3)
    assert i_sep.tolist() == i_cmb.tolist() == [ref]
