"""bss_info (integer output, so exact equality) against the FP64 oracle's LAPACK-order
info, and the mixed path's refusal to report silent success.

info is the 1-based block at which the sequential block Cholesky of
src/backbone_cpu.jl:5-15 meets its first non-positive pivot (oracle.factorize).  The
failures are planted at block 0, at a chain's first block behind a separator, inside
that chain, on the separator and on the last block (positions read from the plan's own
level-0 chain table), plus a coupled pair across a separator whose recursive-order
failure (the separator) comes before the LAPACK-order one (the block after it).
"""
import numpy as np
import pytest
import torch

from blockstructuredsolvers_amd import (factorize, factorize_solve, initialize_batched,
                                        initialize_batched_group, solve, workloads)
from blockstructuredsolvers_amd import _lib as L
from chain_emulator import coupled_failure, plant_positions
from oracle import oracle

pytestmark = pytest.mark.gpu


def planted(base, k):
    A = base.A.copy(order="F")
    A[:, :, k] = -np.eye(base.A.shape[0])
    return workloads.BlockSystem(A, base.B, base.d)


def cases(base, schedule):
    N, n = base.A.shape[2], base.A.shape[0]
    pos = plant_positions(N, n, {"auto": 0, "reference": 1}[schedule])
    out = [(name, planted(base, k)) for name, k in pos.items()]
    out.append(("coupled_pair", coupled_failure(base, pos["chain_first"])))
    out.append(("spd", base))
    return out


def oracle_info(sysm):
    return oracle.factorize(np.array(sysm.A, dtype=np.float64, order="F"),
                            np.array(sysm.B, dtype=np.float64, order="F"))


def upload(data, sysm, T):
    N = sysm.A.shape[2]
    data.A_tensor[:, :, :N].copy_(torch.from_numpy(np.asfortranarray(sysm.A, dtype=T)))
    data.B_tensor[:, :, :N - 1].copy_(torch.from_numpy(np.asfortranarray(sysm.B, dtype=T)))
    data.d_tensor[:, :, :N].copy_(torch.from_numpy(np.asfortranarray(sysm.d, dtype=T)))


@pytest.mark.parametrize("schedule", ["auto", "reference"])
@pytest.mark.parametrize("mode,T", [("plain", np.float64), ("plain", np.float32),
                                    ("mixed", np.float64)])
def test_info_equals_oracle(mode, T, schedule):
    N, n = 300, 32
    base = workloads.spd_system(N, n, seed=61)
    tt = torch.float64 if T == np.float64 else torch.float32
    data = initialize_batched(N, n, tt, "cuda", mode=mode, schedule=schedule, tuned=False,
                              chain_len=3, base_max=4)
    assert data.levels >= (3 if schedule == "auto" else 2)
    for name, sysm in cases(base, schedule):
        ref = oracle_info(sysm)
        upload(data, sysm, T)
        factorize(data)
        rc = solve(data)
        torch.cuda.synchronize()
        got = int(data.info()[0])
        print(f"{mode} {T.__name__} {schedule} {name}: oracle {ref} device {got} status {rc}")
        assert got == ref, name
        if mode == "mixed":
            # a non-SPD system never reports a converged solve
            assert (rc == L.BSS_OK) == (ref == 0), name


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_info_batch_and_group(mode):
    """Several systems in one plan: each gets its own LAPACK-order index (a batch of
    three with failures in two of them, and an uneven group), through factorize_solve
    as well."""
    N, n = 200, 16
    bases = [workloads.spd_system(N, n, seed=70 + k) for k in range(3)]
    pos = plant_positions(N, n, 0)
    systems = [bases[0], coupled_failure(bases[1], pos["chain_first"]),
               planted(bases[2], pos["separator"])]
    refs = [oracle_info(s) for s in systems]
    assert refs[0] == 0 and refs[1] == pos["chain_first"] + 1
    data = initialize_batched(N, n, torch.float64, "cuda", nsys=3, mode=mode, tuned=False,
                              chain_len=3, base_max=4)
    for s, sysm in enumerate(systems):
        data.A_tensor[..., s].copy_(torch.from_numpy(sysm.A))
        data.B_tensor[..., s].copy_(torch.from_numpy(sysm.B))
        data.d_tensor[..., s].copy_(torch.from_numpy(sysm.d))
    factorize_solve(data)
    torch.cuda.synchronize()
    assert data.info().tolist() == refs
    # uneven group: N = 3, 500, 1200, a failure on a separator of the middle system
    Ns = (3, 500, 1200)
    gs = [workloads.spd_system(Nk, 64, seed=400 + Nk) for Nk in Ns]
    k = plant_positions(500, 64, 0)["separator"]
    gs[1] = planted(gs[1], k)
    g = initialize_batched_group(list(Ns), 64, torch.float64, "cuda", mode=mode, tuned=False)
    for s, sysm in enumerate(gs):
        g.A_tensors[s].copy_(torch.from_numpy(sysm.A))
        g.B_tensors[s].copy_(torch.from_numpy(sysm.B))
        g.d_tensors[s].copy_(torch.from_numpy(sysm.d))
    factorize(g)
    solve(g)
    torch.cuda.synchronize()
    assert g.info().tolist() == [0, k + 1, 0] == [oracle_info(s) for s in gs]


@pytest.mark.parametrize("n", [96, 128])
def test_info_big_blocks(n):
    """The staged factor kernel (n = 96, 128) through the same re-check."""
    N = 300
    base = workloads.spd_system(N, n, seed=80 + n)
    pos = plant_positions(N, n, 0)
    for name in ("inside_chain", "separator"):
        sysm = planted(base, pos[name])
        data = initialize_batched(N, n, torch.float64, "cuda", tuned=False)
        upload(data, sysm, np.float64)
        factorize(data)
        assert int(data.info()[0]) == oracle_info(sysm) == pos[name] + 1, name
        data.close()


def test_info_in_place_mode_flags_failure():
    """With the input overwritten (preserve_input=False, the reference's in-place
    semantics) there is nothing to re-check from: a failed system is still flagged,
    with a block of the recursive elimination order."""
    N, n = 300, 32
    base = workloads.spd_system(N, n, seed=61)
    k = plant_positions(N, n, 0)["inside_chain"]
    sysm = planted(base, k)
    data = initialize_batched(N, n, torch.float64, "cuda", tuned=False, preserve_input=False)
    upload(data, sysm, np.float64)
    factorize(data)
    assert int(data.info()[0]) > 0


@pytest.mark.parametrize("combined", [False, True])
def test_mixed_fallback_ill_conditioned(combined):
    """An input beyond FP32 refinement (SPD by construction, alpha = 1e-8, condition
    number ~5e8): refinement stops short, the plain FP64 fallback solves it to the
    target, the status is BSS_OK and the fallback is reported; later solves of the same
    factorisation use the fallback directly.  With the fallback off the status is
    BSS_NOT_CONVERGED: never a silent success."""
    N, n = 2000, 64
    sysm = workloads.spd_bidiag_system(N, n, seed=12, alpha=1e-8)
    data = initialize_batched(N, n, torch.float64, "cuda", mode="mixed", tuned=False)
    upload(data, sysm, np.float64)
    if combined:
        rc = factorize_solve(data)
    else:
        factorize(data)
        rc = solve(data)
    torch.cuda.synchronize()
    x = data.d_tensor.cpu().numpy()
    berr = oracle.backward_error(sysm.A, sysm.B, x, sysm.d).max()
    print(f"fallback: status {rc} fallback {data.fallback} iterations {data.iterations} "
          f"berr {berr:.2e} device {data.backward_error().max():.2e}")
    assert rc == L.BSS_OK and data.fallback and data.converged
    assert berr <= 1e-12 and data.backward_error().max() <= 1e-12
    assert data.info()[0] == 0
    d2 = np.asfortranarray(np.random.default_rng(3).uniform(size=sysm.d.shape))
    data.d_tensor.copy_(torch.from_numpy(d2))
    assert solve(data) == L.BSS_OK and data.fallback
    torch.cuda.synchronize()
    assert oracle.backward_error(sysm.A, sysm.B, data.d_tensor.cpu().numpy(), d2).max() <= 1e-12
    # a well-conditioned input on the same plan leaves the fallback again
    good = workloads.spd_system(N, n, seed=13)
    upload(data, good, np.float64)
    factorize(data)
    assert solve(data) == L.BSS_OK and not data.fallback
    # fallback off: the unconverged status, and the device backward error says why
    off = initialize_batched(N, n, torch.float64, "cuda", mode="mixed", tuned=False,
                             fallback=False)
    upload(off, sysm, np.float64)
    factorize(off)
    rc = solve(off)
    torch.cuda.synchronize()
    assert rc == L.BSS_NOT_CONVERGED and not off.converged and not off.fallback
    b = off.backward_error().max()
    assert not (b <= 1e-12)
