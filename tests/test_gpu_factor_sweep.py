"""BSS_OPT_FACTOR_SWEEP (mixed mode, one right-hand side): bss_factorize also runs the
first solve's x = 0 pass and level-0 forward sweep on the d it finds and keeps a copy of
d; bss_solve compares d with that copy on the device and skips those two steps only if
every bit is unchanged.  The answer, info, iteration count and backward errors must be
bitwise those with the option off, whether d is unchanged, changed in one bit of one
element, replaced, or solved a second time, for batches, a group, graph capture and
replay, an uneven batch across block sizes and the FP64-separator path."""
import ctypes

import numpy as np
import pytest
import torch

from blockstructuredsolvers_amd import (_lib as L, factorize, initialize_batch,
                                        initialize_batched, initialize_batched_group, solve,
                                        workloads)
from oracle import oracle

pytestmark = pytest.mark.gpu

K_FORWARD = 2  # BSS_K_FORWARD of bss.h: forward substitution sweeps


def make(systems, group=False, **kw):
    n = systems[0].n
    if group:
        data = initialize_batched_group([s.N for s in systems], n, torch.float64, "cuda",
                                        nrhs=1, mode="mixed", tuned=False, **kw)
        for k, s in enumerate(systems):
            data.A_tensors[k][:, :, :s.N].copy_(torch.from_numpy(s.A))
            if s.N > 1:
                data.B_tensors[k][:, :, :s.N - 1].copy_(torch.from_numpy(s.B))
            data.d_tensors[k][:, :, :s.N].copy_(torch.from_numpy(s.d))
        return data
    N, S = systems[0].N, len(systems)
    data = initialize_batched(N, n, torch.float64, "cuda", nrhs=1, nsys=S, mode="mixed",
                              tuned=False, **kw)
    for k, s in enumerate(systems):
        sl = (Ellipsis, k) if S > 1 else (Ellipsis,)
        data.A_tensor[sl].copy_(torch.from_numpy(s.A))
        if N > 1:
            data.B_tensor[sl].copy_(torch.from_numpy(s.B))
        data.d_tensor[sl].copy_(torch.from_numpy(s.d))
    return data


def d_views(data):
    return list(data.d_tensors) if hasattr(data, "d_tensors") else [data.d_tensor]


def run(systems, on, mutate=None, solves=1, group=False, **kw):
    data = make(systems, group=group, **kw)
    L.call("bss_set_option", data._plan, L.BSS_OPT_FACTOR_SWEEP, 1.0 if on else 0.0)
    factorize(data)
    outs = []
    for k in range(solves):
        if mutate is not None:
            mutate(d_views(data)[-1], k)
        rc = solve(data)
        torch.cuda.synchronize()
        x = np.concatenate([v.cpu().numpy().ravel(order="F") for v in d_views(data)])
        outs.append((x, rc, data.iterations, data.backward_error().copy()))
    info = np.atleast_1d(data.info()).copy()
    data.close()
    return outs, info


def assert_same(a, b):
    (oa, ia), (ob, ib) = a, b
    assert np.array_equal(ia, ib)
    assert len(oa) == len(ob)
    for (xa, ra, ita, ba), (xb, rb, itb, bb) in zip(oa, ob):
        assert np.array_equal(xa, xb)
        assert ra == rb and ita == itb
        assert np.array_equal(ba, bb)


def systems_for(n, N, nsys, seed):
    return [workloads.spd_system(N, n, nrhs=1, seed=seed + k) for k in range(nsys)]


@pytest.mark.parametrize("n,N,nsys", [(16, 300, 1), (32, 257, 2), (64, 600, 1), (64, 1000, 3),
                                      (96, 200, 1)])
def test_unchanged_d_bitwise(n, N, nsys):
    systems = systems_for(n, N, nsys, 5000 + n + N)
    on = run(systems, True, chain_len=7)
    assert_same(on, run(systems, False, chain_len=7))
    x = on[0][0][0].reshape((n, 1, N, nsys), order="F")
    for k, s in enumerate(systems):
        assert oracle.backward_error(s.A, s.B, np.asfortranarray(x[..., k]), s.d).max() <= 1e-12


def flip_one_bit(d, k):
    # one element (the middle index of every dimension) moves to its next double up
    idx = tuple(sz // 2 for sz in d.shape)
    d[idx] = float(np.nextafter(d[idx].item(), np.inf))


def replace(d, k):
    g = torch.Generator(device="cpu").manual_seed(77 + k)
    d.copy_(torch.rand(d.shape, generator=g, dtype=torch.float64))


@pytest.mark.parametrize("mutate", [flip_one_bit, replace])
def test_changed_d_bitwise(mutate):
    systems = systems_for(64, 700, 2, 5100)
    assert_same(run(systems, True, mutate=mutate, chain_len=5),
                run(systems, False, mutate=mutate, chain_len=5))


def test_second_solve_and_new_rhs():
    systems = systems_for(48, 400, 1, 5200)

    def new_rhs_second(d, k):
        if k == 1:
            replace(d, k)

    assert_same(run(systems, True, mutate=new_rhs_second, solves=3, chain_len=3),
                run(systems, False, mutate=new_rhs_second, solves=3, chain_len=3))


def test_fp64_separator_path():
    """pivot_tol above 1 makes every factorisation re-factor its separators in FP64:
    the first solve is redone from d on them, with the x = 0 pass and level-0 sweep run
    again regardless of the option."""
    systems = systems_for(64, 500, 1, 5300)
    on = run(systems, True, chain_len=3, pivot_tol=2.0)
    assert_same(on, run(systems, False, chain_len=3, pivot_tol=2.0))
    x = np.asfortranarray(on[0][0][0].reshape((64, 1, 500), order="F"))
    assert oracle.backward_error(systems[0].A, systems[0].B, x, systems[0].d).max() <= 1e-12


def test_group_bitwise():
    systems = [workloads.spd_system(N, 64, nrhs=1, seed=5400 + N) for N in (3, 200, 777)]
    assert_same(run(systems, True, group=True), run(systems, False, group=True))


def test_graph_replays_bitwise():
    """With hipGraph replay the factorisation's captured sequence holds the d snapshot,
    the x = 0 pass and the fused level-0 kernel; three factorize/solve steps (the first
    captures, the others replay) with d restored, then one with a new d, must give the
    option-off bits every time."""
    systems = systems_for(64, 900, 2, 5600)

    def steps(on):
        data = make(systems, graph=True, chain_len=15)
        L.call("bss_set_option", data._plan, L.BSS_OPT_FACTOR_SWEEP, 1.0 if on else 0.0)
        d0 = data.d_tensor.clone()
        outs = []
        for k in range(4):
            data.d_tensor.copy_(d0 if k < 3 else d0 * 0.5)
            factorize(data)
            outs.append((solve(data), data.iterations))
            torch.cuda.synchronize()
            outs.append(data.d_tensor.cpu().numpy().copy())
        data.close()
        return outs

    on, off = steps(True), steps(False)
    for a, b in zip(on, off):
        if isinstance(a, np.ndarray):
            assert np.array_equal(a, b)
        else:
            assert a == b


def test_uneven_batch_bitwise():
    """An uneven batch (one group plan per block size, factorisation as one graph):
    option on against off, bitwise, for two steps with d restored."""
    Ns, ns = [300, 257, 40, 500, 3], [64, 32, 64, 16, 96]
    systems = [workloads.spd_system(N, n, nrhs=1, seed=5700 + k)
               for k, (N, n) in enumerate(zip(Ns, ns))]

    def steps(on):
        data = initialize_batch(Ns, ns, torch.float64, "cuda", mode="mixed")
        L.call("bss_batch_set_option", data._batch, L.BSS_OPT_FACTOR_SWEEP, 1.0 if on else 0.0)
        for k, s in enumerate(systems):
            data.A_tensors[k][:, :, :s.N].copy_(torch.from_numpy(s.A))
            if s.N > 1:
                data.B_tensors[k][:, :, :s.N - 1].copy_(torch.from_numpy(s.B))
            data.d_tensors[k][:, :, :s.N].copy_(torch.from_numpy(s.d))
        d0 = [d.clone() for d in data.d_tensors]
        outs = []
        for _ in range(2):
            for d, v in zip(data.d_tensors, d0):
                d.copy_(v)
            factorize(data)
            rc = solve(data)
            torch.cuda.synchronize()
            outs.append((rc, [d.cpu().numpy().copy() for d in data.d_tensors],
                         data.info().tolist()))
        data.close()
        return outs

    on, off = steps(True), steps(False)
    for (ra, xa, ia), (rb, xb, ib) in zip(on, off):
        assert ra == rb and ia == ib
        assert all(np.array_equal(a, b) for a, b in zip(xa, xb))
    for k, s in enumerate(systems):
        x = np.asfortranarray(on[-1][1][k][:, :, :s.N])
        assert oracle.backward_error(s.A, s.B, x, s.d).max() <= 1e-12


def test_sweep_skipped_when_d_unchanged():
    """The level-0 forward launches of the first solve: with the option on and d
    unchanged, the first correction's sweep leaves at once, so the two corrections'
    level-0 forward time is about one sweep instead of two."""
    systems = systems_for(64, 8192, 4, 5500)
    times = {}
    for on in (False, True):
        data = make(systems)
        L.call("bss_set_option", data._plan, L.BSS_OPT_FACTOR_SWEEP, 1.0 if on else 0.0)
        factorize(data)
        solve(data)  # warm-up
        factorize(data)
        d0 = data.d_tensor.clone()
        torch.cuda.synchronize()
        factorize(data)
        data.d_tensor.copy_(d0)  # the same bits: still unchanged
        L.call("bss_profile_enable", data._plan, 1)
        solve(data)
        torch.cuda.synchronize()
        c, t = ctypes.c_int(), ctypes.c_double()
        L.call("bss_profile_read", data._plan, K_FORWARD, 0, ctypes.byref(c), ctypes.byref(t))
        L.call("bss_profile_enable", data._plan, 0)
        assert data.iterations == 2 and c.value == 2
        times[on] = t.value
        data.close()
    print(f"level-0 forward, two corrections: off {times[False]:.3f} ms, on {times[True]:.3f} ms")
    assert times[True] < 0.7 * times[False]
