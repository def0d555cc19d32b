"""GPU parity of the libbss solver against the FP64 CPU oracle (restatement of the
reference's cpu.jl path) on identical seeded inputs.

Bars: plain FP64 -- relative solution difference <= 1e-12 and
normwise backward error <= 1e-13; plain FP32 -- the reference tests' bound
||x - x_ref|| <= 10 sqrt(eps32) ||x_ref|| (test/test_rocm.jl:118); mixed -- backward
error <= 1e-12.
"""
import numpy as np
import pytest
import torch

from blockstructuredsolvers_amd import (factorize, factorize_solve, initialize_batched, solve,
                                        workloads)
from oracle import oracle

pytestmark = pytest.mark.gpu

F64_REL = 1e-12
F64_BERR = 1e-13


def run_gpu(sysm, T=torch.float64, **kw):
    N, n, nrhs = sysm.N, sysm.n, sysm.nrhs
    data = initialize_batched(N, n, T, "cuda", nrhs=nrhs, **kw)
    data.A_tensor[:, :, :N].copy_(torch.from_numpy(sysm.A))
    data.B_tensor[:, :, :N - 1].copy_(torch.from_numpy(sysm.B))
    data.d_tensor[:, :, :N].copy_(torch.from_numpy(sysm.d))
    factorize(data)
    solve(data)
    torch.cuda.synchronize()
    x = data.d_tensor[:, :, :N].double().cpu().numpy()
    return x, data


def check_f64(sysm, x):
    xo, info = oracle.solve_system(sysm.A, sysm.B, sysm.d)
    assert info == 0
    rel = np.abs(x - xo).max() / np.abs(xo).max()
    berr = oracle.backward_error(sysm.A, sysm.B, x, sysm.d).max()
    assert rel <= F64_REL, f"rel diff {rel:.3e}"
    assert berr <= F64_BERR, f"backward error {berr:.3e}"
    return rel, berr


@pytest.mark.parametrize("n", [16, 32, 48, 64, 96, 128])
@pytest.mark.parametrize("schedule", ["auto", "reference"])
def test_f64_small(n, schedule):
    sysm = workloads.spd_system(50, n, nrhs=1, seed=7 + n)
    x, data = run_gpu(sysm, schedule=schedule)
    print(f"n={n} schedule={schedule} levels={data.levels}", check_f64(sysm, x))
    assert data.info()[0] == 0


@pytest.mark.parametrize("N", [1, 2, 3, 5, 9, 17, 100, 257])
def test_f64_sizes(N):
    sysm = workloads.spd_system(N, 32, nrhs=2, seed=N)
    for schedule in ("auto", "reference"):
        x, _ = run_gpu(sysm, schedule=schedule)
        check_f64(sysm, x)


def test_f64_config1_reference_schedule():
    """Config 1 (N=100, n=32, FP64, seed 42): one level (100, 8, 12) + base of 12."""
    sysm = workloads.spd_system(100, 32, nrhs=1, seed=42)
    x, data = run_gpu(sysm, schedule="reference")
    assert data.level_sizes() == [100, 12]
    check_f64(sysm, x)


def test_f64_multilevel_padding():
    """N=2000 reference schedule pads to N_use and recurses several levels."""
    sysm = workloads.spd_system(2000, 16, nrhs=1, seed=3)
    x, data = run_gpu(sysm, schedule="reference")
    assert data.N >= 2000 and data.levels >= 3
    check_f64(sysm, x)
    # the user's d blocks beyond N stay zero-solution for the identity padding
    pad = data.d_tensor[:, :, 2000:].cpu().numpy()
    assert np.all(pad == 0)


@pytest.mark.parametrize("chain_len", [1, 2, 3, 5, 16])
def test_f64_chain_lengths(chain_len):
    sysm = workloads.spd_system(300, 32, nrhs=3, seed=11)
    x, _ = run_gpu(sysm, chain_len=chain_len, base_max=4)
    check_f64(sysm, x)


@pytest.mark.parametrize("n", [64, 128])
def test_f64_multi_rhs16(n):
    sysm = workloads.spd_system(200, n, nrhs=16, seed=5)
    x, _ = run_gpu(sysm)
    check_f64(sysm, x)


@pytest.mark.parametrize("n", [96, 128])
def test_big_blocks_multilevel_and_info(n):
    """n = 96 / 128 (staged factor kernel): several levels, chain lengths
    1..4, a non-SPD block reported at its index."""
    sysm = workloads.spd_system(700, n, nrhs=3, seed=40 + n)
    for cl in (1, 3, 4):
        x, data = run_gpu(sysm, chain_len=cl, base_max=2, tuned=False)
        assert data.levels >= 3
        check_f64(sysm, x)
    A = sysm.A.copy(order="F")
    A[:, :, 333] = -np.eye(n)
    _, data = run_gpu(workloads.BlockSystem(A, sysm.B, sysm.d), tuned=False)
    assert data.info()[0] > 0


@pytest.mark.parametrize("n", [32, 64, 96, 128])
def test_f32_matches_reference_tolerance(n):
    sysm = workloads.spd_system(50, n, nrhs=1, seed=21, manufactured=True)
    s32 = workloads.BlockSystem(sysm.A.astype(np.float32, order="F"),
                                sysm.B.astype(np.float32, order="F"),
                                sysm.d.astype(np.float32, order="F"))
    x, _ = run_gpu(s32, T=torch.float32)
    xo, _ = oracle.solve_system(sysm.A, sysm.B, sysm.d)
    err = oracle.bss_norm(x - xo)
    assert err <= 10 * np.sqrt(np.finfo(np.float32).eps) * oracle.bss_norm(xo), err


def test_repeated_solves_new_rhs():
    """solve! may be repeated after one factorize! (factors are not modified)."""
    sysm = workloads.spd_system(120, 32, nrhs=1, seed=9)
    x, data = run_gpu(sysm)
    check_f64(sysm, x)
    rng = np.random.default_rng(1)
    for _ in range(3):
        d2 = np.asfortranarray(rng.uniform(size=sysm.d.shape))
        data.d_tensor[:, :, :sysm.N].copy_(torch.from_numpy(d2))
        solve(data)
        torch.cuda.synchronize()
        x2 = data.d_tensor.cpu().numpy()
        check_f64(workloads.BlockSystem(sysm.A, sysm.B, d2), x2)


def test_non_spd_reports_info():
    sysm = workloads.spd_system(40, 16, nrhs=1, seed=2)
    A = sysm.A.copy(order="F")
    A[:, :, 17] = -np.eye(16)  # block 18 (1-based) is negative definite
    bad = workloads.BlockSystem(A, sysm.B, sysm.d)
    _, data = run_gpu(bad)
    info = data.info()[0]
    assert info > 0
    ref_info = oracle.factorize(A.copy(order="F"), sysm.B.copy(order="F"))
    assert ref_info > 0


def run_gpu_batch(systems, T=torch.float64, **kw):
    N, n, nrhs = systems[0].N, systems[0].n, systems[0].nrhs
    data = initialize_batched(N, n, T, "cuda", nrhs=nrhs, nsys=len(systems), **kw)
    for s, sysm in enumerate(systems):
        data.A_tensor[:, :, :N, s].copy_(torch.from_numpy(sysm.A))
        data.B_tensor[:, :, :N - 1, s].copy_(torch.from_numpy(sysm.B))
        data.d_tensor[:, :, :N, s].copy_(torch.from_numpy(sysm.d))
    factorize(data)
    solve(data)
    torch.cuda.synchronize()
    return [data.d_tensor[:, :, :N, s].double().cpu().numpy() for s in range(len(systems))], data


@pytest.mark.parametrize("n,nrhs", [(16, 1), (32, 1), (48, 1), (64, 1), (96, 1), (64, 16), (16, 3),
                                    (96, 2), (128, 1), (128, 16)])
def test_mixed_refined_backward_error(n, nrhs):
    """FP32 MFMA factorisation + FP64 refinement: normwise backward error <= 1e-12 as
    measured by the FP64 oracle, agreeing with the device's own measurement."""
    sysm = workloads.spd_system(300, n, nrhs=nrhs, seed=100 + n)
    x, data = run_gpu(sysm, mode="mixed")
    berr = oracle.backward_error(sysm.A, sysm.B, x, sysm.d)
    assert berr.max() <= 1e-12, berr
    dev = data.backward_error()[0]
    assert dev.max() <= 1e-12
    assert np.allclose(dev, berr, rtol=0.5, atol=1e-16)
    assert 1 <= data.iterations <= 10
    xo, _ = oracle.solve_system(sysm.A, sysm.B, sysm.d)
    assert np.abs(x - xo).max() / np.abs(xo).max() <= 1e-10
    # inputs untouched in mixed mode
    assert np.array_equal(data.A_tensor[:, :, :sysm.N].cpu().numpy(), sysm.A)


def test_mixed_repeated_solves_and_schedules():
    sysm = workloads.spd_system(257, 32, nrhs=2, seed=17)
    for schedule in ("auto", "reference"):
        x, data = run_gpu(sysm, mode="mixed", schedule=schedule)
        assert oracle.backward_error(sysm.A, sysm.B, x, sysm.d).max() <= 1e-12
        rng = np.random.default_rng(3)
        d2 = np.asfortranarray(rng.uniform(size=sysm.d.shape))
        data.d_tensor[:, :, :sysm.N].copy_(torch.from_numpy(d2))
        solve(data)
        torch.cuda.synchronize()
        x2 = data.d_tensor[:, :, :sysm.N].cpu().numpy()
        assert oracle.backward_error(sysm.A, sysm.B, x2, d2).max() <= 1e-12


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_batch_equals_alone(mode):
    """A system's arithmetic does not depend on its co-residents: solving it inside a
    batch of three gives bitwise the same answer as solving it alone."""
    systems = [workloads.spd_system(400, 32, nrhs=1, seed=200 + k) for k in range(3)]
    xs, data = run_gpu_batch(systems, mode=mode)
    for k, sysm in enumerate(systems):
        xa, _ = run_gpu(sysm, mode=mode)
        if mode == "plain":
            assert np.array_equal(xs[k], xa), k
        else:
            # the iteration count is shared by the batch (stop when all converge), so a
            # system may get one more refinement step than alone: equal to the tolerance
            assert np.abs(xs[k] - xa).max() <= 1e-12 * np.abs(xa).max()
        assert oracle.backward_error(sysm.A, sysm.B, xs[k], sysm.d).max() <= (
            1e-13 if mode == "plain" else 1e-12)


@pytest.mark.parametrize("mode,T", [("plain", torch.float32), ("mixed", torch.float64)])
def test_factor_launch_variants_agree(mode, T):
    """At n = 64 the FP32 factorisation has two launch variants: two chains per CU (a
    level with more chains than CUs) and one chain per CU with prefetch (fewer).  A
    system factored alone (75 top-level chains) and inside a batch of eight (600) runs
    its top level through different variants with the same arithmetic."""
    systems = [workloads.spd_system(300, 64, nrhs=1, seed=500 + k) for k in range(8)]
    if T == torch.float32:
        systems = [workloads.BlockSystem(s.A.astype(np.float32, order="F"),
                                         s.B.astype(np.float32, order="F"),
                                         s.d.astype(np.float32, order="F")) for s in systems]
    xs, _ = run_gpu_batch(systems, T=T, mode=mode, chain_len=3, factor_waves=8, tuned=False)
    for k in (0, 5):
        xa, _ = run_gpu(systems[k], T=T, mode=mode, chain_len=3, factor_waves=8,
                        tuned=False)
        if mode == "plain":
            assert np.array_equal(xs[k], xa), k
        else:
            assert np.abs(xs[k] - xa).max() <= 1e-12 * np.abs(xa).max()
            assert oracle.backward_error(systems[k].A, systems[k].B, xs[k],
                                         systems[k].d).max() <= 1e-12


def run_gpu_group(systems, T=torch.float64, combined=False, **kw):
    from blockstructuredsolvers_amd import initialize_batched_group
    n, nrhs = systems[0].n, systems[0].nrhs
    data = initialize_batched_group([s.N for s in systems], n, T, "cuda", nrhs=nrhs, **kw)
    for s, sysm in enumerate(systems):
        N = sysm.N
        data.A_tensors[s][:, :, :N].copy_(torch.from_numpy(sysm.A))
        data.B_tensors[s][:, :, :N - 1].copy_(torch.from_numpy(sysm.B))
        data.d_tensors[s][:, :, :N].copy_(torch.from_numpy(sysm.d))
    if combined:
        factorize_solve(data)
    else:
        factorize(data)
        solve(data)
    torch.cuda.synchronize()
    return [data.d_tensors[s][:, :, :sysm.N].double().cpu().numpy()
            for s, sysm in enumerate(systems)], data


UNEVEN_N = [1, 2, 7, 100, 257, 1000]


@pytest.mark.parametrize("schedule", ["auto", "reference"])
def test_group_uneven_f64(schedule):
    """An uneven batch (different N per system, one plan) matches the oracle per system
    and each system's answer is bitwise the one it gets alone."""
    systems = [workloads.spd_system(N, 32, nrhs=2, seed=300 + N) for N in UNEVEN_N]
    xs, data = run_gpu_group(systems, schedule=schedule, tuned=False)
    assert np.all(data.info() == 0)
    for sysm, x in zip(systems, xs):
        check_f64(sysm, x)
        xa, _ = run_gpu(sysm, schedule=schedule, tuned=False)
        assert np.array_equal(x, xa), sysm.N


def test_group_uneven_mixed_and_info():
    systems = [workloads.spd_system(N, 64, nrhs=1, seed=400 + N) for N in (3, 500, 1200)]
    xs, data = run_gpu_group(systems, mode="mixed", tuned=False)
    for sysm, x in zip(systems, xs):
        assert oracle.backward_error(sysm.A, sysm.B, x, sysm.d).max() <= 1e-12
    # a failing block in the middle system is reported for that system only
    bad = systems[1]
    A = bad.A.copy(order="F")
    A[:, :, 37] = -np.eye(64)
    systems[1] = workloads.BlockSystem(A, bad.B, bad.d)
    _, data = run_gpu_group(systems, tuned=False)
    info = data.info()
    assert info[0] == 0 and info[2] == 0 and info[1] > 0


@pytest.mark.parametrize("mode,T", [("plain", torch.float64), ("plain", torch.float32),
                                    ("mixed", torch.float64)])
def test_graph_replay_matches_eager(mode, T):
    """With hipGraph replay on, every replay (new A, B and d each time, with unrelated
    eager device work and synchronisations between replays) is bitwise the eager
    result of the same plan configuration."""
    N, n = 300, 32
    plans = {g: initialize_batched(N, n, T, "cuda", nrhs=2, mode=mode, tuned=False, graph=g)
             for g in (False, True)}
    junk = torch.randn(512, 512, device="cuda")
    for rep in range(4):
        sysm = workloads.spd_system(N, n, nrhs=2, seed=500 + rep)
        out = {}
        for g, data in plans.items():
            data.A_tensor.copy_(torch.from_numpy(sysm.A))
            data.B_tensor.copy_(torch.from_numpy(sysm.B))
            data.d_tensor.copy_(torch.from_numpy(sysm.d))
            factorize(data)
            solve(data)
            junk = torch.tanh(junk @ junk.T * 1e-3)  # eager work between replays
            torch.cuda.synchronize()
            out[g] = data.d_tensor.double().cpu().numpy()
        assert np.array_equal(out[True], out[False]), rep
        if T == torch.float64:
            bound = 1e-13 if mode == "plain" else 1e-12
            assert oracle.backward_error(sysm.A, sysm.B, out[True], sysm.d).max() <= bound
    # repeated solves on the same factors replay the solve graph only
    data = plans[True]
    d2 = np.asfortranarray(np.random.default_rng(9).uniform(size=(n, 2, N)))
    for _ in range(3):
        data.d_tensor.copy_(torch.from_numpy(d2))
        solve(data)
        torch.cuda.synchronize()
    x2 = data.d_tensor.double().cpu().numpy()
    plans[False].d_tensor.copy_(torch.from_numpy(d2))
    solve(plans[False])
    torch.cuda.synchronize()
    assert np.array_equal(x2, plans[False].d_tensor.double().cpu().numpy())


@pytest.mark.parametrize("n,nrhs", [(32, 2), (64, 1), (16, 16)])
def test_mixed_fp64_separator_refactor(n, nrhs):
    """Mixed mode re-factors the separator levels in FP64 when an FP32 separator pivot
    falls below the threshold; a threshold above every possible U_jj^2 / A_jj (2.0)
    forces that path.  The refined answer meets the same bars, and a well-conditioned
    input under the default threshold keeps FP32 separators."""
    sysm = workloads.spd_system(700, n, nrhs=nrhs, seed=600 + n)
    x, data = run_gpu(sysm, mode="mixed", pivot_tol=2.0, tuned=False)
    assert data.levels > 1 and data.refactored == data.levels - 1
    assert oracle.backward_error(sysm.A, sysm.B, x, sysm.d).max() <= 1e-12
    xo, _ = oracle.solve_system(sysm.A, sysm.B, sysm.d)
    assert np.abs(x - xo).max() / np.abs(xo).max() <= 1e-10
    assert data.info()[0] == 0
    it64 = data.iterations
    # repeated solve on the FP64-separator factors
    d2 = np.asfortranarray(np.random.default_rng(4).uniform(size=sysm.d.shape))
    data.d_tensor.copy_(torch.from_numpy(d2))
    solve(data)
    torch.cuda.synchronize()
    assert oracle.backward_error(sysm.A, sysm.B, data.d_tensor.cpu().numpy(), d2).max() <= 1e-12
    x32, d32 = run_gpu(sysm, mode="mixed", tuned=False)
    assert d32.refactored == 0
    assert it64 <= d32.iterations


def test_mixed_fp64_separator_refactor_group_graph():
    """The FP64-separator path in an uneven group with graph replay."""
    systems = [workloads.spd_system(N, 32, nrhs=1, seed=700 + N) for N in (2, 90, 600)]
    for graph in (False, True):
        xs, data = run_gpu_group(systems, mode="mixed", pivot_tol=2.0, tuned=False, graph=graph)
        assert data.refactored == data.levels - 1
        for sysm, x in zip(systems, xs):
            assert oracle.backward_error(sysm.A, sysm.B, x, sysm.d).max() <= 1e-12



@pytest.mark.parametrize("mode,T,N,n,nrhs,pivot_tol",
                         [("plain", torch.float64, 1500, 64, 1, None),
                          ("plain", torch.float32, 1500, 64, 2, None),
                          ("mixed", torch.float64, 1500, 64, 1, None),
                          ("mixed", torch.float64, 1500, 64, 3, 2.0),
                          ("mixed", torch.float64, 400, 128, 1, None),
                          ("plain", torch.float64, 700, 16, 16, None)])
def test_factorize_solve_matches_separate_calls(mode, T, N, n, nrhs, pivot_tol):
    """factorize_solve (first solve's level-0 forward sweep on a side stream beside the
    separator levels' factorisation) is bitwise factorize then solve, on new inputs
    each time, with eager device work between calls, and leaves the same factors for
    a later solve; pivot_tol = 2.0 forces the FP64 separator re-factorisation."""
    kw = dict(nrhs=nrhs, mode=mode, tuned=False, nsys=2)
    if pivot_tol is not None:
        kw["pivot_tol"] = pivot_tol
    plans = {c: initialize_batched(N, n, T, "cuda", **kw) for c in (False, True)}
    assert plans[True].levels > 2
    junk = torch.randn(512, 512, device="cuda")
    for rep in range(3):
        systems = [workloads.spd_system(N, n, nrhs=nrhs, seed=900 + 10 * rep + s) for s in range(2)]
        out = {}
        for c, data in plans.items():
            for s, sysm in enumerate(systems):
                data.A_tensor[..., s].copy_(torch.from_numpy(sysm.A))
                data.B_tensor[..., s].copy_(torch.from_numpy(sysm.B))
                data.d_tensor[..., s].copy_(torch.from_numpy(sysm.d))
            if c:
                factorize_solve(data)
            else:
                factorize(data)
                solve(data)
            junk = torch.tanh(junk @ junk.T * 1e-3)
            torch.cuda.synchronize()
            out[c] = data.d_tensor.double().cpu().numpy()
        assert np.array_equal(out[True], out[False]), rep
        if mode == "mixed":
            assert plans[True].refactored == plans[False].refactored
            if pivot_tol:
                assert plans[True].refactored == plans[True].levels - 1
        assert plans[True].info().tolist() == plans[False].info().tolist() == [0, 0]
        if mode == "mixed":
            assert plans[True].iterations == plans[False].iterations
        if T == torch.float64:
            bound = 1e-13 if mode == "plain" else 1e-12
            for s, sysm in enumerate(systems):
                assert oracle.backward_error(sysm.A, sysm.B, out[True][..., s], sysm.d).max() <= bound
    # a later solve on the factors left by factorize_solve
    d2 = np.asfortranarray(np.random.default_rng(5).uniform(size=tuple(plans[True].d_tensor.shape)))
    x2 = {}
    for c, data in plans.items():
        data.d_tensor.copy_(torch.from_numpy(d2))
        solve(data)
        torch.cuda.synchronize()
        x2[c] = data.d_tensor.double().cpu().numpy()
    assert np.array_equal(x2[True], x2[False])


def test_factorize_solve_group_and_graph():
    """factorize_solve on an uneven group (mixed mode), and with graph replay on (where
    it is factorize then solve): bitwise the separate calls."""
    systems = [workloads.spd_system(N, 32, nrhs=1, seed=950 + N) for N in (1, 90, 700)]
    for graph in (False, True):
        xs, _ = run_gpu_group(systems, mode="mixed", tuned=False, graph=graph)
        ys, data = run_gpu_group(systems, mode="mixed", tuned=False, graph=graph, combined=True)
        for x, y in zip(xs, ys):
            assert np.array_equal(x, y)
        assert data.info().tolist() == [0, 0, 0]


def test_seq_solver_matches_oracle():
    """initialize_seq / factorize / solve (src/gpu_seq.jl): one chain, no recursion,
    filled block by block through the per-block views as the reference's callers do."""
    from blockstructuredsolvers_amd import initialize_seq
    sysm = workloads.spd_system(40, 32, nrhs=1, seed=77)
    data = initialize_seq(40, 32, torch.float64, "cuda")
    assert data.levels == 1 and data.N == 40
    assert data.A_vec.shape == (40 * 32 * 32, 1) and data.d_vec.shape == (40 * 32, 1)
    ptrs = data.A_ptrs.cpu().numpy()
    assert all(int(ptrs[i]) == data.A_list[i].data_ptr() for i in (0, 1, 39))
    assert int(data.B_ptrs.cpu()[38]) == data.B_list[38].data_ptr()
    for i in range(40):
        data.A_list[i].copy_(torch.from_numpy(sysm.A[:, :, i]))
        data.d_list[i].copy_(torch.from_numpy(sysm.d[:, :, i]))
        if i < 39:
            data.B_list[i].copy_(torch.from_numpy(sysm.B[:, :, i]))
    factorize(data)
    solve(data)
    torch.cuda.synchronize()
    check_f64(sysm, data.d_tensor.cpu().numpy())
    assert data.info()[0] == 0


def test_mixed_first_pass_check_depends_on_tolerance():
    """The refinement loop skips the host's backward-error check after the first FP32
    pass only when the tolerance lies below what one FP32 solve can reach (1e-9): a
    loose tolerance (1e-5) stops after one correction, the default (1e-12) needs at
    least two, and both meet their tolerance by the oracle's measure."""
    sysm = workloads.spd_system(400, 64, nrhs=1, seed=901)
    x, data = run_gpu(sysm, mode="mixed", tol=1e-5)
    assert data.iterations == 1
    assert oracle.backward_error(sysm.A, sysm.B, x, sysm.d).max() <= 1e-5
    x, data = run_gpu(sysm, mode="mixed")
    assert data.iterations >= 2
    assert oracle.backward_error(sysm.A, sysm.B, x, sysm.d).max() <= 1e-12


@pytest.mark.parametrize("n,nrhs", [(16, 1), (64, 1), (96, 1), (128, 1), (96, 2), (128, 3)])
def test_mixed_device_backward_error_after_one_pass(n, nrhs):
    """After one FP32 pass the residual is about 1e-7 of ||M|| ||x|| and FP64 evaluates
    it to about 1e-9 of itself, so the device's backward error must match the oracle's
    closely: this pins the ||M||_inf row sums (|A_i|, |B_i|, and the |B_{i-1}^T| column
    sums that the one-column kernels from n = 96 on reduce across lanes) and the carry
    B_{i-1}^T x_{i-1} in the residual itself."""
    sysm = workloads.spd_system(300, n, nrhs=nrhs, seed=700 + n + nrhs)
    x, data = run_gpu(sysm, mode="mixed", tol=1e-5)
    assert data.iterations == 1
    berr = oracle.backward_error(sysm.A, sysm.B, x, sysm.d)
    dev = data.backward_error()[0]
    assert berr.max() <= 1e-5
    assert np.allclose(dev, berr, rtol=1e-6, atol=0), (dev, berr)
