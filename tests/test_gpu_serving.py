"""Request loop on the GPU: requests that arrive while others are in flight are
admitted by HBM footprint and overlap on the device; every served answer equals the
same request solved alone (bitwise for plain FP64, within the refinement tolerance for
mixed) and meets the oracle's backward-error bar."""
import numpy as np
import pytest
import torch

from blockstructuredsolvers_amd import serving
from oracle import oracle

pytestmark = pytest.mark.gpu


def small_stream(count, rate, seed, device_blocks=False):
    return serving.smoother_requests(count, rate, n_choices=(16, 32, 64), N_range=(40, 700),
                                     rhs_range=(1, 4), seed=seed, device_blocks=device_blocks)


@pytest.mark.parametrize("mode,device_blocks", [("plain", False), ("mixed", False),
                                                ("plain", True)])
def test_served_equals_alone(mode, device_blocks):
    reqs = small_stream(14, rate=2000.0, seed=11, device_blocks=device_blocks)
    srv = serving.Server(mode=mode, workers=4, tuned=False)
    res = srv.serve(reqs)
    assert srv.stats["max_inflight"] >= 2  # requests overlapped
    summ = serving.summarize(res)
    assert summ["latency_ms_p99"] >= summ["latency_ms_p50"] > 0
    for r, out in zip(reqs, res):
        assert out.rid == r.rid and out.info == 0 and len(out.x) == len(r.rhs)
        alone = serving.solve_alone(r, mode=mode, tuned=False)
        for k, (xs, xa) in enumerate(zip(out.x, alone)):
            if mode == "plain":
                assert np.array_equal(xs, xa), (r.rid, k)
            else:
                assert np.abs(xs - xa).max() <= 1e-12 * np.abs(xa).max(), (r.rid, k)
            A, B = r.dense()
            berr = oracle.backward_error(A, B, xs, r.rhs[k]).max()
            assert berr <= (1e-13 if mode == "plain" else 1e-12), (r.rid, k, berr)


def test_admission_respects_hbm_budget():
    """With a budget that fits only the two largest requests, the loop never holds
    more than the budget in flight and still serves everything."""
    reqs = small_stream(10, rate=0.0, seed=5)  # all arrive at once
    fps = sorted(serving.request_footprint(r.N, r.n, serving.solve_width(r.width),
                                           schedule=serving.SERVE_SCHEDULE) for r in reqs)
    budget = fps[-1] + fps[-2]
    srv = serving.Server(workers=8, hbm_budget=budget, tuned=False)
    assert sorted(srv.footprint(r) for r in reqs) == fps
    res = srv.serve(reqs, realtime=False)
    assert srv.stats["max_inflight_bytes"] <= budget
    assert len(res) == len(reqs)
    for r, out in zip(reqs, res):
        xo, info = oracle.solve_system(r.A, r.B, r.rhs[0])
        assert info == 0
        assert np.abs(out.x[0] - xo).max() <= 1e-12 * np.abs(xo).max()
    with pytest.raises(ValueError):
        serving.Server(hbm_budget=fps[0] - 1).serve(reqs[:1])


@pytest.mark.parametrize("n,T", [(16, torch.float64), (64, torch.float64), (128, torch.float32)])
def test_device_smoother_blocks_match_host(n, T):
    """bss_build_smoother: the device-built normal-equation blocks equal the host
    formula (A = Q^-1 + F^T Q^-1 F + H^T R^-1 H, B = -Q^-1 F) in every block."""
    from blockstructuredsolvers_amd import build_smoother, factorize, initialize_batched, solve
    from blockstructuredsolvers_amd.workloads import smoother_model
    N = 37
    model = smoother_model(n, np.random.default_rng(n))
    Ai, Bi = model.blocks()
    data = initialize_batched(N, n, T, "cuda", tuned=False)
    build_smoother(data, model)
    torch.cuda.synchronize()
    A = data.A_tensor.double().cpu().numpy()
    B = data.B_tensor.double().cpu().numpy()
    tol = 1e-13 if T == torch.float64 else 1e-6
    for i in range(N):
        assert np.abs(A[:, :, i] - Ai).max() <= tol * np.abs(Ai).max(), i
    for i in range(N - 1):
        assert np.abs(B[:, :, i] - Bi).max() <= tol * np.abs(Bi).max(), i
    if T == torch.float64:
        d = np.asfortranarray(np.random.default_rng(1).uniform(size=(n, 1, N)))
        data.d_tensor.copy_(torch.from_numpy(d))
        factorize(data)
        solve(data)
        torch.cuda.synchronize()
        Ah = np.asfortranarray(np.repeat(Ai[:, :, None], N, axis=2))
        Bh = np.asfortranarray(np.repeat(Bi[:, :, None], N - 1, axis=2))
        assert oracle.backward_error(Ah, Bh, data.d_tensor.cpu().numpy(), d).max() <= 1e-13


def test_serve_frames_round_trip():
    """Requests sent as wire frames (dense and model-carrying) come back as result frames
    whose solutions equal the same requests solved alone, bitwise (plain FP64)."""
    from blockstructuredsolvers_amd import wire
    reqs = small_stream(4, rate=0.0, seed=21) + small_stream(3, rate=0.0, seed=22,
                                                             device_blocks=True)
    for k, r in enumerate(reqs):
        r.rid = 1000 + k
    frames = [wire.encode_request(r) for r in reqs]
    srv = serving.Server(mode="plain", workers=4, tuned=False)
    out_frames, records = wire.serve_frames(srv, frames, realtime=False)
    assert [rec["rid"] for rec in records] == [r.rid for r in reqs]
    for r, fr in zip(reqs, out_frames):
        rid, info, xs = wire.decode_result(fr)
        assert rid == r.rid and info == 0 and len(xs) == len(r.rhs)
        alone = serving.solve_alone(r, mode="plain", tuned=False)
        for xs_k, xa in zip(xs, alone):
            assert np.array_equal(xs_k, xa), r.rid


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_coalesced_groups_equal_alone(mode):
    """Requests waiting together are coalesced into group plans (same n and column
    class; one factorisation and one solve of all columns for the whole group), and
    every answer still equals the request solved alone."""
    reqs = small_stream(16, rate=0.0, seed=31, device_blocks=True)  # all arrive at once
    srv = serving.Server(mode=mode, workers=2, tuned=False)
    res = srv.serve(reqs, realtime=False)
    assert srv.stats["max_group"] >= 3 and srv.stats["groups"] < len(reqs)
    for r, out in zip(reqs, res):
        assert out.info == 0 and len(out.x) == len(r.rhs)
        alone = serving.solve_alone(r, mode=mode, tuned=False)
        for k, (xs, xa) in enumerate(zip(out.x, alone)):
            if mode == "plain":
                assert np.array_equal(xs, xa), (r.rid, k)
            else:
                assert np.abs(xs - xa).max() <= 1e-12 * np.abs(xa).max(), (r.rid, k)


@pytest.mark.parametrize("T", [torch.float64, torch.float32])
@pytest.mark.parametrize("n", [16, 64, 128])
def test_columns_do_not_depend_on_width(n, T):
    """The multi-column sweeps compute every column on its own: the first two columns
    get the same bits from solves 2, 3, 8 and 16 columns wide.  The server relies on
    this when it solves a request inside a wider group plan."""
    from blockstructuredsolvers_amd import factorize, initialize_batched, solve, workloads
    N = 150
    npt = np.float64 if T == torch.float64 else np.float32
    sysm = workloads.spd_system(N, n, nrhs=16, seed=5 + n)
    outs = []
    for P in (2, 3, 8, 16):
        data = initialize_batched(N, n, T, "cuda", nrhs=P, tuned=False)
        data.A_tensor.copy_(torch.from_numpy(sysm.A.astype(npt)))
        data.B_tensor.copy_(torch.from_numpy(sysm.B.astype(npt)))
        data.d_tensor.copy_(torch.from_numpy(np.asfortranarray(sysm.d[:, :P].astype(npt))))
        factorize(data)
        solve(data)
        torch.cuda.synchronize()
        outs.append(data.d_tensor.cpu().numpy()[:, :2].copy())
        data.close()
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_wide_requests_and_padding(mode):
    """Requests of 1 to 20 columns (several right-hand sides of 1 to 3 columns each):
    one-column requests in one-column plans, wider ones padded into plans up to 16
    columns wide and solved in rounds of 16 beyond that; every answer equals the
    request solved alone and meets the oracle's backward-error bar."""
    rng = np.random.default_rng(41)
    reqs = []
    for k, (reps, w) in enumerate([(1, 1), (1, 1), (2, 1), (3, 2), (1, 3), (7, 1), (5, 3),
                                   (4, 3), (1, 1), (6, 3)]):
        base = serving.smoother_requests(1, 0.0, n_choices=(32,), N_range=(60, 200),
                                         rhs_range=(reps, reps), nrhs=w,
                                         seed=int(rng.integers(1 << 30)))[0]
        base.rid = k
        reqs.append(base)
    assert max(r.width for r in reqs) > 16
    srv = serving.Server(mode=mode, workers=2, tuned=False)
    res = srv.serve(reqs, realtime=False)
    assert srv.stats["groups"] < len(reqs)
    for r, out in zip(reqs, res):
        assert out.info == 0 and len(out.x) == len(r.rhs) and len(out.solve_ms) == len(r.rhs)
        alone = serving.solve_alone(r, mode=mode, tuned=False)
        A, B = r.dense()
        for q, (xs, xa) in enumerate(zip(out.x, alone)):
            assert xs.shape == (r.n, r.nrhs, r.N)
            if mode == "plain":
                assert np.array_equal(xs, xa), (r.rid, q)
            else:
                assert np.abs(xs - xa).max() <= 1e-12 * np.abs(xa).max(), (r.rid, q)
            berr = oracle.backward_error(A, B, np.asfortranarray(xs), r.rhs[q]).max()
            assert berr <= (1e-13 if mode == "plain" else 1e-12), (r.rid, q, berr)
