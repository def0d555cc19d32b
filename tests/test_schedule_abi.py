"""Host-side logic of libbss on the CPU: the library loads and exports every symbol
include/bss.h declares, argument errors come back as statuses, the reference schedule
it computes matches the restatement of find_sequence_upper, and the auto schedule
tiles every level into chains and separators."""
import ctypes
import os
import re

import numpy as np
import pytest

import blockstructuredsolvers_amd as bss
from blockstructuredsolvers_amd import _lib as L
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "bss.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bss_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    syms = header_symbols()
    assert len(syms) >= 13
    lib = L.lib()
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(L.SIGNATURES)


def test_python_constants_match_header_enums():
    """Every BSS_* enumerator of include/bss.h that the Python mirror names has the
    header's value there (options, statuses, modes, kernel kinds, queries)."""
    txt = open(os.path.join(ROOT, "include", "bss.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    enums = {k: int(v) for k, v in re.findall(r"\b(BSS_[A-Z0-9_]+)\s*=\s*(-?\d+)", txt)}
    assert enums["BSS_OPT_FACTOR_SWEEP"] == 11
    shared = [k for k in enums if hasattr(L, k)]
    assert len(shared) >= 12
    for k in shared:
        assert getattr(L, k) == enums[k], k


def test_argument_errors():
    p = ctypes.c_void_p()
    rc = L.lib().bss_plan_create(10, 24, 1, 1, 1, 0, 0, ctypes.byref(p))
    assert rc == -5 and "n=24" in L.last_error()
    rc = L.lib().bss_plan_create(10, 32, 1, 1, 0, 1, 0, ctypes.byref(p))
    assert rc == -1 and "mixed" in L.last_error()
    rc = L.lib().bss_plan_create(0, 32, 1, 1, 1, 0, 0, ctypes.byref(p))
    assert rc == -1
    rc = L.lib().bss_plan_create(10, 32, 17, 1, 1, 0, 0, ctypes.byref(p))
    assert rc == -5
    assert L.lib().bss_factorize(None, None) == -1
    assert L.lib().bss_factorize_solve(None, None) == -1
    L.call("bss_plan_create", 10, 32, 1, 1, 1, 0, 0, ctypes.byref(p))
    try:
        assert L.lib().bss_factorize(p, None) == -4  # not bound
        assert L.lib().bss_solve(p, None) == -4      # not factored
        assert L.lib().bss_factorize_solve(p, None) == -4  # not bound
        assert L.lib().bss_set_option(p, 99, 1.0) == -1
    finally:
        L.lib().bss_plan_destroy(p)


@pytest.mark.parametrize("N", [1, 2, 3, 5, 9, 16, 50, 100, 257, 1000, 4096, 8192, 10007])
def test_reference_schedule_matches_restatement(N):
    n_use, levels = bss.reference_schedule(N)
    u, P, Nl, ml = oracle.ref_find_sequence_upper(N)
    assert n_use == u
    assert levels == list(zip(Nl, ml, P))[::-1]


def plan_chains(N, schedule=0, **opt):
    p = ctypes.c_void_p()
    L.call("bss_plan_create", N, 64, 1, 1, 1, 0, schedule, ctypes.byref(p))
    try:
        for k, v in opt.items():
            L.call("bss_set_option", p, {"chain_len": L.BSS_OPT_CHAIN_LEN,
                                         "base_max": L.BSS_OPT_BASE_MAX}[k], float(v))
        v = ctypes.c_int64()
        L.call("bss_plan_get", p, L.BSS_Q_LEVELS, ctypes.byref(v))
        out = []
        for k in range(v.value):
            L.call("bss_plan_get", p, L.BSS_Q_LEVEL_N + k, ctypes.byref(v))
            NL = v.value
            buf = (ctypes.c_int * (4 * (NL + 2)))()
            nch = L.call("bss_plan_level_chains", p, k, buf, NL + 2)
            out.append((NL, [tuple(buf[4 * c:4 * c + 4]) for c in range(nch)]))
        L.call("bss_plan_get", p, L.BSS_Q_N_USE, ctypes.byref(v))
        return v.value, out
    finally:
        L.lib().bss_plan_destroy(p)


@pytest.mark.parametrize("N", [1, 2, 3, 4, 7, 8, 9, 10, 31, 100, 1000, 8192])
@pytest.mark.parametrize("chain_len,base_max", [(3, 8), (1, 2), (2, 4), (7, 16), (64, 8)])
def test_auto_schedule_tiles_levels(N, chain_len, base_max):
    n_use, levels = plan_chains(N, chain_len=chain_len, base_max=base_max)
    assert n_use == N
    assert levels[0][0] == N
    for k, (NL, chains) in enumerate(levels):
        covered = np.zeros(NL, dtype=int)
        seps = []
        for (start, ln, hl, hr) in chains:
            assert ln >= 1
            covered[start:start + ln] += 1
            if hr:
                seps.append(start + ln)
            if hl:
                assert start - 1 >= 0
        for s_ in seps:
            covered[s_] += 1
        assert np.all(covered == 1), (k, NL)
        if k + 1 < len(levels):
            assert levels[k + 1][0] == len(seps)
            lens = [c[1] for c in chains]
            assert max(lens) - min(lens) <= 1  # balanced chains
            assert max(lens) <= 2 * max(chain_len, 1) + 1
    assert len(levels[-1][1]) == 1  # base: one chain, no separators
    assert levels[-1][0] <= max(base_max, 2) or len(levels) == 1


def test_reference_schedule_levels_layout():
    """Reference layout: separators at blocks k(m+1), chains of m between them."""
    n_use, levels = plan_chains(100, schedule=1)
    assert n_use == 100 and [lv[0] for lv in levels] == [100, 12]
    NL, chains = levels[0]
    assert len(chains) == 11
    assert all(c == (k * 9 + 1, 8, 1, 1) for k, c in enumerate(chains))
    assert levels[1][1] == [(0, 12, 0, 0)]


def test_group_plan_sizes_and_levels():
    """An uneven group: each system keeps the schedule it has alone (same level sizes),
    the group has as many levels as its deepest system and one workgroup per chain of
    every system that has a level."""
    Ns = [1, 7, 100, 1000, 8192]
    p = ctypes.c_void_p()
    arr = (ctypes.c_int * len(Ns))(*Ns)
    L.call("bss_plan_create_group", len(Ns), arr, 64, 1, 1, 0, 0, ctypes.byref(p))
    try:
        nu = (ctypes.c_int * len(Ns))()
        assert L.call("bss_plan_system_sizes", p, nu) == len(Ns)
        assert list(nu) == Ns
        v = ctypes.c_int64()
        L.call("bss_plan_get", p, L.BSS_Q_LEVELS, ctypes.byref(v))
        levels = v.value
        alone = [plan_chains(N) for N in Ns]
        assert levels == max(len(a[1]) for a in alone)
        for k in range(levels):
            L.call("bss_plan_get", p, L.BSS_Q_LEVEL_CHAINS + k, ctypes.byref(v))
            assert v.value == sum(len(a[1][k][1]) for a in alone if k < len(a[1]))
    finally:
        L.lib().bss_plan_destroy(p)
    # reference schedule pads each system on its own
    L.call("bss_plan_create_group", len(Ns), arr, 32, 1, 1, 0, 1, ctypes.byref(p))
    try:
        nu = (ctypes.c_int * len(Ns))()
        L.call("bss_plan_system_sizes", p, nu)
        assert list(nu) == [bss.reference_schedule(N)[0] for N in Ns]
    finally:
        L.lib().bss_plan_destroy(p)


def test_group_argument_errors_and_options():
    p = ctypes.c_void_p()
    bad = (ctypes.c_int * 3)(5, 0, 7)
    assert L.lib().bss_plan_create_group(3, bad, 32, 1, 1, 0, 0, ctypes.byref(p)) == -1
    assert "system 1" in L.last_error()
    assert L.lib().bss_plan_create_group(0, bad, 32, 1, 1, 0, 0, ctypes.byref(p)) == -1
    ok = (ctypes.c_int * 2)(5, 9)
    L.call("bss_plan_create_group", 2, ok, 32, 1, 1, 1, 0, ctypes.byref(p))
    try:
        L.call("bss_set_option", p, L.BSS_OPT_GRAPH, 1.0)
        L.call("bss_set_option", p, L.BSS_OPT_PIVOT_TOL, 2.0)
        assert L.lib().bss_set_option(p, L.BSS_OPT_PIVOT_TOL, -1.0) == -1
        # profiler selection: a mask over the kernel kinds and a top level
        L.call("bss_set_option", p, L.BSS_OPT_PROFILE_KINDS, float(1 << L.KERNEL_KINDS["residual"]))
        L.call("bss_set_option", p, L.BSS_OPT_PROFILE_MAX_LEVEL, 0.0)
        assert L.lib().bss_set_option(p, L.BSS_OPT_PROFILE_KINDS, float(1 << 20)) == -1
        assert L.lib().bss_set_option(p, L.BSS_OPT_PROFILE_KINDS, 2.5) == -1
        assert L.lib().bss_set_option(p, L.BSS_OPT_PROFILE_MAX_LEVEL, -1.0) == -1
        v = ctypes.c_int64()
        L.call("bss_plan_get", p, L.BSS_Q_WORKSPACE_BYTES, ctypes.byref(v))
        assert v.value > 0
        # schedule options are frozen once the layout exists
        assert L.lib().bss_set_option(p, L.BSS_OPT_CHAIN_LEN, 2.0) == -4
    finally:
        L.lib().bss_plan_destroy(p)


def test_serving_workload_and_footprint():
    """Served-stream generator (Poisson arrivals, sizes in range, smoother blocks) and the
    admission footprint (host-only plan query, grows with N and n)."""
    from blockstructuredsolvers_amd import serving
    reqs = serving.smoother_requests(60, rate=100.0, n_choices=(16, 32), N_range=(20, 60),
                                     rhs_range=(1, 8), seed=3)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0)
    assert 0.3 < t[-1] / (60 / 100.0) < 3.0  # mean inter-arrival ~ 1/rate
    dev = serving.smoother_requests(60, rate=100.0, n_choices=(16, 32), N_range=(20, 60),
                                    rhs_range=(1, 8), seed=3, device_blocks=True)
    for r, m in zip(reqs, dev):  # same stream either way; the model expands to the blocks
        assert m.A is None and m.model is not None and m.N == r.N and m.arrival == r.arrival
        A, B = m.dense()
        assert np.array_equal(A, r.A) and np.array_equal(B, r.B)
    for r in reqs:
        assert r.n in (16, 32) and 20 <= r.N <= 60 and 1 <= len(r.rhs) <= 8
        assert r.A.shape == (r.n, r.n, r.N) and r.B.shape == (r.n, r.n, r.N - 1)
        assert np.allclose(r.A[:, :, 0], r.A[:, :, 0].T)
        # smoother blocks are SPD: A_i = Q^-1 + F^T Q^-1 F + H^T H
        assert np.linalg.eigvalsh(r.A[:, :, 0]).min() > 0
    f = serving.request_footprint
    assert f(100, 32, 1) < f(200, 32, 1) < f(200, 64, 1)
    assert f(100, 32, 1, mode="mixed") < f(200, 32, 1, mode="mixed") < f(200, 64, 1, mode="mixed")


def test_group_footprint_bounded_by_sum():
    """The request loop admits a coalesced group by the sum of its systems' single-system
    footprints at the group's width, without laying out the group plan.  That sum must
    bound the group plan's own footprint (host-only plan queries), for random groups over
    every block size, width, precision, mode and schedule (the library default and the
    request loop's); one system gives equality."""
    from blockstructuredsolvers_amd import serving
    import torch
    rng = np.random.default_rng(11)
    for trial in range(120):
        G = int(rng.integers(1, 24))
        n = int(rng.choice([16, 32, 48, 64, 96, 128]))
        nrhs = int(rng.integers(1, 17))
        mode = ("plain", "mixed")[trial % 2]
        dt = torch.float32 if (mode == "plain" and trial % 4 == 0) else torch.float64
        sched = serving.SERVE_SCHEDULE if trial % 3 == 0 else None
        Ns = [int(v) for v in rng.integers(1, 3000, size=G)]
        g = serving.group_footprint(Ns, n, nrhs, dt, mode, sched)
        bound = sum(serving.request_footprint(N, n, nrhs, dt, mode, sched) for N in Ns)
        assert g <= bound, (Ns, n, nrhs, mode, dt, sched)
        if G == 1:
            assert g == bound
    # the schedule changes the workspace (fewer, smaller separator levels with long chains)
    assert (serving.request_footprint(4000, 64, 1, schedule=serving.SERVE_SCHEDULE)
            != serving.request_footprint(4000, 64, 1))


@pytest.mark.parametrize("schedule", [0, 1])
@pytest.mark.parametrize("N", [2, 3, 5, 9, 17, 50, 100, 257, 2000, 8192, 10007])
def test_every_separator_has_one_owner(N, schedule):
    """The solve sweeps do the level hand-off themselves: the chain whose right
    separator it is, or for a separator at block 0 the chain right of it, assembles and
    later writes back each separator.  Every separator of every level must have exactly
    one such owner, in both schedules."""
    p = ctypes.c_void_p()
    L.call("bss_plan_create", N, 32, 1, 1, 1, 0, schedule, ctypes.byref(p))
    try:
        v = ctypes.c_int64()
        L.call("bss_plan_get", p, L.BSS_Q_LEVELS, ctypes.byref(v))
        for k in range(v.value):
            L.call("bss_plan_get", p, L.BSS_Q_LEVEL_N + k, ctypes.byref(v))
            NL = v.value
            buf = (ctypes.c_int * (4 * (NL + 2)))()
            nch = L.call("bss_plan_level_chains", p, k, buf, NL + 2)
            chains = [tuple(buf[4 * c:4 * c + 4]) for c in range(nch)]
            seps = {st + ln for (st, ln, hl, hr) in chains if hr}
            seps |= {st - 1 for (st, ln, hl, hr) in chains if hl}
            owners = {}
            for (st, ln, hl, hr) in chains:
                if hr:
                    owners[st + ln] = owners.get(st + ln, 0) + 1
                if hl and st == 1:
                    owners[0] = owners.get(0, 0) + 1
            assert set(owners) == seps and all(c == 1 for c in owners.values()), (k, owners)
    finally:
        L.lib().bss_plan_destroy(p)
