"""The level algorithm the HIP kernels implement (one-sided panel elimination of each
chain, reduced separator system, recursion), executed in numpy on the chain tables
libbss produces, agrees with the FP64 oracle -- for the auto schedule at several
chain lengths and for the reference schedule with padding.  This separates design
errors from device-code errors."""
import numpy as np
import pytest

from blockstructuredsolvers_amd import workloads
from oracle import oracle

from chain_emulator import solve_system


@pytest.mark.parametrize("N,n,nrhs,kw", [
    (1, 16, 1, {}),
    (2, 16, 2, {}),
    (3, 16, 1, {"base_max": 1}),
    (50, 16, 1, {}),
    (50, 16, 3, {"schedule": 1}),
    (120, 16, 2, {"chain_len": 1, "base_max": 2}),
    (120, 16, 1, {"chain_len": 5, "base_max": 4}),
    (257, 16, 1, {"chain_len": 2}),
    (2000, 16, 1, {"schedule": 1}),
])
def test_emulated_levels_match_oracle(N, n, nrhs, kw):
    s = workloads.spd_system(N, n, nrhs=nrhs, seed=N + nrhs)
    xo, info = oracle.solve_system(s.A, s.B, s.d)
    assert info == 0
    x = solve_system(s.A, s.B, s.d, **kw)
    assert np.abs(x - xo).max() <= 1e-12 * np.abs(xo).max()
    assert oracle.backward_error(s.A, s.B, x, s.d).max() <= 1e-14


@pytest.mark.parametrize("schedule", [0, 1])
def test_info_order_recursive_vs_sequential(schedule):
    """bss_info's design, on the CPU: the recursive elimination order can meet a failure
    at a separator before the block where LAPACK order fails (a strongly coupled pair
    across the separator), so the library re-factors a flagged system as a single chain
    from block 0; that chain's first failure is the oracle's info for every planted
    position."""
    from chain_emulator import blocks_of, coupled_failure, emulate_info, plan_levels, \
        plant_positions
    N, n = 60, 16
    base = workloads.spd_system(N, n, seed=31)
    n_use, rec_levels = plan_levels(N, n, schedule=schedule)
    _, seq_levels = plan_levels(N, n, base_max=N)
    assert len(seq_levels) == 1 and len(seq_levels[0][1]) == 1
    cases = plant_positions(N, n, schedule)
    if schedule == 1:
        assert cases["separator"] == 0 or cases["separator"] % 9 == 0
    for name, k in cases.items():
        A = base.A.copy(order="F")
        A[:, :, k] = -np.eye(n)
        ref = oracle.factorize(A.copy(order="F"), base.B.copy(order="F"))
        assert ref == k + 1, name
        assert emulate_info(*blocks_of(A, base.B), seq_levels) == ref, name
        assert emulate_info(*blocks_of(A, base.B, n_use), rec_levels) > 0, name
    k = cases["chain_first"]
    bad = coupled_failure(base, k)
    ref = oracle.factorize(bad.A.copy(order="F"), bad.B.copy(order="F"))
    assert ref == k + 1
    # the recursive order fails first at the separator k - 1 (1-based k)
    assert emulate_info(*blocks_of(bad.A, bad.B, n_use), rec_levels) == k
    assert emulate_info(*blocks_of(bad.A, bad.B), seq_levels) == ref
