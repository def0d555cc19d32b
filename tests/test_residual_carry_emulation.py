"""The residual kernel's cross-lane column sums (red16 / red16_first in
blockstructuredsolvers_amd/csrc/mixed.hip), emulated lane for lane in numpy: for G
slots per lane and the 16 lanes of a row group, halving exchanges over lane masks 8, 4,
... while the slot count stays even, then whole exchanges for the masks left, must
leave in every lane the sums over the 16 lanes of the slots red16_first names, bitwise
equal among the lanes that share them, for the group sizes the kernel uses (8 at
n = 64 and 128, 4 at n = 96).  A planted index mistake (the wrong half kept) fails."""
import numpy as np
import pytest


def halvings(G):
    v, h = G, 0
    while h < 4 and v % 2 == 0:
        v //= 2
        h += 1
    return h


def red16(vals, keep_wrong_half=False):
    """vals: (16, G) per-lane slot values; returns the per-lane results (16, CNT)."""
    G = vals.shape[1]
    H = halvings(G)
    v = vals.copy()
    V = G
    for i in range(H):
        m = 8 >> i
        half = V // 2
        new = v.copy()
        for lr in range(16):
            up = (lr & m) != 0
            if keep_wrong_half:
                up = not up
            partner = lr ^ m
            pup = (partner & m) != 0
            if keep_wrong_half:
                pup = not pup
            for j in range(half):
                recv = v[partner, j] if pup else v[partner, half + j]  # partner's send
                own = v[lr, half + j] if up else v[lr, j]
                new[lr, j] = own + recv
        v = new
        V = half
    cnt = G >> H
    for i in range(H, 4):
        m = 8 >> i
        v = v.copy()
        v[:, :cnt] = v[:, :cnt] + v[np.arange(16) ^ m, :cnt]
    return v[:, :cnt]


def first_slot(G, lr):
    f, V = 0, G
    for i in range(halvings(G)):
        V //= 2
        if lr & (8 >> i):
            f += V
    return f


@pytest.mark.parametrize("G", [4, 8])
def test_red16_sums_land_on_named_slots(G):
    rng = np.random.default_rng(G)
    vals = rng.standard_normal((16, G))
    out = red16(vals)
    cnt = G >> halvings(G)
    colsum = vals.sum(axis=0)
    seen = set()
    for lr in range(16):
        f = first_slot(G, lr)
        for c in range(cnt):
            assert np.isclose(out[lr, c], colsum[f + c], rtol=1e-12, atol=1e-12)
            seen.add(f + c)
    assert seen == set(range(G))  # every slot's sum is written by some lane
    # lanes naming the same slot hold the same bits (all of them store it)
    by_slot = {}
    for lr in range(16):
        for c in range(cnt):
            by_slot.setdefault(first_slot(G, lr) + c, set()).add(out[lr, c].tobytes())
    assert all(len(b) == 1 for b in by_slot.values())


def test_planted_half_mistake_fails():
    rng = np.random.default_rng(1)
    vals = rng.standard_normal((16, 8))
    out = red16(vals, keep_wrong_half=True)
    colsum = vals.sum(axis=0)
    ok = all(np.isclose(out[lr, 0], colsum[first_slot(8, lr)]) for lr in range(16))
    assert not ok
