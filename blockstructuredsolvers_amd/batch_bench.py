"""Uneven-batch benchmark: 512 independent block-tridiagonal SPD systems with block
counts N_k = round(exp(U(ln 16, ln 4096))) and block sizes n_k uniform over
{16, 32, 48, 64, 96, 128}, the same A/B/d generator per system (A_i = G G^T + n I,
B_i ~ N(0,1), d ~ U(0,1)), generated on the device.

The systems are grouped by block size into one group plan per n (every system keeps
its own level schedule; each level of a group is one launch over all its systems) and
the six groups run concurrently on six HIP streams.  One step = factorize + solve of
all 512 systems; the plain FP64 path and the mixed path (FP32 factor + FP64
refinement to 1e-12) are timed.  A sample of systems gets its normwise backward error
||d - M x|| / (||M|| ||x|| + ||d||) measured in FP64 on the device.  Prints one JSON
line.

    python -m blockstructuredsolvers_amd.batch_bench --systems 512 --steps 3
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import factorize, initialize_batched_group, solve

SIZES = (16, 32, 48, 64, 96, 128)

# Level schedule of every system: level-0 chains of 15 blocks, 3 below, base 4.  On this
# batch it is the fastest of the schedules tried, within 2% of each block size's own
# best (MI355X, 512 systems: plain 7,451 -> 8,065 and mixed 8,077 -> 8,573 systems/s
# against the library's default of chains of 3 and base 8).
SCHEDULE = dict(chain_len_levels=[15] + [3] * 15, base_max=4)


def draw(count: int, seed: int):
    rng = np.random.default_rng(seed)
    Ns = np.rint(np.exp(rng.uniform(np.log(16), np.log(4096), size=count))).astype(int)
    ns = rng.choice(SIZES, size=count)
    return [int(v) for v in Ns], [int(v) for v in ns]


def fill_group(data, Ns, n, seed: int):
    """Device-side generator of the seeded SPD blocks for every system of a group
    (torch on the device: data preparation, not part of the timed work)."""
    g = torch.Generator(device=data.device).manual_seed(seed)
    dt = data.dtype
    for s, N in enumerate(Ns):
        G = torch.randn((N, n, n), generator=g, device=data.device, dtype=torch.float64)
        A = torch.bmm(G, G.transpose(1, 2).contiguous())
        A += n * torch.eye(n, device=data.device, dtype=torch.float64)
        # (block, row, col) -> the views' (row, col, block)
        data.A_tensors[s].copy_(A.to(dt).permute(1, 2, 0))
        if N > 1:
            B = torch.randn((N - 1, n, n), generator=g, device=data.device, dtype=torch.float64)
            data.B_tensors[s].copy_(B.to(dt).permute(1, 2, 0))
        d = torch.rand((N, 1, n), generator=g, device=data.device, dtype=torch.float64)
        data.d_tensors[s].copy_(d.to(dt).permute(2, 1, 0))


def run(mode: str, groups, steps: int, warmup: int, seed: int, check: int, picks=None):
    """Time `steps` batch steps; returns (plans, ms per step, sampled right-hand sides):
    `check` systems drawn at random, or the (group, system) pairs of `picks`."""
    plans, streams, dsave, samples = [], [], [], []
    rng = np.random.default_rng(seed)
    for gi, (n, Ns) in enumerate(groups):
        data = initialize_batched_group(Ns, n, torch.float64, "cuda", mode=mode,
                                        tuned=False, **SCHEDULE)
        fill_group(data, Ns, n, seed + gi)
        plans.append(data)
        streams.append(torch.cuda.Stream())
        dsave.append(data._d.clone())
        # a few systems of each group for the backward-error check: their inputs now
        chosen = ([s for g, s in picks if g == gi] if picks is not None else
                  [int(s) for s in rng.permutation(len(Ns))[:max(1, check // len(groups))]])
        for s in chosen:
            samples.append((gi, s, data.d_tensors[s].double().cpu().numpy()))
    torch.cuda.synchronize()

    def step():
        for data, st, d0 in zip(plans, streams, dsave):
            with torch.cuda.stream(st):
                data._d.copy_(d0)
                factorize(data)
                solve(data)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    return plans, (time.perf_counter() - t0) / steps * 1e3, samples


def run_unified(mode: str, groups, steps: int, warmup: int, seed: int, picks=None,
                graph: bool = True):
    """The same batch as `run` through one uneven-batch object (bss_batch_*): one call
    per step for all block sizes, each call one hipGraph launch (plain) or one graph
    launch for the factorisation plus each group's refinement on its own thread
    (mixed).  The systems get the same seeded blocks as in `run`, so plain-mode answers
    are bitwise those of `run`.  Returns (data, ms per step, sampled right-hand sides)."""
    from . import initialize_batch
    Ns = [N for _, g in groups for N in g]
    ns = [n for n, g in groups for _ in g]
    data = initialize_batch(Ns, ns, torch.float64, "cuda", mode=mode, graph=graph, **SCHEDULE)
    first = np.cumsum([0] + [len(g) for _, g in groups])
    for gi, (n, g) in enumerate(groups):
        gen = torch.Generator(device=data.device).manual_seed(seed + gi)
        for k, N in enumerate(g):
            s = int(first[gi]) + k
            G = torch.randn((N, n, n), generator=gen, device=data.device, dtype=torch.float64)
            A = torch.bmm(G, G.transpose(1, 2).contiguous())
            A += n * torch.eye(n, device=data.device, dtype=torch.float64)
            data.A_tensors[s].copy_(A.permute(1, 2, 0))
            if N > 1:
                B = torch.randn((N - 1, n, n), generator=gen, device=data.device,
                                dtype=torch.float64)
                data.B_tensors[s].copy_(B.permute(1, 2, 0))
            d = torch.rand((N, 1, n), generator=gen, device=data.device, dtype=torch.float64)
            data.d_tensors[s].copy_(d.permute(2, 1, 0))
    samples = []
    for gi, k in (picks or []):
        s = int(first[gi]) + k
        samples.append((s, data.d_tensors[s].double().cpu().numpy()))
    d0 = data._d.clone()
    torch.cuda.synchronize()

    def step():
        data._d.copy_(d0)
        factorize(data)
        solve(data)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    return data, (time.perf_counter() - t0) / steps * 1e3, samples


def backward_error(A, B, x, d):
    """max over columns of ||d - M x||_inf / (||M||_inf ||x||_inf + ||d||_inf) in FP64,
    M = tridiag(B^T, A, B), blocks in the views' (row, col, block) layout."""
    At, Bt = A.permute(2, 0, 1), B.permute(2, 0, 1)
    xt, dt = x.permute(2, 0, 1), d.permute(2, 0, 1)
    y = torch.bmm(At, xt)
    rs = At.abs().sum(dim=2)
    if Bt.shape[0]:
        y[:-1] += torch.bmm(Bt, xt[1:].contiguous())
        y[1:] += torch.bmm(Bt.transpose(1, 2).contiguous(), xt[:-1].contiguous())
        rs[:-1] += Bt.abs().sum(dim=2)
        rs[1:] += Bt.abs().sum(dim=1)
    r = (dt - y).abs().amax(dim=(0, 1))
    den = rs.max() * xt.abs().amax(dim=(0, 1)) + dt.abs().amax(dim=(0, 1))
    return float((r / den).max())


def check_samples(plans, samples):
    """Worst backward error of the sampled systems (their original right-hand sides)."""
    worst = 0.0
    for gi, s, d0 in samples:
        data = plans[gi]
        A = data.A_tensors[s].double()
        B = data.B_tensors[s].double()
        x = data.d_tensors[s].double()
        d = torch.from_numpy(d0).to(x.device)
        worst = max(worst, backward_error(A, B, x, d))
    return worst


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=512)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--check", type=int, default=12)
    a = ap.parse_args(argv)
    if not torch.cuda.is_available():
        print("batch_bench needs a GPU", file=sys.stderr)
        return 2
    Ns, ns = draw(a.systems, a.seed)
    groups = [(n, [N for N, m in zip(Ns, ns) if m == n]) for n in SIZES]
    groups = [(n, g) for n, g in groups if g]
    out = {"systems": a.systems, "blocks_total": int(sum(Ns)),
           "groups": {str(n): len(g) for n, g in groups},
           "bytes_A_B_fp64": int(sum(8 * n * n * (2 * N - 1) for N, n in zip(Ns, ns)))}
    for mode in ("plain", "mixed"):
        plans, ms, samples = run(mode, groups, a.steps, a.warmup, a.seed, a.check)
        berr = check_samples(plans, samples)
        out[mode] = {"ms_per_step": ms, "systems_per_s": a.systems / ms * 1e3,
                     "checked_systems": len(samples), "max_backward_error": berr}
        if mode == "mixed":
            out[mode]["iterations"] = [int(d.iterations) for d in plans]
        for d in plans:
            d.close()
        del plans
        torch.cuda.empty_cache()
        print(f"batch_bench: {mode} {out[mode]}", file=sys.stderr, flush=True)
    out["mixed_over_plain"] = out["mixed"]["systems_per_s"] / out["plain"]["systems_per_s"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
