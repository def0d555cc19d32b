"""Schedule autotuner for libbss plans.

A plan's speed depends on its level schedule (chain length per level, size of the
single-chain base level) and, at n = 64, on the factor workgroup size (4 or 8
waves).  The reference fixes its schedule by an integer search
(src/gpu_batched.jl:112-162); here candidates are timed on the device and the fastest
is cached in ``tuned.json`` next to this file.

Cache keys are (N, n, element type, mode) plus, for schedules tuned on a batch, the
batch size bucket (":x8" = tuned on 8 systems per plan): the best level-0 chain length
depends on how many chains the batch offers the device (long chains mean less work on
the latency-bound lower levels, but fewer workgroups).  A plan of S systems uses the
entry of the largest tuned bucket <= S, else the single-system entry, so a system's
schedule -- and therefore its arithmetic -- is the same alone and in a batch of fewer
than the smallest tuned batch size; explicit schedule options, or tuned=False, fix it
everywhere.
"""
from __future__ import annotations

import itertools
import json
import os
import time
from typing import Dict, Iterable, List, Optional

import numpy as np
import torch

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")


def key(N: int, n: int, dtype, mode: str, nsys: int = 1) -> str:
    t = "f64" if dtype in (torch.float64, np.float64, "float64") else "f32"
    return f"N{N}:n{n}:{t}:{mode}" + (f":x{nsys}" if nsys > 1 else "")


def load_cache(path: str = CACHE) -> Dict[str, dict]:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def _bucket(k: str) -> int:
    parts = k.split(":")
    return int(parts[4][1:]) if len(parts) > 4 else 1


def lookup(N: int, n: int, dtype, mode: str, path: str = CACHE, nsys: int = 1) -> Optional[dict]:
    """The entry of the largest tuned batch-size bucket <= nsys: exact (N, n, type,
    mode), else the one with the same n/type/mode whose N is closest on a log scale
    (schedules transfer across nearby sizes)."""
    cache = load_cache(path)
    base = key(N, n, dtype, mode)
    suffix = base.split(":", 1)[1]
    best, rank = None, None
    for kk, v in cache.items():
        b = _bucket(kk)
        if b > nsys or ":".join(kk.split(":")[1:4]) != suffix:
            continue
        NN = int(kk.split(":")[0][1:])
        dd = abs(np.log(NN) - np.log(N))
        if dd >= np.log(2.0):
            continue
        r = (-b, dd)  # the largest bucket first, then the nearest N
        if rank is None or r < rank:
            best, rank = v["config"], r
    return best


def candidates(n: int, nsys: int = 1) -> List[dict]:
    """Top-level chain lengths up to 63 (long chains pay when many systems share the
    plan), lower levels 2-4, base 2-8; at n = 64 both factor workgroup sizes (one
    system) or 8 waves (batches, where 8 won every earlier tune)."""
    waves = ([4, 8] if nsys == 1 else [8]) if n == 64 else [None]
    out = []
    for w, top, low, base in itertools.product(waves, [2, 3, 4, 6, 7, 15, 31, 63], [2, 3, 4],
                                               [2, 4, 8]):
        out.append({"factor_waves": w, "chain_len_levels": [top] + [low] * 15,
                    "base_max": base})
    return out


def _device_systems(systems, dtype):
    """The systems' A, B, d stacked on the device once (system axis last)."""
    st = lambda key: torch.stack([torch.from_numpy(getattr(s, key)).to(dtype) for s in systems],
                                 dim=-1).cuda()
    return st("A"), st("B"), st("d")


def _time_config(cfg, systems, dtype, mode, reps, bss, dev=None):
    from .solver import initialize_batched
    s0 = systems[0]
    N, n, nrhs, S = s0.N, s0.n, s0.nrhs, len(systems)
    data = initialize_batched(N, n, dtype, "cuda", nrhs=nrhs, nsys=S, mode=mode,
                              **cfg)
    A, B, d0 = dev if dev is not None else _device_systems(systems, dtype)
    if S == 1:
        A, B, d0 = A[..., 0], B[..., 0], d0[..., 0]
    data.A_tensor.copy_(A)
    data.B_tensor.copy_(B)
    ts = []
    for r in range(reps + 1):
        data.d_tensor.copy_(d0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bss.factorize(data)
        bss.solve(data)
        torch.cuda.synchronize()
        if r:
            ts.append(time.perf_counter() - t0)
    data.close()
    return float(np.median(ts)) * 1e3


def tune(N: int, n: int, nrhs: int = 1, dtype=torch.float64, mode: str = "plain",
         reps: int = 3, configs: Optional[Iterable[dict]] = None, save: bool = True,
         path: str = CACHE, log=print, nsys: int = 1) -> dict:
    """Time every candidate schedule on `nsys` seeded systems in one plan (the batch the
    schedule will serve) and cache the fastest."""
    import blockstructuredsolvers_amd as bss
    from .workloads import spd_system
    systems = [spd_system(N, n, nrhs=nrhs, seed=1234 + k) for k in range(nsys)]
    dev = _device_systems(systems, dtype)
    results = []
    for cfg in (configs or candidates(n, nsys)):
        c = {k: v for k, v in cfg.items() if v is not None}
        ms = _time_config(c, systems, dtype, mode, reps, bss, dev)
        results.append((ms, c))
        log(f"autotune {key(N, n, dtype, mode, nsys)} {c} -> {ms:.3f} ms")
    results.sort(key=lambda t: t[0])
    best_ms, best = results[0]
    entry = {"config": best, "ms": best_ms,
             "default_ms": next((m for m, c in results if c.get("factor_waves", 4) == 4 and
                                 c["chain_len_levels"][0] == 3 and c["chain_len_levels"][1] == 3
                                 and c["base_max"] == 8), None)}
    if save:
        cache = load_cache(path)
        cache[key(N, n, dtype, mode, nsys)] = entry
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return entry


def main(argv=None) -> int:
    """python -m blockstructuredsolvers_amd.autotune --N 8192 --n 64 --nsys 8
    --modes plain,mixed: tune and cache the schedule of each mode (FP64 input)."""
    import argparse
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=8192)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--nsys", type=int, default=1)
    ap.add_argument("--modes", default="plain,mixed")
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args(argv)
    for mode in a.modes.split(","):
        e = tune(a.N, a.n, a.nrhs, torch.float64, mode, reps=a.reps, nsys=a.nsys,
                 log=lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps({"key": key(a.N, a.n, torch.float64, mode, a.nsys), **e}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
