"""Schedule autotuner for libbss plans.

A plan's speed depends on its level schedule (chain length per level, size of the
single-chain base level) and, at n = 64, on the factor workgroup size (4 or 8
waves).  The reference fixes its schedule by an integer search
(src/gpu_batched.jl:112-162); here candidates are timed on the device and the fastest
is cached in ``tuned.json`` next to this file.

Cache keys are (N, n, element type, mode) and never include the number of systems in
a batch, so a system's schedule -- and therefore its arithmetic -- is the same alone
and inside any batch.
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


def key(N: int, n: int, dtype, mode: str) -> str:
    t = "f64" if dtype in (torch.float64, np.float64, "float64") else "f32"
    return f"N{N}:n{n}:{t}:{mode}"


def load_cache(path: str = CACHE) -> Dict[str, dict]:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def lookup(N: int, n: int, dtype, mode: str, path: str = CACHE) -> Optional[dict]:
    """Exact (N, n, type, mode) entry, else the entry with the same n/type/mode whose N
    is closest on a log scale (schedules transfer across nearby sizes)."""
    cache = load_cache(path)
    k = key(N, n, dtype, mode)
    if k in cache:
        return cache[k]["config"]
    suffix = k.split(":", 1)[1]
    best, dist = None, None
    for kk, v in cache.items():
        if kk.split(":", 1)[1] != suffix:
            continue
        NN = int(kk.split(":")[0][1:])
        dd = abs(np.log(NN) - np.log(N))
        if dist is None or dd < dist:
            best, dist = v["config"], dd
    return best if dist is not None and dist < np.log(2.0) else None


def candidates(n: int) -> List[dict]:
    waves = [4, 8] if n == 64 else [None]
    out = []
    for w, top, low, base in itertools.product(waves, [2, 3, 4, 6], [1, 2, 3, 4], [2, 4, 8]):
        out.append({"factor_waves": w, "chain_len_levels": [top] + [low] * 15,
                    "base_max": base})
    return out


def _time_config(cfg, systems, dtype, mode, reps, bss):
    from .solver import initialize_batched
    s0 = systems[0]
    N, n, nrhs, S = s0.N, s0.n, s0.nrhs, len(systems)
    data = initialize_batched(N, n, dtype, "cuda", nrhs=nrhs, nsys=S, mode=mode,
                              **cfg)
    for k, sk in enumerate(systems):
        view = (lambda t: t) if S == 1 else (lambda t, k=k: t[..., k])
        view(data.A_tensor).copy_(torch.from_numpy(sk.A).to(dtype))
        view(data.B_tensor).copy_(torch.from_numpy(sk.B).to(dtype))
    d0 = torch.stack([torch.from_numpy(sk.d).to(dtype) for sk in systems], dim=-1).cuda()
    if S == 1:
        d0 = d0[..., 0]
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
    results = []
    for cfg in (configs or candidates(n)):
        c = {k: v for k, v in cfg.items() if v is not None}
        ms = _time_config(c, systems, dtype, mode, reps, bss)
        results.append((ms, c))
        log(f"autotune {key(N, n, dtype, mode)} {c} -> {ms:.3f} ms")
    results.sort(key=lambda t: t[0])
    best_ms, best = results[0]
    entry = {"config": best, "ms": best_ms,
             "default_ms": next((m for m, c in results if c.get("factor_waves", 4) == 4 and
                                 c["chain_len_levels"][0] == 3 and c["chain_len_levels"][1] == 3
                                 and c["base_max"] == 8), None)}
    if save:
        cache = load_cache(path)
        cache[key(N, n, dtype, mode)] = entry
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return entry
