"""Served-stream benchmark: Poisson arrivals of state-space smoother systems (state size
n in {16, 32, 64}, horizon N ~ U[128, 4096], 1..8 right-hand sides per factorisation)
through the request loop, at a sweep of arrival rates.  Requests carry the model and
the device builds their blocks (``--upload-blocks`` sends dense blocks instead).
Prints one JSON line per rate with p50/p99 request latency, p50/p99 solve latency and
throughput.

    python -m blockstructuredsolvers_amd.serve_bench --rates 50,200,800 --count 64
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import torch

from . import serving


def run(rate: float, count: int, mode: str, workers: int, seed: int, N_max: int,
        device_blocks: bool = True, max_group: int = 64, group_plan: bool = False) -> dict:
    """Serve `count` requests arriving at `rate`/s, or all at t = 0 for ``rate <= 0`` (the
    requests of the 3,200/s stream, so that the lines compare like with like); with
    ``group_plan`` also the rate of the same requests run at once as one group plan per
    block size and column class (``serving.group_plan_rate``), the saturation bound of
    the loop."""
    reqs = serving.smoother_requests(count, rate if rate > 0 else 3200.0, N_range=(128, N_max),
                                     seed=seed, device_blocks=device_blocks)
    if rate <= 0:
        for r in reqs:
            r.arrival = 0.0
    srv = serving.Server(mode=mode, workers=workers, max_group=max_group)
    srv.serve(reqs[:min(4, count)], realtime=False)  # warm-up: kernels loaded, pools primed
    srv.stats.update(groups=0, max_group=0, max_inflight=0, max_inflight_bytes=0)
    t0 = time.perf_counter()
    res = srv.serve(reqs)
    wall = time.perf_counter() - t0
    out = serving.summarize(res)
    out.update(rate_per_s=rate, mode=mode, workers=workers, wall_s=wall,
               blocks="device-built from the model" if device_blocks else "uploaded",
               offered_span_s=reqs[-1].arrival, max_inflight=srv.stats["max_inflight"],
               groups=srv.stats["groups"], max_group=srv.stats["max_group"],
               max_inflight_gb=srv.stats["max_inflight_bytes"] / 1e9,
               rhs_total=sum(len(r.rhs) for r in reqs), failed=sum(r.info != 0 for r in res))
    if group_plan:
        out["group_plan_systems_per_s"] = serving.group_plan_rate(srv, reqs)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", default="50,200,800",
                    help="arrival rates per second; 0: every request at t = 0")
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--mode", default="plain", choices=["plain", "mixed"])
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--N-max", type=int, default=4096)
    ap.add_argument("--max-group", type=int, default=64,
                    help="requests coalesced into one group plan at most (1: none)")
    ap.add_argument("--group-plan", action="store_true",
                    help="also time the same requests as one group plan per class")
    ap.add_argument("--upload-blocks", action="store_true",
                    help="send dense A, B with each request instead of the model")
    a = ap.parse_args(argv)
    if not torch.cuda.is_available():
        print("serve_bench needs a GPU", file=sys.stderr)
        return 2
    for r in [float(x) for x in a.rates.split(",")]:
        print(json.dumps(run(r, a.count, a.mode, a.workers, a.seed, a.N_max,
                             not a.upload_blocks, a.max_group, a.group_plan)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
