"""Headline benchmark: systems/s of the mixed-precision refined solver vs the plain FP64
GPU path at normwise backward error <= 1e-12, plus p50/p99 solve latency.

Workload: block-tridiagonal SPD systems of N=8192 blocks of n=64 (A_i = G G^T + n I,
B_i ~ N(0,1), d ~ U(0,1), one seed per system), one right-hand side, the size the
mixed-precision target is quoted on.  A step factorizes and solves `--nsys` resident
systems (default 16, where throughput saturates: 2, 8 and 16 systems per step give
about 1,000, 1,190 and 1,240 mixed systems/s) -- mixed: FP32 MFMA factorisation + FP64
refinement until every system's backward error is <= 1e-12; plain: one FP64
factorize + solve.  Both paths use the autotuned schedule for that batch size.  Inputs are resident
in HBM before timing starts; the plans keep A and B unchanged so every step does the
full work on the same inputs; each step restores the right-hand side (a 4 MB device
copy per system) before solving.

One JSON line on stdout.  Only one GPU is used: --gpus > 1 exits with a message.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_systems(N, n, nrhs, nsys, seed):
    from blockstructuredsolvers_amd import workloads
    return [workloads.spd_system(N, n, nrhs=nrhs, seed=seed + s) for s in range(nsys)]


def upload(data, systems, torch):
    nsys = len(systems)
    for s, sysm in enumerate(systems):
        if nsys == 1:
            data.A_tensor.copy_(torch.from_numpy(sysm.A))
            data.B_tensor.copy_(torch.from_numpy(sysm.B))
            data.d_tensor.copy_(torch.from_numpy(sysm.d))
        else:
            data.A_tensor[..., s].copy_(torch.from_numpy(sysm.A))
            data.B_tensor[..., s].copy_(torch.from_numpy(sysm.B))
            data.d_tensor[..., s].copy_(torch.from_numpy(sysm.d))


def kernel_totals(data):
    """{(kind, level): (count, total_ms)} recorded by the plan's profiler."""
    from blockstructuredsolvers_amd import _lib as L
    out = {}
    for name, kind in L.KERNEL_KINDS.items():
        for level in range(data.levels):
            c, t = ctypes.c_int(), ctypes.c_double()
            L.call("bss_profile_read", data._plan, kind, level, ctypes.byref(c), ctypes.byref(t))
            if c.value:
                out[(name, level)] = (c.value, t.value)
    return out


# kernels whose top-level launches are timed with events inside the timed region
REPORTED_KINDS = ("factor", "forward", "backward", "residual")


def profile(data, on, kinds=None, max_level=None):
    """Switch the plan's event profiler; `kinds` / `max_level` restrict it to the launches
    of those kernel kinds at levels <= max_level (None: all)."""
    from blockstructuredsolvers_amd import _lib as L
    mask = (1 << len(L.KERNEL_KINDS)) - 1 if kinds is None else \
        sum(1 << L.KERNEL_KINDS[k] for k in kinds)
    L.call("bss_set_option", data._plan, L.BSS_OPT_PROFILE_KINDS, float(mask))
    L.call("bss_set_option", data._plan, L.BSS_OPT_PROFILE_MAX_LEVEL,
           float(1 << 30 if max_level is None else max_level))
    L.call("bss_profile_enable", data._plan, 1 if on else 0)


def run_path(mode, systems, args, torch, bss):
    """Build a plan, warm up, time `steps` steps; returns a result dict."""
    N, n, nrhs, nsys = args.N, args.n, args.nrhs, len(systems)
    from blockstructuredsolvers_amd import _lib as L
    # plain: factors to the workspace so A, B stay resident inputs across steps
    data = bss.initialize_batched(N, n, torch.float64, "cuda", nrhs=nrhs, nsys=nsys, mode=mode,
                                  schedule=args.schedule, chain_len=args.chain_len,
                                  base_max=args.base_max,
                                  tol=args.tol if mode == "mixed" else None,
                                  chain_len_levels=args.chain_lens, factor_waves=args.factor_waves,
                                  graph=bool(args.graph),
                                  pivot_tol=args.pivot_tol if mode == "mixed" else None)
    upload(data, systems, torch)
    d0 = data._d.clone()
    torch.cuda.synchronize()

    def step():
        data._d.copy_(d0)
        bss.factorize(data)
        bss.solve(data)

    # the warmup steps find the kernel kind with the largest total time at level 0; in
    # the timed region HIP events bracket only its top-level launches (every event pair
    # adds stream time between launches)
    if args.warmup:
        profile(data, True, kinds=REPORTED_KINDS, max_level=0)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    timed_kinds = REPORTED_KINDS
    kw = kernel_totals(data) if args.warmup else {}
    if kw:
        timed_kinds = (max(kw.items(), key=lambda kv: kv[1][1])[0][0],)
    profile(data, True, kinds=timed_kinds, max_level=0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kt = kernel_totals(data)
    # per-kernel breakdown of every launch: a few more steps after the timed region
    profile(data, True)
    breakdown_steps = min(args.steps, 3)
    for _ in range(breakdown_steps):
        step()
    torch.cuda.synchronize()
    kt_all = kernel_totals(data)
    profile(data, False)
    iters = data.iterations if mode == "mixed" else 0
    # mixed, one right-hand side: bss_factorize runs the first correction's level-0
    # forward sweep inside its level-0 factor kernel (BSS_OPT_FACTOR_SWEEP, on by default;
    # d is restored before each factorisation, so solve finds it unchanged)
    fused_l0 = mode == "mixed" and nrhs == 1 and n in (16, 32, 48, 64, 96)
    berr_dev = data.backward_error().max() if mode == "mixed" else None
    x0 = (data.d_tensor if nsys == 1 else data.d_tensor[..., 0]).cpu().numpy()

    # the same steps through the combined entry point (first solve's forward sweeps
    # beside the factorisation's small levels); its answer must be bitwise the separate
    # calls'.  Both are timed here with the profiler off (no events at all), so the
    # comparison is like for like.
    def step_combined():
        data._d.copy_(d0)
        bss.factorize_solve(data)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    wall_s = timed(step)
    wall_c = timed(step_combined)
    xc = (data.d_tensor if nsys == 1 else data.d_tensor[..., 0]).cpu().numpy()
    combined = {"systems_per_s": nsys * args.steps / wall_c,
                "ms_per_step": 1e3 * wall_c / args.steps,
                "separate_calls_no_events_systems_per_s": nsys * args.steps / wall_s,
                "bitwise_equal_to_separate_calls": bool(np.array_equal(xc, x0))}

    # latency: repeated solves with resident factors (new RHS each time)
    lat = []
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for k in range(args.lat_solves):
        data._d.copy_(d0)
        torch.cuda.synchronize()
        ev0.record()
        bss.solve(data)
        ev1.record()
        torch.cuda.synchronize()
        lat.append(ev0.elapsed_time(ev1))
    res = {
        "mode": mode,
        "wall_s": wall,
        "ms_per_step": 1e3 * wall / args.steps,
        "systems_per_s": nsys * args.steps / wall,
        "iterations": iters,
        "device_berr": berr_dev,
        "solve_ms_p50": float(np.percentile(lat, 50)) if lat else None,
        "solve_ms_p99": float(np.percentile(lat, 99)) if lat else None,
        "kernels": kt,
        "kernels_all": kt_all,
        "breakdown_steps": breakdown_steps,
        "fused_l0": fused_l0,
        "levels": data.level_sizes(),
        "x0": x0,
        "combined": combined,
    }
    data.close()
    return res


def run_pipelined(mode, systems, args, torch, bss, combined=False):
    """Throughput with consecutive steps alternating between two plans on two HIP
    streams, each driven by its own host thread (the mixed loop's host waits are per
    stream), so one step's latency-bound small levels overlap the next step's top-level
    kernels.  Same per-step work as run_path; no events.  With `combined`, each step is
    one factorize_solve call.  Returns systems/s."""
    import threading
    N, n, nrhs, nsys = args.N, args.n, args.nrhs, len(systems)
    plans = []
    for _ in range(2):
        data = bss.initialize_batched(N, n, torch.float64, "cuda", nrhs=nrhs, nsys=nsys,
                                      mode=mode, schedule=args.schedule,
                                      chain_len=args.chain_len, base_max=args.base_max,
                                      tol=args.tol if mode == "mixed" else None,
                                      chain_len_levels=args.chain_lens,
                                      factor_waves=args.factor_waves)
        upload(data, systems, torch)
        plans.append((data, data._d.clone(), torch.cuda.Stream()))
    torch.cuda.synchronize()

    def worker(k, steps):
        data, d0, st = plans[k]
        with torch.cuda.stream(st):
            for _ in range(steps):
                data._d.copy_(d0)
                if combined:
                    bss.factorize_solve(data)
                else:
                    bss.factorize(data)
                    bss.solve(data)
            st.synchronize()

    def run(steps_each):
        ths = [threading.Thread(target=worker, args=(k, steps_each)) for k in range(2)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()

    run(max(1, args.warmup))
    torch.cuda.synchronize()
    each = max(1, args.steps // 2)
    t0 = time.perf_counter()
    run(each)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    for data, _, _ in plans:
        data.close()
    return nsys * 2 * each / wall


def level0_blocks(N, levels):
    """interior (factored) blocks of the top level = N - separators"""
    return N - (levels[1] if len(levels) > 1 else 0)


def roofline_for(res, args, nsys):
    """Roofline entry for the kernel with the largest total time in the timed region."""
    kt = res["kernels"]
    (name, level), (count, total) = max(kt.items(), key=lambda kv: kv[1][1])
    N, n, nrhs = args.N, args.n, args.nrhs
    mixed = res["mode"] == "mixed"
    sf = 4 if mixed else 8           # factor element size
    NL = res["levels"][level]
    P = res["levels"][level + 1] if level + 1 < len(res["levels"]) else 0
    blocks = (NL - P) * nsys          # blocks eliminated by one launch of this level
    fused = bool(res.get("fused_l0")) and level == 0
    if name == "forward" and fused:
        # the first correction's level-0 sweep ran inside the factorisation: its launch
        # leaves at once (one per step), so the rate is over the launches that swept
        count = max(count - res.get("profiled_steps", 0), 1)
    avg_ms = total / count
    if name == "factor":
        s_in = 8 if level == 0 else sf
        # algorithmic: read A, B; write U, C (4 N n^2 s, with input/factor sizes); with the
        # fused level-0 forward sweep also read r and write y (2 n nrhs sf)
        bytes_per = 2 * n * n * s_in + 2 * n * n * sf + (2 * n * nrhs * sf if fused else 0)
        flops_per = 7.0 / 3.0 * n ** 3
        alg_bytes = blocks * bytes_per
        alg_flops = blocks * flops_per
        peak_tf = FP32_PEAK_TFLOPS if mixed else FP64_PEAK_TFLOPS
        intensity = flops_per / bytes_per
        ridge = peak_tf * 1e12 / (HBM_PEAK_GBS * 1e9)
        if intensity >= ridge:
            ach = alg_flops / (avg_ms * 1e-3) / 1e12
            return dict(kernel=f"chain_factor L{level}", bound="mfma", achieved=ach, peak=peak_tf,
                        unit="TFLOP/s", frac=ach / peak_tf, traffic=None,
                        per_launch_ms=avg_ms, launches=count)
    elif name in ("forward", "backward"):
        # read Uinv (upper triangle), C, W and the vectors; write the vectors
        tri = n * (n + 1) // 2
        bytes_per = (tri + 2 * n * n) * sf + 2 * n * nrhs * sf
        alg_bytes = blocks * bytes_per
    elif name == "residual":
        bytes_per = 2 * n * n * 8 + n * nrhs * (8 * 3 + 4 * 2)
        alg_bytes = NL * nsys * bytes_per
    else:
        alg_bytes = 0
    ach = alg_bytes / (avg_ms * 1e-3) / 1e9
    out = dict(kernel=f"{name} L{level}", bound="hbm", achieved=ach, peak=HBM_PEAK_GBS,
               unit="GB/s", frac=ach / HBM_PEAK_GBS, traffic=None, per_launch_ms=avg_ms,
               launches=count)
    if name == "factor" and fused:
        out["with_forward_sweep"] = True  # the level-0 kernel also runs the first sweep
    return out


def factor_gflops(res, args):
    """Algorithmic factor rate: (7/3) N n^3 flops per system over the summed time of the
    factorisation kernels (all levels, reduced-diagonal assembly included), from the
    breakdown steps."""
    kt = res["kernels_all"]
    steps_ms = sum(t for (k, _), (c, t) in kt.items() if k in ("factor", "assemble_diag"))
    flops = 7.0 / 3.0 * args.N * args.n ** 3 * args.nsys * res["breakdown_steps"]
    return flops / (steps_ms * 1e-3) / 1e9 if steps_ms > 0 else None


def kernel_rooflines(res, args, nsys):
    """Top-level kernels of one path against the HBM roof (algorithmic bytes), timed in
    the breakdown steps (events around every launch)."""
    out = {}
    for (name, level) in sorted(res["kernels_all"]):
        if level != 0 or name not in ("factor", "forward", "backward", "residual"):
            continue
        sub = dict(res)
        sub["kernels"] = {(name, level): res["kernels_all"][(name, level)]}
        sub["profiled_steps"] = res["breakdown_steps"]
        r = roofline_for(sub, args, nsys)
        out[name] = {"per_launch_ms": r["per_launch_ms"], "achieved": r["achieved"],
                     "unit": r["unit"], "frac": r["frac"]}
    return out


def alpha_point(alpha, args, torch, bss, oracle):
    """Mixed path on one system with A_i = G G^T + alpha n I: iterations to the backward-
    error target and the error against the FP64 oracle.  The literal generator with
    B ~ 0.5 N(0,1) is not SPD below alpha ~ 1e-1 at n = 64, so the SPD-by-construction
    variant (block-bidiagonal R^T R + alpha n I) is used at every alpha."""
    from blockstructuredsolvers_amd import workloads
    sysm = workloads.spd_bidiag_system(args.N, args.n, nrhs=args.nrhs, seed=args.seed, alpha=alpha)
    data = bss.initialize_batched(args.N, args.n, torch.float64, "cuda", nrhs=args.nrhs,
                                  mode="mixed", tol=args.tol, max_iters=30)
    upload(data, [sysm], torch)
    d0 = data._d.clone()
    ts = []
    for k in range(4):
        data._d.copy_(d0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bss.factorize(data)
        bss.solve(data)
        torch.cuda.synchronize()
        if k:
            ts.append(time.perf_counter() - t0)
    x = data.d_tensor.cpu().numpy()
    out = {"alpha": alpha, "iterations": data.iterations,
           "refactored_levels": data.refactored,
           "device_berr": float(data.backward_error().max()),
           "systems_per_s": 1.0 / float(np.median(ts))}
    data.close()
    if not args.no_oracle:
        xo, info = oracle.solve_system(sysm.A, sysm.B, sysm.d)
        out["oracle_berr"] = float(oracle.backward_error(sysm.A, sysm.B, x, sysm.d).max())
        out["rel_diff_vs_oracle"] = float(np.abs(x - xo).max() / np.abs(xo).max())
    return out


def batch_line(args, torch, oracle):
    """Uneven batch (512 systems, N_k log-uniform in [16, 4096], n_k uniform over
    {16, 32, 48, 64, 96, 128}, generated on the device): one group plan per n, each on
    its own stream; a step factors and solves every system.  Plain FP64 and mixed.  Two
    systems per n (the smallest and one of median size up to 2048 blocks) get their
    backward error from the FP64 oracle, as the checker."""
    from blockstructuredsolvers_amd import batch_bench
    Ns, ns = batch_bench.draw(args.batch_systems, 7)
    groups = [(n, [N for N, m in zip(Ns, ns) if m == n]) for n in batch_bench.SIZES]
    groups = [(n, g) for n, g in groups if g]
    pick = []
    for gi, (n, g) in enumerate(groups):
        order = np.argsort(g)
        small = [int(s) for s in order if g[s] <= 2048]
        pick += [(gi, s) for s in sorted({int(order[0]), small[len(small) // 2]})]
    out = {"systems": args.batch_systems, "blocks_total": int(sum(Ns)),
           "systems_per_n": {str(n): len(g) for n, g in groups},
           "bytes_A_B_fp64": int(sum(8 * n * n * (2 * N - 1) for N, n in zip(Ns, ns)))}
    for mode in ("plain", "mixed"):
        plans, ms, samples = batch_bench.run(mode, groups, steps=args.batch_steps, warmup=1,
                                             seed=7, check=0, picks=pick)
        worst = 0.0
        for gi, s, d0 in samples:
            data = plans[gi]
            A = np.asfortranarray(data.A_tensors[s].double().cpu().numpy())
            B = np.asfortranarray(data.B_tensors[s].double().cpu().numpy())
            x = data.d_tensors[s].double().cpu().numpy()
            worst = max(worst, float(oracle.backward_error(A, B, x, np.asfortranarray(d0)).max()))
        out[mode] = {"ms_per_step": ms, "systems_per_s": args.batch_systems / ms * 1e3,
                     "checked_systems": len(samples),
                     "max_backward_error_oracle": worst}
        if mode == "mixed":
            out[mode]["iterations_per_n"] = {str(n): int(d.iterations)
                                             for (n, _), d in zip(groups, plans)}
        for d in plans:
            d.close()
        del plans
        torch.cuda.empty_cache()
        # the same batch through one uneven-batch object: one call (one graph launch for
        # the plain path) per step for all block sizes
        data, ms_u, samples_u = batch_bench.run_unified(mode, groups, steps=args.batch_steps,
                                                        warmup=1, seed=7, picks=pick)
        worst_u = 0.0
        for s, d0 in samples_u:
            A = np.asfortranarray(data.A_tensors[s].double().cpu().numpy())
            B = np.asfortranarray(data.B_tensors[s].double().cpu().numpy())
            x = data.d_tensors[s].double().cpu().numpy()
            worst_u = max(worst_u,
                          float(oracle.backward_error(A, B, x, np.asfortranarray(d0)).max()))
        out[mode]["one_batch_call"] = {"ms_per_step": ms_u,
                                       "systems_per_s": args.batch_systems / ms_u * 1e3,
                                       "max_backward_error_oracle": worst_u}
        data.close()
        del data
        torch.cuda.empty_cache()
        log(f"bench: batch {mode}: {out[mode]}")
    out["mixed_over_plain"] = out["mixed"]["systems_per_s"] / out["plain"]["systems_per_s"]
    return out


def serve_lines(args):
    """Served stream (Poisson arrivals of smoother systems, n in {16, 32, 64}, N ~
    U[128, 4096], 1-8 right-hand sides per factorisation, blocks built on the device):
    throughput and p50/p99 latency at a light and a saturating arrival rate and with
    every request present at t = 0 (rate 0), plain and mixed.  The rate-0 line (else the
    highest rate's) also carries the group-plan rate of the same requests (all present
    at once, one group plan per block size and column class)."""
    from blockstructuredsolvers_amd import serve_bench
    out = []
    sat = 0.0 if 0.0 in args.serve_rates else max(args.serve_rates)
    for mode in ("plain", "mixed"):
        for rate in args.serve_rates:
            top = rate == sat
            r = serve_bench.run(rate, args.serve_count, mode, workers=8, seed=0, N_max=4096,
                                group_plan=top)
            keep = {k: r[k] for k in ("mode", "rate_per_s", "requests", "systems_per_s",
                                      "latency_ms_p50", "latency_ms_p99", "solve_ms_p50",
                                      "solve_ms_p99", "queue_ms_p50", "failed", "groups",
                                      "max_group", "group_plan_systems_per_s") if k in r}
            out.append(keep)
            log(f"bench: served {keep}")
    return out


def load_traffic(kernel, mode, args):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by a copy of known size) for this workload, if recorded."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    key = f"{kernel}|{mode}|N{args.N}|n{args.n}|nrhs{args.nrhs}|nsys{args.nsys}"
    entry = table.get(key)
    return entry["bytes"] if entry else None


def cpu_baseline(system, seconds, threads):
    """The FP64 oracle (C restatement of the reference's cpu.jl path) solving copies of
    one system, one system per thread, until `seconds` of wall time have passed."""
    from oracle import oracle
    done = [0] * threads
    stop = time.perf_counter() + seconds
    copies = [(system.A.copy(order="F"), system.B.copy(order="F"), system.d.copy(order="F"))
              for _ in range(threads)]

    def work(k):
        A0, B0, d0 = copies[k]
        while time.perf_counter() < stop:
            A, B, d = A0.copy(order="F"), B0.copy(order="F"), d0.copy(order="F")
            oracle.factorize(A, B)
            oracle.solve(A, B, d)
            done[k] += 1

    t0 = time.perf_counter()
    ths = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    el = time.perf_counter() - t0
    return sum(done) / el, sum(done), el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--N", type=int, default=8192)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--nsys", type=int, default=32)
    ap.add_argument("--tol", type=float, default=1e-12)
    ap.add_argument("--schedule", default="auto", choices=["auto", "reference"])
    ap.add_argument("--chain-len", type=int, default=None)
    ap.add_argument("--base-max", type=int, default=None)
    ap.add_argument("--chain-lens", type=lambda v: [int(x) for x in v.split(",")], default=None,
                    help="per-level chain lengths, top level first")
    ap.add_argument("--factor-waves", type=int, default=None)
    ap.add_argument("--graph", type=int, default=0, help="1: replay hipGraph-captured levels")
    ap.add_argument("--pivot-tol", type=float, default=None,
                    help="mixed: FP32 separator pivot threshold (2.0 forces the FP64 "
                         "separator re-factorisation)")
    ap.add_argument("--lat-solves", type=int, default=30)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (default: every core this process "
                         "may run on)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-direct", type=int, default=1,
                    help="1: time the CPU direct solvers (LAPACK banded Cholesky, SuperLU) "
                         "on one assembled system as context")
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--modes", default="plain,mixed")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--alphas", default="1,1e-2,1e-4",
                    help="conditioning sweep of the mixed path (empty: skip)")
    ap.add_argument("--no-pipelined", action="store_true",
                    help="skip the two-stream pipelined throughput measurement")
    ap.add_argument("--batch-systems", type=int, default=512,
                    help="uneven-batch line: systems per step (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--serve-count", type=int, default=128,
                    help="served-stream lines: requests per rate and mode (0: skip)")
    ap.add_argument("--serve-rates", type=lambda v: [float(x) for x in v.split(",")],
                    default=[200.0, 3200.0, 0.0],
                    help="served-stream arrival rates per second (0: every request at t = 0)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 256 MiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the solver runs on one GPU (no multi-GPU sharding); "
                          f"--gpus must be 1 (got {args.gpus})"}))
        sys.exit(2)

    import torch
    import blockstructuredsolvers_amd as bss
    from oracle import oracle

    log(f"bench: generating {args.nsys} systems N={args.N} n={args.n} nrhs={args.nrhs}")
    systems = make_systems(args.N, args.n, args.nrhs, args.nsys, args.seed)
    results = {}
    for mode in args.modes.split(","):
        log(f"bench: {mode} path")
        results[mode] = run_path(mode, systems, args, torch, bss)
        r = results[mode]
        lat = (f"solve p50 {r['solve_ms_p50']:.3f} ms p99 {r['solve_ms_p99']:.3f} ms"
               if r["solve_ms_p50"] is not None else "no latency solves")
        log(f"bench: {mode}: {r['systems_per_s']:.1f} systems/s, {r['ms_per_step']:.3f} ms/step, "
            f"{lat}, iters {r['iterations']}")
        for (k, lvl), (c, t) in sorted(r["kernels_all"].items(), key=lambda kv: -kv[1][1]):
            log(f"    {k:14s} L{lvl}: {c:4d} launches, {t / c:.4f} ms avg, {t:.3f} ms total")
        for (k, lvl), (c, t) in r["kernels"].items():
            log(f"    timed region (events on this kernel only): {k} L{lvl}: {c} launches, "
                f"{t / c:.4f} ms avg")
        if not args.no_pipelined:
            r["pipelined"] = run_pipelined(mode, systems, args, torch, bss)
            log(f"bench: {mode}: two plans on two streams, steps alternating: "
                f"{r['pipelined']:.1f} systems/s")
            r["pipelined_combined"] = run_pipelined(mode, systems, args, torch, bss, combined=True)
            log(f"bench: {mode}: the same through factorize_solve: "
                f"{r['pipelined_combined']:.1f} systems/s")

    acc = {}
    if not args.no_oracle:
        log("bench: FP64 oracle on system 0")
        s0 = systems[0]
        xo, info = oracle.solve_system(s0.A, s0.B, s0.d)
        for mode, r in results.items():
            x = r["x0"]
            acc[mode] = {
                "rel_diff_vs_oracle": float(np.abs(x - xo).max() / np.abs(xo).max()),
                "backward_error": float(oracle.backward_error(s0.A, s0.B, x, s0.d).max()),
            }
        log(f"bench: accuracy {acc}")

    cpu = None
    if not args.no_cpu:
        # every core this process may run on, within the CPU share the host allots a
        # job (OMP_NUM_THREADS, when set)
        host_cores = len(os.sched_getaffinity(0))
        share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or host_cores
        threads = args.cpu_threads or min(host_cores, share)
        log(f"bench: CPU baseline, {threads} threads ({host_cores} cores in this process's "
            f"affinity mask, CPU share {share}), {args.cpu_seconds}s")
        v, count, el = cpu_baseline(systems[0], args.cpu_seconds, threads)
        cpu = {"value": v, "unit": "systems/s", "cores": threads, "host_cores": host_cores,
               "cpu_share": share, "kind": "port",
               "sample": f"{count} FP64 factorize+solve of one N={args.N}, n={args.n}, "
                         f"nrhs={args.nrhs} system in {el:.1f} s, one system per thread on "
                         f"{threads} of the {host_cores} cores this process may use (CPU "
                         f"share {share}) (C restatement of "
                         "cpu.jl, oracle/bss_oracle.c, unblocked loops, no BLAS)"}

    direct = None
    if not args.no_cpu and args.cpu_direct:
        # the reference's other CPU comparison (benchmark/benchmark_cpu.jl: sparse direct
        # solvers on the assembled matrix), with the direct solvers this image has
        from blockstructuredsolvers_amd import cpu_direct
        log("bench: CPU direct solvers on the assembled matrix")
        direct = cpu_direct.run(args.N, args.n, seed=args.seed)
        log(f"    {direct}")

    if args.calib_copy:
        # a plain copy of known size: HBM counters of the kernels above are scaled by
        # what FETCH_SIZE / WRITE_SIZE read for it
        src = torch.empty(256 << 20, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    sweep = []
    if args.alphas and "mixed" in args.modes.split(","):
        for a in [float(v) for v in args.alphas.split(",") if v]:
            log(f"bench: conditioning sweep alpha={a:g}")
            sweep.append(alpha_point(a, args, torch, bss, oracle))
            log(f"    {sweep[-1]}")

    batch = batch_line(args, torch, oracle) if args.batch_systems else None
    served = serve_lines(args) if args.serve_count else None

    head = results.get("mixed") or results.get("plain")
    plain = results.get("plain")
    roof = roofline_for(dict(head, profiled_steps=args.steps), args, args.nsys)
    traffic = load_traffic(roof["kernel"], head["mode"], args)
    if traffic is not None:
        roof["traffic"] = traffic
    out = {
        "metric": "systems/s, mixed vs plain FP64 at <=1e-12 backward err; p50/p99 solve ms",
        "value": head["systems_per_s"],
        "unit": "systems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factor + f64 refinement" if head["mode"] == "mixed" else "f64",
        "data": "synthetic (seeded A_i = G G^T + n I, B_i ~ N(0,1), d ~ U(0,1))",
        "config": {"workload": f"block-tridiagonal SPD, N={args.N} blocks of n={args.n}, "
                               f"nrhs={args.nrhs}, {args.nsys} systems per step",
                   "N": args.N, "n": args.n, "nrhs": args.nrhs, "systems_per_step": args.nsys,
                   "schedule": args.schedule, "levels": head["levels"], "tol": args.tol},
        "mixed_systems_per_s": results["mixed"]["systems_per_s"] if "mixed" in results else None,
        "plain_fp64_systems_per_s": plain["systems_per_s"] if plain else None,
        "mixed_over_plain": (results["mixed"]["systems_per_s"] / plain["systems_per_s"]
                             if plain and "mixed" in results else None),
        "refinement_iterations": results["mixed"]["iterations"] if "mixed" in results else None,
        "two_stream_pipelined_systems_per_s": {m: r.get("pipelined") for m, r in results.items()},
        "two_stream_pipelined_factorize_solve_systems_per_s":
            {m: r.get("pipelined_combined") for m, r in results.items()},
        "factorize_solve_call": {m: r["combined"] for m, r in results.items()},
        "solve_ms_p50": {m: r["solve_ms_p50"] for m, r in results.items()},
        "solve_ms_p99": {m: r["solve_ms_p99"] for m, r in results.items()},
        "accuracy_vs_fp64_oracle": acc,
        "factor_gflops": {m: factor_gflops(r, args) for m, r in results.items()},
        "kernel_rooflines": {m: kernel_rooflines(r, args, args.nsys) for m, r in results.items()},
        "conditioning_sweep": sweep,
        "uneven_batch": batch,
        "served_stream": served,
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_direct_solvers": direct,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
