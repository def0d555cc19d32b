"""Driver of the update-product probe (update_probe.hip) for MFMA counter runs:
`python3 tests/gpu_probe/update_probe_bench.py --blocks 12288 --reps 1 --dtype f64`
launches the probe over that many 64x64 blocks (the interior blocks of one level-0
factor launch of two N = 8192 systems) a few times and prints the event-timed duration
and the executed MFMA rate (224 16x16x16 tile products per block and repetition)."""
import argparse
import ctypes
import os
import sys

import torch

HERE = os.path.dirname(os.path.abspath(__file__))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--blocks", type=int, default=12288)
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32"])
    ap.add_argument("--launches", type=int, default=3)
    a = ap.parse_args()
    lib = ctypes.CDLL(os.path.join(HERE, "libtileprobe.so"))
    lib.update_probe.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 5
    lib.update_probe.restype = ctypes.c_int
    dt = torch.float64 if a.dtype == "f64" else torch.float32
    n, nb = 64, a.blocks
    g = torch.Generator(device="cuda").manual_seed(1)
    # (block, col, row) storage: its tril is an upper-triangular matrix
    U = torch.tril(torch.randn((nb, n, n), generator=g, device="cuda", dtype=dt) * 0.1)
    BZ = torch.randn((nb, 2 * n, n), generator=g, device="cuda", dtype=dt)
    A = torch.randn((nb, n, n), generator=g, device="cuda", dtype=dt)
    out = torch.zeros((nb, 5 * n, n), device="cuda", dtype=dt)
    s = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times = []
    for _ in range(a.launches):
        ev0.record(s)
        rc = lib.update_probe(1 if a.dtype == "f64" else 0, nb, a.reps, U.data_ptr(),
                              BZ.data_ptr(), A.data_ptr(), out.data_ptr(), s.cuda_stream)
        ev1.record(s)
        torch.cuda.synchronize()
        if rc != 0:
            print(f"update_probe failed: {rc}", file=sys.stderr)
            return 1
        times.append(ev0.elapsed_time(ev1))
    ms = min(times)
    flops = 224 * 2 * 16 ** 3 * nb * a.reps
    peak = 78.6 if a.dtype == "f64" else 157.3
    print(f"update_probe {a.dtype} blocks={nb} reps={a.reps}: {ms:.3f} ms, "
          f"{flops / ms / 1e9:.2f} TFLOP/s executed MFMA ({100 * flops / ms / 1e9 / peak:.1f}% "
          f"of the {peak} TFLOP/s peak)", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
