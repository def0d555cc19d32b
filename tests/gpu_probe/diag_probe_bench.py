"""Driver of the diagonal-tile probe (diag_probe.hip):
`python3 tests/gpu_probe/diag_probe_bench.py` prints, for one and two workgroups per CU
and each kind of work on the other waves, the cycles wave 0 takes per 16x16 diagonal
tile (Cholesky and inverse, FP32) and whether wave 4 shares wave 0's SIMD."""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MODES = {0: "others idle", 1: "MFMA (LDS operands)", 2: "VALU FMAs",
         3: "MFMA off wave 0's SIMD", 4: "MFMA (register operands)", 5: "LDS reads"}


def main() -> int:
    lib = ctypes.CDLL(os.path.join(HERE, "libtileprobe.so"))
    lib.diag_probe.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    lib.diag_probe.restype = ctypes.c_int
    reps = 64
    print("| workgroups per CU | other waves | cycles per diagonal tile | wave 4 on wave 0's SIMD |")
    print("|---|---|---|---|")
    for grid in (256, 512):
        for mode in MODES:
            buf = np.zeros(grid * 16, dtype=np.uint64)
            rc = lib.diag_probe(mode, grid, reps, buf.ctypes.data_as(ctypes.c_void_p))
            if rc != 0:
                print(f"diag_probe failed: {rc}", file=sys.stderr)
                return 1
            buf = buf.reshape(grid, 8, 2).astype(np.float64)
            per_tile = buf[:, 0, 0] / reps
            same = np.mean(buf[:, 4, 1] == buf[:, 0, 1])
            print(f"| {grid // 256} | {MODES[mode]} | {per_tile.mean():.0f} | {same:.2f} |", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
