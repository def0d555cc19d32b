"""GPU checks below the solver API: the MFMA tile product the kernels are built from,
and the per-block factors one chain produces (compared with a numpy restatement of
the same chain elimination).  These localise a failure that the end-to-end parity
tests would only report as a wrong solution."""
import ctypes
import os

import numpy as np
import pytest
import torch

from blockstructuredsolvers_amd import factorize, initialize_batched, workloads

pytestmark = pytest.mark.gpu

PROBE = os.path.join(os.path.dirname(__file__), "gpu_probe", "libtileprobe.so")


def _probe():
    lib = ctypes.CDLL(PROBE)
    lib.tile_probe.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p, ctypes.c_int,
                                                    ctypes.c_void_p, ctypes.c_int,
                                                    ctypes.c_void_p, ctypes.c_int]
    lib.tile_probe.restype = ctypes.c_int
    return lib


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("ta", [0, 1])
@pytest.mark.parametrize("tb", [0, 1])
@pytest.mark.parametrize("neg", [0, 1])
def test_tile_mma_layout(dtype, ta, tb, neg):
    """D += op(X) op(Y) for one 16x16 tile with K = 48, every transpose combination."""
    K, ld = 48, 53
    g = torch.Generator().manual_seed(ta * 4 + tb * 2 + neg)
    # column-major storage: element (r, c) at c*ld + r; torch (cols, ld) row-major
    X = torch.randn(64, ld, generator=g, dtype=torch.float64)
    Y = torch.randn(64, ld, generator=g, dtype=torch.float64)
    D0 = torch.randn(16, 16, generator=g, dtype=torch.float64)
    Xm = X.t()  # Xm[r, c] = X storage (c, r)
    Ym = Y.t()
    opX = Xm[:K, :16].t() if ta else Xm[:16, :K]
    opY = Ym[:16, :K].t() if tb else Ym[:K, :16]
    ref = D0.t() + (-1 if neg else 1) * (opX @ opY)  # D stored column-major
    Xd, Yd = X.to(dtype).cuda(), Y.to(dtype).cuda()
    Dd = D0.to(dtype).cuda().contiguous()  # D0 storage (c, r) -> element (r, c) = D0[c, r]
    rc = _probe().tile_probe(1 if dtype == torch.float64 else 0, ta, tb, neg, Xd.data_ptr(), ld,
                             Yd.data_ptr(), ld, Dd.data_ptr(), K)
    assert rc == 0
    got = Dd.double().cpu().t()  # element (r, c)
    tol = 1e-12 if dtype == torch.float64 else 1e-4
    assert torch.allclose(got, ref, atol=tol * 10, rtol=tol), (got - ref).abs().max()


def _chain_factors(A, B):
    """Sequential chain elimination restated in numpy: Uinv_i = chol(A_i')^{-1} (upper),
    C_i = Uinv_i^T B_i."""
    n, _, N = A.shape
    Uinv, C = [], []
    Chat = np.zeros((n, n))
    for i in range(N):
        U = np.linalg.cholesky(A[:, :, i] - Chat).T
        Ui = np.linalg.inv(U)
        Uinv.append(Ui)
        if i + 1 < N:
            Ci = Ui.T @ B[:, :, i]
            C.append(Ci)
            Chat = Ci.T @ Ci
    return Uinv, C


@pytest.mark.parametrize("n", [16, 32, 48, 64, 96, 128])
def test_single_chain_factors(n):
    """N <= base size: the whole system is one chain; after factorize (plain FP64) the A
    slots hold Uinv_i (upper 16x16 tiles) and the B slots C_i."""
    N = 4
    s = workloads.spd_system(N, n, nrhs=1, seed=n)
    data = initialize_batched(N, n, torch.float64, "cuda", preserve_input=False)
    data.A_tensor.copy_(torch.from_numpy(s.A))
    data.B_tensor.copy_(torch.from_numpy(s.B))
    factorize(data)
    torch.cuda.synchronize()
    assert data.levels == 1
    Uinv, C = _chain_factors(s.A, s.B)
    A_out = data.A_tensor.cpu().numpy()
    B_out = data.B_tensor.cpu().numpy()
    tiles = np.kron(np.triu(np.ones((n // 16, n // 16))), np.ones((16, 16))).astype(bool)
    for i in range(N):
        err = np.abs(A_out[:, :, i][tiles] - Uinv[i][tiles]).max() / np.abs(Uinv[i]).max()
        assert err < 1e-12, f"block {i}: Uinv error {err:.3e}"
    for i in range(N - 1):
        err = np.abs(B_out[:, :, i] - C[i]).max() / np.abs(C[i]).max()
        assert err < 1e-12, f"block {i}: C error {err:.3e}"


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_update_probe_products(dtype):
    """The update-product probe (the factor kernel's trsm and Schur products on their
    own, used for their MFMA counters) computes [C|W] = Uinv^T [B|Z], S = W^T W,
    Zn = -C^T W and A' = A - C^T C (upper tiles of S and A')."""
    lib = ctypes.CDLL(PROBE)
    lib.update_probe.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 5
    lib.update_probe.restype = ctypes.c_int
    n, nb = 64, 3
    rng = np.random.default_rng(5)
    # (block, col, row) storage: tril of the storage is an upper-triangular matrix
    U = np.tril(rng.standard_normal((nb, n, n)) * 0.1)
    BZ = rng.standard_normal((nb, 2 * n, n))
    A = rng.standard_normal((nb, n, n))
    dev = [torch.from_numpy(x).to(dtype).cuda() for x in (U, BZ, A)]
    out = torch.zeros((nb, 5 * n, n), dtype=dtype, device="cuda")
    rc = lib.update_probe(1 if dtype == torch.float64 else 0, nb, 1, dev[0].data_ptr(),
                          dev[1].data_ptr(), dev[2].data_ptr(), out.data_ptr(),
                          torch.cuda.current_stream().cuda_stream)
    assert rc == 0
    torch.cuda.synchronize()
    got = out.double().cpu().numpy()
    tol = 1e-12 if dtype == torch.float64 else 2e-5
    upper = np.kron(np.triu(np.ones((n // 16, n // 16))), np.ones((16, 16))).astype(bool)
    for b in range(nb):
        # storage (col, row) -> matrices
        Um, BZm, Am = U[b].T, BZ[b].T, A[b].T
        CW = Um.T @ BZm
        C, W = CW[:, :n], CW[:, n:]
        o = got[b]
        scale = np.abs(CW).max()
        assert np.abs(o[:2 * n].T - CW).max() <= tol * scale * n
        S = W.T @ W
        assert np.abs((o[2 * n:3 * n].T - S)[upper]).max() <= tol * np.abs(S).max() * n
        Zn = -C.T @ W
        assert np.abs(o[3 * n:4 * n].T - Zn).max() <= tol * np.abs(Zn).max() * n
        Ap = Am - C.T @ C
        assert np.abs((o[4 * n:].T - Ap)[upper]).max() <= tol * np.abs(Ap).max() * n
