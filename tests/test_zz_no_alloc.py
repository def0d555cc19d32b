"""bss_factorize_solve_noalloc (no device allocation; workspace in the library's
per-device scratch array) against the plan path it is a variant of: the same kernels and
schedule, so the answers are bitwise equal, and A, B stay unchanged.  Two calls in a row,
the second on a smaller system (the scratch keeps the first call's contents), for FP64
and FP32; a system larger than the scratch is refused with BSS_E_ARG."""
import ctypes

import numpy as np
import pytest
import torch

from blockstructuredsolvers_amd import _lib as L
from blockstructuredsolvers_amd import factorize, initialize_batched, solve, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


def storage(a):
    """Fortran (n, m, N) block array -> device tensor of (N, m, n) C-order storage, which
    is the reference layout (column-major blocks, contiguous)."""
    return torch.from_numpy(np.ascontiguousarray(np.transpose(a, (2, 1, 0)))).cuda()


def noalloc_call(N, n, nrhs, T, A, B, d):
    code = L.BSS_F64 if T == torch.float64 else L.BSS_F32
    stream = torch.cuda.current_stream().cuda_stream
    return L.lib().bss_factorize_solve_noalloc(N, n, nrhs, code, A.data_ptr(),
                                               B.data_ptr() if B.numel() else None,
                                               d.data_ptr(), stream)


def plan_solve(sysm, T):
    N, n, nrhs = sysm.A.shape[2], sysm.A.shape[0], sysm.d.shape[1]
    data = initialize_batched(N, n, T, "cuda", nrhs=nrhs, tuned=False)
    data.A_tensor.copy_(torch.from_numpy(sysm.A))
    data.B_tensor.copy_(torch.from_numpy(sysm.B))
    data.d_tensor.copy_(torch.from_numpy(sysm.d))
    factorize(data)
    solve(data)
    torch.cuda.synchronize()
    x = data.d_tensor.cpu().numpy()
    data.close()
    return x


@pytest.mark.parametrize("T", [torch.float64, torch.float32])
def test_noalloc_matches_plan_path(T):
    npT = np.float64 if T == torch.float64 else np.float32
    for k, (N, n, nrhs) in enumerate([(500, 32, 2), (97, 16, 1)]):
        s64 = workloads.spd_system(N, n, nrhs=nrhs, seed=800 + k)
        sysm = workloads.BlockSystem(*(np.asfortranarray(a, dtype=npT)
                                       for a in (s64.A, s64.B, s64.d)))
        print(f"noalloc T={T} N={N} n={n} nrhs={nrhs}", flush=True)
        x_plan = plan_solve(sysm, T)
        A, B, d = storage(sysm.A), storage(sysm.B), storage(sysm.d)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        rc = noalloc_call(N, n, nrhs, T, A, B, d)
        assert rc == L.BSS_OK, L.last_error()
        torch.cuda.synchronize()
        if k > 0:  # after the first call on the device, nothing is allocated at all
            assert torch.cuda.mem_get_info()[0] == free0
        x = d.cpu().numpy().transpose(2, 1, 0)
        assert np.array_equal(x, x_plan), (N, n)
        assert np.array_equal(A.cpu().numpy().transpose(2, 1, 0), sysm.A)
        assert np.array_equal(B.cpu().numpy().transpose(2, 1, 0), sysm.B)
        if T == torch.float64:
            assert oracle.backward_error(s64.A, s64.B, x, s64.d).max() <= 1e-13


def test_noalloc_rejects_oversize():
    N, n = 4096, 64
    A = torch.zeros((N, n, n), dtype=torch.float64, device="cuda")
    B = torch.zeros((N - 1, n, n), dtype=torch.float64, device="cuda")
    d = torch.zeros((N, 1, n), dtype=torch.float64, device="cuda")
    rc = noalloc_call(N, n, 1, torch.float64, A, B, d)
    assert rc == L.BSS_E_ARG
    assert "allocation-free" in L.last_error()
    assert torch.count_nonzero(d).item() == 0
