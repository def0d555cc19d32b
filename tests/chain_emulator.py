"""numpy emulation of libbss's level algorithm (test helper, CPU only).

It executes the same recursion as the HIP kernels -- the chain layout is read from
libbss itself (bss_plan_level_chains), and every chain is eliminated with the
one-sided panel scheme of chain_factor.hpp / chain_solve.hpp:

  factor:  Z_1 = B_L^T; per block  A_i -= C_{i-1}^T C_{i-1}; Uinv_i = chol(A_i)^{-1};
           C_i = Uinv_i^T B_i; W_i = Uinv_i^T Z_i; Z_{i+1} = -C_i^T W_i; S_L += W_i^T W_i
           A'_p = A[sep_p] - C_last^T C_last (chain left of p) - S_L (chain right of p)
           B'_{p,p+1} = -W_last^T C_last
  solve:   y_i = Uinv_i^T (d_i - C_{i-1}^T y_{i-1}); d'_p = d[sep_p] - C_last^T y_last - sum W^T y
           x_i = Uinv_i (y_i - C_i x_{i+1} - W_i x_L)

so a CPU test can show the design itself reproduces the oracle (independently of
the device code), for every schedule the library produces.
"""
import ctypes

import numpy as np

from blockstructuredsolvers_amd import _lib as L


def plan_levels(N, n, nrhs=1, schedule=0, chain_len=None, base_max=None):
    """Chain tables of every level of a libbss plan (host-only queries)."""
    p = ctypes.c_void_p()
    L.call("bss_plan_create", N, n, nrhs, 1, L.BSS_F64, L.BSS_MODE_PLAIN, schedule,
           ctypes.byref(p))
    try:
        if chain_len is not None:
            L.call("bss_set_option", p, L.BSS_OPT_CHAIN_LEN, float(chain_len))
        if base_max is not None:
            L.call("bss_set_option", p, L.BSS_OPT_BASE_MAX, float(base_max))
        v = ctypes.c_int64()
        L.call("bss_plan_get", p, L.BSS_Q_N_USE, ctypes.byref(v))
        n_use = v.value
        L.call("bss_plan_get", p, L.BSS_Q_LEVELS, ctypes.byref(v))
        nlev = v.value
        levels = []
        for k in range(nlev):
            L.call("bss_plan_get", p, L.BSS_Q_LEVEL_N + k, ctypes.byref(v))
            NL = v.value
            buf = (ctypes.c_int * (4 * (NL + 2)))()
            nch = L.call("bss_plan_level_chains", p, k, buf, NL + 2)
            chains = [tuple(buf[4 * c:4 * c + 4]) for c in range(nch)]
            levels.append((NL, chains))
        return n_use, levels
    finally:
        L.lib().bss_plan_destroy(p)


def _seps(NL, chains):
    """Separator block positions of a level, in order, from its chain table."""
    seps = []
    for (start, ln, hl, hr) in chains:
        if hr:
            seps.append(start + ln)
    # separators with an empty chain on their left are the starts-1 of chains with hl
    for (start, ln, hl, hr) in chains:
        if hl and (start - 1) not in seps:
            seps.append(start - 1)
    return sorted(set(seps))


def emulate(A, B, d, levels):
    """A: list of n x n, B: list of n x n (len N-1), d: list of n x nrhs; returns x."""
    NL, chains = levels[0]
    n = A[0].shape[0]
    seps = _seps(NL, chains)
    sidx = {s: k for k, s in enumerate(seps)}
    P = len(seps)
    Uinv = [None] * NL
    C = [None] * NL
    W = [None] * NL
    SR = [np.zeros((n, n)) for _ in range(P)]
    SL = [np.zeros((n, n)) for _ in range(P)]
    Bn = [np.zeros((n, n)) for _ in range(max(P - 1, 0))]
    info = []
    for (start, ln, hl, hr) in chains:
        Chat = None
        Z = B[start - 1].T.copy() if hl else None
        S = np.zeros((n, n))
        for i in range(ln):
            blk = start + i
            last = i == ln - 1
            Ai = A[blk] - (Chat if Chat is not None else 0)
            try:
                U = np.linalg.cholesky(Ai).T
            except np.linalg.LinAlgError:
                info.append(blk)
                U = np.eye(n)
            Uinv[blk] = np.linalg.inv(U)
            hasC = (not last) or hr
            if hasC:
                C[blk] = Uinv[blk].T @ B[blk]
                Chat = C[blk].T @ C[blk]
            if hl:
                W[blk] = Uinv[blk].T @ Z
                S += W[blk].T @ W[blk]
                if not last:
                    Z = -C[blk].T @ W[blk]
        last = start + ln - 1
        if hr:
            SR[sidx[start + ln]] = C[last].T @ C[last]
            if hl:
                Bn[sidx[start - 1]] = -W[last].T @ C[last]
        if hl:
            SL[sidx[start - 1]] = S
    # forward
    y = [None] * NL
    rR = [np.zeros_like(d[0]) for _ in range(P)]
    rL = [np.zeros_like(d[0]) for _ in range(P)]
    for (start, ln, hl, hr) in chains:
        acc = np.zeros_like(d[0])
        for i in range(ln):
            blk = start + i
            r = d[blk] - (C[blk - 1].T @ y[blk - 1] if i > 0 else 0)
            y[blk] = Uinv[blk].T @ r
            if hl:
                acc = acc + W[blk].T @ y[blk]
        if hr:
            rR[sidx[start + ln]] = C[start + ln - 1].T @ y[start + ln - 1]
        if hl:
            rL[sidx[start - 1]] = acc
    x = [None] * NL
    if P > 0:
        An = [A[s] - SR[k] - SL[k] for k, s in enumerate(seps)]
        dn = [d[s] - rR[k] - rL[k] for k, s in enumerate(seps)]
        xs = emulate(An, Bn, dn, levels[1:])
        for k, s in enumerate(seps):
            x[s] = xs[k]
    for (start, ln, hl, hr) in chains:
        xn = x[start + ln] if hr else None
        xL = x[start - 1] if hl else None
        for i in range(ln - 1, -1, -1):
            blk = start + i
            t = y[blk].copy()
            if i < ln - 1 or hr:
                t -= C[blk] @ xn
            if hl:
                t -= W[blk] @ xL
            x[blk] = Uinv[blk] @ t
            xn = x[blk]
    return x


def solve_system(A, B, d, **plan_kw):
    """Emulated libbss solve of Fortran arrays A (n,n,N), B (n,n,N-1), d (n,nrhs,N)."""
    n, _, N = A.shape
    nrhs = d.shape[1]
    n_use, levels = plan_levels(N, n, nrhs, **plan_kw)
    Al = [np.eye(n) for _ in range(n_use)]
    Bl = [np.zeros((n, n)) for _ in range(max(n_use - 1, 0))]
    dl = [np.zeros((n, nrhs)) for _ in range(n_use)]
    for i in range(N):
        Al[i] = np.array(A[:, :, i], dtype=np.float64)
        dl[i] = np.array(d[:, :, i], dtype=np.float64)
    for i in range(N - 1):
        Bl[i] = np.array(B[:, :, i], dtype=np.float64)
    xl = emulate(Al, Bl, dl, levels)
    x = np.zeros((n, nrhs, N), order="F")
    for i in range(N):
        x[:, :, i] = xl[i]
    return x


def emulate_info(A, B, levels, orig=None):
    """First failing block (1-based, original numbering) in libbss's recursive
    elimination order: the chains of a level in order, then the separator levels.  A
    failed block continues with U = I (the emulation stays finite); 0 = no failure.
    A, B: lists of n x n blocks of one level; levels as plan_levels returns them."""
    NL, chains = levels[0]
    n = A[0].shape[0]
    orig = list(range(NL)) if orig is None else orig
    seps = _seps(NL, chains)
    sidx = {s: k for k, s in enumerate(seps)}
    P = len(seps)
    SR = [np.zeros((n, n)) for _ in range(P)]
    SL = [np.zeros((n, n)) for _ in range(P)]
    Bn = [np.zeros((n, n)) for _ in range(max(P - 1, 0))]
    fails = []
    for (start, ln, hl, hr) in chains:
        Chat = None
        Z = B[start - 1].T.copy() if hl else None
        S = np.zeros((n, n))
        Cl = Wl = None
        for i in range(ln):
            blk = start + i
            Ai = A[blk] - (Chat if Chat is not None else 0)
            try:
                U = np.linalg.cholesky(Ai).T
            except np.linalg.LinAlgError:
                fails.append(orig[blk] + 1)
                U = np.eye(n)
            Uinv = np.linalg.inv(U)
            if i < ln - 1 or hr:
                Cl = Uinv.T @ B[blk]
                Chat = Cl.T @ Cl
            if hl:
                Wl = Uinv.T @ Z
                S += Wl.T @ Wl
                if i < ln - 1:
                    Z = -Cl.T @ Wl
        if hr:
            SR[sidx[start + ln]] = Cl.T @ Cl
            if hl:
                Bn[sidx[start - 1]] = -Wl.T @ Cl
        if hl:
            SL[sidx[start - 1]] = S
    if P > 0:
        An = [A[s] - SR[k] - SL[k] for k, s in enumerate(seps)]
        sub = emulate_info(An, Bn, levels[1:], [orig[s] for s in seps])
        if sub:
            fails.append(sub)
    return min(fails) if fails else 0


def blocks_of(A, B, n_use=None):
    """Fortran (n, n, N) / (n, n, N-1) arrays -> lists of FP64 blocks, padded to n_use
    blocks with identity diagonal and zero coupling blocks (the reference's padding)."""
    n, _, N = A.shape
    n_use = N if n_use is None else n_use
    Al = [np.array(A[:, :, i], dtype=np.float64) for i in range(N)]
    Bl = [np.array(B[:, :, i], dtype=np.float64) for i in range(B.shape[2])]
    Al += [np.eye(n) for _ in range(n_use - N)]
    Bl += [np.zeros((n, n)) for _ in range(max(n_use - 1, 0) - len(Bl))]
    return Al, Bl


def plant_positions(N, n, schedule=0):
    """Failing-block positions (0-based) that exercise every place of a level of a
    libbss plan: block 0, a chain's first block behind a separator, a block inside that
    chain, the separator itself and the last block."""
    _, levels = plan_levels(N, n, schedule=schedule)
    chains = levels[0][1]
    start, ln, hl, hr = next(c for c in chains if c[2] and c[1] >= 2)
    return {"block0": 0, "chain_first": start, "inside_chain": start + 1,
            "separator": start - 1, "last": N - 1}


def coupled_failure(sysm, k):
    """A copy of `sysm` whose LAPACK-order failure is block k (0-based, k >= 1) although
    A_k alone is SPD: A_{k-1} = A_k = I, B_{k-2} = 0 (when k >= 2) and B_{k-1} = 2 I, so
    the pair [[I, 2I], [2I, I]] is indefinite and the sequential pivot at k is I - 4 I."""
    from blockstructuredsolvers_amd import workloads
    n = sysm.A.shape[0]
    A = sysm.A.copy(order="F")
    B = sysm.B.copy(order="F")
    A[:, :, k - 1] = np.eye(n)
    A[:, :, k] = np.eye(n)
    if k >= 2:
        B[:, :, k - 2] = 0.0
    B[:, :, k - 1] = 2.0 * np.eye(n)
    return workloads.BlockSystem(A, B, sysm.d)
