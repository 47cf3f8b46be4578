"""GPU tests of the dense linear-algebra kernels through the C ABI."""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sla
import torch

pytestmark = pytest.mark.gpu


def _lib():
    from updes_amd import _lib as L
    L.require_gpu()
    return L


@pytest.mark.parametrize("trans", [0, 1])
@pytest.mark.parametrize("m,n,lda,alpha,beta", [(1, 1, 1, 1.0, 0.0), (5, 40006, 40006, -1.0, 1.0),
                                                (7, 301, 301, 2.0, 0.5), (1000, 257, 264, 1.0, 0.0),
                                                (3001, 3001, 3008, -1.0, 1.0), (2049, 130, 131, 1.0, 0.0),
                                                (64, 129, 130, 1.0, -1.0)])
def test_gemv_matches_numpy(trans, m, n, lda, alpha, beta):
    """rbfx_gemv_f64, both orientations: the row-per-wave product and the
    transposed product over row chunks (paired 16-byte columns where lda is
    even, single columns otherwise, odd column counts, row counts that are
    not multiples of the four waves), against NumPy in FP64."""
    L = _lib()
    rng = np.random.default_rng(m + 3 * n + lda)
    A = rng.standard_normal((m, lda))
    x = rng.standard_normal(m if trans else n)
    y0 = rng.standard_normal(n if trans else m)
    dA = torch.as_tensor(A).cuda()
    dx = torch.as_tensor(x).cuda()
    dy = torch.as_tensor(y0).cuda()
    ws = L.workspace(L.lib().rbfx_gemv_workspace(trans, m, n), "cuda")
    L.check(L.lib().rbfx_gemv_f64(trans, m, n, alpha, L.ptr(dA), lda, L.ptr(dx), beta, L.ptr(dy), L.ptr(ws),
                                  ws.numel(), L.stream()))
    Am = A[:, :n]
    ref = alpha * ((Am.T @ x) if trans else (Am @ x)) + beta * y0
    got = dy.cpu().numpy()
    scale = np.abs(Am).sum(axis=0 if trans else 1).max() * np.abs(x).max() * abs(alpha) + np.abs(y0).max()
    assert np.abs(got - ref).max() <= 1e-14 * scale * max(1, (m if trans else n) ** 0.5)


@pytest.mark.parametrize("m,n,lda,nrs", [(1, 1, 1, 1), (300, 257, 264, 200), (5000, 301, 301, 5000), (64, 64, 64, 0)])
def test_frobenius_norms_match_numpy(m, n, lda, nrs):
    """rbfx_frobenius_f64 (rows i < n_scaled scaled) and the one-pass pair
    (scaled, unscaled) against NumPy."""
    L = _lib()
    from updes_amd import engine as E
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, lda))
    rs = rng.uniform(0.5, 2.0, nrs)
    dA = torch.as_tensor(A).cuda()
    drs = torch.as_tensor(rs).cuda()
    scale = np.ones(m)
    scale[:nrs] = rs
    ref_s = np.linalg.norm(A[:, :n] * scale[:, None])
    ref = np.linalg.norm(A[:, :n])
    got_s, got = E.frobenius_pair(dA, m, n, drs)
    assert abs(got_s - ref_s) <= 1e-13 * ref_s and abs(got - ref) <= 1e-13 * ref
    assert E.frobenius(dA, m, n, drs) == got_s and E.frobenius(dA, m, n) == got


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("m,n,k,beta", [(1, 1, 1, 1.0), (130, 70, 17, 1.0), (300, 257, 64, 1.0),
                                         (512, 384, 256, 1.0), (300, 258, 64, 1.0), (130, 72, 16, 0.0),
                                         (1000, 520, 512, 1.0), (257, 8, 8, 0.0), (129, 136, 32, 1.0),
                                         (2944, 2952, 64, 1.0)])
def test_gemm_matches_torch(dtype, m, n, k, beta):
    """All GEMM paths: LDS-DMA with 128 x 128 tiles (products with at least
    2 x CU-count such tiles, the last shape) or 64 x 64 tiles (smaller FP64
    products), both needing k a multiple of the k-tile and n of 16 bytes, and
    register-staged (the rest), with partial tiles on every edge."""
    L = _lib()
    g = torch.Generator().manual_seed(m * 7 + n + k)
    A = torch.randn(m, k, generator=g, dtype=torch.float64)
    B = torch.randn(k, n, generator=g, dtype=torch.float64)
    C = torch.randn(m, n, generator=g, dtype=torch.float64)
    ref = beta * C - A @ B
    dA, dB, dC = (t.to(dtype).cuda().contiguous() for t in (A, B, C))
    fn = L.lib().rbfx_gemm_f64 if dtype == torch.float64 else L.lib().rbfx_gemm_f32
    L.check(fn(m, n, k, -1.0, L.ptr(dA), k, L.ptr(dB), n, beta, L.ptr(dC), n, L.stream()))
    torch.cuda.synchronize()
    tol = 1e-12 if dtype == torch.float64 else 2e-4
    err = (dC.double().cpu() - ref).abs().max().item() / max(1.0, ref.abs().max().item())
    assert err < tol, err


@pytest.mark.parametrize("alpha", [0.5, 1.0])
def test_gemm_alpha(alpha):
    """alpha = +-1 takes the LDS-DMA kernel, other values the register-staged one."""
    L = _lib()
    m, n, k = 260, 136, 64
    g = torch.Generator().manual_seed(5)
    A, B, C = (torch.randn(*sh, generator=g, dtype=torch.float64) for sh in ((m, k), (k, n), (m, n)))
    ref = C + alpha * (A @ B)
    dA, dB, dC = (t.cuda().contiguous() for t in (A, B, C))
    L.check(L.lib().rbfx_gemm_f64(m, n, k, alpha, L.ptr(dA), k, L.ptr(dB), n, 1.0, L.ptr(dC), n, L.stream()))
    torch.cuda.synchronize()
    assert (dC.cpu() - ref).abs().max().item() < 1e-12 * ref.abs().max().item()


def _random_well(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    return a


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("n", [1, 7, 64, 200, 600, 1301])
def test_getrf_getrs(dtype, n):
    from updes_amd import engine as E
    _lib()
    a = _random_well(n, n)
    ld = E._ld(n)
    m = torch.zeros((n, ld), dtype=dtype)
    m[:, :n] = torch.as_tensor(a).to(dtype)
    m = m.cuda()
    lu = E.LU(m, n)
    assert lu.singular == 0
    ipiv = lu.ipiv.cpu().numpy()
    LU = lu.a[:, :n].double().cpu().numpy()
    Lm = np.tril(LU, -1) + np.eye(n)
    Um = np.triu(LU)
    pa = a.astype(np.float64 if dtype == torch.float64 else np.float32).astype(np.float64).copy()
    for i in range(n):
        p = ipiv[i]
        assert i <= p < n
        pa[[i, p]] = pa[[p, i]]
    rel = np.abs(Lm @ Um - pa).max() / np.abs(a).max()
    tol = 1e-13 * n if dtype == torch.float64 else 1e-6 * n
    assert rel < tol, rel
    assert np.abs(np.tril(LU, -1)).max() <= 1.0 + 1e-12   # partial pivoting bound
    b = np.random.default_rng(1).standard_normal(n)
    for trans in (False, True):
        x = lu.solve(torch.as_tensor(b).cuda(), trans=trans).cpu().numpy()
        aa = a if dtype == torch.float64 else a.astype(np.float32).astype(np.float64)
        res = (aa.T if trans else aa) @ x - b
        bound = (1e-11 if dtype == torch.float64 else 1e-3) * np.linalg.norm(aa, 1) * np.linalg.norm(x, 1) * max(1, n / 100)
        assert np.abs(res).max() < bound + 1e-12, (trans, np.abs(res).max(), bound)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("n", [511, 1301, 2100, 4099])
def test_getrs_applies_stored_factors(dtype, n):
    """The superblock sweeps apply the stored factors (FP32 ones included) in
    FP64: x matches an FP64 substitution with the same factors (SciPy) to
    near machine precision, for ragged last superblocks and empty chunks."""
    import scipy.linalg as sl
    from updes_amd import engine as E
    _lib()
    a = _random_well(n, 7 * n)
    m = torch.zeros((n, E._ld(n)), dtype=dtype)
    m[:, :n] = torch.as_tensor(a).to(dtype)
    lu = E.LU(m.cuda(), n)
    LU = lu.a[:, :n].double().cpu().numpy()
    piv = lu.ipiv.cpu().numpy()
    b = np.random.default_rng(n).standard_normal(n)
    for trans in (False, True):
        x = lu.solve(torch.as_tensor(b).cuda(), trans=trans).cpu().numpy()
        ref = sl.lu_solve((LU, piv), b, trans=1 if trans else 0)
        rel = np.linalg.norm(x - ref) / np.linalg.norm(ref)
        assert rel < 1e-11, (trans, rel)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("n,groups,narrow", [(500, 1, 0), (700, 1, 0), (1301, 2, 0), (1301, 3, 0),
                                             (1301, 0, 1), (700, 1, 1), (1301, 3, 600)])
def test_getrf_multirow_leaves(dtype, n, groups, narrow):
    """Leaves with two or three rows per thread (panel launches capped to few
    workgroups) and 16-wide leaves (narrow: panels of at least that many rows;
    0 = default, none at these sizes): same pivots as LAPACK and a small
    P A - L U residual."""
    from updes_amd import engine as E
    L = _lib()
    a = _random_well(n, 11)
    ld = E._ld(n)
    m = torch.zeros((n, ld), dtype=dtype)
    m[:, :n] = torch.as_tensor(a).to(dtype)
    L.check(L.lib().rbfx_set_panel_groups(groups))
    L.check(L.lib().rbfx_set_narrow_leaf_rows(narrow))
    try:
        lu = E.LU(m.cuda(), n)
        torch.cuda.synchronize()
    finally:
        L.check(L.lib().rbfx_set_panel_groups(0))
        L.check(L.lib().rbfx_set_narrow_leaf_rows(0))
    assert lu.singular == 0
    ipiv = lu.ipiv.cpu().numpy()
    if dtype == torch.float64:
        _, piv_ref = sla.lu_factor(a)
        assert np.array_equal(ipiv, piv_ref)
    LU = lu.a[:, :n].double().cpu().numpy()
    pa = a.astype(np.float64 if dtype == torch.float64 else np.float32).astype(np.float64).copy()
    for i in range(n):
        pa[[i, ipiv[i]]] = pa[[ipiv[i], i]]
    rel = np.abs((np.tril(LU, -1) + np.eye(n)) @ np.triu(LU) - pa).max() / np.abs(a).max()
    assert rel < (1e-13 if dtype == torch.float64 else 1e-6) * n, rel


def test_getrf_f32_leaf_geometry_does_not_change_the_factors():
    """The FP32 leaves run on at most a quarter of the CUs by default (two or
    three rows per thread on tall panels); the factors and pivots are the
    same, bit for bit, as with one row per thread on up to 256 workgroups:
    every row's arithmetic is the same whichever workgroup holds it."""
    from updes_amd import engine as E
    L = _lib()
    n = 17000  # the first panels have more than 16,384 rows: two rows per thread
    g = torch.Generator().manual_seed(17)
    a = torch.randn(n, n, generator=g, dtype=torch.float32) + 4.0 * torch.eye(n, dtype=torch.float32)
    m = torch.zeros((n, E._ld(n)), dtype=torch.float32)
    m[:, :n] = a
    default = E.LU(m.cuda(), n)
    torch.cuda.synchronize()
    L.check(L.lib().rbfx_set_panel_groups(256))
    try:
        wide = E.LU(m.cuda(), n)
        torch.cuda.synchronize()
    finally:
        L.check(L.lib().rbfx_set_panel_groups(0))
    assert default.singular == 0 and wide.singular == 0
    assert torch.equal(default.ipiv, wide.ipiv)
    assert torch.equal(default.a[:, :n], wide.a[:, :n])


def test_getrf_matches_lapack_pivots():
    """Same pivot sequence as LAPACK getrf on a matrix without ties."""
    from updes_amd import engine as E
    _lib()
    n = 300
    a = _random_well(n, 5)
    ld = E._ld(n)
    m = torch.zeros((n, ld), dtype=torch.float64)
    m[:, :n] = torch.as_tensor(a)
    lu = E.LU(m.cuda(), n)
    lu_ref, piv_ref = sla.lu_factor(a)
    assert np.array_equal(lu.ipiv.cpu().numpy(), piv_ref)
    assert np.abs(lu.a[:, :n].cpu().numpy() - lu_ref).max() < 1e-10


def test_getrf_pivot_ties_take_the_lowest_row():
    """The pivot rule in its documented form: partial pivoting on |value|
    with the 20 low mantissa bits of FP64 values cleared, ties to the lowest
    row.  An exact tie and a near-tie (relative difference 2^-40, which
    LAPACK resolves to the larger value) both go to the lower row; the
    factorisation stays backward stable."""
    from updes_amd import engine as E
    _lib()
    n = 300
    for near in (False, True):
        a = _random_well(n, 17)
        a[:, 0] *= 0.4 / np.abs(a[:, 0]).max()
        a[3, 0] = 1.0
        a[5, 0] = 1.0 + (2.0 ** -40 if near else 0.0)
        m = torch.zeros((n, E._ld(n)), dtype=torch.float64)
        m[:, :n] = torch.as_tensor(a)
        lu = E.LU(m.cuda(), n)
        ipiv = lu.ipiv.cpu().numpy()
        assert ipiv[0] == 3, (near, ipiv[0])
        assert sla.lu_factor(a)[1][0] == (5 if near else 3)
        LU = lu.a[:, :n].cpu().numpy()
        Lm = np.tril(LU, -1) + np.eye(n)
        pa = a.copy()
        for i in range(n):
            pa[[i, ipiv[i]]] = pa[[ipiv[i], i]]
        assert np.abs(Lm @ np.triu(LU) - pa).max() < 1e-13 * n * np.abs(a).max()


def test_expired_sweep_wait_fails_fast_and_the_next_solve_recovers():
    """Failure handling of the triangular sweeps, forced from the host: with a
    wait bound of 0 the waits that are not satisfied at once expire, the
    solve's kernels exit, NaN is written, and LU.solve raises RbfxError
    (RBFX_ERR_TIMEOUT) within seconds.  With the default bound restored the
    same factors solve normally (each solve clears its status word)."""
    import time
    from updes_amd import engine as E
    L = _lib()
    lib = L.lib()
    n = 20000
    a = _random_well(n, 11)
    m = torch.zeros((n, E._ld(n)), dtype=torch.float32)
    m[:, :n] = torch.as_tensor(a, dtype=torch.float32)
    lu = E.LU(m.cuda(), n)
    b = torch.as_tensor(np.random.default_rng(3).standard_normal(n), device="cuda")
    L.check(lib.rbfx_set_wait_seconds(0.0))
    try:
        t0 = time.perf_counter()
        with pytest.raises(L.RbfxError) as ei:
            lu.solve(b)
        assert ei.value.status == L.RBFX_ERR_TIMEOUT
        assert time.perf_counter() - t0 < 10.0
    finally:
        L.check(lib.rbfx_set_wait_seconds(-1.0))
    x = lu.solve(b).cpu().numpy()
    aa = a.astype(np.float32).astype(np.float64)
    res = aa @ x - b.cpu().numpy()
    assert np.isfinite(x).all()
    assert np.abs(res).max() < 1e-3 * np.linalg.norm(aa, 1) * np.linalg.norm(x, 1) * n / 100
