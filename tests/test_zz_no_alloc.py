"""The allocation-free solve (rbfx_getrs_noalloc_f64) through the C ABI
against rbfx_getrs_f64 with a workspace: same kernels, same inputs, so the
results are bit-identical.  Two factorisations in a row, the second smaller,
so scratch contents left by the first call cannot leak into the second."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _factor(n, seed):
    from updes_amd import engine as E
    a = np.random.default_rng(seed).standard_normal((n, n))
    m = torch.zeros((n, E._ld(n)), dtype=torch.float64)
    m[:, :n] = torch.as_tensor(a)
    return E.LU(m.cuda(), n)


def test_noalloc_solve_matches_workspace_solve():
    from updes_amd import _lib as L
    L.require_gpu()
    lib = L.lib()
    for n in (3000, 1301):
        lu = _factor(n, n)
        b = torch.as_tensor(np.random.default_rng(n + 1).standard_normal(n), device="cuda")
        for trans in (False, True):
            ref = lu.solve(b, trans)
            x = b.clone()
            L.check(lib.rbfx_getrs_noalloc_f64(int(trans), n, L.ptr(lu.a), lu.ld, L.ptr(lu.perm), L.ptr(lu.dinv),
                                               L.ptr(x), L.stream()))
            torch.cuda.synchronize()
            assert torch.isfinite(x).all()
            assert torch.equal(x, ref), (n, trans, float((x - ref).abs().max()))


def test_noalloc_rejects_n_beyond_its_scratch():
    from updes_amd import _lib as L
    L.require_gpu()
    lib = L.lib()
    t = torch.zeros(4, dtype=torch.float64, device="cuda")
    n = L.GETRS_NOALLOC_MAX_N + 1
    rc = lib.rbfx_getrs_noalloc_f64(0, n, L.ptr(t), n, L.ptr(t), L.ptr(t), L.ptr(t), L.stream())
    assert rc == L.RBFX_ERR_ARG
    assert b"exceeds its fixed scratch" in lib.rbfx_last_error()
