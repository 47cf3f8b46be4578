"""GPU parity of the FP32 preconditioner's local cardinal basis
(updes_amd/csrc/cardinal.hip, through the C ABI): the coefficient kernel and
the formation of K~ = diag(rs) D C against the NumPy restatement
(tests/_cardinal_ref.py), and flexible GMRES-IR in that basis, forward and
transposed, against the FP64 LU of the oracle's K."""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sla
import torch

from oracle import assembly as OA, cloud as OC, rbf as R
from tests import _cardinal_ref as CR
from tests import _cases as C

pytestmark = pytest.mark.gpu


def _setup(kind, param, M, nx=12, ny=11, facets=None, seed=2, k=20):
    import updes_amd as u
    from updes_amd import engine as E
    facets = facets or C.FACET_SETS[1]
    J = C.jitter(nx, ny, 0.2, seed)
    oc = OC.SquareCloud(nx, ny, facets, jitter=J)
    uc = u.SquareCloud(Nx=nx, Ny=ny, facet_types=facets, jitter=J)
    desc = E.OpDesc.builtin(3)  # Laplacian rows
    prob = E.Problem(E.device_cloud(uc), kind, param, M, desc)
    basis = E.CardinalBasis(prob, k)
    return oc, uc, prob, basis


@pytest.mark.parametrize("kind,param", C.KINDS)
@pytest.mark.parametrize("M", [3, 6])
@pytest.mark.parametrize("k", [20, 15])
def test_cardinal_coefficients_match_numpy(kind, param, M, k):
    """Local cardinal coefficients against NumPy, for local systems of more
    than 32 rows (k = 20: 29 stencil entries with the special points) and of
    at most 32 (k = 15: 24 entries), which use a smaller LDS layout."""
    oc, uc, prob, basis = _setup(kind, param, M, k=k)
    st = basis.stencil.cpu().numpy().reshape(oc.N, basis.k)
    beta_ref, alpha_ref = CR.coefficients(oc.sorted_nodes, st, kind, param, M)
    beta = basis.beta.cpu().numpy().reshape(oc.N, basis.k)
    alpha = basis.alpha.cpu().numpy()[:oc.N * M].reshape(oc.N, M)
    sb = np.abs(beta_ref).max(axis=1, keepdims=True)
    assert (np.abs(beta - beta_ref) / sb).max() <= 1e-8
    sa = np.maximum(np.abs(alpha_ref).max(axis=1, keepdims=True), 1e-300)
    assert (np.abs(alpha - alpha_ref) / sa).max() <= 1e-8


def _widen_unions(basis, N):
    """Union tables whose every block lists all N node ids (a valid superset
    of each block's stencil ids, position = id), so that no block fits the
    LDS staging and the kernels take their global-memory paths."""
    ncb = -(-N // 256)
    st = basis.stencil.reshape(N, basis.k)
    ptr = torch.arange(ncb + 1, dtype=torch.int64) * N
    idx = torch.arange(N, dtype=torch.int32).repeat(ncb)
    basis.union = (ptr.cuda(), idx.cuda(), st.to(torch.int16).contiguous())


@pytest.mark.parametrize("kind,param,nx,union,k,M", [
    (R.POLYHARMONIC, 1.0, 12, True, 20, 6), (R.GAUSSIAN, 5.0, 12, True, 20, 6), (R.THIN_PLATE, 1.0, 12, True, 20, 6),
    (R.POLYHARMONIC, 1.0, 12, False, 20, 6), (R.POLYHARMONIC, 1.0, 40, True, 20, 6),
    (R.POLYHARMONIC, 1.0, 40, True, 15, 6), (R.GAUSSIAN, 5.0, 12, True, 15, 3), (R.POLYHARMONIC, 1.0, 40, True, 7, 6),
    (R.POLYHARMONIC, 1.0, 40, True, 15, 10), (R.POLYHARMONIC, 1.0, 40, "wide", 15, 6),
    (R.POLYHARMONIC, 1.0, 52, "wide", 20, 6)])
def test_cardinal_form_matches_numpy(kind, param, nx, union, k, M):
    """Every formation kernel: stencils of at most 24 entries with at most 6
    monomials through the LDS ring (producer waves streaming the rows of
    K[:, U], consumer waves contracting them; k = 7 and 15 give 16 and 24
    entries with the 9 special points), larger stencils or more monomials
    through the batched LDS staging (k = 20: 29 entries; M = 10), both over
    several 256-column blocks and row ranges at 40 x 40; blocks whose union
    does not fit LDS ("wide": every id listed) from global memory in either
    kernel; and the gather kernel without union tables."""
    from updes_amd import engine as E, _lib as L
    oc, uc, prob, basis = _setup(kind, param, M, nx=nx, ny=nx - 1, facets=C.FACET_SETS[2], k=k)
    if union == "wide":
        _widen_unions(basis, oc.N)
    else:
        basis.use_union = union
    K64 = E.assemble(prob, L.MAT_K, torch.float64)
    rs = torch.empty(oc.N, dtype=torch.float64, device="cuda")
    Kt = basis.form(K64, rs)[:, :oc.N].cpu().numpy()
    st = basis.stencil.cpu().numpy().reshape(oc.N, basis.k)
    Cm = CR.basis_matrix(st, basis.beta.cpu().numpy().reshape(oc.N, basis.k),
                         basis.alpha.cpu().numpy()[:oc.N * M].reshape(oc.N, M))
    rs_ref, Kt_ref = CR.form(K64[:, :prob.n].cpu().numpy(), Cm, st)
    np.testing.assert_allclose(rs.cpu().numpy(), rs_ref, rtol=1e-11)
    # FP32 output of an FP64 sum: within a couple of FP32 ulps of the row scale
    scale = np.abs(Kt_ref).max(axis=1, keepdims=True)
    assert (np.abs(Kt - Kt_ref) / scale).max() <= 2.5e-7


@pytest.mark.parametrize("trans", [0, 1])
def test_gmres_ir_cardinal_matches_fp64_lu(trans):
    """rbfx_gmres_ir_cardinal on a 30 x 30 jittered Poisson system with a
    Neumann facet (r^3, degree 2): converges without stalling to the FP64
    LU's solution (forward: the nodal values; transposed: the adjoint of q)."""
    from updes_amd import engine as E, _lib as L
    M = 6
    ft = {"North": "d", "South": "n", "East": "d", "West": "d"}
    oc, uc, prob, basis = _setup(R.POLYHARMONIC, 1.0, M, nx=30, ny=30, facets=ft, seed=4)
    fac = E.KFactor(prob, "fp32_ir", basis=basis)
    assert fac.mode == "cardinal"
    K = OA.assemble_K(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, M)
    rng = np.random.default_rng(5 + trans)
    b = np.concatenate([rng.standard_normal(oc.N), np.zeros(M) if not trans else rng.standard_normal(M)])
    ref = sla.lu_solve(sla.lu_factor(K), b, trans=trans)
    x = fac.solve(torch.as_tensor(b, device="cuda"), trans=bool(trans)).cpu().numpy()
    st = fac.last_stats.as_dict()
    assert not st["fell_back"] and st["basis"] == "cardinal", st
    assert st["ir_inner"] <= 40, st
    if not trans:
        # nodal values within the 1e-8 bar of the FP64 LU's (random, rough
        # right-hand side: the worst case for the forward error of two
        # backward-stable solves)
        A = OA.assemble_A(oc, R.POLYHARMONIC, 1.0, M)
        u, u_ref = A[:oc.N] @ x, A[:oc.N] @ ref
        e = np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)
    else:
        # the adjoint of q (the first N entries) is what gradients use
        e = np.linalg.norm(x[:oc.N] - ref[:oc.N]) / np.linalg.norm(ref[:oc.N])
    print("trans %d: rel err vs FP64 LU %.2e, stats %s" % (trans, e, st))
    assert e <= 1e-8, (e, st)
    # backward stable: normwise backward error of the FP64 residual at the
    # refinement's stopping level (2^-54)
    r = b - (K.T @ x if trans else K @ x)
    eta = np.linalg.norm(r) / (np.linalg.norm(K) * np.linalg.norm(x) + np.linalg.norm(b))
    assert eta <= 2.0 ** -53, eta


def test_cardinal_basis_rejects_bad_stencils():
    from updes_amd import engine as E, _lib as L
    oc, uc, prob, basis = _setup(R.POLYHARMONIC, 1.0, 3)
    bad = basis.stencil.clone()
    bad[5] = -3
    c = basis.struct(None)
    c.d_stencil = bad.data_ptr()
    ws = L.workspace(256, "cuda")
    st = L.lib().rbfx_cardinal_coeffs(prob.ref(), ctypes.byref(c), L.ptr(ws), ws.numel(), L.stream())
    assert st == -1 and b"outside" in L.lib().rbfx_last_error()
