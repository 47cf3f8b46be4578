"""GPU parity with local supports (``support_size < N``): assembly of K and A,
the matrix-free products, the solve and the adjoint, against the oracle's
masked restatement (updes/assembly.py:31-32, 127-128, 171-258) on the same
neighbour table.  The table is the cloud's own (the reference's BallTree
query, checked against brute force on the CPU), handed to the oracle by
original id so that equidistant neighbours resolve identically."""
import numpy as np
import pytest
import torch

from oracle import assembly as OA, cloud as OC, rbf as R, solver as OS
from tests import _cases as C

pytestmark = pytest.mark.gpu


def _orig_supports(uc):
    tab = uc.sorted_local_supports.numpy()
    return uc._order[tab[uc._new_id]]


def _pair(nx, ny, facets, k, jit=0.2, seed=0):
    import updes_amd as u
    J = C.jitter(nx, ny, jit, seed) if jit else None
    uc = u.SquareCloud(Nx=nx, Ny=ny, facet_types=facets, jitter=J, support_size=k)
    oc = OC.SquareCloud(nx, ny, facets, jitter=J, support_size=k, supports=_orig_supports(uc))
    return uc, oc


def _lap(x, center=None, rbf=None, monomial=None, fields=None):
    import updes_amd as u
    return u.nodal_laplacian(x, center, rbf, monomial)


def _rowscaled_err(a, b):
    scale = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-300)
    return float((np.abs(a - b) / scale).max())


@pytest.mark.parametrize("kind,param", [C.KINDS[0], C.KINDS[2], C.KINDS[3]])
@pytest.mark.parametrize("fs", range(len(C.FACET_SETS)))
@pytest.mark.parametrize("k", [1, 11, 40])
def test_local_assembly_K_matches_oracle(kind, param, fs, k):
    import updes_amd as u
    facets = C.FACET_SETS[fs]
    uc, oc = _pair(9, 8, facets, k, seed=fs)
    M = 6
    betas = {f: np.linspace(0.5, 1.5, len(oc.facet_nodes[f])) for f, t in oc.facet_types.items() if t == "r"}
    rb = OA.robin_coefficients(oc, betas)
    dt, nu, vel = 0.1, 0.3, (1.0, -0.5)
    Kref = OA.assemble_K(C.adv_diff_desc(dt, nu, vel), oc, kind, param, M, rb)
    rc, _ = u.duplicate_robin_coeffs({f: (0.0, betas[f]) if t == "r" else 0.0 for f, t in uc.facet_types.items()}, uc)

    def op(x, center=None, rbf=None, monomial=None, fields=None):
        val = u.nodal_value(x, center, rbf, monomial)
        grad = u.nodal_gradient(x, center, rbf, monomial)
        lap = u.nodal_laplacian(x, center, rbf, monomial)
        return val + dt * torch.dot(torch.tensor(vel, dtype=torch.float64), grad) - dt * nu * lap

    print("local K: kind %d fs %d k %d" % (kind, fs, k), flush=True)
    K = u.assemble_K(op, uc, C.torch_rbf(kind, param), M, None, rc).cpu().numpy()
    assert K.shape == Kref.shape
    assert _rowscaled_err(K, Kref) < 1e-12
    # entries the supports remove are exact zeros
    od = OC.SquareCloud(9, 8, facets, jitter=C.jitter(9, 8, 0.2, fs))
    Kd = OA.assemble_K(C.adv_diff_desc(dt, nu, vel), od, kind, param, M, rb)
    outside = (Kref == 0) & (Kd != 0)
    assert not np.any(K[outside])


@pytest.mark.parametrize("fs", [0, 4])
def test_local_assembly_on_regular_grid_with_ties(fs):
    import updes_amd as u
    uc, oc = _pair(10, 7, C.FACET_SETS[fs], 6, jit=0)
    K = u.assemble_K(_lap, uc, u.polyharmonic, 3).cpu().numpy()
    Kref = OA.assemble_K(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 3)
    assert _rowscaled_err(K, Kref) < 1e-12
    A = u.assemble_A(uc, u.polyharmonic, 3).cpu().numpy()
    assert _rowscaled_err(A, OA.assemble_A(oc, R.POLYHARMONIC, 1.0, 3)) < 1e-12


@pytest.mark.parametrize("kind,param", [C.KINDS[0], C.KINDS[2]])
def test_local_assembly_A_matches_oracle(kind, param):
    import updes_amd as u
    uc, oc = _pair(12, 10, C.README_FACETS, 15)
    A = u.assemble_A(uc, C.torch_rbf(kind, param), 3).cpu().numpy()
    Aref = OA.assemble_A(oc, kind, param, 3)
    assert _rowscaled_err(A, Aref) < 1e-12
    assert not np.any(A[:oc.N, :oc.N][~OA.support_mask(oc)])


@pytest.mark.parametrize("k", [17, "max"])
@pytest.mark.parametrize("fs", [1, 3, 4, 5])
def test_local_matvecs_match_assembled(fs, k):
    """y = K x, K^T x, A x, A^T x recomputed matrix-free with the support test
    (and, with every node as support, A^T x through the symmetric A)."""
    import updes_amd as u
    from updes_amd import _lib as L, engine as E
    uc = u.SquareCloud(Nx=33, Ny=31, facet_types=C.FACET_SETS[fs], jitter=C.jitter(33, 31, 0.2, fs),
                       support_size=k)
    dc = E.device_cloud(uc)
    beta = torch.linspace(0.5, 1.5, dc.Nr, dtype=torch.float64) if dc.Nr else None
    prob = E.Problem(dc, L.POLYHARMONIC, 1.0, 6, E.OpDesc.builtin(L.LAP), beta)
    g = torch.Generator().manual_seed(fs)
    x = torch.randn(prob.n, generator=g, dtype=torch.float64).cuda()
    for mat in (L.MAT_K, L.MAT_A):
        Kd = E.assemble(prob, mat, torch.float64)[:, :prob.n]
        for trans in (False, True):
            y = E.matvec(prob, x, mat, trans=trans)
            ref = (Kd.T if trans else Kd) @ x
            err = (y - ref).abs().max().item() / ref.abs().max().item()
            print("fs %d mat %d trans %d: %.2e" % (fs, mat, trans, err), flush=True)
            assert err < 1e-13


@pytest.mark.parametrize("kind,param", C.KINDS)
@pytest.mark.parametrize("M", [0, 3, 6])
def test_value_matvec_matches_assembled(kind, param, M):
    """A x and A^T x with every node as support (the value-row kernel, whose
    r comes from the square-root sequence without range scaling where the
    wave allows, then the P^T rows) for every kernel family, with and
    without monomials, against the assembled A."""
    import updes_amd as u
    from updes_amd import _lib as L, engine as E
    uc = u.SquareCloud(Nx=29, Ny=27, facet_types=C.FACET_SETS[1], jitter=C.jitter(29, 27, 0.2, M),
                       support_size="max")
    dc = E.device_cloud(uc)
    prob = E.Problem(dc, kind, param, M, E.OpDesc.builtin(L.LAP), None)
    x = torch.randn(prob.n, generator=torch.Generator().manual_seed(M + 7 * kind), dtype=torch.float64).cuda()
    Ad = E.assemble(prob, L.MAT_A, torch.float64)[:, :prob.n]
    for trans in (False, True):
        y = E.matvec(prob, x, L.MAT_A, trans=trans)
        ref = (Ad.T if trans else Ad) @ x
        err = (y - ref).abs().max().item() / ref.abs().max().item()
        assert err < 1e-13, (trans, err)


@pytest.mark.parametrize("fs", [0, 1, 2, 3, 5])
@pytest.mark.parametrize("k", [10, 25])
def test_local_pde_solver_matches_reference_algorithm(fs, k):
    """Solution values against the restated reference B-path with the same
    local supports (inv(A), D inv(A), QR)."""
    import updes_amd as u
    nx = 14
    facets = C.FACET_SETS[fs]
    uc, oc = _pair(nx, nx, facets, k, seed=fs)
    bc_np = {f: (np.zeros(len(ids)) if oc.facet_types[f][0] == "p" else np.cos(np.arange(len(ids))))
             for f, ids in oc.facet_nodes.items()}
    betas = {f: np.ones(len(oc.facet_nodes[f])) for f, t in oc.facet_types.items() if t == "r"}
    q = OA.assemble_q(oc, np.ones(oc.Ni), bc_np)
    rb = OA.robin_coefficients(oc, betas)
    u_ref, c_ref, _ = OS.pde_solver_bpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 3, q, rb)
    bcs = {}
    for f, t in uc.facet_types.items():
        v = torch.as_tensor(bc_np[f])
        bcs[f] = (v, torch.as_tensor(betas[f])) if t == "r" else v
    u.clear_cache()
    sol = u.pde_solver_jit(_lap, lambda x, ce, r, fl: 1.0, uc, bcs, u.polyharmonic, 1)
    vals = sol.vals.cpu().numpy()
    rel = np.linalg.norm(vals - u_ref) / np.linalg.norm(u_ref)
    print("fs %d k %d: rel-L2 %.2e" % (fs, k, rel), flush=True)
    assert rel < 1e-9
    A = OA.assemble_A(oc, R.POLYHARMONIC, 1.0, 3)
    assert np.linalg.norm(A[:oc.N] @ sol.coeffs.cpu().numpy() - u_ref) <= 1e-9 * np.linalg.norm(u_ref)


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_local_poisson_40x40_both_precisions(precision):
    import updes_amd as u
    nx, k = 40, 30
    ft = {"North": "d", "South": "d", "East": "d", "West": "d"}
    uc, oc = _pair(nx, nx, ft, k)
    ex = lambda x: torch.sin(torch.pi * x[0]) * torch.sin(torch.pi * x[1])
    rhs = lambda x, centers, rbf, fields: -2 * torch.pi ** 2 * torch.sin(torch.pi * x[0]) * torch.sin(torch.pi * x[1])
    u.clear_cache()
    sol = u.pde_solver_jit(_lap, rhs, uc, {f: ex for f in ft}, u.polyharmonic, 2, precision=precision)
    x, y = oc.sorted_nodes.T
    bc = {f: np.sin(np.pi * x[ids]) * np.sin(np.pi * y[ids]) for f, ids in oc.facet_nodes.items()}
    q = OA.assemble_q(oc, -2 * np.pi ** 2 * np.sin(np.pi * x[:oc.Ni]) * np.sin(np.pi * y[:oc.Ni]), bc)
    u_ref, _, _ = OS.pde_solver_kpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 6, q)
    rel = np.linalg.norm(sol.vals.cpu().numpy() - u_ref) / np.linalg.norm(u_ref)
    st = u.last_solve_stats()
    print("%s: rel-L2 %.2e stats %s" % (precision, rel, st), flush=True)
    assert rel < 1e-8
    if precision == "fp32_ir":
        # local supports: the FP32 LU of K itself preconditions (the cardinal
        # basis needs every node in every support)
        assert st["basis"] == "lu", st


def test_local_laplace_control_gradient():
    """dL/d(North values) against the oracle's B^-T dL/du with the same supports."""
    import updes_amd as u
    nx, k = 16, 20
    ft = {"North": "d", "South": "d", "East": "d", "West": "d"}
    uc, oc = _pair(nx, nx, ft, k)
    rng = np.random.default_rng(1)
    nN = len(uc.facet_nodes["North"])
    c0 = torch.as_tensor(rng.normal(0, 0.1, nN), device="cuda")
    xi = uc.sorted_nodes[:uc.Ni].cuda()
    target = torch.sin(torch.pi * xi[:, 0]) * xi[:, 1]
    u.clear_cache()
    c = c0.clone().requires_grad_(True)
    bcs = {"North": c, "South": 0.0, "East": 0.0, "West": 0.0}
    sol = u.pde_solver(_lap, lambda x, ce, r, f: 0.0, uc, bcs, u.polyharmonic, 1)
    loss = ((sol.vals[:uc.Ni] - target) ** 2).mean()
    loss.backward()
    g = c.grad.detach().cpu().numpy()
    zeros = {f: np.zeros(len(ids)) for f, ids in oc.facet_nodes.items()}
    zeros["North"] = c0.cpu().numpy()
    q = OA.assemble_q(oc, 0.0, zeros)
    u_ref, _, B = OS.pde_solver_bpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 3, q)
    gu = np.zeros(oc.N)
    gu[:oc.Ni] = 2 * (u_ref[:oc.Ni] - target.cpu().numpy()) / oc.Ni
    g_ref = np.linalg.solve(B.T, gu)[oc.facet_nodes["North"]]
    assert np.linalg.norm(g - g_ref) <= 1e-8 * np.linalg.norm(g_ref)


def test_large_poisson_40k_local_supports():
    """N = 40,000 with supports of 20 nodes: sampled rows of K hold exactly the
    support columns (19 per Dirichlet row, r^3 being nonzero between distinct
    nodes) plus the monomials, and the FP64 solve has a small normwise
    backward error in the augmented system (matrix-free FP64 residual with the
    same support test, ||K||_F from random probes)."""
    import updes_amd as u
    import bench
    from updes_amd import _lib as L, engine as E
    k = 20
    cloud, bcs, b, basis = bench.fourier_problem(200, seed=0, support_size=k)
    tab = cloud.sorted_local_supports.numpy()
    u.clear_cache()
    vals, loss, g = bench.solve_and_grad(cloud, bcs, b, basis, "fp64")
    fac = list(E._K_CACHE.d.values())[-1]
    n, N = fac.n, cloud.N
    ld = (n + 63) // 64 * 64
    for r0 in (0, cloud.Ni - 8, cloud.Ni + 100, N - 8):
        rows = torch.empty((8, ld), dtype=torch.float64, device="cuda")
        L.check(L.lib().rbfx_assemble(fac.problem.ref(), L.MAT_K, r0, 8, L.ptr(rows), ld, None, ld, L.stream()))
        nz = (rows[:, :N] != 0).cpu().numpy()
        for q in range(8):
            cols = np.flatnonzero(nz[q])
            assert set(cols) <= set(tab[r0 + q]), r0 + q
            if cloud.Ni <= r0 + q < cloud.Ni + cloud.Nd:
                assert np.array_equal(cols, np.sort(tab[r0 + q]))
    with torch.no_grad():
        f_int = basis @ b
        q = u.assemble_q(lambda x, centers, rbf, fields: -f_int, bcs, cloud, u.polyharmonic, 6, None)
        rhs = torch.cat([q, torch.zeros(6, dtype=torch.float64, device=q.device)])
        c = fac.solve(rhs)
        r = E.matvec(fac.problem, c) - rhs
        gen = torch.Generator("cuda").manual_seed(0)
        kf = np.sqrt(np.mean([torch.linalg.norm(E.matvec(fac.problem, torch.randn(n, dtype=torch.float64,
                                                                           device="cuda", generator=gen))).item() ** 2
                              for _ in range(4)]))
    eta = torch.linalg.norm(r).item() / (kf * torch.linalg.norm(c).item() + torch.linalg.norm(rhs).item())
    print("N=40k, support 20: backward error %.3e, loss %.6e" % (eta, loss.item()), flush=True)
    assert np.isfinite(loss.item())
    assert eta <= n * np.finfo(float).eps * 100


def test_local_time_step_and_gradient():
    """One implicit advection-diffusion step with rhs value(x, coeffs(u^n))
    on local supports: the coefficients come from A with the local Phi
    (``assembly.py:404-410``), the rhs evaluation sums every centre
    (``operators.py:118-150``).  Values and d(w . u^{n+1})/du^n against the
    oracle's B-path with the same supports."""
    import functools
    import updes_amd as u
    nx, k, M = 16, 20, 3
    ft = {"North": "d", "South": "d", "West": "p1", "East": "p1"}
    uc, oc = _pair(nx, nx, ft, k, seed=7)
    dt, nu, vel = 1e-2, 0.05, (1.0, 0.5)
    V = torch.tensor(vel, dtype=torch.float64)

    def diff(x, center=None, rbf=None, monomial=None, fields=None):
        return (u.nodal_value(x, center, rbf, monomial) + dt * torch.dot(V, u.nodal_gradient(x, center, rbf, monomial))
                - dt * nu * u.nodal_laplacian(x, center, rbf, monomial))

    def rhs(x, centers=None, rbf=None, fields=None):
        return u.value(x, fields[:, 0], centers, rbf)

    bcs = {f: 0.0 for f in ft}
    xy = uc.sorted_nodes.cuda()
    u0 = torch.exp(-((xy[:, 0] - 0.3) ** 2 + (xy[:, 1] - 0.5) ** 2) / (2 * 0.15 ** 2))
    w = torch.as_tensor(np.random.default_rng(3).standard_normal(uc.N), device="cuda")
    u.clear_cache()
    a = u0.clone().requires_grad_(True)
    u1 = u.pde_solver(diff, rhs, uc, bcs, u.polyharmonic, 1, rhs_args=[a]).vals
    (u1 * w).sum().backward()
    # oracle: u1 = B^-1 q, q_int = Eval(x_int) A^-1 [u0; 0], boundary rows 0
    N, Ni = oc.N, oc.Ni
    desc = C.adv_diff_desc(dt, nu, vel)
    B = OA.assemble_B(desc, oc, R.POLYHARMONIC, 1.0, M)
    invA = np.linalg.inv(OA.assemble_A(oc, R.POLYHARMONIC, 1.0, M))
    Ev = np.stack([OS.value(oc.sorted_nodes[:Ni], e, oc.sorted_nodes, R.POLYHARMONIC, 1.0)
                   for e in np.eye(N + M)], axis=1)                # (Ni, N+M)
    Q = np.zeros((N, N))
    Q[:Ni] = Ev @ invA[:, :N]                                      # dq/du0
    u0n = u0.cpu().numpy()
    u_ref = OS.qr_solve(B, Q @ u0n)
    g_ref = Q.T @ np.linalg.solve(B.T, w.cpu().numpy())
    tol = max(1e-9, 10 * np.linalg.cond(B) * np.finfo(float).eps)
    e_u = np.linalg.norm(u1.detach().cpu().numpy() - u_ref) / np.linalg.norm(u_ref)
    e_g = np.linalg.norm(a.grad.cpu().numpy() - g_ref) / np.linalg.norm(g_ref)
    print("local time step: vals %.2e grad %.2e (tol %.1e)" % (e_u, e_g, tol), flush=True)
    assert e_u <= tol and e_g <= tol
