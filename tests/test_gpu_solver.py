"""GPU parity of assembly, solve and evaluation against the CPU oracle."""
import numpy as np
import pytest
import torch

from oracle import assembly as OA, cloud as OC, rbf as R, solver as OS
from tests import _cases as C

pytestmark = pytest.mark.gpu


def _laplace_op(x, center=None, rbf=None, monomial=None, fields=None):
    import updes_amd as u
    return u.nodal_laplacian(x, center, rbf, monomial)


def _value_op(x, center=None, rbf=None, monomial=None, fields=None):
    import updes_amd as u
    return u.nodal_value(x, center, rbf, monomial)


def _rowscaled_err(a, b):
    scale = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-300)
    return float((np.abs(a - b) / scale).max())


@pytest.mark.parametrize("kind,param", C.KINDS)
@pytest.mark.parametrize("fs", range(len(C.FACET_SETS)))
def test_assembly_K_matches_oracle(kind, param, fs):
    import updes_amd as u
    facets = C.FACET_SETS[fs]
    nx, ny = 9, 8
    J = C.jitter(nx, ny, 0.2, seed=fs)
    oc = OC.SquareCloud(nx, ny, facets, jitter=J)
    uc = u.SquareCloud(Nx=nx, Ny=ny, facet_types=facets, jitter=J)
    M = 6
    betas = {f: np.linspace(0.5, 1.5, len(oc.facet_nodes[f])) for f, t in oc.facet_types.items() if t == "r"}
    rb = OA.robin_coefficients(oc, betas)
    Kref = OA.assemble_K(C.adv_diff_desc(0.1, 0.3, (1.0, -0.5)), oc, kind, param, M, rb)
    rc, _ = u.duplicate_robin_coeffs({f: (0.0, betas[f]) if t == "r" else 0.0 for f, t in uc.facet_types.items()}, uc)
    dt, nu, vel = 0.1, 0.3, (1.0, -0.5)

    def op(x, center=None, rbf=None, monomial=None, fields=None):
        val = u.nodal_value(x, center, rbf, monomial)
        grad = u.nodal_gradient(x, center, rbf, monomial)
        lap = u.nodal_laplacian(x, center, rbf, monomial)
        return val + dt * torch.dot(torch.tensor(vel, dtype=torch.float64), grad) - dt * nu * lap

    K = u.assemble_K(op, uc, C.torch_rbf(kind, param), M, None, rc).cpu().numpy()
    assert K.shape == Kref.shape
    assert _rowscaled_err(K, Kref) < 1e-12


@pytest.mark.parametrize("kind,param", C.KINDS[:3])
def test_assembly_A_matches_oracle(kind, param):
    import updes_amd as u
    oc = OC.SquareCloud(12, 10, C.README_FACETS)
    uc = u.SquareCloud(Nx=12, Ny=10, facet_types=C.README_FACETS)
    A = u.assemble_A(uc, C.torch_rbf(kind, param), 3).cpu().numpy()
    Aref = OA.assemble_A(oc, kind, param, 3)
    assert _rowscaled_err(A, Aref) < 1e-12


def test_readme_laplace_matches_reference_algorithm():
    """README example (N=600): vals within 1e-8 rel-L2 of the restated reference
    (inv(A), D inv(A), QR); coefficients agree to the conditioning of A."""
    import updes_amd as u
    ft = C.README_FACETS
    oc = OC.SquareCloud(30, 20, ft)
    cloud = u.SquareCloud(Nx=30, Ny=20, facet_types=ft)
    sine = lambda coord: torch.sin(torch.pi * coord[0])
    zero = lambda coord: 0.0
    bcs = {"South": zero, "West": zero, "North": sine, "East": zero}
    sol = u.pde_solver_jit(diff_operator=_laplace_op, rhs_operator=lambda x, centers, rbf, fields: 0.0,
                           cloud=cloud, boundary_conditions=bcs, rbf=u.polyharmonic, max_degree=1)
    bc_arr = {f: np.sin(np.pi * oc.sorted_nodes[oc.facet_nodes[f], 0]) if f == "North" else
              np.zeros(len(oc.facet_nodes[f])) for f in ft}
    q = OA.assemble_q(oc, 0.0, bc_arr)
    u_ref, c_ref, B = OS.pde_solver_bpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 3, q)
    vals = sol.vals.cpu().numpy()
    rel = np.linalg.norm(vals - u_ref) / np.linalg.norm(u_ref)
    assert rel < 1e-8, rel
    # coefficients: ill-determined (kappa(A) ~ 2e7); compare through the values they produce
    A = OA.assemble_A(oc, R.POLYHARMONIC, 1.0, 3)
    assert np.linalg.norm(A[:600] @ sol.coeffs.cpu().numpy() - u_ref) / np.linalg.norm(u_ref) < 1e-8
    assert np.linalg.norm(sol.coeffs.cpu().numpy() - c_ref) / np.linalg.norm(c_ref) < 1e-4


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_poisson_small_both_precisions(precision):
    import updes_amd as u
    nx = 30
    ft = {"North": "d", "South": "d", "East": "d", "West": "d"}
    J = C.jitter(nx, nx, 0.2, 0)
    oc = OC.SquareCloud(nx, nx, ft, jitter=J)
    cloud = u.SquareCloud(Nx=nx, Ny=nx, facet_types=ft, jitter=J)
    ex = lambda x: torch.sin(torch.pi * x[0]) * torch.sin(torch.pi * x[1])
    rhs = lambda x, centers, rbf, fields: -2 * torch.pi ** 2 * torch.sin(torch.pi * x[0]) * torch.sin(torch.pi * x[1])
    sol = u.pde_solver_jit(_laplace_op, rhs, cloud, {f: ex for f in ft}, u.polyharmonic, 2, precision=precision)
    x, y = oc.sorted_nodes.T
    q = np.where(np.arange(oc.N) < oc.Ni, -2 * np.pi ** 2 * np.sin(np.pi * x) * np.sin(np.pi * y),
                 np.sin(np.pi * x) * np.sin(np.pi * y))
    u_ref, _, _ = OS.pde_solver_bpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 6, q)
    vals = sol.vals.cpu().numpy()
    assert np.linalg.norm(vals - u_ref) / np.linalg.norm(u_ref) < 1e-8
    st = u.last_solve_stats()
    assert st["precision"] == precision
    if precision == "fp32_ir":
        assert st["fell_back"] is False and st["basis"] == "cardinal", st


def test_eval_value_gradient_laplacian():
    import updes_amd as u
    oc = OC.SquareCloud(10, 9, C.README_FACETS, jitter=C.jitter(10, 9, 0.1, 3))
    for kind, param in C.KINDS:
        rng = np.random.default_rng(kind)
        coeffs = rng.standard_normal(oc.N + 6)
        pts = np.concatenate([rng.uniform(0, 1, (17, 2)), oc.sorted_nodes[:5]])
        ref = OS.eval_ops(pts, coeffs, oc.sorted_nodes, kind, param)
        rbf = C.torch_rbf(kind, param)
        c = torch.as_tensor(coeffs).cuda()
        ctr = torch.as_tensor(oc.sorted_nodes).cuda()
        p = torch.as_tensor(pts).cuda()
        v = u.value_vec(p, c, ctr, rbf).cpu().numpy()
        g = u.gradient_vec(p, c, ctr, rbf).cpu().numpy()
        l = u.laplacian_vec(p, c, ctr, rbf).cpu().numpy()
        sc = np.abs(ref).max(axis=1)
        assert np.abs(v - ref[R.VAL]).max() <= 1e-12 * max(sc[R.VAL], 1)
        assert np.abs(g[:, 0] - ref[R.DX]).max() <= 1e-12 * max(sc[R.DX], 1)
        assert np.abs(g[:, 1] - ref[R.DY]).max() <= 1e-12 * max(sc[R.DY], 1)
        assert np.abs(l - ref[R.LAP]).max() <= 1e-12 * max(sc[R.LAP], 1)


def test_integrate_field_matches_oracle_and_reference_test():
    """updes/tests/test_integrals.py on the GPU path: r^11, degree 3, 12 x 12,
    x^2/(1+y^2): same integral as the oracle, and |I - pi/12| < 0.1."""
    import functools
    import updes_amd as u
    ft = {"North": "d", "South": "d", "East": "d", "West": "d"}
    cloud = u.SquareCloud(Nx=12, Ny=12, facet_types=ft)
    rbf = functools.partial(u.polyharmonic, a=5)
    xy = cloud.sorted_nodes.cuda()
    s = xy[:, 0] ** 2 / (1 + xy[:, 1] ** 2)
    c = u.get_field_coefficients(s, cloud, rbf, 3)
    val = float(u.integrate_field(c, cloud, rbf, 3))
    assert abs(val - np.pi / 12) < 1e-1
    # against the oracle: with r^11 and degree 3, kappa(A) ~ 2e12 and two FP64
    # solves differ by ~kappa eps (1e-6 in this integral), so the entrywise
    # comparison uses the well-conditioned r^3 / degree-1 interpolant
    oc = OC.SquareCloud(12, 12, ft)
    c3 = u.get_field_coefficients(s, cloud, u.polyharmonic, 1)
    val3 = float(u.integrate_field(c3, cloud, u.polyharmonic, 1))
    c_ref = OA.core_compute_coefficients(s.cpu().numpy(), oc, R.POLYHARMONIC, 1.0, 3)
    ref = OS.integrate_field(c_ref, oc, R.POLYHARMONIC, 1.0)
    kappa = np.linalg.cond(OA.assemble_A(oc, R.POLYHARMONIC, 1.0, 3))
    assert abs(val3 - ref) <= max(1e-12, 10 * kappa * np.finfo(float).eps) * abs(ref)
    # constant field: the rule's weights exactly
    one = float(u.integrate_field(u.get_field_coefficients(torch.ones(cloud.N, dtype=torch.float64, device="cuda"),
                                                           cloud, rbf, 3), cloud, rbf, 3))
    assert abs(one - (11 * 11 + 10 + 10 + 1) / 121) < 1e-9


def test_pde_multi_solver_picard_matches_oracle():
    """pde_multi_solver (operators.py:696-771): Picard iteration of two coupled
    fields; the diff operator of u reads v's nodal values through `fields`
    (per-row coefficients, assembly.py:114), so the matrix changes with every
    iterate.  Replayed with the oracle's augmented-system solves."""
    import updes_amd as u
    nx = 12
    ft = {"North": "d", "South": "d", "East": "d", "West": "d"}
    cloud = u.SquareCloud(Nx=nx, Ny=nx, facet_types=ft)
    oc = OC.SquareCloud(nx, nx, ft)

    def op_u(x, center=None, rbf=None, monomial=None, fields=None):
        return u.nodal_laplacian(x, center, rbf, monomial) - fields[1] * u.nodal_value(x, center, rbf, monomial)

    def op_v(x, center=None, rbf=None, monomial=None, fields=None):
        return u.nodal_laplacian(x, center, rbf, monomial)

    rhs_u = lambda x, centers=None, rbf=None, fields=None: -1.0
    rhs_v = lambda x, centers=None, rbf=None, fields=None: 2.0
    bc_u = {f: 0.0 for f in ft}
    bc_v = {f: (lambda p: p[0]) for f in ft}
    z = torch.zeros(cloud.N, dtype=torch.float64, device="cuda")
    iters = 4
    sols = u.pde_multi_solver([op_u, op_v], [rhs_u, rhs_v], cloud, [bc_u, bc_v], u.polyharmonic, 1,
                              nb_iters=iters, diff_args=[[z, z], [z, z]], rhs_args=[None, None])
    # oracle replay
    x = oc.sorted_nodes[:, 0]
    bcu = {f: np.zeros(len(oc.facet_nodes[f])) for f in ft}
    bcv = {f: oc.sorted_nodes[oc.facet_nodes[f], 0] for f in ft}
    vals = [np.zeros(oc.N), np.zeros(oc.N)]
    for _ in range(iters):
        v_rows = vals[1][:oc.Ni]
        desc_u = {"rbf": [-v_rows, 0, 0, 1, 0, 0, 0], "mon": [-v_rows, 0, 0, 1, 0, 0, 0]}
        qu = OA.assemble_q(oc, -np.ones(oc.Ni), bcu)
        qv = OA.assemble_q(oc, 2.0 * np.ones(oc.Ni), bcv)
        uu, _, _ = OS.pde_solver_kpath(desc_u, oc, R.POLYHARMONIC, 1.0, 3, qu)
        vv, _, _ = OS.pde_solver_kpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 3, qv)
        vals = [uu, vv]
    for i in range(2):
        got = sols[i].vals.cpu().numpy()
        assert np.linalg.norm(got - vals[i]) <= 1e-10 * np.linalg.norm(vals[i]), i
    # the coupling is active: u differs from the uncoupled (v = 0) solve
    u0, _, _ = OS.pde_solver_kpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 3, OA.assemble_q(oc, -np.ones(oc.Ni), bcu))
    assert np.linalg.norm(vals[0] - u0) > 1e-3 * np.linalg.norm(u0)


@pytest.mark.parametrize("fs", [0, 1, 2, 3, 5])
def test_boundary_conditions_end_to_end(fs):
    """pde_solver with Dirichlet / Neumann / Robin (value, beta) / periodic
    facets given as callables, arrays and tuples (operators.py:512-555,
    628-647) against the reference algorithm.  (The all-periodic facet set is
    singular for the Laplacian and is covered by the assembly test only.)"""
    import updes_amd as u
    facets = C.FACET_SETS[fs]
    nx, ny = 10, 9
    oc = OC.SquareCloud(nx, ny, facets)
    uc = u.SquareCloud(Nx=nx, Ny=ny, facet_types=facets)
    fd = lambda p: torch.sin(torch.pi * p[0]) + p[1]
    fn = lambda p: torch.cos(p[0]) * 0.5
    fr = lambda p: p[0] + 1.0
    bcs, arrs, betas = {}, {}, {}
    for f, t in facets.items():
        pts = oc.sorted_nodes[oc.facet_nodes[f]]
        if t == "d":
            bcs[f], arrs[f] = fd, np.sin(np.pi * pts[:, 0]) + pts[:, 1]
        elif t == "n":
            bcs[f], arrs[f] = fn, np.cos(pts[:, 0]) * 0.5
        elif t == "r":
            b = np.linspace(0.5, 1.5, len(pts))
            bcs[f], arrs[f], betas[f] = (fr, torch.as_tensor(b)), pts[:, 0] + 1.0, b
        else:
            bcs[f], arrs[f] = 0.0, np.zeros(len(pts))
    rhs = lambda x, centers=None, rbf=None, fields=None: x[0] * x[1] + 1.0
    sol = u.pde_solver(_laplace_op, rhs, uc, bcs, u.polyharmonic, 1)
    xi = oc.sorted_nodes[:oc.Ni]
    q = OA.assemble_q(oc, xi[:, 0] * xi[:, 1] + 1.0, arrs)
    rb = OA.robin_coefficients(oc, betas) if betas else None
    u_ref, _, B = OS.pde_solver_bpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 3, q, rb)
    tol = max(1e-8, 10 * np.linalg.cond(B) * np.finfo(float).eps)
    got = sol.vals.cpu().numpy()
    assert np.linalg.norm(got - u_ref) <= tol * np.linalg.norm(u_ref), (fs, np.linalg.norm(got - u_ref) / np.linalg.norm(u_ref))


def test_solves_on_successive_clouds_without_clearing_the_cache():
    """Factorisations are cached per cloud object; clouds created and dropped
    in a loop (CPython reuses their ids) must never pick up another cloud's
    factors.  Each solve is checked against the oracle."""
    import gc
    import updes_amd as u
    ft = C.README_FACETS
    for nx, ny in [(10, 8), (14, 9), (10, 8), (12, 11), (14, 9)]:
        cloud = u.SquareCloud(Nx=nx, Ny=ny, facet_types=ft)
        bcs = {"South": 0.0, "West": 0.0, "North": lambda p: torch.sin(torch.pi * p[0]), "East": 0.0}
        sol = u.pde_solver(_laplace_op, lambda x, centers, rbf, fields: 1.0, cloud, bcs, u.polyharmonic, 1)
        oc = OC.SquareCloud(nx, ny, ft)
        arr = {f: np.sin(np.pi * oc.sorted_nodes[oc.facet_nodes[f], 0]) if f == "North" else
               np.zeros(len(oc.facet_nodes[f])) for f in ft}
        q = OA.assemble_q(oc, 1.0, arr)
        u_ref, _, _ = OS.pde_solver_kpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 3, q)
        got = sol.vals.cpu().numpy()
        assert got.shape == u_ref.shape
        assert np.linalg.norm(got - u_ref) <= 1e-9 * np.linalg.norm(u_ref), (nx, ny)
        del cloud, sol
        gc.collect()


@pytest.mark.parametrize("nx,ny,ft,deg", [(30, 20, C.README_FACETS, 1),
                                          (60, 60, {"North": "d", "South": "d", "East": "d", "West": "d"}, 1),
                                          (14, 12, C.FACET_SETS[2], 2), (14, 12, C.FACET_SETS[3], 1)])
def test_assemble_B_and_steady_sol_mat_match_oracle(nx, ny, ft, deg):
    """B = (D inv(A))[:, :N] (assembly.py:366-401) and SteadySol.mat against
    the oracle's assemble_B at the README (30 x 20) and adjoint-control
    (60 x 60) sizes, plus Robin and periodic facet sets; inv(A) against the
    oracle's inverse."""
    import updes_amd as u
    oc = OC.SquareCloud(nx, ny, ft)
    uc = u.SquareCloud(Nx=nx, Ny=ny, facet_types=ft)
    M = 3 if deg == 1 else 6
    betas = {f: np.linspace(0.5, 1.5, len(oc.facet_nodes[f])) for f, t in oc.facet_types.items() if t == "r"}
    rb = OA.robin_coefficients(oc, betas)
    rc, _ = u.duplicate_robin_coeffs({f: (0.0, betas[f]) if t == "r" else 0.0 for f, t in uc.facet_types.items()}, uc)
    B = u.assemble_B(_laplace_op, uc, u.polyharmonic, M, None, rc).cpu().numpy()
    Bref = OA.assemble_B(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, M, rb)
    kappa = np.linalg.cond(OA.assemble_A(oc, R.POLYHARMONIC, 1.0, M))
    # two FP64 inverses of A agree to ~kappa(A) eps; B = D inv(A) inherits that
    tol = max(1e-10, 10 * kappa * np.finfo(float).eps)
    assert np.linalg.norm(B - Bref) <= tol * np.linalg.norm(Bref), (np.linalg.norm(B - Bref) / np.linalg.norm(Bref), tol)
    iA = u.assemble_invert_A(uc, u.polyharmonic, M).cpu().numpy()
    A = OA.assemble_A(oc, R.POLYHARMONIC, 1.0, M)
    assert np.linalg.norm(A @ iA - np.eye(oc.N + M)) <= 100 * kappa * np.finfo(float).eps * np.sqrt(oc.N + M)
    # SteadySol.mat is the same B
    bcs = {f: 0.0 for f in ft}
    if any(t == "r" for t in uc.facet_types.values()):
        bcs.update({f: (0.0, torch.as_tensor(betas[f])) for f, t in uc.facet_types.items() if t == "r"})
    sol = u.pde_solver(_laplace_op, lambda x, centers, rbf, fields: 1.0, uc, bcs, u.polyharmonic, deg)
    np.testing.assert_allclose(sol.mat.cpu().numpy(), B, rtol=0, atol=1e-12 * np.abs(B).max())


@pytest.mark.parametrize("fs", [2, 3, 5])
def test_rhs_abi_matches_oracle(fs):
    """rbfx_rhs through the C ABI: interior rows = constant + row array +
    weighted operator values of fields given by their coefficients (the
    oracle's coefficients, so the check isolates the right-hand side), the
    boundary rows the given values in node order, the monomial rows 0
    (assembly.py:434-485); and the package's assemble_q (which now assembles
    through it) against the oracle's, with Robin tuples and periodic facets."""
    import updes_amd as u
    from updes_amd import engine as E, _lib as L
    from updes_amd.operators import assemble_q
    facets = C.FACET_SETS[fs]
    nx, ny = 11, 10
    oc = OC.SquareCloud(nx, ny, facets)
    uc = u.SquareCloud(Nx=nx, Ny=ny, facet_types=facets)
    X = oc.sorted_nodes
    M = 6
    f1 = np.sin(2 * X[:, 0]) + X[:, 1] ** 2
    f2 = np.cos(X[:, 1]) * X[:, 0]
    c1 = OA.core_compute_coefficients(f1, oc, R.POLYHARMONIC, 1.0, M)
    c2 = OA.core_compute_coefficients(f2, oc, R.POLYHARMONIC, 1.0, M)
    xi = X[:oc.Ni]
    row = 1.0 + xi[:, 0]
    crow = np.sin(xi[:, 0])
    ops1 = OS.eval_ops(xi, c1, X, R.POLYHARMONIC, 1.0)
    ops2 = OS.eval_ops(xi, c2, X, R.POLYHARMONIC, 1.0)
    bc = np.random.default_rng(fs).standard_normal(oc.N - oc.Ni)
    expect = np.concatenate([0.7 + crow + 2.0 * row * ops1[R.LAP] - 0.5 * ops1[R.DX] + ops2[R.VAL], bc, np.zeros(M)])
    prob = E.Problem(E.device_cloud(uc), L.POLYHARMONIC, 1.0, M)
    dev = torch.device("cuda")
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    terms = [(L.LAP, 2.0, t(row), t(c1)), (L.DX, -0.5, None, t(c1)), (L.VAL, 1.0, None, t(c2))]
    q = E.rhs(prob, t(0.7 + crow), t(bc), terms).cpu().numpy()
    assert np.abs(q - expect).max() <= 1e-12 * np.abs(expect).max()
    # the package's assemble_q (callables and arrays; Robin values as
    # pde_solver passes them after splitting the (value, beta) tuples)
    fd = lambda p: torch.sin(torch.pi * p[0]) + p[1]
    bcs, arrs = {}, {}
    for f, ty in facets.items():
        pts = X[oc.facet_nodes[f]]
        if ty == "d":
            bcs[f], arrs[f] = fd, np.sin(np.pi * pts[:, 0]) + pts[:, 1]
        elif ty == "r":
            v = pts[:, 1] - 0.3
            bcs[f], arrs[f] = torch.as_tensor(v), v
        elif ty == "n":
            v = np.linspace(-1, 1, len(pts))
            bcs[f], arrs[f] = torch.as_tensor(v), v
        else:
            # periodic facets: zeros, as pde_solver passes them (operators.py:546-555)
            bcs[f], arrs[f] = 0.0, np.zeros(len(pts))
    rhs = lambda x, centers=None, rbf=None, fields=None: x[0] * x[1] + 1.0
    got = assemble_q(rhs, bcs, uc, u.polyharmonic, M, None).cpu().numpy()
    ref = OA.assemble_q(oc, xi[:, 0] * xi[:, 1] + 1.0, arrs)
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
