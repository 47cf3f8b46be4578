"""Pins of the CPU oracle (no GPU): closed forms against SymPy/mpmath,
a brute-force high-precision restatement built independently from the
reference's loop semantics, invariants the mathematics fixes, and planted
mistakes that the pins must catch."""
import functools

import mpmath as mp
import numpy as np
import pytest
import sympy as sp

from oracle import assembly as OA, cloud as OC, rbf as R, solver as OS
from tests import _cases as C

mp.mp.dps = 40

X, Y, CX, CY = sp.symbols("x y cx cy", real=True)


def _sym_kernel(kind, param):
    r = sp.sqrt((X - CX) ** 2 + (Y - CY) ** 2)
    p = sp.nsimplify(param)
    if kind == R.POLYHARMONIC:
        return r ** (2 * p + 1)
    if kind == R.GAUSSIAN:
        return sp.exp(-(p * r) ** 2)
    if kind == R.THIN_PLATE:
        return sp.log(r) * r ** (2 * p)
    if kind == R.MULTIQUADRIC:
        return sp.sqrt(1 + (p * r) ** 2)
    return 1 / sp.sqrt(1 + (p * r) ** 2)


@functools.lru_cache(maxsize=None)
def _sym_ops(kind, param):
    f = _sym_kernel(kind, param)
    dxx = sp.diff(f, X, 2)
    dyy = sp.diff(f, Y, 2)
    exprs = [f, sp.diff(f, X), sp.diff(f, Y), dxx + dyy, dxx, dyy, sp.diff(f, X, Y)]
    return [sp.lambdify((X, Y, CX, CY), e, "mpmath") for e in exprs]


def _mp_ops(kind, param, x, y, cx, cy):
    """High-precision operator values with the reference's r = 0 rule."""
    if x == cx and y == cy:
        return [mp.mpf(R.phi_zero(kind, param))] + [mp.mpf(0)] * 6
    fs = _sym_ops(kind, param)
    return [f(mp.mpf(x), mp.mpf(y), mp.mpf(cx), mp.mpf(cy)) for f in fs]


@pytest.mark.parametrize("kind,param", C.KINDS + [(R.THIN_PLATE, 2.0), (R.GAUSSIAN, 1.5)])
def test_closed_forms_match_sympy(kind, param):
    rng = np.random.default_rng(kind * 10 + int(param))
    for _ in range(25):
        x, y, cx, cy = rng.uniform(-1, 1.5, 4)
        got = R.rbf_ops(kind, param, x, y, cx, cy)
        ref = _mp_ops(kind, param, x, y, cx, cy)
        for t in range(R.NOPS):
            assert abs(got[t] - float(ref[t])) <= 1e-13 * max(1.0, abs(float(ref[t]))), (t, got[t], ref[t])
    # r = 0: phi(0) and zero derivatives (reference nan_to_num at r = 0)
    z = R.rbf_ops(kind, param, 0.3, 0.4, 0.3, 0.4)
    assert z[R.VAL] == R.phi_zero(kind, param) and np.all(z[1:] == 0)


def test_monomials_match_sympy():
    rng = np.random.default_rng(0)
    for k, (a, b) in enumerate(R.MONOMIAL_EXPONENTS):
        p = X ** a * Y ** b
        ex = [p, sp.diff(p, X), sp.diff(p, Y), sp.diff(p, X, 2) + sp.diff(p, Y, 2), sp.diff(p, X, 2),
              sp.diff(p, Y, 2), sp.diff(p, X, Y)]
        for _ in range(5):
            x, y = rng.uniform(-2, 2, 2)
            got = R.monomial_ops(k, x, y)
            for t in range(R.NOPS):
                assert abs(got[t] - float(ex[t].subs({X: x, Y: y}))) < 1e-12
    assert [R.compute_nb_monomials(d) for d in range(5)] == [1, 3, 6, 10, 15]


# ---- brute-force restatement (independent loops over the reference semantics) ----
def _brute_matrices(oc, kind, param, M, desc, robin_betas):
    """A and D built entry by entry in mpmath following updes/assembly.py's loops."""
    N, Ni, Nd, Nn, Nr, Np = oc.N, oc.Ni, oc.Nd, oc.Nn, oc.Nr, oc.Np
    nodes = [(float(a), float(b)) for a, b in oc.sorted_nodes]
    nrm = [(float(a), float(b)) for a, b in oc.sorted_outward_normals]
    ops = lambda i, j: _mp_ops(kind, param, nodes[i][0], nodes[i][1], nodes[j][0], nodes[j][1])
    mono = lambda k, i: [mp.mpf(v) for v in R.monomial_ops(k, nodes[i][0], nodes[i][1])]
    support = lambda i: [j for j in range(N) if j != i]        # cloud.py:110-111
    A = mp.zeros(N + M, N + M)
    for i in range(N):
        for j in support(i):
            A[i, j] = ops(i, j)[0]
        for k in range(M):
            A[i, N + k] = mono(k, i)[0]
            A[N + k, i] = mono(k, i)[0]
    D = mp.zeros(N, N + M)
    for i in range(Ni):
        for j in support(i):
            o = ops(i, j)
            D[i, j] = sum(mp.mpf(desc["rbf"][t]) * o[t] for t in range(7))
        for k in range(M):
            m = mono(k, i)
            D[i, N + k] = sum(mp.mpf(desc["mon"][t]) * m[t] for t in range(7))
    for i in range(Ni, Ni + Nd):
        for j in support(i):
            D[i, j] = ops(i, j)[0]
        for k in range(M):
            D[i, N + k] = mono(k, i)[0]
    for i in range(Ni + Nd, Ni + Nd + Nn):
        n = nrm[i - Ni - Nd]
        for j in support(i):
            o = ops(i, j)
            D[i, j] = o[1] * n[0] + o[2] * n[1]
        for k in range(M):
            m = mono(k, i)
            D[i, N + k] = m[1] * n[0] + m[2] * n[1]
    for i in range(Ni + Nd + Nn, Ni + Nd + Nn + Nr):
        t = i - Ni - Nd - Nn
        nb = nrm[t]                       # assembly.py:206 (index shifted by Nn)
        ng = nrm[i - Ni - Nd]             # assembly.py:303 (node's own normal)
        beta = mp.mpf(robin_betas[t])
        for j in support(i):
            o = ops(i, j)
            D[i, j] = beta * o[0] + o[1] * nb[0] + o[2] * nb[1]
        for k in range(M):
            m = mono(k, i)
            D[i, N + k] = beta * m[0] + m[1] * ng[0] + m[2] * ng[1]
    start, jump, half = Ni + Nd + Nn + Nr, 0, sum(Np) // 2
    for nbp in Np:
        nc = nbp // 2
        for i in range(start, start + nc):
            i2 = i + nc
            row_d, row_n = i - jump, i - jump + half
            n1, n2 = nrm[i - Ni - Nd], nrm[i2 - Ni - Nd]
            for j in sorted(set(support(i)) | set(support(i2))):
                o1, o2 = ops(i, j), ops(i2, j)
                D[row_d, j] = o1[0] - o2[0]
                D[row_n, j] = (o1[1] * n1[0] + o1[2] * n1[1]) - (o2[1] * -n2[0] + o2[2] * -n2[1])
            for k in range(M):
                m1, m2 = mono(k, i), mono(k, i2)
                D[row_d, N + k] = m1[0] - m2[0]
                D[row_n, N + k] = (m1[1] * n1[0] + m1[2] * n1[1]) - (m2[1] * -n2[0] + m2[2] * -n2[1])
        start += nbp
        jump += nc
    return A, D


def _brute_solve(A, D, q, N, M):
    Ainv = mp.inverse(A)
    B = (D * Ainv)[:, :N]
    u = mp.lu_solve(B, mp.matrix([mp.mpf(v) for v in q]))
    c = Ainv * mp.matrix([u[i] for i in range(N)] + [0] * M)
    return np.array([float(v) for v in u]), np.array([float(c[i]) for i in range(N + M)]), B


BRUTE_CASES = [
    ({"North": "d", "South": "d", "East": "d", "West": "d"}, R.POLYHARMONIC, 1.0, 1, C.LAPLACIAN),
    ({"South": "n", "West": "d", "North": "d", "East": "d"}, R.POLYHARMONIC, 1.0, 1, C.LAPLACIAN),
    ({"South": "r", "West": "n", "North": "d", "East": "r"}, R.GAUSSIAN, 2.0, 1, C.adv_diff_desc(0.1, 0.5, (1, .5))),
    ({"South": "d", "West": "p1", "North": "d", "East": "p1"}, R.GAUSSIAN, 3.0, 1, C.adv_diff_desc(0.05, 1, (1, 0))),
    ({"South": "n", "North": "r", "West": "p1", "East": "p1"}, R.THIN_PLATE, 1.0, 2, C.LAPLACIAN),
]


def _oracle_case(i, nx=5, ny=4):
    facets, kind, param, deg, desc = BRUTE_CASES[i]
    J = C.jitter(nx, ny, 0.15, i)
    oc = OC.SquareCloud(nx, ny, facets, jitter=J)
    M = R.compute_nb_monomials(deg)
    betas = {f: np.linspace(0.5, 1.0, len(oc.facet_nodes[f])) for f, t in oc.facet_types.items() if t == "r"}
    rb = OA.robin_coefficients(oc, betas)
    rng = np.random.default_rng(i)
    bc = {f: rng.standard_normal(len(oc.facet_nodes[f])) for f in oc.facet_types}
    q = OA.assemble_q(oc, rng.standard_normal(oc.Ni), bc)
    return oc, kind, param, M, desc, rb, q


@pytest.mark.parametrize("case", range(len(BRUTE_CASES)))
def test_oracle_matches_bruteforce(case):
    oc, kind, param, M, desc, rb, q = _oracle_case(case)
    A_mp, D_mp, B_mp = None, None, None
    A_mp, D_mp = _brute_matrices(oc, kind, param, M, desc, rb if rb is not None else [])
    A = OA.assemble_A(oc, kind, param, M)
    D = OA.assemble_D(desc, oc, kind, param, M, rb)
    A_ref = np.array(A_mp.tolist(), dtype=float)
    D_ref = np.array(D_mp.tolist(), dtype=float)
    assert np.abs(A - A_ref).max() <= 1e-13 * np.abs(A_ref).max()
    rs = np.maximum(np.abs(D_ref).max(axis=1, keepdims=True), 1e-300)
    assert (np.abs(D - D_ref) / rs).max() <= 1e-13
    u_ref, c_ref, _ = _brute_solve(A_mp, D_mp, q, oc.N, M)
    u, c, _ = OS.pde_solver_bpath(desc, oc, kind, param, M, q, rb)
    assert np.linalg.norm(u - u_ref) <= 1e-9 * np.linalg.norm(u_ref)
    assert np.linalg.norm(c - c_ref) <= 1e-7 * np.linalg.norm(c_ref)
    uk, ck, _ = OS.pde_solver_kpath(desc, oc, kind, param, M, q, rb)
    assert np.linalg.norm(uk - u_ref) <= 1e-9 * np.linalg.norm(u_ref)


# ---- invariants ------------------------------------------------------------------
def test_linear_reproduction_laplace():
    """deg >= 1 reproduces linear Dirichlet data exactly (Laplace, u = a + bx + cy)."""
    ft = {"North": "d", "South": "d", "East": "d", "West": "d"}
    oc = OC.SquareCloud(12, 11, ft, jitter=C.jitter(12, 11, 0.2, 1))
    x, y = oc.sorted_nodes.T
    ex = 0.3 + 1.7 * x - 0.6 * y
    q = np.where(np.arange(oc.N) < oc.Ni, 0.0, ex)
    u, c, B = OS.pde_solver_bpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 3, q)
    assert np.abs(u - ex).max() < 1e-9
    assert np.abs(c[:oc.N]).max() < 1e-7                      # lambda = 0, field is the polynomial
    assert np.allclose(c[oc.N:], [0.3, 1.7, -0.6], atol=1e-8)
    # Dirichlet rows of B are identity rows (up to kappa * eps)
    Bd = B[oc.Ni:]
    eye = np.zeros_like(Bd)
    eye[np.arange(oc.N - oc.Ni), np.arange(oc.Ni, oc.N)] = 1.0
    assert np.abs(Bd - eye).max() < 1e-8


def test_quadratic_reproduction_poisson():
    """deg = 2 reproduces u = x^2 + y^2 for Lap u = 4 (with a Neumann side)."""
    ft = {"North": "d", "South": "n", "East": "d", "West": "d"}
    oc = OC.SquareCloud(11, 12, ft, jitter=C.jitter(11, 12, 0.2, 2))
    x, y = oc.sorted_nodes.T
    ex = x ** 2 + y ** 2
    q = ex.copy()
    q[:oc.Ni] = 4.0
    nn = oc.Ni + oc.Nd
    q[nn:nn + oc.Nn] = -2 * y[nn:nn + oc.Nn]                   # d/dn with n = (0, -1)
    u, c, _ = OS.pde_solver_bpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 6, q)
    assert np.abs(u - ex).max() < 1e-9
    assert np.allclose(c[oc.N:], [0, 0, 0, 1, 0, 1], atol=1e-7)


def _robin_linear_case(nx, ny, beta_fn, a=0.4, b=1.3, c=-0.8):
    """Laplace with u = a + b x + c y: Robin on South and East (Nn = 0, so the
    Robin phi rows' normal index and the node's own normal agree), Dirichlet
    on North and West; g = beta u + grad u . n with each node's own normal."""
    ft = {"North": "d", "South": "r", "East": "r", "West": "d"}
    oc = OC.SquareCloud(nx, ny, ft, jitter=C.jitter(nx, ny, 0.2, 5))
    assert oc.Nn == 0 and oc.Nr > 0
    X = oc.sorted_nodes
    u = lambda p: a + b * p[..., 0] + c * p[..., 1]
    betas_by_facet, bc = {}, {}
    for f, t in oc.facet_types.items():
        ids = np.array(oc.facet_nodes[f])
        if t == "r":
            beta = beta_fn(X[ids])
            n = np.array([oc.outward_normals[i] for i in ids])
            betas_by_facet[f] = beta
            bc[f] = (beta * u(X[ids]) + b * n[:, 0] + c * n[:, 1], beta)
        else:
            bc[f] = u(X[ids])
    q = OA.assemble_q(oc, 0.0, bc)
    return oc, q, OA.robin_coefficients(oc, betas_by_facet), u(X)


def test_linear_reproduction_robin_pins_beta_order_and_normals():
    """deg = 1 reproduces a linear field exactly with Robin data whose beta
    varies node by node: the Robin rows (assembly.py:192-213, 302-315), the
    beta <-> node order of the sorted Robin coefficients (assembly.py:192,
    operators.py:512-541) and the outward normals all enter the check."""
    beta_fn = lambda p: 1.0 + p[:, 0] + 2.0 * p[:, 1]
    oc, q, rb, ex = _robin_linear_case(11, 12, beta_fn)
    u, c, _ = OS.pde_solver_bpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 3, q, rb)
    assert np.abs(u - ex).max() < 1e-9
    assert np.allclose(c[oc.N:], [0.4, 1.3, -0.8], atol=1e-8)
    # the same data with the betas in another node order is not reproduced
    u2, _, _ = OS.pde_solver_bpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 3, q, rb[::-1].copy())
    assert np.abs(u2 - ex).max() > 1e-3
    # nor with the Robin rows' normals flipped (g computed with the true normals)
    ft = {"North": "d", "South": "r", "East": "r", "West": "d"}
    oc_flip = OC.SquareCloud(11, 12, ft, jitter=C.jitter(11, 12, 0.2, 5))
    oc_flip.sorted_outward_normals = -oc_flip.sorted_outward_normals
    oc_flip.outward_normals = {k: -v for k, v in oc_flip.outward_normals.items()}
    u3, _, _ = OS.pde_solver_bpath(C.LAPLACIAN, oc_flip, R.POLYHARMONIC, 1.0, 3, q, rb)
    assert np.abs(u3 - ex).max() > 1e-3


def _periodic_case(n):
    """Laplace, W/E periodic, Dirichlet N/S, exact u = sin(2 pi x + 0.7)
    sinh(2 pi y) / sinh(2 pi) + y: harmonic, 1-periodic in x with values and
    x-derivatives that match across x = 0 and x = 1 (and do not vanish)."""
    ft = {"North": "d", "South": "d", "West": "p", "East": "p"}
    oc = OC.SquareCloud(n, n, ft)
    X = oc.sorted_nodes
    u = lambda p: np.sin(2 * np.pi * p[..., 0] + 0.7) * np.sinh(2 * np.pi * p[..., 1]) / np.sinh(2 * np.pi) + p[..., 1]
    bc = {f: (u(X[np.array(oc.facet_nodes[f])]) if t == "d" else np.zeros(len(oc.facet_nodes[f])))
          for f, t in oc.facet_types.items()}
    return oc, OA.assemble_q(oc, 0.0, bc), u(X)


def test_periodic_exact_solution_to_discretisation_accuracy():
    """An x-periodic harmonic field is reproduced with W/E periodic facets
    (pd rows u1 - u2 = 0, pn rows grad u1 . n1 - grad u2 . (-n2) = 0;
    assembly.py:215-267, 317-359) to discretisation accuracy: the error falls
    at second order as the cloud is refined (r^3, degree 2; max error 11% at
    12 x 12 and 2.8% at 24 x 24, against 3.5% and 0.84% with Dirichlet data on
    all four sides).  A periodic row that did not hold for the exact field
    (wrong pairing, sign or normal) would leave an O(1) error that does not
    fall."""
    errs = []
    for n in (12, 24):
        oc, q, ex = _periodic_case(n)
        assert sum(oc.Np) == 2 * (n - 2)
        u, _, _ = OS.pde_solver_bpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 6, q)
        errs.append(np.abs(u - ex).max() / np.abs(ex).max())
        # the periodic rows hold for the computed field at the paired nodes
        s = oc.Ni + oc.Nd
        h = sum(oc.Np) // 2
        assert np.abs(u[s:s + h] - u[s + h:s + 2 * h]).max() < 1e-8
    assert errs[1] < 0.05, errs
    assert errs[1] < errs[0] / 3.0, errs


def test_value_at_nodes_includes_self_term():
    """value() sums over all centres: u_i for r^3 (phi(0)=0), u_i + lambda_i for
    the Gaussian (phi(0)=1) -- operators.py:131-135."""
    oc = OC.SquareCloud(8, 7, C.README_FACETS)
    f = np.sin(oc.sorted_nodes[:, 0]) + oc.sorted_nodes[:, 1]
    for kind, param in [(R.POLYHARMONIC, 1.0), (R.GAUSSIAN, 3.0)]:
        c = OA.core_compute_coefficients(f, oc, kind, param, 3)
        v = OS.value(oc.sorted_nodes, c, oc.sorted_nodes, kind, param)
        expect = f + R.phi_zero(kind, param) * c[:oc.N]
        assert np.abs(v - expect).max() < 1e-9 * max(1, np.abs(expect).max())


def test_gaussian_zero_diagonal_conditioning():
    """The zero self-interaction (every support omits the node itself,
    ``cloud.py:110-111``) keeps the Gaussian system well posed."""
    oc = OC.SquareCloud(20, 20, {"North": "d", "South": "d", "East": "p1", "West": "p1"})
    A = OA.assemble_A(oc, R.GAUSSIAN, 5.0, 3)
    assert np.linalg.cond(A) < 1e8
    A_text = A.copy()
    A_text[np.arange(oc.N), np.arange(oc.N)] = 1.0
    assert np.linalg.cond(A_text) > 1e4 * np.linalg.cond(A)


def test_renumbering_permutation_like_reference_test():
    """updes/tests/test_interpolation.py:60-63 on the oracle cloud."""
    ft1 = {"South": "n", "West": "d", "North": "n", "East": "d"}
    ft2 = {"South": "d", "West": "n", "North": "d", "East": "n"}
    c1, c2 = OC.SquareCloud(8, 8, ft1), OC.SquareCloud(8, 8, ft2)
    field = np.arange(c1.N)
    new = OC.interpolate_field(field, c1, c2)
    assert np.allclose(field[:c1.Ni], new[:c2.Ni], atol=1e-12)


def test_square_cloud_counts_of_benchmark_configs():
    """Row-class sizes quoted for the five workloads."""
    cases = [((30, 20), C.README_FACETS, (504, 66, 30, [])),
             ((100, 100), {"North": "d", "South": "d", "East": "d", "West": "d"}, (9604, 396, 0, [])),
             ((60, 60), {"North": "d", "South": "d", "East": "d", "West": "d"}, (3364, 236, 0, [])),
             ((80, 80), {"North": "d", "South": "d", "West": "p1", "East": "p1"}, (6084, 160, 0, [156])),
             ((200, 200), {"North": "d", "South": "n", "East": "d", "West": "d"}, (39204, 596, 200, []))]
    for (nx, ny), ft, (ni, nd, nn, np_) in cases:
        if nx * ny > 10000:
            continue   # the oracle cloud is a Python loop; the drop-in cloud test covers 200x200
        oc = OC.SquareCloud(nx, ny, ft)
        assert (oc.Ni, oc.Nd, oc.Nn, oc.Np) == (ni, nd, nn, np_)


# ---- planted mistakes ------------------------------------------------------------
def _planted(monkeypatch, target, name, wrapper):
    orig = getattr(target, name)
    monkeypatch.setattr(target, name, wrapper(orig))


def test_planted_dropped_factor_in_laplacian_is_caught(monkeypatch):
    def wrap(orig):
        def f(kind, param, x, y, cx, cy):
            o = orig(kind, param, x, y, cx, cy)
            o[R.LAP] = o[R.LAP] - o[R.DYY]          # dropped d2/dy2 term
            return o
        return f
    _planted(monkeypatch, R, "rbf_ops", wrap)
    with pytest.raises(AssertionError):
        test_closed_forms_match_sympy(R.POLYHARMONIC, 1.0)


def test_planted_wrong_normal_sign_is_caught(monkeypatch):
    orig = OA.assemble_bd_Phi_P

    def wrong(cloud, kind, param, M, robin_betas=None):
        bdPhi, bdP = orig(cloud, kind, param, M, robin_betas)
        s = slice(cloud.Nd, cloud.Nd + cloud.Nn)
        bdPhi[s] *= -1.0
        return bdPhi, bdP
    monkeypatch.setattr(OA, "assemble_bd_Phi_P", wrong)
    with pytest.raises(AssertionError):
        test_oracle_matches_bruteforce(1)


def test_planted_self_interaction_is_caught(monkeypatch):
    orig = OA.assemble_Phi

    def textbook(cloud, kind, param):
        P = orig(cloud, kind, param)
        P[np.arange(cloud.N), np.arange(cloud.N)] = R.phi_zero(kind, param)
        return P
    monkeypatch.setattr(OA, "assemble_Phi", textbook)
    with pytest.raises(AssertionError):
        test_oracle_matches_bruteforce(2)


def test_planted_periodic_pairing_is_caught(monkeypatch):
    orig = OA.assemble_bd_Phi_P

    def swapped(cloud, kind, param, M, robin_betas=None):
        bdPhi, bdP = orig(cloud, kind, param, M, robin_betas)
        nb = cloud.Nd + cloud.Nn + cloud.Nr
        half = sum(cloud.Np) // 2
        bdPhi[nb:nb + half] *= -1.0          # phi(x_i2) - phi(x_i1)
        return bdPhi, bdP
    monkeypatch.setattr(OA, "assemble_bd_Phi_P", swapped)
    with pytest.raises(AssertionError):
        test_oracle_matches_bruteforce(3)


def _quirky_rule_closed_forms(Nx, Ny):
    """Integral of f = 1 and f = x under integrate_field's rule (operators.py:381-450),
    summed by hand: weights 1 / 1/2 / 1/4 times the cell area; interior points are
    consecutive pairs of the flattened meshgrid stack, so their x-coordinates are
    the even entries of concat(gx.ravel(), gy.ravel())."""
    hx, hy = 1.0 / (Nx - 1), 1.0 / (Ny - 1)
    area = hx * hy
    ones = (Nx - 1) * (Ny - 1) + 0.5 * (2 * (Ny - 2) + 2 * (Nx - 2)) + 0.25 * 4
    qx = [hx + i * (1 - 2 * hx) / (Nx - 2) for i in range(Nx - 1)]
    qy = [hy + j * (1 - 2 * hy) / (Ny - 2) for j in range(Ny - 1)]
    flat = [qx[i] for j in range(Ny - 1) for i in range(Nx - 1)] + [qy[j] for j in range(Ny - 1) for i in range(Nx - 1)]
    inner_x = sum(flat[0::2])
    ex = [hx + i * (1 - 2 * hx) / (Nx - 3) for i in range(Nx - 2)]
    # left x=0, right x=1 (Ny-2 points each), bottom/top at ex; corners (0,0),(1,0),(1,0),(1,1)
    xs = inner_x + 0.5 * ((Ny - 2) * 1.0 + 2 * sum(ex)) + 0.25 * 3.0
    return ones * area, xs * area


def test_integrate_field_rule_closed_forms():
    """f = 1 (pins the weights) and f = x (pins the point layout, including the
    top-left corner evaluated at (1, 0)) are reproduced exactly by deg >= 1."""
    Nx, Ny = 9, 7
    oc = OC.SquareCloud(Nx, Ny, {"North": "d", "South": "d", "East": "d", "West": "d"})
    one, xint = _quirky_rule_closed_forms(Nx, Ny)
    x = oc.sorted_nodes[:, 0]
    for kind, param in [(R.POLYHARMONIC, 1.0), (R.THIN_PLATE, 1.0)]:
        c1 = OA.core_compute_coefficients(np.ones(oc.N), oc, kind, param, 3)
        cx = OA.core_compute_coefficients(x, oc, kind, param, 3)
        assert abs(OS.integrate_field(c1, oc, kind, param) - one) < 1e-10
        assert abs(OS.integrate_field(cx, oc, kind, param) - xint) < 1e-10
    # the rule is not a plain midpoint rule: a corrected top-left corner differs
    pts, w = OS.integrate_field_points(Nx, Ny)
    assert np.allclose(pts[-2], [1.0, 0.0])


def test_integral_like_reference_test():
    """updes/tests/test_integrals.py: r^11 (a=5), max_degree 3, 12 x 12, field
    x^2/(1+y^2); the reference asserts |I - pi/12| < 0.1 (test_integrals.py:83-84)."""
    oc = OC.SquareCloud(12, 12, {"North": "d", "South": "d", "East": "d", "West": "d"})
    xy = oc.sorted_nodes
    s = xy[:, 0] ** 2 / (1 + xy[:, 1] ** 2)
    c = OA.core_compute_coefficients(s, oc, R.POLYHARMONIC, 5.0, 10)
    v = OS.value(xy, c, xy, R.POLYHARMONIC, 5.0)
    # kappa(A) ~ 2e12 for r^11 with degree-3 monomials: reconstruction to ~1e-6
    assert np.mean(np.abs(v - s)) < 1e-5
    val = OS.integrate_field(c, oc, R.POLYHARMONIC, 5.0)
    assert abs(val - np.pi / 12) < 1e-1
