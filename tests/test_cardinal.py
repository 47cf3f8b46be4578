"""Host side of the FP32 preconditioner's local cardinal basis: stencils, the
reverse stencil table, and the method itself on the CPU (NumPy restatement
of the kernels, tests/_cardinal_ref.py): cardinality, moment conditions, and
GMRES on K preconditioned by the FP32 LU of K in that basis reaching FP64
accuracy in a few iterations where the FP32 LU of K itself does not."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import assembly as OA, cloud as OC, rbf as R
from tests import _cardinal_ref as CR
from tests import _cases as C


class _DC:
    def __init__(self, nodes):
        import torch
        self.nodes = torch.as_tensor(nodes)
        self.N = nodes.shape[0]


def _stencil(nodes, k, g=3):
    from updes_amd import engine as E
    return E.cardinal_stencil(_DC(nodes), k, g)


def test_stencil_is_the_k_nearest_nodes_in_ascending_order():
    oc = OC.SquareCloud(15, 12, C.README_FACETS, jitter=C.jitter(15, 12, 0.2, 1))
    X = oc.sorted_nodes
    k = 20
    st = _stencil(X, k, g=0)
    assert st.shape == (oc.N, k) and st.dtype == np.int32
    d = np.sqrt(((X[:, None, :] - X[None, :, :]) ** 2).sum(-1))
    for i in range(oc.N):
        assert np.all(np.diff(st[i]) > 0)
        assert i in st[i]
        # same distances as a brute-force k-nearest query (ties may pick either node)
        np.testing.assert_allclose(np.sort(d[i, st[i]]), np.sort(d[i])[:k], rtol=0, atol=1e-15)


def test_stencil_with_special_points():
    """Every stencil holds the special points (the nodes nearest a 3 x 3 grid
    over the bounding box) and then the node's nearest other nodes."""
    from updes_amd import engine as E
    oc = OC.SquareCloud(15, 12, C.README_FACETS, jitter=C.jitter(15, 12, 0.2, 1))
    X = oc.sorted_nodes
    sp = E.special_points(X, 3)
    assert sp.size == 9
    grid = np.array([[x, y] for x in (1 / 6, 1 / 2, 5 / 6) for y in (1 / 6, 1 / 2, 5 / 6)])
    dist = np.sqrt(((X[None, :, :] - grid[:, None, :]) ** 2).sum(-1))
    assert set(sp.tolist()) == set(np.argmin(dist, axis=1).tolist())
    k = 20
    st = _stencil(X, k, g=3)
    assert st.shape == (oc.N, k + 9)
    d = np.sqrt(((X[:, None, :] - X[None, :, :]) ** 2).sum(-1))
    for i in range(oc.N):
        row = set(st[i].tolist())
        assert np.all(np.diff(st[i]) > 0) and i in row and set(sp.tolist()) <= row
        rest = np.array(sorted(row - set(sp.tolist())))
        others = np.setdiff1d(np.arange(oc.N), sp)
        # the non-special entries are the nearest non-special nodes
        np.testing.assert_allclose(np.sort(d[i, rest]), np.sort(d[i, others])[:rest.size], rtol=0, atol=1e-15)


def test_reverse_stencil_table():
    from updes_amd import engine as E
    rng = np.random.default_rng(0)
    N, k = 50, 7
    st = np.sort(np.stack([rng.choice(N, k, replace=False) for _ in range(N)]), axis=1).astype(np.int32)
    ptr, idx = E.reverse_stencil(st)
    assert ptr[0] == 0 and ptr[-1] == N * k and np.all(np.diff(ptr) >= 0)
    for l in range(N):
        e = idx[ptr[l]:ptr[l + 1]]
        assert np.all(st.reshape(-1)[e] == l)
        assert np.all(np.diff(e // k) > 0)       # ascending j
    assert sorted(idx.tolist()) == list(range(N * k))


@pytest.mark.parametrize("kind,param,M", [(R.POLYHARMONIC, 1.0, 6), (R.POLYHARMONIC, 1.0, 3), (R.GAUSSIAN, 5.0, 3),
                                          (R.THIN_PLATE, 1.0, 6)])
def test_cardinal_functions_are_cardinal_and_in_the_rbf_space(kind, param, M):
    oc = OC.SquareCloud(12, 11, C.FACET_SETS[1], jitter=C.jitter(12, 11, 0.2, 2))
    st = _stencil(oc.sorted_nodes, 16)
    beta, alpha = CR.coefficients(oc.sorted_nodes, st, kind, param, M)
    Cm = CR.basis_matrix(st, beta, alpha)
    A = OA.assemble_A(oc, kind, param, M)
    Psi = A[:oc.N] @ Cm                       # Psi[i, j] = psi_j(x_i)
    for j in range(oc.N):
        np.testing.assert_allclose(Psi[st[j], j], (st[j] == j).astype(float), atol=1e-7)
    # moment conditions P^T beta_j = 0 (every column lies in the RBF space of the solve)
    mom = A[oc.N:, :oc.N] @ Cm[:oc.N]
    assert np.abs(mom).max() <= 1e-12 * np.abs(Cm[:oc.N]).max()


def _gmres_iters(apply_pair, b, tol, max_it, restart):
    """Flexible restarted GMRES; apply_pair(v) -> (z = M^-1 v, K z).  Returns
    (x, iterations, true relative residual)."""
    n = b.shape[0]
    x = np.zeros(n)
    r = b.copy()
    its = 0
    bn = np.linalg.norm(b)
    while its < max_it:
        beta = np.linalg.norm(r)
        if beta <= tol * bn:
            break
        V = np.zeros((n, restart + 1))
        Z = np.zeros((n, restart))
        H = np.zeros((restart + 1, restart))
        V[:, 0] = r / beta
        for j in range(restart):
            Z[:, j], w = apply_pair(V[:, j])
            for _ in range(2):
                h = V[:, :j + 1].T @ w
                w -= V[:, :j + 1] @ h
                H[:j + 1, j] += h
            H[j + 1, j] = np.linalg.norm(w)
            V[:, j + 1] = w / H[j + 1, j]
            its += 1
            e1 = np.zeros(j + 2)
            e1[0] = beta
            y = np.linalg.lstsq(H[:j + 2, :j + 1], e1, rcond=None)[0]
            if np.linalg.norm(H[:j + 2, :j + 1] @ y - e1) <= 0.1 * tol * bn or its >= max_it:
                break
        x = x + Z[:, :j + 1] @ y
        r = b - apply_pair.K @ x
    return x, its, np.linalg.norm(b - apply_pair.K @ x) / bn


def test_fp32_lu_in_the_cardinal_basis_preconditions_gmres_to_fp64_accuracy():
    """Poisson, r^3, degree 2, jittered 40 x 40 with a Neumann facet: FP32 LU of
    K~ = diag(rs) D C as the preconditioner of the row-scaled system reaches
    rel-L2 1e-10 against the FP64 LU in at most 8 iterations; the FP32 LU of
    K itself needs more."""
    nx = 40
    ft = {"North": "d", "South": "n", "East": "d", "West": "d"}
    oc = OC.SquareCloud(nx, nx, ft, jitter=C.jitter(nx, nx, 0.2, 3))
    N, M = oc.N, 6
    K = OA.assemble_K(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, M)
    x, y = oc.sorted_nodes.T
    q = np.where(np.arange(N) < oc.Ni, -2 * np.pi ** 2 * np.sin(np.pi * x) * np.sin(np.pi * y), y + 0.5 * x)
    b = np.concatenate([q, np.zeros(M)])
    c64 = sla.lu_solve(sla.lu_factor(K), b)
    st = _stencil(oc.sorted_nodes, 20)
    beta, alpha = CR.coefficients(oc.sorted_nodes, st, R.POLYHARMONIC, 1.0, M)
    Cm = CR.basis_matrix(st, beta, alpha)
    rs, Kt = CR.form(K, Cm, st)
    lu = sla.lu_factor(Kt.astype(np.float32))
    lu = (lu[0].astype(np.float64), lu[1])
    P = K[N:, :N].T
    G = np.linalg.inv(P.T @ P)

    # row-scaled system diag(s) K x = diag(s) b, s = [rs; 1]: its residual v
    # feeds K~ directly (diag(rs) (v / rs) = v)
    s = np.concatenate([rs, np.ones(M)])
    Ks = s[:, None] * K

    def card(v):
        yv = sla.lu_solve(lu, v[:N])
        z = Cm @ yv
        z[:N] += P @ (G @ v[N:])
        return z, Ks @ z
    card.K = Ks
    c, its, rel = _gmres_iters(card, s * b, 1e-11, 40, 40)
    A = OA.assemble_A(oc, R.POLYHARMONIC, 1.0, M)
    u64, u = A[:N] @ c64, A[:N] @ c
    err = np.linalg.norm(u - u64) / np.linalg.norm(u64)
    assert its <= 8 and rel <= 1e-11 and err <= 1e-10, (its, rel, err)

    lu_k = sla.lu_factor(K.astype(np.float32))
    lu_k = (lu_k[0].astype(np.float64), lu_k[1])

    def plain(v):
        z = sla.lu_solve(lu_k, v)
        return z, K @ z
    plain.K = K
    _, its_k, _ = _gmres_iters(plain, b, 1e-11, 40, 40)
    assert its_k > its
