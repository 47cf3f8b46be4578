"""Local supports (``support_size < N``; updes/cloud.py:83-112 and the support
scatters of updes/assembly.py:31-32, 127-128, 171-258), on the CPU:

* the cloud's neighbour table (scikit-learn BallTree, as the reference queries
  it) against the oracle's brute-force nearest-neighbour statement: identical,
  nearest first, on jittered clouds (no equidistant neighbours); on regular
  grids, where neighbours tie, a nearest set (same distances);
* argument checks of ``support_size`` as in ``cloud.py:97-100``;
* the oracle's masked assembly against the dense oracle (itself pinned by the
  golden vectors): equal on the support pattern, which the test rebuilds row by
  row from the reference's row rules, and zero elsewhere;
* the oracle's augmented solve against its restatement of the reference's
  B-path (inv(A), D inv(A), QR) with local supports.
"""
import numpy as np
import pytest

from oracle import assembly as OA, cloud as OC, rbf as R, solver as OS
from tests import _cases as C


def _orig_supports(uc):
    """The product cloud's table by original id (for an explicit oracle table)."""
    tab = uc.sorted_local_supports.numpy()
    return uc._order[tab[uc._new_id]]


@pytest.mark.parametrize("fs", range(len(C.FACET_SETS)))
@pytest.mark.parametrize("k", [2, 9, 20])
def test_support_table_matches_brute_force_knn(fs, k):
    import updes_amd as u
    nx, ny = 11, 9
    J = C.jitter(nx, ny, 0.3, seed=10 + fs)
    uc = u.SquareCloud(Nx=nx, Ny=ny, facet_types=C.FACET_SETS[fs], jitter=J, support_size=k)
    oc = OC.SquareCloud(nx, ny, C.FACET_SETS[fs], jitter=J, support_size=k)
    assert uc.support_size == k
    tab = uc.sorted_local_supports.numpy()
    assert tab.shape == (uc.N, k - 1)
    assert not np.any(tab == np.arange(uc.N)[:, None])
    # d/n/r nodes are not jittered, so boundary neighbours can still tie: the
    # sets must agree wherever the (k-1)-th and k-th distances differ, and the
    # distances everywhere; both lists are nearest first
    x = uc.sorted_nodes.numpy()
    untied = 0
    for i in range(uc.N):
        d2 = ((x - x[i]) ** 2).sum(1)
        assert np.array_equal(d2[tab[i]], d2[oc.supports[i]])
        assert np.all(np.diff(d2[tab[i]]) >= 0)
        ds = np.sort(np.delete(d2, i))
        if k - 1 < uc.N - 1 and ds[k - 2] < ds[k - 1]:
            assert np.array_equal(np.sort(tab[i]), np.sort(oc.supports[i]))
            untied += 1
    assert untied >= uc.N // 2
    # dict view keyed by new id, nearest first
    ls = uc.local_supports
    assert ls[3] == [int(j) for j in tab[3]]


@pytest.mark.parametrize("k", [3, 5, 6, 10, 13])
def test_support_table_on_regular_grid_is_a_nearest_set(k):
    import updes_amd as u
    uc = u.SquareCloud(Nx=10, Ny=8, facet_types=C.FACET_SETS[3], support_size=k)
    x = uc.sorted_nodes.numpy()
    tab = uc.sorted_local_supports.numpy()
    for i in range(uc.N):
        d2 = ((x - x[i]) ** 2).sum(1)
        others = np.delete(np.arange(uc.N), i)
        assert len(set(tab[i])) == k - 1 and i not in tab[i]
        assert np.array_equal(np.sort(d2[tab[i]]), np.sort(d2[others])[:k - 1])


def test_support_size_arguments():
    import updes_amd as u
    ft = C.FACET_SETS[0]
    for s in ("max", None, 30):
        c = u.SquareCloud(Nx=6, Ny=5, facet_types=ft, support_size=s)
        assert c.support_size == 30
        # every other node, nearest first, built on demand; no table for the kernels
        tab = c.sorted_local_supports.numpy()
        assert tab.shape == (30, 29)
        assert np.array_equal(np.sort(tab, axis=1), np.array([np.delete(np.arange(30), i) for i in range(30)]))
        from updes_amd import engine as E
        assert E.support_table(c) is None
    with pytest.raises(AssertionError):
        u.SquareCloud(Nx=6, Ny=5, facet_types=ft, support_size=0)
    with pytest.raises(AssertionError):
        u.SquareCloud(Nx=6, Ny=5, facet_types=ft, support_size=31)
    c = u.SquareCloud(Nx=6, Ny=5, facet_types=ft, support_size=1)
    assert c.sorted_local_supports.shape == (30, 0)


def _pattern(oc, M):
    """Nonzero pattern of K by the reference's row rules, row by row."""
    N, Ni, Nd, Nn, Nr = oc.N, oc.Ni, oc.Nd, oc.Nn, oc.Nr
    n = N + M
    pat = np.zeros((n, n), dtype=bool)
    for r in range(Ni + Nd + Nn + Nr):          # one node per row (assembly.py:127, 171-210)
        pat[r, oc.supports[r]] = True
    start, jump, half = Ni + Nd + Nn + Nr, 0, sum(oc.Np) // 2
    for nbp in oc.Np:                           # periodic pairs (assembly.py:215-258)
        nc = nbp // 2
        for t in range(nc):
            i1, i2 = start + t, start + t + nc
            for r in (i1 - Ni - jump + Ni, i1 - Ni - jump + Ni + half):
                pat[r, oc.supports[i1]] = True
                pat[r, oc.supports[i2]] = True
        start += nbp
        jump += nc
    pat[:N, N:] = True                          # monomial columns
    pat[N:, :N] = True                          # P^T rows
    return pat


@pytest.mark.parametrize("fs", range(len(C.FACET_SETS)))
def test_masked_oracle_is_dense_oracle_on_the_support(fs):
    nx, ny, k, M = 9, 8, 11, 6
    facets = C.FACET_SETS[fs]
    J = C.jitter(nx, ny, 0.2, seed=fs)
    oc_s = OC.SquareCloud(nx, ny, facets, jitter=J, support_size=k)
    oc_d = OC.SquareCloud(nx, ny, facets, jitter=J)
    betas = {f: np.linspace(0.5, 1.5, len(oc_d.facet_nodes[f])) for f, t in oc_d.facet_types.items() if t == "r"}
    rb = OA.robin_coefficients(oc_d, betas)
    desc = C.adv_diff_desc(0.1, 0.3, (1.0, -0.5))
    Ks = OA.assemble_K(desc, oc_s, R.POLYHARMONIC, 1.0, M, rb)
    Kd = OA.assemble_K(desc, oc_d, R.POLYHARMONIC, 1.0, M, rb)
    pat = _pattern(oc_s, M)
    assert np.array_equal(Ks[pat], Kd[pat])
    assert not np.any(Ks[~pat])
    A_s = OA.assemble_A(oc_s, R.POLYHARMONIC, 1.0, M)
    A_d = OA.assemble_A(oc_d, R.POLYHARMONIC, 1.0, M)
    m = np.zeros_like(A_s, dtype=bool)
    m[np.repeat(np.arange(oc_s.N), k - 1), oc_s.supports.ravel()] = True
    m[:oc_s.N, oc_s.N:] = True
    m[oc_s.N:, :oc_s.N] = True
    assert np.array_equal(A_s[m], A_d[m]) and not np.any(A_s[~m])
    # r^3 is nonzero between distinct nodes: Phi has exactly k-1 nonzeros a row
    assert np.all((A_s[:oc_s.N, :oc_s.N] != 0).sum(1) == k - 1)


@pytest.mark.parametrize("fs", [0, 1, 2, 3, 5])
@pytest.mark.parametrize("k", [10, 25])
def test_oracle_kpath_equals_bpath_with_local_supports(fs, k):
    nx = ny = 14
    J = C.jitter(nx, ny, 0.2, seed=fs)
    oc = OC.SquareCloud(nx, ny, C.FACET_SETS[fs], jitter=J, support_size=k)
    bc = {f: (np.zeros(len(ids)) if oc.facet_types[f][0] == "p" else np.cos(np.arange(len(ids))))
          for f, ids in oc.facet_nodes.items()}
    q = OA.assemble_q(oc, np.ones(oc.Ni), bc)
    rb = OA.robin_coefficients(oc, {f: np.ones(len(oc.facet_nodes[f]))
                                    for f, t in oc.facet_types.items() if t == "r"})
    ub, cb, _ = OS.pde_solver_bpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 3, q, rb)
    uk, ck, _ = OS.pde_solver_kpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 3, q, rb)
    assert np.linalg.norm(uk - ub) <= 1e-9 * np.linalg.norm(ub)


def test_explicit_table_reproduces_the_cloud_on_a_regular_grid():
    """Ties: the oracle fed the BallTree table builds the same supports."""
    import updes_amd as u
    uc = u.SquareCloud(Nx=8, Ny=7, facet_types=C.FACET_SETS[4], support_size=6)
    oc = OC.SquareCloud(8, 7, C.FACET_SETS[4], support_size=6, supports=_orig_supports(uc))
    assert np.array_equal(oc.supports, uc.sorted_local_supports.numpy())
