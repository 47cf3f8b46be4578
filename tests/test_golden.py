"""The oracle reproduces the committed golden vectors (CPU)."""
import os

import numpy as np
import pytest

from oracle import assembly as OA, cloud as OC, rbf as R, solver as OS
from tests import _cases as C
from tests.test_oracle_pins import BRUTE_CASES, _oracle_case

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_readme_laplace_golden():
    g = np.load(os.path.join(G, "readme_laplace.npz"))
    ft = C.README_FACETS
    oc = OC.SquareCloud(30, 20, ft)
    assert np.array_equal(oc.sorted_nodes, g["nodes"])
    u, c, B = OS.pde_solver_bpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 3, g["q"])
    assert np.linalg.norm(u - g["vals"]) <= 1e-12 * np.linalg.norm(g["vals"])
    # golden B rows: Dirichlet diagonal entries are ~1 (identity rows)
    assert np.abs(g["B_diag"][oc.Ni:oc.Ni + oc.Nd] - 1).max() < 1e-8


@pytest.mark.parametrize("case", range(len(BRUTE_CASES)))
def test_small_cases_golden(case):
    g = np.load(os.path.join(G, "small_cases.npz"))
    oc, kind, param, M, desc, rb, q = _oracle_case(case)
    assert np.array_equal(g["case%d_q" % case], q)
    K = OA.assemble_K(desc, oc, kind, param, M, rb)
    assert np.abs(K - g["case%d_K" % case]).max() <= 1e-14 * np.abs(K).max()
    u, c, _ = OS.pde_solver_bpath(desc, oc, kind, param, M, q, rb)
    assert np.linalg.norm(u - g["case%d_vals" % case]) <= 1e-12 * np.linalg.norm(u)


def test_eval_golden():
    g = np.load(os.path.join(G, "eval_ops.npz"))
    for kind, param in C.KINDS:
        ref = g["ops_%d_%g" % (kind, param)]
        got = OS.eval_ops(g["points"], g["coeffs"], g["centers"], kind, param)
        assert np.abs(got - ref).max() <= 1e-13 * max(1, np.abs(ref).max())
