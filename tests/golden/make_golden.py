"""Writes the golden vectors under tests/golden from the pinned oracle.

The reference's own tests hold no vectors for pde_solver (their tightest
solve tolerance is 1e-2), and the reference (JAX) cannot run here,
so these fixtures are produced by the oracle after it has passed its pins
(tests/test_oracle_pins.py: SymPy/mpmath closed forms, an independent
mpmath brute force of the reference loops, invariants).  Run:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import assembly as OA, cloud as OC, rbf as R, solver as OS  # noqa: E402
from tests import _cases as C  # noqa: E402
from tests.test_oracle_pins import BRUTE_CASES, _oracle_case  # noqa: E402


def readme_laplace():
    ft = C.README_FACETS
    oc = OC.SquareCloud(30, 20, ft)
    bc = {f: (np.sin(np.pi * oc.sorted_nodes[oc.facet_nodes[f], 0]) if f == "North"
              else np.zeros(len(oc.facet_nodes[f]))) for f in ft}
    q = OA.assemble_q(oc, 0.0, bc)
    u, c, B = OS.pde_solver_bpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 3, q)
    np.savez_compressed(os.path.join(HERE, "readme_laplace.npz"), q=q, vals=u, coeffs=c,
                        B_diag=np.diag(B), B_row0=B[0], nodes=oc.sorted_nodes)


def small_cases():
    out = {}
    for i in range(len(BRUTE_CASES)):
        oc, kind, param, M, desc, rb, q = _oracle_case(i)
        u, c, _ = OS.pde_solver_bpath(desc, oc, kind, param, M, q, rb)
        K = OA.assemble_K(desc, oc, kind, param, M, rb)
        A = OA.assemble_A(oc, kind, param, M)
        out["case%d_K" % i] = K
        out["case%d_A" % i] = A
        out["case%d_q" % i] = q
        out["case%d_vals" % i] = u
        out["case%d_coeffs" % i] = c
        out["case%d_nodes" % i] = oc.sorted_nodes
        out["case%d_beta" % i] = rb if rb is not None else np.zeros(0)
    np.savez_compressed(os.path.join(HERE, "small_cases.npz"), **out)


def eval_fixture():
    oc = OC.SquareCloud(10, 9, C.README_FACETS, jitter=C.jitter(10, 9, 0.1, 3))
    rng = np.random.default_rng(11)
    coeffs = rng.standard_normal(oc.N + 6)
    pts = np.concatenate([rng.uniform(0, 1, (17, 2)), oc.sorted_nodes[:5]])
    out = {"centers": oc.sorted_nodes, "coeffs": coeffs, "points": pts}
    for kind, param in C.KINDS:
        out["ops_%d_%g" % (kind, param)] = OS.eval_ops(pts, coeffs, oc.sorted_nodes, kind, param)
    np.savez_compressed(os.path.join(HERE, "eval_ops.npz"), **out)


def poisson_100():
    """Poisson 100 x 100, interior jitter U(-0.2h, 0.2h) (seed 0), r^3, degree 2,
    all Dirichlet from u = sin(pi x) sin(pi y); FP64 augmented-system solve."""
    nx = 100
    ft = {"North": "d", "South": "d", "East": "d", "West": "d"}
    oc = OC.SquareCloud(nx, nx, ft, jitter=C.jitter(nx, nx, 0.2, 0))
    x, y = oc.sorted_nodes.T
    ex = np.sin(np.pi * x) * np.sin(np.pi * y)
    q = np.where(np.arange(oc.N) < oc.Ni, -2 * np.pi ** 2 * ex, ex)
    u, c, _ = OS.pde_solver_kpath(C.LAPLACIAN, oc, R.POLYHARMONIC, 1.0, 6, q)
    np.savez_compressed(os.path.join(HERE, "poisson_100.npz"), nodes=oc.sorted_nodes, q=q, vals=u, exact=ex)


def advdiff_80(steps=5):
    """Advection-diffusion 80 x 80, periodic West/East, Dirichlet 0 North/South,
    degree 1, dt=1e-3, nu=0.01, v=(1, 0.5); rhs of each implicit Euler step =
    value(x_i, coeffs(u^n)).  Kernel r^3 (the reference demo's own,
    demos/Advection/01_adv_diff_periodic.py:33): with the gaussian eps=5 the
    periodic system is numerically singular under the reference's semantics
    (cond(B) = 2.8e17 at 20 x 20, 1.7e19 at 80 x 80), so no FP64 solution of it
    is reproducible.  Stored with the reference (B, QR) solve of each step."""
    nx = 80
    ft = {"North": "d", "South": "d", "West": "p1", "East": "p1"}
    oc = OC.SquareCloud(nx, nx, ft)
    xy = oc.sorted_nodes
    un = np.exp(-((xy[:, 0] - 0.3) ** 2 + (xy[:, 1] - 0.5) ** 2) / (2 * 0.1 ** 2))
    out = {"nodes": xy, "u0": un.copy()}
    desc = C.adv_diff_desc()
    bc = {f: np.zeros(len(oc.facet_nodes[f])) for f in ft}
    for s in range(steps):
        cn = OS.field_coefficients(un, oc, R.POLYHARMONIC, 1.0, 3)
        qi = OS.value(xy[:oc.Ni], cn, xy, R.POLYHARMONIC, 1.0)
        q = OA.assemble_q(oc, qi, bc)
        ub, _, _ = OS.pde_solver_bpath(desc, oc, R.POLYHARMONIC, 1.0, 3, q)
        uk, _, _ = OS.pde_solver_kpath(desc, oc, R.POLYHARMONIC, 1.0, 3, q)
        out["u%d" % (s + 1)] = ub.copy()
        out["spread%d" % (s + 1)] = np.linalg.norm(ub - uk) / np.linalg.norm(ub)
        un = ub
    out["cond_A"] = np.array(np.linalg.cond(OA.assemble_A(oc, R.POLYHARMONIC, 1.0, 3)))
    np.savez_compressed(os.path.join(HERE, "advdiff_80.npz"), **out)


if __name__ == "__main__":
    readme_laplace()
    small_cases()
    eval_fixture()
    poisson_100()
    advdiff_80()
    print("written to", HERE)
