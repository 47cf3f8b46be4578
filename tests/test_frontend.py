"""CPU tests of the drop-in front end: cloud numbering, operator lowering and
boundary-condition glue (no GPU needed)."""
import functools

import numpy as np
import pytest
import torch

import updes_amd as u
from updes_amd import _lib as L
from updes_amd._symbolic import NotLowerable
from updes_amd.operators import trace_diff_operator
from oracle import cloud as OC
from tests import _cases as C


@pytest.mark.parametrize("fs", range(len(C.FACET_SETS)))
@pytest.mark.parametrize("nx,ny", [(5, 4), (9, 8), (30, 20)])
def test_cloud_matches_oracle(fs, nx, ny):
    ft = C.FACET_SETS[fs]
    J = C.jitter(nx, ny, 0.2, fs)
    oc = OC.SquareCloud(nx, ny, ft, jitter=J)
    uc = u.SquareCloud(Nx=nx, Ny=ny, facet_types=ft, jitter=J)
    assert (uc.N, uc.Ni, uc.Nd, uc.Nn, uc.Nr, uc.Np) == (oc.N, oc.Ni, oc.Nd, oc.Nn, oc.Nr, oc.Np)
    assert uc.facet_types == oc.facet_types
    assert uc.facet_nodes == oc.facet_nodes
    assert uc.renumbering_map == oc.renumbering_map
    assert uc.node_types == oc.node_types
    assert np.array_equal(uc.sorted_nodes.cpu().numpy(), oc.sorted_nodes)
    assert np.array_equal(uc.sorted_outward_normals.cpu().numpy(), oc.sorted_outward_normals)


def test_large_cloud_counts_fast():
    import time
    t = time.time()
    c = u.SquareCloud(Nx=200, Ny=200, facet_types={"North": "d", "South": "n", "East": "d", "West": "d"})
    assert time.time() - t < 5.0
    assert (c.Ni, c.Nd, c.Nn, c.Np) == (39204, 596, 200, [])
    c = u.SquareCloud(Nx=80, Ny=80, facet_types={"North": "d", "South": "d", "West": "p1", "East": "p1"})
    assert (c.Ni, c.Nd, c.Np) == (6084, 160, [156])


def test_interpolation_like_reference_test():
    """updes/tests/test_interpolation.py:60-63 with the drop-in clouds."""
    c1 = u.SquareCloud(Nx=8, Ny=8, facet_types={"South": "n", "West": "d", "North": "n", "East": "d"})
    c2 = u.SquareCloud(Nx=8, Ny=8, facet_types={"South": "d", "West": "n", "North": "d", "East": "n"})
    field = torch.arange(0, c1.N)
    new = u.interpolate_field(field, c1, c2)
    assert torch.allclose(field[:c1.Ni].double(), new[:c2.Ni].double(), atol=1e-12)


def _x(n=5):
    return torch.stack([torch.linspace(0.1, 0.9, n, dtype=torch.float64),
                        torch.linspace(0.2, 0.8, n, dtype=torch.float64)])


def test_lowering_builtin_and_combinations():
    x = _x()
    lap = lambda x, center=None, rbf=None, monomial=None, fields=None: u.nodal_laplacian(x, center, rbf, monomial)
    d = trace_diff_operator(lap, x, u.polyharmonic)
    assert d.rbf_const == [0, 0, 0, 1.0, 0, 0, 0] and d.mon_const == d.rbf_const

    DT, K = 1e-3, 0.08
    VEL = np.array([100.0, 0.5])

    def adv(x, center=None, rbf=None, monomial=None, fields=None):  # demos/Advection/01_adv_diff_periodic.py:66-70
        val = u.nodal_value(x, center, rbf, monomial)
        grad = u.nodal_gradient(x, center, rbf, monomial)
        lap = u.nodal_laplacian(x, center, rbf, monomial)
        return (val / DT) + np.dot(VEL, grad) - K * lap

    d = trace_diff_operator(adv, x, u.polyharmonic)
    assert np.allclose(d.rbf_const, [1 / DT, 100.0, 0.5, -K, 0, 0, 0])
    # torch.dot and matmul forms
    adv2 = lambda x, center=None, rbf=None, monomial=None, fields=None: \
        torch.tensor(VEL) @ u.nodal_gradient(x, center, rbf, monomial)
    assert np.allclose(trace_diff_operator(adv2, x, u.polyharmonic).rbf_const, [0, 100.0, 0.5, 0, 0, 0, 0])


def test_lowering_per_row_coefficients_and_fields():
    x = _x()
    perm = torch.linspace(1, 2, 5, dtype=torch.float64)

    def darcy(x, center=None, rbf=None, monomial=None, fields=None):  # demos/Darcy/00_darcy_flow.py:85-89
        p = fields[0]
        return u.nodal_div_grad(x, center, rbf, monomial, (p, p))

    d = trace_diff_operator(darcy, x, u.polyharmonic, diff_args=[perm])
    assert torch.allclose(d.rbf_rows[L.DXX], perm) and torch.allclose(d.rbf_rows[L.DYY], perm)

    def xdep(x, center=None, rbf=None, monomial=None, fields=None):
        return x[0] * u.nodal_value(x, center, rbf, monomial) + u.nodal_laplacian(x, center, rbf, monomial)

    d = trace_diff_operator(xdep, x, u.polyharmonic)
    assert torch.allclose(d.rbf_rows[L.VAL], x[0]) and d.rbf_const[L.LAP] == 1.0


def test_lowering_rejects_nonlinear_operators():
    x = _x()

    def bad(x, center=None, rbf=None, monomial=None, fields=None):
        v = u.nodal_value(x, center, rbf, monomial)
        return v * v

    with pytest.raises(NotLowerable):
        trace_diff_operator(bad, x, u.polyharmonic)

    def affine(x, center=None, rbf=None, monomial=None, fields=None):
        return u.nodal_value(x, center, rbf, monomial) + 1.0

    with pytest.raises(NotLowerable):
        trace_diff_operator(affine, x, u.polyharmonic)


def test_rbf_spec():
    assert u.rbf_spec(u.polyharmonic) == (L.POLYHARMONIC, 1.0)
    assert u.rbf_spec(functools.partial(u.gaussian, eps=5)) == (L.GAUSSIAN, 5.0)
    assert u.rbf_spec(functools.partial(u.thin_plate, a=3)) == (L.THIN_PLATE, 3.0)
    with pytest.raises(ValueError):
        u.rbf_spec(lambda x, c: 0.0)
    # numeric evaluation matches the closed form
    x = torch.tensor([0.3, 0.2], dtype=torch.float64)
    c = torch.tensor([0.1, 0.5], dtype=torch.float64)
    r = float(torch.linalg.norm(x - c))
    assert abs(float(u.polyharmonic(x, c)) - r ** 3) < 1e-15
    assert abs(float(u.gaussian(x, c, eps=2.0)) - np.exp(-4 * r * r)) < 1e-15


def test_robin_and_periodic_glue():
    c = u.SquareCloud(Nx=6, Ny=5, facet_types={"South": "r", "West": "p1", "North": "d", "East": "p1"})
    bcs = {"South": (lambda x: x[0], lambda x: 2.0 + 0 * x[0]), "West": 1.0, "North": 0.0, "East": 5.0}
    rc, nb = u.duplicate_robin_coeffs(bcs, c)
    assert sorted(rc) == sorted(c.facet_nodes["South"]) and all(float(v) == 2.0 for v in rc.values())
    nb = u.zerofy_periodic_cond(nb, c)
    assert float(torch.as_tensor(nb["West"]).abs().sum()) == 0.0 and float(torch.as_tensor(nb["East"]).abs().sum()) == 0.0
    arr = u.boundary_conditions_func_to_arr({"South": lambda x: x[0] * 2, "North": 3.0}, c)
    xs = c.sorted_nodes[c.facet_nodes["South"], 0]
    assert torch.allclose(arr["South"].cpu(), 2 * xs.cpu())


def test_gpu_entry_points_fail_loudly_without_gpu():
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    c = u.SquareCloud(Nx=5, Ny=5, facet_types={"North": "d", "South": "d", "East": "d", "West": "d"})
    lap = lambda x, center=None, rbf=None, monomial=None, fields=None: u.nodal_laplacian(x, center, rbf, monomial)
    with pytest.raises(RuntimeError, match="requires a ROCm GPU"):
        u.pde_solver(lap, lambda x, c, r, f: 0.0, c, {k: 0.0 for k in c.facet_types}, u.polyharmonic, 1)
