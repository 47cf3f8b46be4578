"""GPU parity of the second-derivative operator terms (the branch of the
assembly and evaluation kernels that nodal_div_grad lowers to,
updes/operators.py:88-111): assembled K rows with d2/dx2, d2/dy2 and d2/dxdy
coefficients, constant and per row, and the DXX / DYY / DXY outputs of the
field evaluation, against the oracle for every kernel, r = 0 self terms
included."""
import numpy as np
import pytest
import torch

from oracle import assembly as OA, cloud as OC, rbf as R, solver as OS
from tests import _cases as C

pytestmark = pytest.mark.gpu


def _rowscaled_err(a, b):
    scale = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-300)
    return float((np.abs(a - b) / scale).max())


@pytest.mark.parametrize("kind,param", C.KINDS)
def test_div_grad_with_per_row_args_matches_oracle(kind, param):
    """nodal_div_grad(args=fields): fields are the per-node values of
    diff_args at the internal nodes (assembly.py:114, 128)."""
    import updes_amd as u
    facets = C.FACET_SETS[1]
    nx, ny = 9, 8
    J = C.jitter(nx, ny, 0.2, seed=7)
    oc = OC.SquareCloud(nx, ny, facets, jitter=J)
    uc = u.SquareCloud(Nx=nx, Ny=ny, facet_types=facets, jitter=J)
    xs = oc.sorted_nodes
    a0 = 1.0 + xs[:, 0] ** 2          # per-node diff_args (N values each)
    a1 = 2.0 - xs[:, 1]

    def op(x, center=None, rbf=None, monomial=None, fields=None):
        return u.nodal_div_grad(x, center, rbf, monomial, args=fields) + 0.25 * u.nodal_value(x, center, rbf, monomial)

    args = [torch.as_tensor(a0, device="cuda"), torch.as_tensor(a1, device="cuda")]
    K = u.assemble_K(op, uc, C.torch_rbf(kind, param), 6, args, None).cpu().numpy()
    Ni = oc.Ni
    desc = {"rbf": [0.25, 0, 0, 0, a0[:Ni], a1[:Ni], 0], "mon": [0.25, 0, 0, 0, a0[:Ni], a1[:Ni], 0]}
    Kref = OA.assemble_K(desc, oc, kind, param, 6)
    assert _rowscaled_err(K, Kref) < 1e-12


@pytest.mark.parametrize("kind,param", C.KINDS)
def test_all_seven_operator_terms_in_K_match_oracle(kind, param):
    """Every operator term at once, with the mixed derivative and per-row
    parts on top of constants (engine descriptor, as the lowering builds it)."""
    from updes_amd import engine as E, _lib as L
    import updes_amd as u
    facets = C.FACET_SETS[2]
    nx, ny = 8, 9
    J = C.jitter(nx, ny, 0.2, seed=11)
    oc = OC.SquareCloud(nx, ny, facets, jitter=J)
    uc = u.SquareCloud(Nx=nx, Ny=ny, facet_types=facets, jitter=J)
    Ni = oc.Ni
    rng = np.random.default_rng(kind)
    const = [0.3, -0.2, 0.4, 0.1, 0.7, -0.6, 1.3]
    rows = [rng.standard_normal(Ni) if t in (4, 5, 6) else None for t in range(7)]
    dev = torch.device("cuda")
    desc = E.OpDesc(const, const, [None if r is None else torch.as_tensor(r, device=dev) for r in rows],
                    [None if r is None else torch.as_tensor(r, device=dev) for r in rows])
    betas = {f: np.linspace(0.5, 1.5, len(oc.facet_nodes[f])) for f, t in oc.facet_types.items() if t == "r"}
    rb = OA.robin_coefficients(oc, betas)
    prob = E.Problem(E.device_cloud(uc), kind, param, 6, desc, torch.as_tensor(rb))
    K = E.assemble(prob, L.MAT_K)[:, :prob.n].cpu().numpy()
    cref = [const[t] if rows[t] is None else const[t] + rows[t] for t in range(7)]
    Kref = OA.assemble_K({"rbf": cref, "mon": cref}, oc, kind, param, 6, rb)
    assert _rowscaled_err(K, Kref) < 1e-12


@pytest.mark.parametrize("kind,param", C.KINDS)
def test_eval_second_derivatives_match_oracle(kind, param):
    """rbfx_eval's DXX, DYY and DXY rows (and the other four) at random
    points and at the centres themselves (r = 0: 0 for every derivative)."""
    from updes_amd import engine as E
    oc = OC.SquareCloud(10, 9, C.README_FACETS, jitter=C.jitter(10, 9, 0.1, 5))
    rng = np.random.default_rng(100 + kind)
    coeffs = rng.standard_normal(oc.N + 10)
    pts = np.concatenate([rng.uniform(0, 1, (23, 2)), oc.sorted_nodes[:7]])
    ref = OS.eval_ops(pts, coeffs, oc.sorted_nodes, kind, param)
    got = E.eval_ops(kind, param, torch.as_tensor(oc.sorted_nodes).cuda(), torch.as_tensor(coeffs).cuda(),
                     torch.as_tensor(pts).cuda(), 0x7F).cpu().numpy()
    for t in range(R.NOPS):
        sc = max(np.abs(ref[t]).max(), 1.0)
        assert np.abs(got[t] - ref[t]).max() <= 1e-12 * sc, (t, np.abs(got[t] - ref[t]).max() / sc)
