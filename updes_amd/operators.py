"""Nodal operators, field evaluation and the PDE solver (drop-in for
``updes/operators.py``).

Call signatures, argument meanings and results follow the reference:
``pde_solver(diff_operator, rhs_operator, cloud, boundary_conditions, rbf,
max_degree, diff_args=None, rhs_args=None)`` returns ``SteadySol(vals,
coeffs, mat)`` (``operators.py:559-618``).  Differences a user sees:

* arrays are torch tensors (float64, on the GPU) instead of jax arrays, and
  operators / boundary functions must use torch (or plain arithmetic) where
  the reference's used ``jnp``;
* gradients flow through torch.autograd (``loss.backward()``) instead of
  ``jax.grad``;
* keyword ``precision="fp32_ir"`` selects the FP32-LU + GMRES-IR solve;
* ``mat`` is computed on first access.

Operators must be linear combinations of ``nodal_value``, ``nodal_gradient``,
``nodal_laplacian`` and ``nodal_div_grad`` (coefficients may depend on x and
on ``fields``); anything else raises ``NotLowerable``.
"""
import contextlib

import numpy as np
import torch

from . import _lib as L
from . import engine as E
from . import adjoint as AD
from ._symbolic import (CENTER, MONOMIAL, CoeffFields, CoeffRef, EvalExpr, EvalVec, NotLowerable, OpTerm,
                        OpVec)
from .utils import SteadySol, compute_nb_monomials, make_all_monomials, rbf_spec, phi_zero, MONOMIAL_EXPONENTS

_TRACE = [False]


@contextlib.contextmanager
def _tracing():
    old = _TRACE[0]
    _TRACE[0] = True
    try:
        yield
    finally:
        _TRACE[0] = old


# ---- host evaluation of single basis functions (numeric calls) ----------------
def _host_ops(rbf, x, center):
    """All 7 operator values of rbf(., center) at x (torch, float64, host or device)."""
    kind, param = rbf_spec(rbf)
    x = torch.as_tensor(x, dtype=torch.float64)
    c = torch.as_tensor(center, dtype=torch.float64, device=x.device)
    d = x - c
    dx, dy = d[..., 0], d[..., 1]
    r2 = dx * dx + dy * dy
    r = torch.sqrt(r2)
    zero = r == 0
    rs = torch.where(zero, torch.ones_like(r), r)
    if kind == L.POLYHARMONIC:
        m = 2 * param + 1
        phi, f1, g2 = rs ** m, m * rs ** (m - 2), m * (m - 2) * rs ** (m - 2)
    elif kind == L.GAUSSIAN:
        e2 = param * param
        g = torch.exp(-e2 * rs * rs)
        phi, f1, g2 = g, -2 * e2 * g, 4 * e2 * e2 * rs * rs * g
    elif kind == L.THIN_PLATE:
        a = param
        lr = torch.log(rs)
        phi = rs ** (2 * a) * lr
        f1 = rs ** (2 * a - 2) * (2 * a * lr + 1)
        g2 = rs ** (2 * a - 2) * ((4 * a * a - 4 * a) * lr + 4 * a - 2)
    elif kind == L.MULTIQUADRIC:
        e2 = param * param
        s = torch.sqrt(1 + e2 * rs * rs)
        phi, f1, g2 = s, e2 / s, -e2 * e2 * rs * rs / s ** 3
    else:
        e2 = param * param
        s = 1 / torch.sqrt(1 + e2 * rs * rs)
        phi, f1, g2 = s, -e2 * s ** 3, 3 * e2 * e2 * rs * rs * s ** 5
    z = torch.zeros_like(r)
    phi = torch.where(zero, torch.full_like(r, phi_zero(kind)), phi)
    f1 = torch.where(zero, z, f1)
    g2 = torch.where(zero, z, g2)
    f2 = g2 / torch.where(zero, torch.ones_like(r2), r2)
    return [phi, f1 * dx, f1 * dy, g2 + 2 * f1, f2 * dx * dx + f1, f2 * dy * dy + f1, f2 * dx * dy]


def _host_mono_ops(monomial, x):
    a, b = MONOMIAL_EXPONENTS[monomial.id]
    x = torch.as_tensor(x, dtype=torch.float64)
    X, Y = x[..., 0], x[..., 1]

    def pw(v, e):
        return v ** e if e >= 0 else torch.zeros_like(v)

    val = pw(X, a) * pw(Y, b)
    dxv = a * pw(X, a - 1) * pw(Y, b)
    dyv = b * pw(X, a) * pw(Y, b - 1)
    dxx = a * (a - 1) * pw(X, a - 2) * pw(Y, b)
    dyy = b * (b - 1) * pw(X, a) * pw(Y, b - 2)
    dxy = a * b * pw(X, a - 1) * pw(Y, b - 1)
    return [val, dxv, dyv, dxx + dyy, dxx, dyy, dxy]


def _numeric(x, center, rbf, monomial):
    if center is not None:
        return _host_ops(rbf, x, center)
    if monomial is not None:
        return _host_mono_ops(monomial, x)
    raise ValueError("either center or monomial must be given")


# ---- nodal operators (operators.py:15-111) ------------------------------------
def nodal_value(x, center=None, rbf=None, monomial=None):
    """rbf(x, center) or monomial(x) (``operators.py:15-32``)."""
    if center is CENTER or (center is None and monomial is MONOMIAL):
        return OpTerm({L.VAL: 1.0})
    return _numeric(x, center, rbf, monomial)[L.VAL]


def nodal_gradient(x, center=None, rbf=None, monomial=None):
    """Gradient w.r.t. x, 0 at r = 0 (``operators.py:42-60``)."""
    if center is CENTER or (center is None and monomial is MONOMIAL):
        return OpVec(OpTerm({L.DX: 1.0}), OpTerm({L.DY: 1.0}))
    o = _numeric(x, center, rbf, monomial)
    return torch.stack([o[L.DX], o[L.DY]], dim=-1)


def nodal_laplacian(x, center=None, rbf=None, monomial=None):
    """Laplacian w.r.t. x, 0 at r = 0 (``operators.py:70-85``)."""
    if center is CENTER or (center is None and monomial is MONOMIAL):
        return OpTerm({L.LAP: 1.0})
    return _numeric(x, center, rbf, monomial)[L.LAP]


def nodal_div_grad(x, center=None, rbf=None, monomial=None, args=None):
    """args[0] d2/dx2 + args[1] d2/dy2 (``operators.py:88-111``)."""
    a0, a1 = args[0], args[1]
    if center is CENTER or (center is None and monomial is MONOMIAL):
        return OpTerm({L.DXX: 1.0}) * a0 + OpTerm({L.DYY: 1.0}) * a1
    o = _numeric(x, center, rbf, monomial)
    return a0 * o[L.DXX] + a1 * o[L.DYY]


# ---- field evaluation (operators.py:118-371) ----------------------------------
def _eval_points(x):
    """Points as (P, 2) and a function restoring the caller's layout."""
    x = torch.as_tensor(x, dtype=torch.float64)
    if _TRACE[0] and x.dim() == 2 and x.shape[0] == 2:
        return x.T, (lambda v: v)            # vectorised trace: x is (2, rows)
    if x.dim() == 1:
        return x.reshape(1, 2), (lambda v: v[0])
    return x.reshape(-1, 2), (lambda v: v)


def _field_eval(x, field, centers, rbf, op):
    kind, param = rbf_spec(rbf)
    if isinstance(field, CoeffRef):
        return EvalExpr({(op, field.k): 1.0})
    L.require_gpu()
    pts, back = _eval_points(x)
    dev = torch.device("cuda")
    field = torch.as_tensor(field, dtype=torch.float64).to(dev)
    centers = torch.as_tensor(centers, dtype=torch.float64).to(dev)
    return back(AD.EvalOp.apply(field, kind, param, centers, pts.to(dev), op))


def value(x, field, centers, rbf=None, clip_val=None):
    """Field value at x from its coefficients, summed over all centres (``operators.py:118-147``)."""
    v = _field_eval(x, field, centers, rbf, L.VAL)
    return torch.clamp(v, -clip_val, clip_val) if clip_val and torch.is_tensor(v) else v


def value_vec(x, field, centers, rbf=None):
    return value(torch.as_tensor(x).reshape(-1, 2), field, centers, rbf)


def gradient(x, field, centers, rbf=None, clip_val=None):
    """Field gradient at x (``operators.py:156-184``)."""
    gx = _field_eval(x, field, centers, rbf, L.DX)
    gy = _field_eval(x, field, centers, rbf, L.DY)
    if isinstance(gx, EvalExpr):
        return EvalVec(gx, gy)
    g = torch.stack([gx, gy], dim=-1)
    return torch.clamp(g, -clip_val, clip_val) if clip_val else g


def gradient_vec(x, field, centers, rbf=None):
    return gradient(torch.as_tensor(x).reshape(-1, 2), field, centers, rbf)


def laplacian(x, field, centers, rbf=None, clip_val=None):
    """Field Laplacian at x (``operators.py:319-348``)."""
    v = _field_eval(x, field, centers, rbf, L.LAP)
    return torch.clamp(v, -clip_val, clip_val) if clip_val and torch.is_tensor(v) else v


def laplacian_vec(x, field, centers, rbf=None):
    return laplacian(torch.as_tensor(x).reshape(-1, 2), field, centers, rbf)


def divergence(x, field, centers, rbf=None, clip_val=None):
    """d/dx field[...,0] + d/dy field[...,1] (``operators.py:294-312``)."""
    if isinstance(field, (list, tuple)) and isinstance(field[0], CoeffRef):
        f0, f1 = field[0], field[1]
    else:
        field = torch.as_tensor(field)
        f0, f1 = field[..., 0], field[..., 1]
    d = gradient(x, f0, centers, rbf)[..., 0] + gradient(x, f1, centers, rbf)[..., 1]
    return torch.clamp(d, -clip_val, clip_val) if clip_val and torch.is_tensor(d) else d


def divergence_vec(x, field, centers, rbf=None):
    return divergence(torch.as_tensor(x).reshape(-1, 2), field, centers, rbf)


# ---- coefficients (assembly.py:404-430) ------------------------------------------
def _dev():
    L.require_gpu()
    return torch.device("cuda")


def core_compute_coefficients(field, cloud, rbf, nb_monomials):
    """inv(A) [field; 0] (``assembly.py:404-410``), differentiable in ``field``."""
    kind, param = rbf_spec(rbf)
    dev = _dev()
    field = torch.as_tensor(field, dtype=torch.float64).to(dev)
    fac = E.a_factor(cloud, kind, param, nb_monomials)
    rhs = torch.cat([field, torch.zeros(nb_monomials, dtype=torch.float64, device=dev)])
    return AD.ASolve.apply(rhs, fac)


def compute_coefficients(field, cloud, rbf, max_degree):
    return core_compute_coefficients(field, cloud, rbf, compute_nb_monomials(max_degree, cloud.dim))


get_field_coefficients = compute_coefficients


def gradient_vals(x, field, cloud, rbf, max_degree):
    return gradient(x, compute_coefficients(field, cloud, rbf, max_degree), cloud.sorted_nodes, rbf)


def laplacian_vals(x, field, cloud, rbf, max_degree):
    return laplacian(x, compute_coefficients(field, cloud, rbf, max_degree), cloud.sorted_nodes, rbf)


# ---- boundary-condition glue (operators.py:512-555, 628-647) ----------------------
def _facet_points(cloud, f_id):
    ids = torch.as_tensor(cloud.facet_nodes[f_id], dtype=torch.int64)
    nodes = torch.as_tensor(cloud.sorted_nodes, dtype=torch.float64)
    return nodes[ids.to(nodes.device)]


def _call_vectorised(fn, pts):
    """Call a per-point function on all points at once (x[0], x[1] are arrays)."""
    with _tracing():
        v = fn(pts.T)
    v = torch.as_tensor(v, dtype=torch.float64, device=pts.device) if not torch.is_tensor(v) else v
    return v.to(torch.float64).expand(pts.shape[0]) if v.dim() == 0 else v.reshape(pts.shape[0])


def duplicate_robin_coeffs(boundary_conditions, cloud):
    """Split Robin (value, beta) pairs; betas keyed by node id (``operators.py:512-541``)."""
    robin_coeffs = {}
    new_bc = {}
    for f_id, f_type in cloud.facet_types.items():
        if f_type == "r":
            node_ids = cloud.facet_nodes[f_id]
            bc = boundary_conditions[f_id]
            if isinstance(bc, tuple):
                new_bc[f_id] = bc[0]
                betas = bc[1]
                if callable(betas):
                    betas = _call_vectorised(betas, _facet_points(cloud, f_id))
                betas = torch.as_tensor(betas, dtype=torch.float64).reshape(-1)
            else:
                import warnings
                warnings.warn("Did not provide beta coefficients for Robin BC. Using zeros ...")
                new_bc[f_id] = bc
                betas = torch.zeros(len(node_ids), dtype=torch.float64)
            for i in node_ids:
                robin_coeffs[i] = betas[min(i - node_ids[0], betas.numel() - 1)]
        else:
            new_bc[f_id] = boundary_conditions[f_id]
    return robin_coeffs, new_bc


def zerofy_periodic_cond(boundary_conditions, cloud):
    """Periodic facets get zero data (``operators.py:546-555``)."""
    for f_id, f_type in cloud.facet_types.items():
        if f_type[0] == "p":
            boundary_conditions[f_id] = torch.zeros(len(cloud.facet_nodes[f_id]), dtype=torch.float64)
    return boundary_conditions


def boundary_conditions_func_to_arr(boundary_conditions, cloud):
    """Callables evaluated on their facet nodes (``operators.py:628-647``)."""
    out = {}
    for f_id, bc in boundary_conditions.items():
        if callable(bc):
            out[f_id] = _call_vectorised(bc, _facet_points(cloud, f_id))
        elif isinstance(bc, tuple):
            v, b = bc
            pts = _facet_points(cloud, f_id)
            out[f_id] = (_call_vectorised(v, pts) if callable(v) else v,
                         _call_vectorised(b, pts) if callable(b) else b)
        else:
            out[f_id] = bc
    return out


# ---- lowering ------------------------------------------------------------------
def _coef_tensor(v, rows, dev):
    if isinstance(v, (int, float)):
        return float(v), None
    v = torch.as_tensor(v, dtype=torch.float64).to(dev)
    if v.numel() == 1:
        return float(v.reshape(())), None
    return 0.0, v.reshape(rows)


def lower_diff_operator(diff_operator, cloud, rbf, diff_args=None):
    """Trace ``diff_operator`` into an engine.OpDesc (per-row coefficients over Ni rows)."""
    _dev()
    dc = E.device_cloud(cloud)
    return trace_diff_operator(diff_operator, dc.nodes[:dc.Ni].T, rbf, diff_args)


def trace_diff_operator(diff_operator, x, rbf, diff_args=None):
    """Descriptor of ``diff_operator`` at the collocation points x (2, Ni) on x's device.

    The reference evaluates the operator per (node, centre) pair with
    ``fields[i]`` = the diff_args values at node i, or ones when there are
    none (``updes/assembly.py:114-135``); here ``fields[k]`` is the (Ni,)
    vector of the k-th argument.
    """
    dev = x.device
    Ni = x.shape[1]
    if diff_args:
        fields = torch.stack([torch.as_tensor(a, dtype=torch.float64).to(dev)[:Ni] for a in diff_args], dim=0)
    else:
        fields = torch.ones((1, Ni), dtype=torch.float64, device=dev)
    with _tracing():
        t_rbf = diff_operator(x, CENTER, rbf, None, fields)
        t_mon = diff_operator(x, None, rbf, MONOMIAL, fields)
    desc = E.OpDesc()
    for t, consts, rows in ((t_rbf, desc.rbf_const, desc.rbf_rows), (t_mon, desc.mon_const, desc.mon_rows)):
        if isinstance(t, OpVec):
            raise NotLowerable("diff_operator returned a vector (gradient); it must return a scalar operator")
        if not isinstance(t, OpTerm):
            if _is_zero_number(t):
                continue
            raise NotLowerable("diff_operator must return a linear combination of nodal operators, got %r" % type(t))
        for op, c in t.c.items():
            cst, row = _coef_tensor(c, Ni, dev)
            consts[op] += cst
            if row is not None:
                rows[op] = row if rows[op] is None else rows[op] + row
    return desc


def _is_zero_number(v):
    try:
        return float(v) == 0.0
    except Exception:
        return False


def rhs_values(rhs_operator, cloud, rbf, nb_monomials, rhs_args=None):
    """Interior rhs q[:Ni] (``assembly.py:434-470``), differentiable in rhs_args."""
    kind, param = rbf_spec(rbf)
    dev = _dev()
    dc = E.device_cloud(cloud)
    Ni, N = dc.Ni, dc.N
    x = dc.nodes[:Ni].T
    fields = CoeffFields(len(rhs_args)) if rhs_args is not None else None
    with _tracing():
        res = rhs_operator(x, dc.nodes, rbf, fields)
    if isinstance(res, EvalVec):
        raise NotLowerable("rhs_operator returned a vector")
    if not isinstance(res, EvalExpr):
        res = EvalExpr({}, res)
    const = res.const
    q = torch.as_tensor(const, dtype=torch.float64).to(dev) if not torch.is_tensor(const) else const.to(dev, torch.float64)
    q = q.expand(Ni) if q.dim() == 0 else q.reshape(Ni)
    coeffs = {}
    for (op, k), w in res.terms.items():
        if k not in coeffs:
            f = torch.as_tensor(rhs_args[k], dtype=torch.float64).to(dev)
            coeffs[k] = core_compute_coefficients(f, cloud, rbf, nb_monomials) if f.shape[0] == N else f
        vals = AD.EvalOp.apply(coeffs[k], kind, param, dc.nodes, x.T, op)
        w = torch.as_tensor(w, dtype=torch.float64).to(dev) if not torch.is_tensor(w) else w
        q = q + w * vals
    return q


def _bc_vector(cloud, bcs, N, dev):
    """q's boundary part: values scattered to facet nodes (``assembly.py:472-483``)."""
    ids_all, vals_all = [], []
    for f_id in cloud.facet_types:
        assert f_id in bcs, "facets and boundary functions don't match ids"
        bc = bcs[f_id]
        ids = torch.as_tensor(cloud.facet_nodes[f_id], dtype=torch.int64, device=dev)
        if callable(bc):
            v = _call_vectorised(bc, _facet_points(cloud, f_id).to(dev))
        else:
            v = torch.as_tensor(bc, dtype=torch.float64).to(dev)
            v = v.expand(ids.numel()) if v.dim() == 0 else v.reshape(ids.numel())
        ids_all.append(ids)
        vals_all.append(v.to(dev))
    return ids_all, vals_all


def assemble_q(operator, boundary_conditions, cloud, rbf, nb_monomials, rhs_args):
    """q (N,) (``assembly.py:434-485``): the interior values of the lowered
    rhs operator and the facets' values in node order, assembled by rbfx_rhs."""
    dev = _dev()
    N, Ni = cloud.N, cloud.Ni
    q_int = rhs_values(operator, cloud, rbf, nb_monomials, rhs_args)
    ids, vals = _bc_vector(cloud, boundary_conditions, N, dev)
    bc = torch.zeros(N - Ni, dtype=torch.float64, device=dev)
    if ids:
        bc = bc.index_put((torch.cat(ids) - Ni,), torch.cat(vals).to(torch.float64))
    kind, param = rbf_spec(rbf)
    prob = E.Problem(E.device_cloud(cloud), kind, param, nb_monomials)
    return AD.RhsAssemble.apply(q_int.to(torch.float64), bc, prob)


def _robin_beta_vector(cloud, robin_coeffs):
    if not robin_coeffs:
        return None
    return torch.stack([torch.as_tensor(robin_coeffs[k], dtype=torch.float64) for k in sorted(robin_coeffs)])


# ---- solver (operators.py:559-683) --------------------------------------------------
def _mat_thunk(cloud, kind, param, M, fac, dc):
    def build():
        # B = (D inv(A))[:, :N] (updes/assembly.py:399): inv(A) by one
        # multi-right-hand-side solve, then one MFMA GEMM
        return E.b_matrix(fac.problem, E.a_factor(cloud, kind, param, M))
    return build


def pde_solver(diff_operator, rhs_operator, cloud, boundary_conditions, rbf, max_degree, diff_args=None,
               rhs_args=None, precision="fp64", **ir_opts):
    """Solve the PDE by RBF collocation; returns SteadySol(vals, coeffs, mat)."""
    dev = _dev()
    kind, param = rbf_spec(rbf)
    robin_coeffs, bcs = duplicate_robin_coeffs(boundary_conditions, cloud)
    bcs = zerofy_periodic_cond(bcs, cloud)
    M = compute_nb_monomials(max_degree, cloud.dim)
    dc = E.device_cloud(cloud)
    desc = lower_diff_operator(diff_operator, cloud, rbf, diff_args)
    beta = _robin_beta_vector(cloud, robin_coeffs)
    fac = E.k_factor(cloud, kind, param, M, desc, beta, precision, **ir_opts)
    q = assemble_q(rhs_operator, bcs, cloud, rbf, M, rhs_args)
    b = torch.cat([q, torch.zeros(M, dtype=torch.float64, device=dev)])
    c = AD.KSolve.apply(b, fac)
    prob_a = E.Problem(dc, kind, param, M)
    u = AD.AValues.apply(c, prob_a)
    sol = SteadySol(u, c, _mat_thunk(cloud, kind, param, M, fac, dc))
    return sol


def pde_solver_jit(diff_operator, rhs_operator, cloud, boundary_conditions, rbf, max_degree, diff_args=None,
                   rhs_args=None, precision="fp64", **ir_opts):
    """``operators.py:650-683``: boundary callables are evaluated to arrays first."""
    bcs = boundary_conditions_func_to_arr(boundary_conditions, cloud)
    return pde_solver(diff_operator, rhs_operator, cloud, bcs, rbf, max_degree, diff_args, rhs_args,
                      precision, **ir_opts)


pde_solver_jit_with_bc = pde_solver


def last_solve_stats():
    """Precision / refinement statistics of the most recent collocation solve."""
    for fac in reversed(list(E._K_CACHE.d.values())):
        return fac.last_stats.as_dict()
    return None


def pde_multi_solver(diff_operators, rhs_operators, cloud, boundary_conditions, rbf, max_degree, nb_iters=10,
                     tol=1e-6, diff_args=None, rhs_args=None):
    """Picard iteration over coupled scalar PDEs (``operators.py:696-771``)."""
    n = len(diff_operators)
    assert n == len(rhs_operators) == len(boundary_conditions), \
        "The number of differential operators must match the number of right-hand side operators"
    sols_vals = [u for u in diff_args[0][:n]]
    sols = None
    for _ in range(nb_iters):
        sols = [pde_solver_jit(diff_operators[i], rhs_operators[i], cloud, boundary_conditions[i], rbf, max_degree,
                               diff_args=sols_vals + list(diff_args[i][n:]), rhs_args=rhs_args[i]) for i in range(n)]
        sols_vals = [s.vals for s in sols]
    return sols


# ---- post-processing helpers used by the reference tests ------------------------------
def interpolate_field(field, cloud1, cloud2):
    """Reorder a field between two numberings (``operators.py:457-480``)."""
    assert cloud1.N == cloud2.N, "the two clouds do not contain the same number of nodes"
    map1 = cloud1.renumbering_map
    indexer1 = torch.as_tensor([map1[k] for k in sorted(map1)])
    field = torch.as_tensor(field)
    field_orig = field[indexer1.to(field.device)]
    indexer2 = torch.as_tensor(list(cloud2.renumbering_map.keys()))
    return field_orig[indexer2.to(field.device)]


def integrate_field(field, cloud, rbf, max_degree):
    """Midpoint-rule integral over the unit square (``operators.py:381-450``).

    The top-left corner term evaluates at the bottom-right corner, exactly as
    the reference does (it reuses ``quad_points`` at ``operators.py:441``).
    """
    from .cloud import SquareCloud
    assert isinstance(cloud, SquareCloud), "The cloud must be a SquareCloud instance"
    Nx, Ny = cloud.Nx, cloud.Ny
    area = (1 / (Nx - 1)) * (1 / (Ny - 1))
    c = cloud.sorted_nodes
    qx = torch.linspace(1 / (Nx - 1), 1 - 1 / (Nx - 1), Nx - 1, dtype=torch.float64)
    qy = torch.linspace(1 / (Ny - 1), 1 - 1 / (Ny - 1), Ny - 1, dtype=torch.float64)
    gx, gy = torch.meshgrid(qx, qy, indexing="xy")
    # the reference reshapes the (2, ny, nx) stack straight to (nb_squares, 2)
    # (operators.py:404), pairing consecutive entries; kept as is
    inner = torch.stack([gx, gy]).reshape(-1, 2).contiguous()
    ly = torch.linspace(1 / (Ny - 1), 1 - 1 / (Ny - 1), Ny - 2, dtype=torch.float64)
    lx = torch.linspace(1 / (Nx - 1), 1 - 1 / (Nx - 1), Nx - 2, dtype=torch.float64)
    left = torch.stack([torch.zeros(Ny - 2, dtype=torch.float64), ly], dim=-1)
    right = torch.stack([torch.ones(Ny - 2, dtype=torch.float64), ly], dim=-1)
    bottom = torch.stack([lx, torch.zeros(Nx - 2, dtype=torch.float64)], dim=-1)
    top = torch.stack([lx, torch.ones(Nx - 2, dtype=torch.float64)], dim=-1)
    corners = torch.tensor([[0.0, 0.0], [1.0, 0.0], [1.0, 0.0], [1.0, 1.0]], dtype=torch.float64)
    pts = torch.cat([inner, left, right, bottom, top, corners])
    v = value_vec(pts, field, c, rbf)
    n0 = inner.shape[0]
    n1 = n0 + 2 * (Ny - 2) + 2 * (Nx - 2)
    total = v[:n0].sum() + 0.5 * v[n0:n1].sum() + 0.25 * v[n1:].sum()
    return total * area


dot_vec = lambda a, b: (torch.as_tensor(a) * torch.as_tensor(b)).sum(-1)
dot_mat = lambda J, v: torch.einsum("...ij,...j->...i", torch.as_tensor(J), torch.as_tensor(v))
