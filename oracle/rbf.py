"""Closed forms of the RBF kernels, their derivatives, and the monomial basis.

Oracle (test infrastructure only).  Restates:

* ``updes/utils.py:19-22``  distance r = |x - c|
* ``updes/utils.py:30-42``  (inverse) multiquadric  sqrt(1+(eps r)^2), 1/sqrt(1+(eps r)^2)
* ``updes/utils.py:44-48``  gaussian  exp(-(eps r)^2)
* ``updes/utils.py:50-55``  polyharmonic  r^(2a+1)
* ``updes/utils.py:63-69``  thin_plate  nan_to_num(log(r) r^(2a))  (0 at r = 0)
* ``updes/utils.py:92-144`` monomials id 0..14, M = C(deg+2, deg)
* ``updes/operators.py:34-111`` the autodiff operators nodal_gradient /
  nodal_laplacian / nodal_div_grad, whose NaN at r = 0 (reverse mode through
  ``sqrt``) is mapped to 0 by ``jnp.nan_to_num`` (``operators.py:58,83,109``).

For a radial kernel phi(r) with d = x - c, r = |d|, write
    f1 = phi'(r) / r,   f2 = (phi''(r) - phi'(r)/r) / r^2.
Then grad phi = f1 d, Hess phi = f2 d d^T + f1 I and Lap phi = f2 r^2 + 2 f1.
At r == 0 the reference yields phi(0) for the value and 0 for every
derivative; this module returns the same.
"""
import math

import numpy as np

POLYHARMONIC, GAUSSIAN, THIN_PLATE, MULTIQUADRIC, INV_MULTIQUADRIC = 0, 1, 2, 3, 4
KIND_NAMES = {
    POLYHARMONIC: "polyharmonic",
    GAUSSIAN: "gaussian",
    THIN_PLATE: "thin_plate",
    MULTIQUADRIC: "multiquadric",
    INV_MULTIQUADRIC: "inverse_multiquadric",
}
DEFAULT_PARAM = {POLYHARMONIC: 1.0, GAUSSIAN: 1.0, THIN_PLATE: 1.0,
                 MULTIQUADRIC: 1.0, INV_MULTIQUADRIC: 1.0}

# Operator basis used by the oracle's operator descriptors.
VAL, DX, DY, LAP, DXX, DYY, DXY = range(7)
NOPS = 7

# Monomial exponents, ids 0..14 (updes/utils.py:103-132).
MONOMIAL_EXPONENTS = [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1),
                      (1, 2), (0, 3), (4, 0), (3, 1), (2, 2), (1, 3), (0, 4)]


def compute_nb_monomials(max_degree, problem_dimension=2):
    """``updes/utils.py:142-144``: C(max_degree + dim, max_degree)."""
    return math.comb(max_degree + problem_dimension, max_degree)


def phi_zero(kind, param):
    """Value of the kernel at r = 0 as the reference computes it."""
    if kind == POLYHARMONIC:
        return 0.0
    if kind == GAUSSIAN:
        return 1.0
    if kind == THIN_PLATE:
        return 0.0        # nan_to_num(log(0) * 0) = 0   (utils.py:65)
    if kind in (MULTIQUADRIC, INV_MULTIQUADRIC):
        return 1.0
    raise ValueError(kind)


def radial_terms(kind, param, r):
    """Return (phi, f1, f2) for r > 0 (arrays, float64)."""
    r = np.asarray(r, dtype=np.float64)
    if kind == POLYHARMONIC:
        m = 2.0 * param + 1.0
        phi = r ** m
        f1 = m * r ** (m - 2.0)
        f2 = m * (m - 2.0) * r ** (m - 4.0)
    elif kind == GAUSSIAN:
        e2 = param * param
        phi = np.exp(-(param * r) ** 2)
        f1 = -2.0 * e2 * phi
        f2 = 4.0 * e2 * e2 * phi
    elif kind == THIN_PLATE:
        a = param
        lr = np.log(r)
        phi = lr * r ** (2.0 * a)
        f1 = r ** (2.0 * a - 2.0) * (2.0 * a * lr + 1.0)
        f2 = r ** (2.0 * a - 4.0) * ((4.0 * a * a - 4.0 * a) * lr + 4.0 * a - 2.0)
    elif kind == MULTIQUADRIC:
        e2 = param * param
        phi = np.sqrt(1.0 + (param * r) ** 2)
        f1 = e2 / phi
        f2 = -e2 * e2 / phi ** 3
    elif kind == INV_MULTIQUADRIC:
        e2 = param * param
        phi = 1.0 / np.sqrt(1.0 + (param * r) ** 2)
        f1 = -e2 * phi ** 3
        f2 = 3.0 * e2 * e2 * phi ** 5
    else:
        raise ValueError(kind)
    return phi, f1, f2


def rbf_ops(kind, param, x, y, cx, cy):
    """All operator values of phi(., c) at points (x, y) -> array (NOPS, ...).

    Broadcasting over x/y vs cx/cy.  Order: VAL, DX, DY, LAP, DXX, DYY, DXY.
    """
    dx = np.asarray(x, np.float64) - np.asarray(cx, np.float64)
    dy = np.asarray(y, np.float64) - np.asarray(cy, np.float64)
    r2 = dx * dx + dy * dy
    r = np.sqrt(r2)
    zero = r == 0.0
    rs = np.where(zero, 1.0, r)
    with np.errstate(all="ignore"):
        phi, f1, f2 = radial_terms(kind, param, rs)
    phi = np.where(zero, phi_zero(kind, param), phi)
    f1 = np.where(zero, 0.0, f1)
    f2 = np.where(zero, 0.0, f2)
    out = np.empty((NOPS,) + np.broadcast(dx, dy).shape)
    out[VAL] = phi
    out[DX] = f1 * dx
    out[DY] = f1 * dy
    out[LAP] = f2 * r2 + 2.0 * f1
    out[DXX] = f2 * dx * dx + f1
    out[DYY] = f2 * dy * dy + f1
    out[DXY] = f2 * dx * dy
    # the reference maps NaN (from d/dr sqrt at 0) to 0 for every derivative
    out[1:] = np.where(zero, 0.0, out[1:])
    return out


def monomial_ops(k, x, y):
    """All operator values of monomial id k at (x, y) -> array (NOPS, ...)."""
    if not 0 <= k < len(MONOMIAL_EXPONENTS):
        raise ValueError("monomial id %d unsupported (reference supports 0..14)" % k)
    a, b = MONOMIAL_EXPONENTS[k]
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)

    def pw(v, e):
        if e < 0:
            return np.zeros_like(v)
        return v ** e

    out = np.empty((NOPS,) + np.broadcast(x, y).shape)
    out[VAL] = pw(x, a) * pw(y, b)
    out[DX] = a * pw(x, a - 1) * pw(y, b)
    out[DY] = b * pw(x, a) * pw(y, b - 1)
    out[DXX] = a * (a - 1) * pw(x, a - 2) * pw(y, b)
    out[DYY] = b * (b - 1) * pw(x, a) * pw(y, b - 2)
    out[DXY] = a * b * pw(x, a - 1) * pw(y, b - 1)
    out[LAP] = out[DXX] + out[DYY]
    return out


def apply_descriptor(ops, coeffs):
    """Sum_t coeffs[t] * ops[t]; coeffs[t] is a scalar or broadcastable array."""
    acc = np.zeros(ops.shape[1:])
    for t in range(NOPS):
        c = coeffs[t]
        if c is None:
            continue
        if np.isscalar(c) and c == 0.0:
            continue
        acc = acc + np.asarray(c)[..., None] * ops[t] if np.ndim(c) == 1 and ops.ndim == 3 \
            else acc + c * ops[t]
    return acc
