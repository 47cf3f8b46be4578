"""Kernels, monomials and the solution record (drop-in for ``updes/utils.py``).

The radial functions keep the reference's names and keyword parameters
(``updes/utils.py:19-69``): ``polyharmonic(x, center, a=1)``,
``gaussian(x, center, eps=1.)``, ``thin_plate(x, center, a=1)``,
``multiquadric`` and ``inverse_multiquadric``.  Called on numbers/tensors
they evaluate on the host with torch; the solver does not call them -- it
identifies the kernel (function + bound keywords, e.g.
``partial(gaussian, eps=5)``) and evaluates closed forms in HIP kernels.
"""
import functools
import math
from collections import namedtuple

import numpy as np
import torch

from . import _lib as L
from ._symbolic import CENTER, MONOMIAL, OpTerm


def _t(v):
    return v if torch.is_tensor(v) else torch.as_tensor(np.asarray(v, dtype=np.float64))


def distance(node1, node2):
    """Euclidean distance between two points (``utils.py:19-22``)."""
    d = _t(node1) - _t(node2)
    return torch.sqrt((d * d).sum(-1))


def multiquadric_func(r, eps):
    return torch.sqrt(1 + (eps * _t(r)) ** 2)


def inv_multiquadric_func(r, eps):
    return 1.0 / torch.sqrt(1 + (eps * _t(r)) ** 2)


def gaussian_func(r, eps):
    return torch.exp(-(eps * _t(r)) ** 2)


def polyharmonic_func(r, a):
    return _t(r) ** (2 * a + 1)


def thin_plate_func(r, a):
    r = _t(r)
    return torch.nan_to_num(torch.log(r) * r ** (2 * a), nan=0.0, posinf=0.0, neginf=0.0)


def _rbf_call(func, x, center, **kw):
    if center is CENTER:
        return OpTerm({L.VAL: 1.0})
    return func(distance(x, center), **kw)


def multiquadric(x, center, eps=1.0):
    """Hardy's multiquadric sqrt(1 + (eps r)^2)."""
    return _rbf_call(multiquadric_func, x, center, eps=eps)


def inverse_multiquadric(x, center, eps=1.0):
    """Inverse multiquadric 1 / sqrt(1 + (eps r)^2)."""
    return _rbf_call(inv_multiquadric_func, x, center, eps=eps)


def gaussian(x, center, eps=1.0):
    """Gaussian exp(-(eps r)^2)."""
    return _rbf_call(gaussian_func, x, center, eps=eps)


def polyharmonic(x, center, a=1):
    """Polyharmonic spline r^(2a+1)."""
    return _rbf_call(polyharmonic_func, x, center, a=a)


def thin_plate(x, center, a=1):
    """Thin-plate spline r^(2a) log r (0 at r = 0)."""
    return _rbf_call(thin_plate_func, x, center, a=a)


_KINDS = {
    polyharmonic: (L.POLYHARMONIC, "a", 1.0),
    gaussian: (L.GAUSSIAN, "eps", 1.0),
    thin_plate: (L.THIN_PLATE, "a", 1.0),
    multiquadric: (L.MULTIQUADRIC, "eps", 1.0),
    inverse_multiquadric: (L.INV_MULTIQUADRIC, "eps", 1.0),
}


class UnsupportedKernel(ValueError):
    pass


def rbf_spec(rbf):
    """(kernel id, parameter) of a built-in kernel or a functools.partial of one."""
    if rbf is None:
        rbf = polyharmonic
    func, kw = rbf, {}
    while isinstance(func, functools.partial):
        if func.args:
            raise UnsupportedKernel("positional arguments bound to the RBF are not supported")
        kw = {**func.keywords, **kw}
        func = func.func
    if func not in _KINDS:
        raise UnsupportedKernel("RBF %r is not one of the built-in kernels %s" %
                                (rbf, [f.__name__ for f in _KINDS]))
    kind, pname, default = _KINDS[func]
    extra = set(kw) - {pname}
    if extra:
        raise UnsupportedKernel("unknown RBF parameters %s" % sorted(extra))
    param = float(kw.get(pname, default))
    if kind in (L.POLYHARMONIC, L.THIN_PLATE) and (param != int(param) or param < 0):
        raise UnsupportedKernel("the exponent a must be a non-negative integer")
    return kind, param


def phi_zero(kind):
    """Kernel value at r = 0 as the reference evaluates it."""
    return 1.0 if kind in (L.GAUSSIAN, L.MULTIQUADRIC, L.INV_MULTIQUADRIC) else 0.0


MONOMIAL_EXPONENTS = [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1),
                      (1, 2), (0, 3), (4, 0), (3, 1), (2, 2), (1, 3), (0, 4)]


class _Monomial:
    """Monomial id 0..14 (``utils.py:92-134``); callable on x = (x0, x1)."""

    def __init__(self, id):
        if not 0 <= id < len(MONOMIAL_EXPONENTS):
            raise ValueError("monomial id %d unsupported (0..14)" % id)
        self.id = id

    def __call__(self, x):
        a, b = MONOMIAL_EXPONENTS[self.id]
        return x[0] ** a * x[1] ** b

    def __repr__(self):
        return "monomial(%d)" % self.id


def make_monomial(x, id):
    """Value of monomial ``id`` at x (``utils.py:92-134``)."""
    return _Monomial(id)(x)


@functools.lru_cache(maxsize=None)
def make_all_monomials(nb_monomials):
    """Monomials 0..nb_monomials-1 (``utils.py:136-139``)."""
    return [_Monomial(j) for j in range(nb_monomials)]


def compute_nb_monomials(max_degree, problem_dimension=2):
    """C(max_degree + dim, max_degree) (``utils.py:142-144``)."""
    return math.comb(max_degree + problem_dimension, max_degree)


class SteadySol(namedtuple("PDESolution", ["vals", "coeffs", "mat"])):
    """(vals, coeffs, mat) as in ``utils.py:148``.

    ``mat`` (B = (D inv(A))[:, :N]) is computed on first access: the solver
    itself works on the augmented system and never forms B.
    """

    def __new__(cls, vals, coeffs, mat):
        return super().__new__(cls, vals, coeffs, mat)

    def __getattribute__(self, name):
        v = super().__getattribute__(name)
        if name == "mat" and callable(v) and not torch.is_tensor(v):
            return v()
        return v

    def __getitem__(self, i):
        v = tuple.__getitem__(self, i)
        if i == 2 and callable(v) and not torch.is_tensor(v):
            return v()
        return v

    def __iter__(self):
        for i in range(3):
            yield self[i]


__all__ = ["distance", "multiquadric", "inverse_multiquadric", "gaussian", "polyharmonic", "thin_plate",
           "multiquadric_func", "inv_multiquadric_func", "gaussian_func", "polyharmonic_func",
           "thin_plate_func", "make_monomial", "make_all_monomials", "compute_nb_monomials", "SteadySol",
           "rbf_spec", "MONOMIAL", "CENTER"]
