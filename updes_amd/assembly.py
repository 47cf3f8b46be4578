"""Assembly entry points with the reference's names (``updes/assembly.py``).

Each returns torch tensors computed by the HIP assembly kernel (K1); they are
what the parity tests compare against the oracle.  The solver itself never
forms B: it factors the augmented collocation matrix K = [[D], [P^T 0]].
"""
import torch

from . import _lib as L
from . import engine as E
from .operators import (core_compute_coefficients, compute_coefficients, get_field_coefficients, assemble_q,
                        lower_diff_operator, duplicate_robin_coeffs, _robin_beta_vector, _dev)
from .utils import rbf_spec


def _problem(cloud, rbf, nb_monomials, desc=None, robin_beta=None):
    _dev()
    kind, param = rbf_spec(rbf)
    return E.Problem(E.device_cloud(cloud), kind, param, nb_monomials, desc, robin_beta)


def assemble_A(cloud, rbf, nb_monomials=2):
    """A = [[Phi, P], [P^T, 0]] (``assembly.py:62-85``)."""
    p = _problem(cloud, rbf, nb_monomials)
    return E.assemble(p, L.MAT_A)[:, :p.n]


def assemble_Phi(cloud, rbf):
    """Phi with zero diagonal (``assembly.py:10-36``)."""
    return assemble_A(cloud, rbf, 0)


def assemble_P(cloud, nb_monomials):
    """P[:, j] = monomial_j(nodes) (``assembly.py:39-59``)."""
    from .utils import polyharmonic
    A = assemble_A(cloud, polyharmonic, nb_monomials)
    return A[:cloud.N, cloud.N:]


def assemble_invert_A(cloud, rbf, nb_monomials):
    """inv(A) (``assembly.py:87-90``): one multi-right-hand-side solve with the
    FP64 LU factors of A."""
    kind, param = rbf_spec(rbf)
    fac = E.a_factor(cloud, kind, param, nb_monomials)
    return E.lu_inverse(fac)[:, :fac.n]


def assemble_K(operator, cloud, rbf, nb_monomials, diff_args=None, robin_coeffs=None):
    """Augmented collocation matrix [[opPhi, opP], [bdPhi, bdP], [P^T, 0]]."""
    desc = lower_diff_operator(operator, cloud, rbf, diff_args)
    beta = _robin_beta_vector(cloud, robin_coeffs or {})
    p = _problem(cloud, rbf, nb_monomials, desc, beta)
    return E.assemble(p, L.MAT_K)[:, :p.n]


def assemble_op_Phi_P(operator, cloud, rbf, nb_monomials, args):
    """(opPhi, opP) over the internal nodes (``assembly.py:93-137``)."""
    K = assemble_K(operator, cloud, rbf, nb_monomials, args)
    return K[:cloud.Ni, :cloud.N], K[:cloud.Ni, cloud.N:]


def assemble_bd_Phi_P(cloud, rbf, nb_monomials, robin_coeffs=None):
    """(bdPhi, bdP) over the boundary nodes (``assembly.py:141-362``)."""
    from .operators import nodal_value
    op = lambda x, center=None, rbf=None, monomial=None, fields=None: nodal_value(x, center, rbf, monomial)
    K = assemble_K(op, cloud, rbf, nb_monomials, None, robin_coeffs)
    return K[cloud.Ni:cloud.N, :cloud.N], K[cloud.Ni:cloud.N, cloud.N:]


def assemble_B(operator, cloud, rbf, nb_monomials, diff_args, robin_coeffs):
    """B = (D inv(A))[:, :N] (``assembly.py:366-401``): inv(A) by one
    multi-right-hand-side solve, D inv(A) by one MFMA GEMM."""
    desc = lower_diff_operator(operator, cloud, rbf, diff_args)
    beta = _robin_beta_vector(cloud, robin_coeffs or {})
    p = _problem(cloud, rbf, nb_monomials, desc, beta)
    kind, param = rbf_spec(rbf)
    return E.b_matrix(p, E.a_factor(cloud, kind, param, nb_monomials))


__all__ = ["assemble_A", "assemble_Phi", "assemble_P", "assemble_invert_A", "assemble_K", "assemble_op_Phi_P",
           "assemble_bd_Phi_P", "assemble_B", "assemble_q", "core_compute_coefficients", "compute_coefficients",
           "get_field_coefficients"]
