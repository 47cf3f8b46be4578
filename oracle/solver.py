"""The reference solve and field evaluation (oracle, test infrastructure only).

* ``pde_solver_bpath`` restates ``updes/operators.py:559-618``: B = (D inv(A))[:, :N],
  Householder QR solve of B u = q (lineax ``QR()``, operators.py:612-613) and
  coefficients inv(A) [u; 0] (operators.py:616).
* ``pde_solver_kpath`` solves the algebraically identical augmented system
  K c = [q; 0] with one LU (scipy getrf/getrs) and recovers u = [Phi P] c
  (zero-diagonal Phi).  Used as the FP64 reference at sizes where the B-path
  is too slow for a test.
* ``value/gradient/laplacian`` restate ``operators.py:118-187, 319-351``: sums
  over all N centres including the evaluation point itself (phi(0) for the
  value, 0 for the derivatives).
"""
import numpy as np
import scipy.linalg as sla

from . import rbf as R
from . import assembly as AS


def eval_ops(points, coeffs, centers, kind, param, M=None):
    """All operator values (NOPS, P) of the field with coefficients ``coeffs``."""
    centers = np.asarray(centers, np.float64)
    N = centers.shape[0]
    coeffs = np.asarray(coeffs, np.float64)
    lam, gam = coeffs[:N], coeffs[N:]
    pts = np.atleast_2d(np.asarray(points, np.float64))
    out = np.zeros((R.NOPS, pts.shape[0]))
    chunk = max(1, 2_000_000 // max(N, 1))
    for s in range(0, pts.shape[0], chunk):
        p = pts[s:s + chunk]
        ops = R.rbf_ops(kind, param, p[:, 0:1], p[:, 1:2], centers[None, :, 0], centers[None, :, 1])
        out[:, s:s + chunk] = ops @ lam
        for k in range(gam.shape[0]):
            out[:, s:s + chunk] += gam[k] * R.monomial_ops(k, p[:, 0], p[:, 1])
    return out


def value(points, coeffs, centers, kind, param):
    return eval_ops(points, coeffs, centers, kind, param)[R.VAL]


def gradient(points, coeffs, centers, kind, param):
    o = eval_ops(points, coeffs, centers, kind, param)
    return np.stack([o[R.DX], o[R.DY]], axis=-1)


def laplacian(points, coeffs, centers, kind, param):
    return eval_ops(points, coeffs, centers, kind, param)[R.LAP]


def qr_solve(B, q):
    """Householder-QR least-squares solve (lineax QR on a square matrix)."""
    Q, Rm = np.linalg.qr(B)
    return sla.solve_triangular(Rm, Q.T @ q)


def pde_solver_bpath(desc, cloud, kind, param, M, q, robin_betas=None):
    """Reference algorithm: returns (vals, coeffs, B)."""
    D = AS.assemble_D(desc, cloud, kind, param, M, robin_betas)
    invA = np.linalg.inv(AS.assemble_A(cloud, kind, param, M))
    B = (D @ invA)[:, :cloud.N]
    u = qr_solve(B, q)
    c = invA @ np.concatenate([u, np.zeros(M)])
    return u, c, B


def pde_solver_kpath(desc, cloud, kind, param, M, q, robin_betas=None):
    """Augmented-system FP64 LU solve: returns (vals, coeffs, K)."""
    K = AS.assemble_K(desc, cloud, kind, param, M, robin_betas)
    lu = sla.lu_factor(K, check_finite=False)
    c = sla.lu_solve(lu, np.concatenate([q, np.zeros(M)]), check_finite=False)
    A = AS.assemble_A(cloud, kind, param, M)
    u = A[:cloud.N] @ c
    return u, c, K


def field_coefficients(field, cloud, kind, param, M):
    """``assembly.py:404-410`` (fields of length N) / pass-through for length N+M."""
    field = np.asarray(field, np.float64)
    if field.shape[0] == cloud.N:
        return AS.core_compute_coefficients(field, cloud, kind, param, M)
    return field


def integrate_field_points(Nx, Ny):
    """Quadrature points and weights of ``integrate_field`` (operators.py:381-450).

    Restated with the reference's quirks: the interior points are
    ``linspace(h, 1-h, Nx-1)`` x ``linspace(h, 1-h, Ny-1)`` stacked as
    ``array(meshgrid(x, y))`` (shape (2, Ny-1, Nx-1)) and reshaped straight to
    (nb_squares, 2), which pairs consecutive entries (operators.py:402-404);
    the edges use Nx-2 / Ny-2 points on the same span (408-429); the top-left
    corner term evaluates at (1, 0) because it reuses ``quad_points``
    (operators.py:439-441).  Weights: 1 inside, 1/2 on the edges, 1/4 at the
    corners, all times the cell area.
    """
    hx, hy = 1.0 / (Nx - 1), 1.0 / (Ny - 1)
    qx = np.linspace(hx, 1 - hx, Nx - 1)
    qy = np.linspace(hy, 1 - hy, Ny - 1)
    inner = np.array(np.meshgrid(qx, qy)).reshape((Nx - 1) * (Ny - 1), 2)
    ey = np.linspace(hy, 1 - hy, Ny - 2)
    ex = np.linspace(hx, 1 - hx, Nx - 2)
    left = np.stack([np.zeros(Ny - 2), ey], -1)
    right = np.stack([np.ones(Ny - 2), ey], -1)
    bottom = np.stack([ex, np.zeros(Nx - 2)], -1)
    top = np.stack([ex, np.ones(Nx - 2)], -1)
    corners = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 0.0], [1.0, 1.0]])
    pts = np.concatenate([inner, left, right, bottom, top, corners])
    w = np.concatenate([np.ones(len(inner)), np.full(2 * (Ny - 2) + 2 * (Nx - 2), 0.5), np.full(4, 0.25)])
    return pts, w * hx * hy


def integrate_field(coeffs, cloud, kind, param):
    """Integral of the field with coefficients ``coeffs`` (operators.py:381-450)."""
    pts, w = integrate_field_points(cloud.Nx, cloud.Ny)
    return float(w @ value(pts, coeffs, cloud.sorted_nodes, kind, param))
