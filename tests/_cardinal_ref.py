"""NumPy restatement of the local cardinal basis that preconditions the FP32
leg of the solve (updes_amd/csrc/cardinal.hip), for the parity tests of its
kernels.  Test infrastructure: the library never imports it.

For node j with stencil S_j (its k nearest nodes, itself included) the local
system [[Phi_S, P_S], [P_S^T, 0]] [beta'; a'] = [e_j; 0] is solved with Phi_S
zero on its diagonal (as the reference's A, updes/assembly.py:31-32) and the
monomials in the local coordinates (x - x_j) / s (s: the stencil radius);
beta' is projected onto the null space of P_S^T, and the local polynomial is
expanded into the global monomials of the problem.
"""
import math

import numpy as np

from oracle import rbf as R


def coefficients(nodes, stencil, kind, param, M):
    """(beta (N, k), alpha (N, M)) of every node's cardinal function."""
    nodes = np.asarray(nodes, np.float64)
    N, k = stencil.shape
    beta = np.zeros((N, k))
    alpha = np.zeros((N, M))
    for j in range(N):
        S = stencil[j]
        c = nodes[j]
        pts = nodes[S]
        d = pts - c
        s = np.sqrt((d ** 2).sum(1)).max()
        s = s if s > 0 else 1.0
        u = d / s
        Phi = R.rbf_ops(kind, param, pts[:, None, 0], pts[:, None, 1], pts[None, :, 0], pts[None, :, 1])[R.VAL]
        np.fill_diagonal(Phi, 0.0)
        P = np.stack([R.monomial_ops(m, u[:, 0], u[:, 1])[R.VAL] for m in range(M)], 1) if M else np.zeros((k, 0))
        A = np.block([[Phi, P], [P.T, np.zeros((M, M))]])
        rhs = np.zeros(k + M)
        rhs[int(np.flatnonzero(S == j)[0])] = 1.0
        sol = np.linalg.solve(A, rhs)
        b = sol[:k]
        if M:
            Q, _ = np.linalg.qr(P)
            b = b - Q @ (Q.T @ b)
            b = b - Q @ (Q.T @ b)
        beta[j] = b
        for t in range(M):
            ti, tj = R.MONOMIAL_EXPONENTS[t]
            acc = 0.0
            for m in range(M):
                ea, eb = R.MONOMIAL_EXPONENTS[m]
                if ea < ti or eb < tj:
                    continue
                acc += sol[k + m] * s ** -(ea + eb) * math.comb(ea, ti) * (-c[0]) ** (ea - ti) * \
                    math.comb(eb, tj) * (-c[1]) ** (eb - tj)
            alpha[j, t] = acc
    return beta, alpha


def basis_matrix(stencil, beta, alpha):
    """Dense C (N + M, N): column j = [beta_j scattered on S_j; alpha_j]."""
    N, k = stencil.shape
    M = alpha.shape[1]
    C = np.zeros((N + M, N))
    for j in range(N):
        C[stencil[j], j] = beta[j]
        C[N:, j] = alpha[j]
    return C


def form(K, C, stencil):
    """(rs, K~ = diag(rs) D C) with rs[i] = 1 / max over j in S_i of |(D C)[i, j]|."""
    N = stencil.shape[0]
    DC = K[:N] @ C
    rs = np.array([1.0 / np.abs(DC[i, stencil[i]]).max() for i in range(N)])
    return rs, rs[:, None] * DC
