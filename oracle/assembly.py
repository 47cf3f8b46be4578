"""Dense collocation assembly (oracle, test infrastructure only).

Restates ``updes/assembly.py``.  Every RBF block is written through the
support mask of ``support_mask``: row i of Phi, op(Phi) and the d/n/r rows of
bd(Phi) keep only the columns of node i's support, the periodic rows the union
of the supports of both nodes of the pair (``assembly.py:31-32, 127-128,
171-258``).  With ``support_size="max"`` a support is every other node
(``cloud.py:110-111``): a zero diagonal, and all N columns in periodic rows.

Operators are given as descriptors instead of Python callables: a descriptor is
``{"rbf": [c_VAL, c_DX, c_DY, c_LAP, c_DXX, c_DYY, c_DXY],
   "mon": [...same for the monomial block...]}``
where each c is None, a scalar or an (Ni,) array of per-row coefficients.  A
built-in operator such as ``nodal_laplacian`` is {"rbf": [0,0,0,1,0,0,0], ...}.
"""
import numpy as np

from . import rbf as R


def _pairs(kind, param, xr, yr, xc, yc):
    return R.rbf_ops(kind, param, xr[:, None], yr[:, None], xc[None, :], yc[None, :])


def _combine(ops, coeffs):
    """ops: (NOPS, rows, cols) ; coeffs: list of NOPS scalars / (rows,) arrays."""
    acc = np.zeros(ops.shape[1:])
    for t, c in enumerate(coeffs):
        if c is None:
            continue
        c = np.asarray(c, dtype=np.float64)
        if c.ndim == 0:
            if c == 0.0:
                continue
            acc += c * ops[t]
        else:
            acc += c.reshape((-1,) + (1,) * (ops.ndim - 2)) * ops[t]
    return acc


def support_mask(cloud):
    """(N, N) bool: column j is in the support of node i (``cloud.py:83-112``)."""
    N = cloud.N
    sup = getattr(cloud, "supports", None)
    if sup is None:
        return ~np.eye(N, dtype=bool)
    m = np.zeros((N, N), dtype=bool)
    m[np.repeat(np.arange(N), sup.shape[1]), np.asarray(sup).ravel()] = True
    return m


def assemble_Phi(cloud, kind, param):
    """``assembly.py:10-36``: Phi[i, supp(i)] = rbf(x_i, x_supp); other entries 0."""
    x, y = cloud.sorted_nodes[:, 0], cloud.sorted_nodes[:, 1]
    Phi = _pairs(kind, param, x, y, x, y)[R.VAL]
    return np.where(support_mask(cloud), Phi, 0.0)


def assemble_P(cloud, M):
    """``assembly.py:39-59``: P[:, j] = monomial_j(nodes)."""
    x, y = cloud.sorted_nodes[:, 0], cloud.sorted_nodes[:, 1]
    return np.stack([R.monomial_ops(k, x, y)[R.VAL] for k in range(M)], axis=1)


def assemble_A(cloud, kind, param, M):
    """``assembly.py:62-85``: A = [[Phi, P], [P^T, 0]]."""
    Phi = assemble_Phi(cloud, kind, param)
    P = assemble_P(cloud, M)
    N = cloud.N
    A = np.zeros((N + M, N + M))
    A[:N, :N] = Phi
    A[:N, N:] = P
    A[N:, :N] = P.T
    return A


def assemble_op_Phi_P(desc, cloud, kind, param, M):
    """``assembly.py:93-137``: op applied at the Ni internal nodes."""
    Ni = cloud.Ni
    x, y = cloud.sorted_nodes[:, 0], cloud.sorted_nodes[:, 1]
    ops = _pairs(kind, param, x[:Ni], y[:Ni], x, y)
    opPhi = np.where(support_mask(cloud)[:Ni], _combine(ops, desc["rbf"]), 0.0)
    opP = np.zeros((Ni, M))
    for k in range(M):
        mops = R.monomial_ops(k, x[:Ni], y[:Ni])
        opP[:, k] = _combine(mops[:, :, None], desc["mon"])[:, 0]
    return opPhi, opP


def assemble_bd_Phi_P(cloud, kind, param, M, robin_betas=None):
    """``assembly.py:141-362``: boundary rows [d | n | r | pd | pn].

    ``robin_betas``: (Nr,) coefficients of the Robin nodes sorted by node id
    (``cloud.sort_dict_by_keys(robin_coeffs)``, assembly.py:192).
    Robin rows of bd(Phi) take their normal at ``sorted_outward_normals[i-Ni-Nd-Nn]``
    exactly as ``assembly.py:206`` does (the bd(P) rows use the node's own
    normal, ``assembly.py:303``); the two agree only when Nn == 0.
    """
    N, Ni, Nd, Nn, Nr, Np = cloud.N, cloud.Ni, cloud.Nd, cloud.Nn, cloud.Nr, cloud.Np
    nodes = cloud.sorted_nodes
    x, y = nodes[:, 0], nodes[:, 1]
    nrm = cloud.sorted_outward_normals
    nb = Nd + Nn + Nr + sum(Np)
    bdPhi = np.zeros((nb, N))
    bdP = np.zeros((nb, M))
    mask = support_mask(cloud)

    def ops_at(i_rows):
        return _pairs(kind, param, x[i_rows], y[i_rows], x, y)

    def mon_at(i_rows):
        return np.stack([R.monomial_ops(k, x[i_rows], y[i_rows]) for k in range(M)], axis=-1)

    # Dirichlet (169-176, 283-287)
    rows = np.arange(Ni, Ni + Nd)
    if Nd:
        bdPhi[rows - Ni] = np.where(mask[rows], ops_at(rows)[R.VAL], 0.0)
        bdP[rows - Ni] = mon_at(rows)[R.VAL]
    # Neumann (178-190, 290-299)
    rows = np.arange(Ni + Nd, Ni + Nd + Nn)
    if Nn:
        o = ops_at(rows)
        n = nrm[rows - Ni - Nd]
        v = o[R.DX] * n[:, 0:1] + o[R.DY] * n[:, 1:2]
        bdPhi[rows - Ni] = np.where(mask[rows], v, 0.0)
        m = mon_at(rows)
        nn = np.array([cloud.outward_normals[i] for i in rows])
        bdP[rows - Ni] = m[R.DX] * nn[:, 0:1] + m[R.DY] * nn[:, 1:2]
    # Robin (194-212, 302-315)
    rows = np.arange(Ni + Nd + Nn, Ni + Nd + Nn + Nr)
    if Nr:
        betas = np.zeros(Nr) if robin_betas is None else np.asarray(robin_betas, np.float64)
        o = ops_at(rows)
        n_bug = nrm[rows - Ni - Nd - Nn]
        v = betas[:, None] * o[R.VAL] + o[R.DX] * n_bug[:, 0:1] + o[R.DY] * n_bug[:, 1:2]
        bdPhi[rows - Ni] = np.where(mask[rows], v, 0.0)
        m = mon_at(rows)
        nn = np.array([cloud.outward_normals[i] for i in rows])
        bdP[rows - Ni] = betas[:, None] * m[R.VAL] + m[R.DX] * nn[:, 0:1] + m[R.DY] * nn[:, 1:2]
    # Periodic (215-267, 317-359)
    half = sum(Np) // 2
    start, jump = Ni + Nd + Nn + Nr, 0
    for nbp in Np:
        nc = nbp // 2
        i1 = np.arange(start, start + nc)
        i2 = i1 + nc
        o1, o2 = ops_at(i1), ops_at(i2)
        m1, m2 = mon_at(i1), mon_at(i2)
        rpd = i1 - Ni - jump
        pmask = mask[i1] | mask[i2]        # concatenated supports (219-221, 242-244)
        bdPhi[rpd] = np.where(pmask, o1[R.VAL] - o2[R.VAL], 0.0)
        bdP[rpd] = m1[R.VAL] - m2[R.VAL]
        n1 = nrm[i1 - Ni - Nd]
        n2 = nrm[i1 - Ni - Nd + nc]
        g1 = o1[R.DX] * n1[:, 0:1] + o1[R.DY] * n1[:, 1:2]
        g2 = o2[R.DX] * (-n2[:, 0:1]) + o2[R.DY] * (-n2[:, 1:2])
        rpn = rpd + half
        bdPhi[rpn] = np.where(pmask, g1 - g2, 0.0)
        nn1 = np.array([cloud.outward_normals[i] for i in i1])
        nn2 = np.array([cloud.outward_normals[i] for i in i2])
        gm1 = m1[R.DX] * nn1[:, 0:1] + m1[R.DY] * nn1[:, 1:2]
        gm2 = m2[R.DX] * (-nn2[:, 0:1]) + m2[R.DY] * (-nn2[:, 1:2])
        bdP[rpn] = gm1 - gm2
        start += nbp
        jump += nc
    return bdPhi, bdP


def assemble_D(desc, cloud, kind, param, M, robin_betas=None):
    """The N x (N+M) operator matrix D = [[opPhi, opP], [bdPhi, bdP]] (assembly.py:384-396)."""
    opPhi, opP = assemble_op_Phi_P(desc, cloud, kind, param, M)
    bdPhi, bdP = assemble_bd_Phi_P(cloud, kind, param, M, robin_betas)
    return np.block([[opPhi, opP], [bdPhi, bdP]])


def assemble_K(desc, cloud, kind, param, M, robin_betas=None):
    """Augmented collocation matrix K = [[D], [P^T, 0]] ((N+M) x (N+M)).

    K c = [q; 0] is algebraically the reference system: with c = A^-1 [u; 0]
    (assembly.py:404-410) the second block row of A gives P^T lambda = 0 and
    B u = D A^-1 [u; 0] = D c = q (assembly.py:399-401).
    """
    D = assemble_D(desc, cloud, kind, param, M, robin_betas)
    P = assemble_P(cloud, M)
    K = np.zeros((cloud.N + M, cloud.N + M))
    K[:cloud.N] = D
    K[cloud.N:, :cloud.N] = P.T
    return K


def assemble_B(desc, cloud, kind, param, M, robin_betas=None):
    """``assembly.py:366-401``: B = (D @ inv(A))[:, :N]."""
    D = assemble_D(desc, cloud, kind, param, M, robin_betas)
    invA = np.linalg.inv(assemble_A(cloud, kind, param, M))
    return (D @ invA)[:, :cloud.N]


def core_compute_coefficients(field, cloud, kind, param, M):
    """``assembly.py:404-410``: inv(A) @ [field; 0]."""
    rhs = np.concatenate([np.asarray(field, np.float64), np.zeros(M)])
    return np.linalg.inv(assemble_A(cloud, kind, param, M)) @ rhs


def robin_coefficients(cloud, betas_by_facet):
    """``operators.py:512-541`` + ``cloud.py:72-81``: Robin betas sorted by node id.

    ``robin_coeffs[i] = betas[i - node_ids[0]]`` for the facet's nodes; the
    index is clamped like JAX's out-of-bounds gather (it only runs past the end
    when two Robin facets interleave their ids).
    """
    rc = {}
    for f, t in cloud.facet_types.items():
        if t == "r":
            ids = cloud.facet_nodes[f]
            b = np.asarray(betas_by_facet.get(f, np.zeros(len(ids))), np.float64).reshape(-1)
            for i in ids:
                rc[i] = b[min(i - ids[0], len(b) - 1)]
    return np.array([rc[k] for k in sorted(rc)]) if rc else None


def assemble_q(cloud, rhs_interior, bc_arrays):
    """``assembly.py:434-485`` for an already-evaluated interior rhs.

    rhs_interior: (Ni,) values (or scalar) of rhs_operator at the internal
    nodes; bc_arrays: {facet: (len(facet_nodes),) values}; periodic facets are
    zeroed (``operators.py:546-555``) and Robin tuples carry their value first.
    """
    q = np.zeros(cloud.N)
    q[:cloud.Ni] = rhs_interior
    for f, t in cloud.facet_types.items():
        assert f in bc_arrays, "facets and boundary functions don't match ids"
        ids = np.array(cloud.facet_nodes[f])
        if t[0] == "p":
            q[ids] = 0.0
        else:
            val = bc_arrays[f][0] if isinstance(bc_arrays[f], tuple) else bc_arrays[f]
            q[ids] = val
    return q
