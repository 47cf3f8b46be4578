"""Device-side engine: clouds and operators as ``rbfx_problem``s, assembly,
LU factorisations, solves and evaluations, all through the HIP library.

Everything here runs on the GPU through ``librbfx.so``; there is no host
fallback (``_lib.require_gpu`` raises when no device is present).
"""
import collections
import ctypes
import hashlib
import itertools

import numpy as np
import torch

from . import _lib as L


def _ld(n):
    return (n + 63) // 64 * 64


# ---- cloud ------------------------------------------------------------------
class DeviceCloud:
    """Node coordinates, normals and periodic pairs of a cloud on the device."""

    def __init__(self, cloud, device):
        self.N, self.Ni, self.Nd, self.Nn, self.Nr = cloud.N, cloud.Ni, cloud.Nd, cloud.Nn, cloud.Nr
        self.Np = list(cloud.Np)
        nodes = torch.as_tensor(cloud.sorted_nodes, dtype=torch.float64)
        self.nodes = nodes.to(device)
        self.x = self.nodes[:, 0].contiguous()
        self.y = self.nodes[:, 1].contiguous()
        nrm = getattr(cloud, "sorted_outward_normals", None)
        nb = self.Nn + self.Nr + sum(self.Np)
        if nb > 0:
            nrm = torch.as_tensor(nrm, dtype=torch.float64).to(device).reshape(-1, 2)
            if nrm.shape[0] != nb:
                raise ValueError("sorted_outward_normals has %d rows, expected %d" % (nrm.shape[0], nb))
            self.nx = nrm[:, 0].contiguous()
            self.ny = nrm[:, 1].contiguous()
            # Robin phi rows read the normal at index i-Ni-Nd-Nn (updes/assembly.py:206)
            if self.Nr:
                self.rnx = self.nx[0:self.Nr].contiguous()
                self.rny = self.ny[0:self.Nr].contiguous()
            else:
                self.rnx = self.rny = None
        else:
            self.nx = self.ny = self.rnx = self.rny = None
        # periodic pairs: group g occupies [start, start+Np[g]), node i pairs i + Np[g]/2
        i1, i2 = [], []
        start = self.Ni + self.Nd + self.Nn + self.Nr
        for nbp in self.Np:
            nc = nbp // 2
            i1.extend(range(start, start + nc))
            i2.extend(range(start + nc, start + 2 * nc))
            start += nbp
        self.Npd = len(i1)
        if self.Npd:
            self.pd1 = torch.tensor(i1, dtype=torch.int64, device=device)
            self.pd2 = torch.tensor(i2, dtype=torch.int64, device=device)
        else:
            self.pd1 = self.pd2 = None
        if self.Ni + self.Nd + self.Nn + self.Nr + 2 * self.Npd != self.N:
            raise ValueError("periodic facets must pair up node for node (odd periodic group size)")
        # local supports (support_size < N): (N, S) ascending neighbour ids
        sup = support_table(cloud)
        self.local = sup is not None
        self.S = 0 if sup is None else int(sup.shape[1])
        self.sup = None
        if self.local and self.S > 0:
            self.sup = torch.as_tensor(sup, dtype=torch.int32).to(device).contiguous()


def support_table(cloud):
    """Ascending (N, support_size-1) int32 neighbour table, or None when every
    node is in every support (``updes/cloud.py:97-112``)."""
    sup = cloud.__dict__.get("_support_sorted", False)
    if sup is not False:
        return sup
    k = getattr(cloud, "support_size", "max")
    if k in ("max", None) or int(k) >= cloud.N:
        return None
    sls = np.asarray(torch.as_tensor(cloud.sorted_local_supports).cpu().numpy(), dtype=np.int64)
    if sls.shape != (cloud.N, int(k) - 1) or sls.min(initial=0) < 0 or sls.max(initial=0) >= cloud.N:
        raise ValueError("sorted_local_supports must be (N, support_size-1) node ids")
    return np.sort(sls, axis=1).astype(np.int32)


_TOKENS = itertools.count(1)


def cloud_key(cloud):
    """Cache key of a cloud: a token owned by the cloud object (never reused,
    unlike ``id()``) plus its row-class sizes and node data."""
    tok = cloud.__dict__.get("_rbfx_token")
    if tok is None:
        tok = next(_TOKENS)
        cloud.__dict__["_rbfx_token"] = tok
    nodes = cloud.sorted_nodes
    data = nodes.data_ptr() if isinstance(nodes, torch.Tensor) else id(nodes)
    return (tok, int(cloud.N), int(cloud.Ni), int(cloud.Nd), int(cloud.Nn), int(cloud.Nr), tuple(cloud.Np),
            data, str(getattr(cloud, "support_size", "max")))


def device_cloud(cloud):
    """The cloud's DeviceCloud on the current device, built once and cached on
    the cloud (with its stencil and support tables)."""
    dev = torch.device("cuda", torch.cuda.current_device())
    dc = cloud.__dict__.get("_rbfx_device_cloud")
    if dc is None or dc.nodes.device != dev or dc.nodes.data_ptr() == 0:
        dc = DeviceCloud(cloud, dev)
        cloud.__dict__["_rbfx_device_cloud"] = dc
    return dc


# ---- operator descriptor ----------------------------------------------------
class OpDesc:
    """Coefficients of a lowered nodal operator (constant part + per-row part)."""

    def __init__(self, rbf_const=None, mon_const=None, rbf_rows=None, mon_rows=None):
        self.rbf_const = list(rbf_const or [0.0] * L.NOPS)
        self.mon_const = list(mon_const or [0.0] * L.NOPS)
        self.rbf_rows = list(rbf_rows or [None] * L.NOPS)
        self.mon_rows = list(mon_rows or [None] * L.NOPS)

    @staticmethod
    def builtin(op):
        c = [0.0] * L.NOPS
        c[op] = 1.0
        return OpDesc(c, list(c))

    def key(self):
        h = hashlib.sha1()
        h.update(np.asarray(self.rbf_const + self.mon_const, dtype=np.float64).tobytes())
        for r in self.rbf_rows + self.mon_rows:
            if r is None:
                h.update(b"-")
            else:
                h.update(r.detach().double().cpu().numpy().tobytes())
        return h.hexdigest()

    def to(self, device):
        cv = lambda r: None if r is None else r.detach().to(device=device, dtype=torch.float64).contiguous()
        return OpDesc(self.rbf_const, self.mon_const, [cv(r) for r in self.rbf_rows], [cv(r) for r in self.mon_rows])


class Problem:
    """A ctypes ``rbfx_problem`` plus the tensors it points to (kept alive)."""

    def __init__(self, dc, kind, param, M, desc=None, robin_beta=None):
        self.dc, self.kind, self.param, self.M = dc, int(kind), float(param), int(M)
        self.n = dc.N + self.M
        p = L.Problem()
        p.N, p.M, p.kernel, p.flags, p.param = dc.N, self.M, self.kind, 0, self.param
        p.d_x, p.d_y = dc.x.data_ptr(), dc.y.data_ptr()
        p.Ni, p.Nd, p.Nn, p.Nr, p.Npd = dc.Ni, dc.Nd, dc.Nn, dc.Nr, dc.Npd
        p.d_nx = dc.nx.data_ptr() if dc.nx is not None else None
        p.d_ny = dc.ny.data_ptr() if dc.ny is not None else None
        p.d_robin_normal_x = dc.rnx.data_ptr() if dc.rnx is not None else None
        p.d_robin_normal_y = dc.rny.data_ptr() if dc.rny is not None else None
        self.beta = None
        if dc.Nr:
            if robin_beta is None:
                robin_beta = torch.zeros(dc.Nr, dtype=torch.float64)
            self.beta = torch.as_tensor(robin_beta, dtype=torch.float64).to(dc.x.device).reshape(-1).contiguous()
            if self.beta.numel() != dc.Nr:
                raise ValueError("expected %d Robin coefficients, got %d" % (dc.Nr, self.beta.numel()))
            p.d_beta = self.beta.data_ptr()
        p.d_pd_i1 = dc.pd1.data_ptr() if dc.pd1 is not None else None
        p.d_pd_i2 = dc.pd2.data_ptr() if dc.pd2 is not None else None
        if dc.local:
            p.flags |= L.FLAG_LOCAL_SUPPORT
            p.support_cols = dc.S
            p.d_support = dc.sup.data_ptr() if dc.sup is not None else None
        self.desc = desc.to(dc.x.device) if desc is not None else OpDesc()
        for t in range(L.NOPS):
            p.op_rbf_const[t] = float(self.desc.rbf_const[t])
            p.op_mon_const[t] = float(self.desc.mon_const[t])
            r, m = self.desc.rbf_rows[t], self.desc.mon_rows[t]
            if r is not None and r.numel() != dc.Ni:
                raise ValueError("per-row operator coefficients must have Ni=%d entries" % dc.Ni)
            if m is not None and m.numel() != dc.Ni:
                raise ValueError("per-row operator coefficients must have Ni=%d entries" % dc.Ni)
            p.d_op_rbf_row[t] = r.data_ptr() if r is not None else None
            p.d_op_mon_row[t] = m.data_ptr() if m is not None else None
        self.p = p

    def ref(self):
        return ctypes.byref(self.p)


# ---- primitives -------------------------------------------------------------
def rhs(problem, q_int, bc, terms=()):
    """[q; 0] (n,) through rbfx_rhs: interior values q_int (Ni,) plus the
    terms (op, weight, weight_row or None, coeffs), boundary values bc
    (N - Ni,) in node order."""
    dc = problem.dc
    dev = dc.x.device
    q = torch.empty(problem.n, dtype=torch.float64, device=dev)
    arr = (L.RhsTerm * max(1, len(terms)))()
    keep = []
    for t, (op, w, row, coeffs) in enumerate(terms):
        coeffs = coeffs.detach().to(dev, torch.float64).contiguous()
        row = None if row is None else row.detach().to(dev, torch.float64).contiguous()
        keep += [coeffs, row]
        arr[t].op, arr[t].reserved, arr[t].weight = int(op), 0, float(w)
        arr[t].d_weight_row = None if row is None else row.data_ptr()
        arr[t].d_coeffs = coeffs.data_ptr()
    lib = L.lib()
    ws = L.workspace(lib.rbfx_rhs_workspace(problem.ref()), dev) if terms else None
    L.check(lib.rbfx_rhs(problem.ref(), 0.0, L.ptr(q_int), len(terms), arr, L.ptr(bc), L.ptr(q), L.ptr(ws),
                         0 if ws is None else ws.numel(), L.stream()))
    return q


def assemble(problem, matrix=L.MAT_K, dtype=torch.float64, also32=False):
    """Dense row-major (n, ld) matrix; with also32 returns (fp64, fp32)."""
    n = problem.n
    ld = _ld(n)
    dev = problem.dc.x.device
    out64 = torch.empty((n, ld), dtype=torch.float64, device=dev) if (dtype == torch.float64 or also32) else None
    out32 = torch.empty((n, ld), dtype=torch.float32, device=dev) if (dtype == torch.float32 or also32) else None
    L.check(L.lib().rbfx_assemble(problem.ref(), matrix, 0, n, L.ptr(out64), ld, L.ptr(out32), ld, L.stream()))
    if also32:
        return out64, out32
    return out64 if dtype == torch.float64 else out32


def matvec(problem, x, matrix=L.MAT_K, trans=False):
    x = x.detach().to(torch.float64).contiguous()
    y = torch.empty_like(x)
    wsb = L.lib().rbfx_matvec_workspace(problem.ref(), int(trans))
    ws = L.workspace(wsb, x.device)
    L.check(L.lib().rbfx_matvec(problem.ref(), matrix, int(trans), L.ptr(x), L.ptr(y), L.ptr(ws), ws.numel(),
                                L.stream()))
    return y


def eval_ops(kind, param, centers, coeffs, points, mask):
    """(7, P) operator values of the field ``coeffs`` at ``points`` (rows of mask set)."""
    centers = centers.detach().to(torch.float64)
    cx, cy = centers[:, 0].contiguous(), centers[:, 1].contiguous()
    pts = points.detach().to(torch.float64).reshape(-1, 2)
    px, py = pts[:, 0].contiguous(), pts[:, 1].contiguous()
    coeffs = coeffs.detach().to(torch.float64).contiguous()
    N = cx.numel()
    M = coeffs.numel() - N
    P = px.numel()
    out = torch.zeros((L.NOPS, P), dtype=torch.float64, device=cx.device)
    if P:
        L.check(L.lib().rbfx_eval(int(kind), float(param), N, M, L.ptr(cx), L.ptr(cy), L.ptr(coeffs), P,
                                  L.ptr(px), L.ptr(py), int(mask), L.ptr(out), L.stream()))
    return out


def eval_t(kind, param, centers, M, points, op, w):
    """Adjoint of eval_ops for one operator: (N+M,) vector."""
    centers = centers.detach().to(torch.float64)
    cx, cy = centers[:, 0].contiguous(), centers[:, 1].contiguous()
    pts = points.detach().to(torch.float64).reshape(-1, 2)
    px, py = pts[:, 0].contiguous(), pts[:, 1].contiguous()
    w = w.detach().to(torch.float64).contiguous()
    N = cx.numel()
    out = torch.empty(N + int(M), dtype=torch.float64, device=cx.device)
    L.check(L.lib().rbfx_eval_t(int(kind), float(param), N, int(M), L.ptr(cx), L.ptr(cy), px.numel(), L.ptr(px),
                                L.ptr(py), int(op), L.ptr(w), L.ptr(out), L.stream()))
    return out


class LU:
    """In-place LU factorisation of a row-major (n, ld) matrix, FP64 or FP32."""

    def __init__(self, mat, n):
        self.n = n
        self.a = mat
        self.ld = mat.shape[1]
        self.dtype = mat.dtype
        dev = mat.device
        lib = L.lib()
        nbytes = 8 if self.dtype == torch.float64 else 4
        self.ipiv = torch.empty(n, dtype=torch.int64, device=dev)
        self.info = torch.zeros(1, dtype=torch.int32, device=dev)
        ws = L.workspace(lib.rbfx_getrf_workspace(n, nbytes), dev)
        fn = lib.rbfx_getrf_f64 if nbytes == 8 else lib.rbfx_getrf_f32
        L.check(fn(n, L.ptr(mat), self.ld, L.ptr(self.ipiv), L.ptr(ws), ws.numel(), L.ptr(self.info), L.stream()))
        info = int(self.info.item())
        if info < 0:
            raise L.RbfxError(-4, "LU factorisation failed (non-finite column or device-side wait expired)")
        self.perm = torch.empty(n, dtype=torch.int64, device=dev)
        self.dinv = torch.empty(lib.rbfx_getrs_dinv_elems(n), dtype=torch.float64, device=dev)
        prep = lib.rbfx_getrs_prepare_f64 if nbytes == 8 else lib.rbfx_getrs_prepare_f32
        L.check(prep(n, L.ptr(mat), self.ld, L.ptr(self.ipiv), L.ptr(self.perm), L.ptr(self.dinv), None, 0,
                     L.stream()))
        self.singular = int(self.info.item())
        self._ws = L.workspace(lib.rbfx_getrs_workspace(n), dev)

    def solve(self, b, trans=False):
        """x with A x = b (or A^T x = b), b an (n,) FP64 tensor (not modified).

        Synchronises: a device-side wait of the triangular sweeps that expired
        raises RbfxError (RBFX_ERR_TIMEOUT) instead of returning NaN."""
        if b.numel() != self.n:
            raise ValueError("right-hand side has %d entries, the factorisation is %d x %d"
                             % (b.numel(), self.n, self.n))
        x = b.detach().to(torch.float64).contiguous().clone()
        lib = L.lib()
        fn = lib.rbfx_getrs_f64 if self.dtype == torch.float64 else lib.rbfx_getrs_f32
        L.check(fn(int(trans), self.n, L.ptr(self.a), self.ld, L.ptr(self.perm), L.ptr(self.dinv), L.ptr(x),
                   L.ptr(self._ws), self._ws.numel(), L.stream()))
        L.check(lib.rbfx_getrs_status(self.n, L.ptr(self._ws), self._ws.numel(), L.stream()))
        return x


def lu_inverse(lu):
    """inv(A) from FP64 LU factors: one multi-right-hand-side solve with the
    identity ((n, ld) row-major; columns past n are zero)."""
    if lu.dtype != torch.float64:
        raise ValueError("lu_inverse needs FP64 factors")
    n = lu.n
    lib = L.lib()
    X = torch.zeros((n, _ld(n)), dtype=torch.float64, device=lu.a.device)
    X[:, :n].fill_diagonal_(1.0)
    ws = L.workspace(lib.rbfx_getrs_multi_workspace(n, n), X.device)
    L.check(lib.rbfx_getrs_multi_f64(n, L.ptr(lu.a), lu.ld, L.ptr(lu.perm), L.ptr(lu.dinv), n, L.ptr(X), X.shape[1],
                                     L.ptr(ws), ws.numel(), L.stream()))
    return X


def gemm(A, B, m, n, k, alpha=1.0):
    """C = alpha A[:m, :k] B[:k, :n] (row-major FP64) with the MFMA kernel; C is (m, ld(n))."""
    C = torch.zeros((m, _ld(n)), dtype=torch.float64, device=A.device)
    L.check(L.lib().rbfx_gemm_f64(m, n, k, float(alpha), L.ptr(A), A.shape[1], L.ptr(B), B.shape[1], 0.0, L.ptr(C),
                                  C.shape[1], L.stream()))
    return C


def b_matrix(problem, a_lu):
    """B = (D inv(A))[:, :N] (updes/assembly.py:366-401): D is the first N rows
    of K, inv(A) one multi-right-hand-side solve with A's factors."""
    N, n = problem.dc.N, problem.n
    K = assemble(problem, L.MAT_K, torch.float64)
    inv_a = lu_inverse(a_lu)
    return gemm(K, inv_a, N, N, n)[:, :N]


class SolveStats:
    def __init__(self):
        self.precision = "fp64"
        self.basis = None
        self.ir_outer = 0
        self.ir_inner = 0
        self.ir_relres = None
        self.fell_back = False

    def as_dict(self):
        return dict(precision=self.precision, basis=self.basis, ir_outer=self.ir_outer, ir_inner=self.ir_inner,
                    ir_relres=self.ir_relres, fell_back=self.fell_back)


# ---- local cardinal basis (the FP32 preconditioner's change of basis) ----------
DEFAULT_STENCIL = 15
DEFAULT_SPECIAL = 3   # a 3 x 3 grid of special points


def special_points(X, g):
    """Node ids nearest to a g x g grid of points spread over the cloud's
    bounding box (distinct ids, ascending)."""
    if g <= 0:
        return np.zeros(0, dtype=np.int64)
    from scipy.spatial import cKDTree
    lo, hi = X.min(axis=0), X.max(axis=0)
    t = (np.arange(g) + 0.5) / g
    pts = np.array([[lo[0] + a * (hi[0] - lo[0]), lo[1] + b * (hi[1] - lo[1])] for a in t for b in t])
    _, ids = cKDTree(X).query(pts, 1)
    return np.unique(np.asarray(ids, dtype=np.int64))


def cardinal_stencil(dc, k, g=DEFAULT_SPECIAL):
    """(N, k + #special) int32 stencils of the local cardinal functions, in
    ascending id order (sorted ids make the gathers of the formation kernel
    coalesce): every node's nearest nodes (itself included) and the special
    points, the nodes nearest a g x g grid over the cloud.  The special points
    hold each cardinal function near 0 far from its node, where its local
    polynomial part would otherwise grow quadratically (Beatson, Cherrie and
    Mouat's approximate cardinal functions for RBF interpolation, 1999).
    Cached on the device cloud."""
    cache = dc.__dict__.setdefault("_stencils", {})
    X = dc.nodes.detach().cpu().numpy()
    sp = special_points(X, g) if dc.N > int(k) + g * g else np.zeros(0, dtype=np.int64)
    kt = min(int(k) + sp.size, dc.N)
    st = cache.get((kt, g))
    if st is None:
        from scipy.spatial import cKDTree
        _, nb = cKDTree(X).query(X, kt)
        nb = np.asarray(nb, dtype=np.int64).reshape(dc.N, kt)
        own = np.arange(dc.N)
        missing = ~(nb == own[:, None]).any(axis=1)
        nb[missing, -1] = own[missing]  # coincident nodes: keep the node itself
        if sp.size:
            # the special points, then the nearest nodes not among them (a
            # stable sort moves the special entries of each row to its end)
            order = np.argsort(np.isin(nb, sp), axis=1, kind="stable")
            extra = np.take_along_axis(nb, order, axis=1)[:, :kt - sp.size]
            nb = np.concatenate([np.broadcast_to(sp, (dc.N, sp.size)), extra], axis=1)
        st = np.sort(nb, axis=1).astype(np.int32)
        cache[(kt, g)] = st
    return st


def reverse_stencil(st):
    """CSR of the stencil table by node id: ptr (N+1) int64 and idx int32 with
    idx[ptr[l]:ptr[l+1]] = the entries j*k + t where st[j, t] == l, ascending."""
    N, k = st.shape
    flat = st.reshape(-1).astype(np.int64)
    order = np.argsort(flat, kind="stable")
    ptr = np.zeros(N + 1, dtype=np.int64)
    np.cumsum(np.bincount(flat, minlength=N), out=ptr[1:])
    return ptr, order.astype(np.int32)


FORM_COLS = 256   # columns per block of the union tables of the formation kernel


def union_tables(st, cols=FORM_COLS):
    """For each block of `cols` consecutive nodes: the sorted distinct ids of
    their stencils (CSR ptr int64 / idx int32) and, per stencil entry, its
    position in its block's list (int16; 0 where a list exceeds int16)."""
    N, kt = st.shape
    ptr = [0]
    idx = []
    pos = np.zeros((N, kt), dtype=np.int16)
    for j0 in range(0, N, cols):
        blk = st[j0:j0 + cols]
        U = np.unique(blk)
        if U.size <= np.iinfo(np.int16).max:
            pos[j0:j0 + cols] = np.searchsorted(U, blk).astype(np.int16)
        idx.append(U.astype(np.int32))
        ptr.append(ptr[-1] + U.size)
    return np.asarray(ptr, dtype=np.int64), np.concatenate(idx), pos


def monomial_matrix(X, M):
    from .utils import MONOMIAL_EXPONENTS
    return np.stack([X[:, 0] ** a * X[:, 1] ** b for a, b in MONOMIAL_EXPONENTS[:M]], axis=1) if M else \
        np.zeros((X.shape[0], 0))


def cloud_tables(dc, k, g, M):
    """Device copies of the stencil table, its reverse table and (P^T P)^-1:
    they depend on the cloud alone (and on the degree), so they are built once
    per device cloud, like its support tables."""
    cache = dc.__dict__.setdefault("_card_tables", {})
    key = (int(k), int(g), int(M))
    t = cache.get(key)
    if t is None:
        dev = dc.x.device
        st = cardinal_stencil(dc, k, g)
        ptr, idx = reverse_stencil(st)
        P = monomial_matrix(dc.nodes.detach().cpu().numpy(), M)
        ptp_inv = np.linalg.inv(P.T @ P) if M else np.zeros((1, 1))
        uptr, uidx, upos = union_tables(st)
        t = (torch.as_tensor(st).to(dev), torch.as_tensor(ptr).to(dev), torch.as_tensor(idx).to(dev),
             torch.as_tensor(np.ascontiguousarray(ptp_inv).reshape(-1)).to(dev),
             (torch.as_tensor(uptr).to(dev), torch.as_tensor(uidx).to(dev), torch.as_tensor(upos).to(dev)))
        cache[key] = t
    return t


class CardinalBasis:
    """Local cardinal functions of one (cloud, kernel, degree): stencils,
    coefficients (GPU kernel), reverse stencil table and (P^T P)^-1."""

    def __init__(self, problem, k=DEFAULT_STENCIL, g=DEFAULT_SPECIAL):
        dc = problem.dc
        if dc.local:
            raise ValueError("the cardinal basis needs every node in every support")
        self.dc, self.N, self.M = dc, dc.N, problem.M
        dev = dc.x.device
        self.stencil, self.rev_ptr, self.rev_idx, self.ptp_inv, self.union = cloud_tables(dc, k, g, self.M)
        self.use_union = True
        self.k = int(self.stencil.shape[1])
        self.beta = torch.empty(self.N * self.k, dtype=torch.float64, device=dev)
        self.alpha = torch.empty(max(self.N * self.M, 1), dtype=torch.float64, device=dev)
        self._ws = L.workspace(256, dev)
        c = self.struct(None)
        L.check(L.lib().rbfx_cardinal_coeffs(problem.ref(), ctypes.byref(c), L.ptr(self._ws), self._ws.numel(),
                                             L.stream()))

    def struct(self, rs):
        c = L.Cardinal()
        c.N, c.M, c.k = self.N, self.M, self.k
        c.d_x, c.d_y = self.dc.x.data_ptr(), self.dc.y.data_ptr()
        c.d_stencil = self.stencil.data_ptr()
        c.d_beta, c.d_alpha = self.beta.data_ptr(), self.alpha.data_ptr()
        c.d_rs = rs.data_ptr() if rs is not None else None
        c.d_rev_ptr, c.d_rev_idx = self.rev_ptr.data_ptr(), self.rev_idx.data_ptr()
        c.d_ptp_inv = self.ptp_inv.data_ptr()
        if self.use_union:
            c.d_union_ptr, c.d_union_idx, c.d_union_pos = (t.data_ptr() for t in self.union)
        return c

    def form(self, K64, rs):
        """K~ = diag(rs) D C (N x ld(N), FP32) from the stored FP64 K; fills rs."""
        N = self.N
        out = torch.empty((N, _ld(N)), dtype=torch.float32, device=K64.device)
        c = self.struct(rs)
        L.check(L.lib().rbfx_cardinal_form(ctypes.byref(c), L.ptr(K64), K64.shape[1], L.ptr(out), out.shape[1],
                                           L.ptr(self._ws), self._ws.numel(), L.stream()))
        return out


def frobenius(mat, rows, cols, row_scale=None):
    """Frobenius norm of mat[:rows, :cols], rows scaled by row_scale (a prefix)."""
    lib = L.lib()
    ws = L.workspace(lib.rbfx_frobenius_workspace(rows), mat.device)
    out = ctypes.c_double(0.0)
    nrs = 0 if row_scale is None else row_scale.numel()
    L.check(lib.rbfx_frobenius_f64(rows, cols, L.ptr(mat), mat.shape[1], L.ptr(row_scale), nrs, ctypes.byref(out),
                                   L.ptr(ws), ws.numel(), L.stream()))
    return out.value


def frobenius_pair(mat, rows, cols, row_scale):
    """(scaled, unscaled) Frobenius norms of mat[:rows, :cols] from one read:
    rows scaled by row_scale (a prefix), and without scaling."""
    lib = L.lib()
    ws = L.workspace(lib.rbfx_frobenius_workspace(rows), mat.device)
    scaled, plain = ctypes.c_double(0.0), ctypes.c_double(0.0)
    L.check(lib.rbfx_frobenius_pair_f64(rows, cols, L.ptr(mat), mat.shape[1], L.ptr(row_scale), row_scale.numel(),
                                        ctypes.byref(scaled), ctypes.byref(plain), L.ptr(ws), ws.numel(),
                                        L.stream()))
    return scaled.value, plain.value


class KFactor:
    """Factorisation of the collocation matrix K for one (cloud, rbf, operator).

    precision "fp64": FP64 LU.  precision "fp32_ir": FP64 K is assembled and
    kept for the residuals, the N x N matrix K~ = diag(rs) D C of K in the
    local cardinal basis C is factored in FP32 on MFMA, and flexible GMRES on
    K preconditioned by those factors refines to FP64 accuracy.  With local
    supports (support_size < N) the FP32 LU of K itself preconditions, with
    matrix-free residuals.  When refinement stalls, an FP64 LU is built and
    used instead (fell_back).
    """

    def __init__(self, problem, precision="fp64", ir_tol=1e-12, ir_restart=40, ir_max_inner=160,
                 ir_stencil=DEFAULT_STENCIL, basis=None):
        self.problem = problem
        self.n = problem.n
        self.precision = precision
        self.ir_tol, self.ir_restart, self.ir_max_inner = ir_tol, ir_restart, ir_max_inner
        self.lu64 = None
        self.lu32 = None
        self.K64 = None
        self.basis = None
        self.mode = "fp64"
        self.last_stats = SolveStats()
        if precision == "fp64":
            self.lu64 = LU(assemble(problem, L.MAT_K, torch.float64), self.n)
        elif precision == "fp32_ir" and problem.dc.local:
            self.mode = "lu"
            self.lu32 = LU(assemble(problem, L.MAT_K, torch.float32), self.n)
        elif precision == "fp32_ir":
            self.mode = "cardinal"
            self.K64 = assemble(problem, L.MAT_K, torch.float64)
            self.basis = basis if basis is not None else CardinalBasis(problem, ir_stencil)
            N = problem.dc.N
            self.rs = torch.empty(N, dtype=torch.float64, device=self.K64.device)
            kt = self.basis.form(self.K64, self.rs)
            # norms of the row-scaled forward system and of the transposed one
            self.knorm = frobenius_pair(self.K64, self.n, self.n, self.rs)
            self.lu32 = LU(kt, N)
        else:
            raise ValueError("precision must be 'fp64' or 'fp32_ir'")

    def _fallback(self):
        if self.lu64 is None:
            if self.K64 is not None:
                # the stored FP64 K is factored in place
                K64, self.K64 = self.K64, None
                self.lu64 = LU(K64, self.n)
            else:
                self.lu64 = LU(assemble(self.problem, L.MAT_K, torch.float64), self.n)
        return self.lu64

    def solve(self, b, trans=False):
        if b.numel() != self.n:
            raise ValueError("right-hand side has %d entries, the system is %d x %d" % (b.numel(), self.n, self.n))
        st = SolveStats()
        st.precision = self.precision
        st.basis = self.mode
        if self.precision == "fp64" or (self.lu64 is not None and self.lu32 is None):
            if self.precision != "fp64":
                # refinement stalled on an earlier solve with these factors: keep
                # its statistics, this solve uses the FP64 factors directly
                prev = self.last_stats
                st.fell_back = True
                st.ir_outer, st.ir_inner, st.ir_relres = prev.ir_outer, prev.ir_inner, prev.ir_relres
            x = self.lu64.solve(b, trans)
        else:
            x = self._ir(b, trans, st)
        self.last_stats = st
        return x

    def _ir(self, b, trans, st):
        lib = L.lib()
        n = self.n
        b = b.detach().to(torch.float64).contiguous()
        x = torch.empty_like(b)
        stats = L.IrStats()
        lu = self.lu32
        if self.mode == "cardinal":
            N, M = self.basis.N, self.basis.M
            ws = L.workspace(lib.rbfx_gmres_ir_cardinal_workspace(N, M, self.ir_restart), b.device)
            c = self.basis.struct(self.rs)
            status = lib.rbfx_gmres_ir_cardinal(int(trans), ctypes.byref(c), L.ptr(self.K64), self.K64.shape[1],
                                                float(self.knorm[1 if trans else 0]), L.ptr(lu.a), lu.ld,
                                                L.ptr(lu.perm),
                                                L.ptr(lu.dinv), L.ptr(b), L.ptr(x), float(self.ir_tol),
                                                int(self.ir_restart), int(self.ir_max_inner), L.ptr(ws),
                                                ws.numel(), ctypes.byref(stats), L.stream())
        else:
            ws = L.workspace(max(lib.rbfx_gmres_ir_workspace(n, self.ir_restart),
                                 lib.rbfx_matvec_workspace(self.problem.ref(), int(trans))), b.device)
            status = lib.rbfx_gmres_ir(int(trans), n, self.problem.ref(), None, 0, L.ptr(lu.a), lu.ld,
                                       L.ptr(lu.perm), L.ptr(lu.dinv), L.ptr(b), L.ptr(x), float(self.ir_tol),
                                       int(self.ir_restart), int(self.ir_max_inner), L.ptr(ws), ws.numel(),
                                       ctypes.byref(stats), L.stream())
        st.ir_outer, st.ir_inner, st.ir_relres = stats.outer_iters, stats.inner_iters, stats.final_relres
        if status == L.RBFX_ERR_STALL:
            # later solves with this factorisation (e.g. the adjoint) go straight
            # to the FP64 factors; the FP32 factors are released
            st.fell_back = True
            self.lu32 = None
            lu64 = self._fallback()
            return lu64.solve(b, trans)
        L.check(status)
        return x


class _LRU:
    def __init__(self, size):
        self.size = size
        self.d = collections.OrderedDict()

    def get(self, key):
        v = self.d.get(key)
        if v is not None:
            self.d.move_to_end(key)
        return v

    def put(self, key, val):
        self.d[key] = val
        self.d.move_to_end(key)
        while len(self.d) > self.size:
            self.d.popitem(last=False)

    def clear(self):
        self.d.clear()


_K_CACHE = _LRU(2)
_A_CACHE = _LRU(2)
CACHE_ENABLED = [True]


def clear_cache():
    """Drop cached factorisations (they can hold many GB at large N)."""
    _K_CACHE.clear()
    _A_CACHE.clear()
    _BASIS_CACHE.clear()
    torch.cuda.empty_cache()


def a_factor(cloud, kind, param, M):
    """FP64 LU of the interpolation matrix A (updes/assembly.py:62-90), cached."""
    key = (cloud_key(cloud), kind, param, M)
    f = _A_CACHE.get(key) if CACHE_ENABLED[0] else None
    if f is None:
        prob = Problem(device_cloud(cloud), kind, param, M)
        f = LU(assemble(prob, L.MAT_A, torch.float64), prob.n)
        f.problem = prob
        if CACHE_ENABLED[0]:
            _A_CACHE.put(key, f)
    return f


_BASIS_CACHE = _LRU(2)


def cardinal_basis(cloud, kind, param, M, k=DEFAULT_STENCIL):
    """Cached local cardinal basis of (cloud, kernel, degree): it does not
    depend on the operator or the boundary data."""
    key = (cloud_key(cloud), kind, param, M, int(k))
    bas = _BASIS_CACHE.get(key) if CACHE_ENABLED[0] else None
    if bas is None:
        bas = CardinalBasis(Problem(device_cloud(cloud), kind, param, M), k)
        if CACHE_ENABLED[0]:
            _BASIS_CACHE.put(key, bas)
    return bas


def k_factor(cloud, kind, param, M, desc, robin_beta, precision, **ir):
    rkey = None if robin_beta is None else hashlib.sha1(
        torch.as_tensor(robin_beta, dtype=torch.float64).cpu().numpy().tobytes()).hexdigest()
    key = (cloud_key(cloud), kind, param, M, desc.key(), rkey, precision, tuple(sorted(ir.items())))
    f = _K_CACHE.get(key) if CACHE_ENABLED[0] else None
    if f is None:
        prob = Problem(device_cloud(cloud), kind, param, M, desc, robin_beta)
        opts = dict(ir)
        if precision == "fp32_ir" and not prob.dc.local:
            opts["basis"] = cardinal_basis(cloud, kind, param, M, opts.get("ir_stencil", DEFAULT_STENCIL))
        f = KFactor(prob, precision, **opts)
        if CACHE_ENABLED[0]:
            _K_CACHE.put(key, f)
    return f
