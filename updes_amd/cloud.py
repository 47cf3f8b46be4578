"""Point clouds (drop-in for ``updes/cloud.py`` ``Cloud`` / ``SquareCloud``).

Same numbering, node types, facet ownership, renumbering and normals as the
reference (``updes/cloud.py:13-49, 72-172, 378-510``), built with vectorised
index arithmetic instead of per-node Python loops and BallTree queries, so a
200 x 200 cloud is built in milliseconds.  The reference's N x (N-1)
``sorted_local_supports`` table (``cloud.py:83-112, 405``) is not stored:
with ``support_size="max"`` it only says "every node but itself", which the
assembly kernels encode directly; ``local_supports`` is produced on demand
for small clouds.  A ``support_size`` below N (RBF-FD style local supports)
is honoured: the neighbour table is built with the reference's own BallTree
query and handed to the kernels, which zero every RBF column outside it.

Jitter: the reference draws it from ``jax.random`` with ``noise_key``
(``cloud.py:431-444``); here ``noise_key`` may be an integer seed (NumPy
generator) and ``jitter`` an explicit (N, 2) array indexed by the original
id.  Either is added to nodes whose type is not d/n/r, as in the reference.
"""
import warnings

import numpy as np
import torch


def default_device():
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


class Cloud:
    """Base class: facet bookkeeping shared by every cloud (``cloud.py:10-49``)."""

    def __init__(self, facet_types, support_size="max"):
        self.N = 0
        self.Ni = 0
        self.Nd = 0
        self.Nr = 0
        self.Nn = 0
        self.Np = []
        self.facet_nodes = {}
        self.support_size = support_size
        self.dim = 2
        self.facet_precedence = {k: i for i, k in enumerate(facet_types)}
        ft = {}
        for i, (k, v) in enumerate(facet_types.items()):
            ft[k] = v + str(i) if v[0] == "p" else v
        self.facet_types = ft
        self._device = default_device()

    # -- reference helpers ---------------------------------------------------
    def sort_dict_by_keys(self, dictionnary):
        """Values of a node-keyed dict stacked in key order (``cloud.py:72-81``)."""
        if not hasattr(self, "renumbering_map"):
            warnings.warn("Nodes have not been renumbered to place internals first, and boundary "
                          "nodes afterwards. This function might not work properly")
        items = sorted(dictionnary.items(), key=lambda kv: kv[0])
        vals = [torch.as_tensor(np.asarray(v, dtype=np.float64)) for _, v in items]
        return torch.stack(vals, dim=-1).T

    def average_spacing(self):
        x = self._nodes_np
        d = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
        iu = np.triu_indices(self.N)
        return float(d[iu].mean())

    @property
    def device(self):
        return self._device

    def to(self, device):
        self._device = torch.device(device)
        self.sorted_nodes = self.sorted_nodes.to(self._device)
        self.sorted_outward_normals = self.sorted_outward_normals.to(self._device)
        return self


class SquareCloud(Cloud):
    """Regular (optionally jittered) cloud on the unit square (``cloud.py:378-510``)."""

    def __init__(self, Nx=7, Ny=5, noise_key=None, jitter=None, **kwargs):
        super().__init__(**kwargs)
        for f in self.facet_types:
            if f not in ("North", "South", "East", "West"):
                raise ValueError("SquareCloud facets are North/South/East/West, got %r" % f)
        self.Nx, self.Ny = int(Nx), int(Ny)
        Nx, Ny = self.Nx, self.Ny
        self.N = N = Nx * Ny
        gid = np.arange(N)
        I, J = gid // Ny, gid % Ny
        self._ij = (I, J)

        # facet of each node with precedence North > South > East > West (cloud.py:455-470)
        fac = np.full(N, -1, dtype=np.int64)
        names = ["North", "South", "East", "West"]
        conds = [J == Ny - 1, J == 0, I == Nx - 1, I == 0]
        for code, c in enumerate(conds):  # earlier condition wins
            fac[(fac < 0) & c] = code
        self._facet_code = fac
        for name in names:
            if name not in self.facet_types and np.any(fac == names.index(name)):
                raise KeyError(name)
        orig_facet_nodes = {f: gid[fac == names.index(f)] for f in self.facet_types}
        typ = np.array(["i"] * N, dtype=object)
        for f, t in self.facet_types.items():
            typ[orig_facet_nodes[f]] = t
        self._types_orig = typ

        self.Nd = int(sum(len(orig_facet_nodes[f]) for f, t in self.facet_types.items() if t == "d"))
        self.Nn = int(sum(len(orig_facet_nodes[f]) for f, t in self.facet_types.items() if t == "n"))
        self.Nr = int(sum(len(orig_facet_nodes[f]) for f, t in self.facet_types.items() if t == "r"))
        npd = {}
        for f, t in self.facet_types.items():
            if t[0] == "p":
                npd[t[:-1]] = npd.get(t[:-1], 0) + len(orig_facet_nodes[f])
        self.Np = [npd[k] for k in sorted(npd)]
        self.Ni = N - self.Nd - self.Nn - self.Nr - sum(self.Np)

        # coordinates (cloud.py:425-446)
        xs = np.linspace(0.0, 1.0, Nx)
        ys = np.linspace(0.0, 1.0, Ny)
        coords = np.stack([xs[I], ys[J]], axis=1)
        movable = np.array([t not in ("d", "n", "r") for t in typ])
        if jitter is not None:
            jitter = np.asarray(jitter, dtype=np.float64).reshape(N, 2)
            coords[movable] += jitter[movable]
        elif noise_key is not None:
            delta = min(xs[1] - xs[0] if Nx > 1 else 1.0, ys[1] - ys[0] if Ny > 1 else 1.0) / 2.0
            rng = np.random.default_rng(int(noise_key))
            noise = rng.uniform(-delta, delta, size=(N, 2))
            coords[movable] += noise[movable]

        # normals for n / r / p nodes, corners -> North / South (cloud.py:493-510)
        nrm = np.zeros((N, 2))
        nrm[fac == 0] = (0.0, 1.0)
        nrm[fac == 1] = (0.0, -1.0)
        nrm[fac == 2] = (1.0, 0.0)
        nrm[fac == 3] = (-1.0, 0.0)
        has_nrm = np.array([t[0] in ("n", "r", "p") for t in typ])

        # renumbering (cloud.py:115-172)
        first = np.array([t[0] for t in typ])
        order = [gid[first == "i"], gid[first == "d"], gid[first == "n"], gid[first == "r"]]
        pkeys = sorted({t for t in typ if t[0] == "p"})
        for k in pkeys:
            order.append(gid[typ == k])
        order = np.concatenate(order).astype(np.int64)
        new_id = np.empty(N, dtype=np.int64)
        new_id[order] = np.arange(N)
        self._order = order            # new id -> original id
        self._new_id = new_id          # original id -> new id
        self._nodes_np = coords[order]
        self._normals_np = nrm[order][self.Ni + self.Nd:]
        assert np.all(has_nrm[order][self.Ni + self.Nd:]) and not np.any(has_nrm[order][:self.Ni + self.Nd])
        self.facet_nodes = {f: [int(v) for v in new_id[orig_facet_nodes[f]]] for f in self.facet_types}
        self.sorted_nodes = torch.as_tensor(self._nodes_np, device=self._device)
        self.sorted_outward_normals = torch.as_tensor(self._normals_np.copy(), device=self._device)
        self._define_local_supports(coords)

    def _define_local_supports(self, coords):
        """The support_size - 1 nearest other nodes of each node (``cloud.py:83-112``).

        "max" / None mean every node (``support_size`` becomes N) and store
        nothing: the kernels exclude the node itself by index.  A smaller size
        queries scikit-learn's BallTree exactly as the reference does (built
        on the coordinates in original id order, ``leaf_size=1``, k =
        support_size, first hit, the node itself, dropped), so ties between
        equidistant neighbours of a regular grid resolve the same way; the
        ids are then renumbered.  ``_support_knn`` keeps the neighbours
        nearest first (new ids, rows in new-id order), ``_support_sorted``
        the same sets ascending, as the device kernels search them.
        """
        if self.support_size == "max" or self.support_size is None:
            self.support_size = self.N
        k = int(self.support_size)
        if k != self.support_size or k <= 0:
            raise AssertionError("Support size must be strictly greater than 0")
        if k > self.N:
            raise AssertionError("Support size must be strictly less than or equal the number of nodes")
        self.support_size = k
        self._support_knn = None
        self._support_sorted = None
        if k == self.N:
            return
        from sklearn.neighbors import BallTree
        tree = BallTree(coords, leaf_size=1, metric="euclidean")
        _, nb = tree.query(coords, k=k)
        nb = nb[:, 1:]
        knn = np.empty_like(nb)
        knn[self._new_id] = self._new_id[nb]
        self._support_knn = knn.astype(np.int64)
        self._support_sorted = np.sort(knn, axis=1).astype(np.int32)

    @property
    def sorted_local_supports(self):
        """(N, support_size-1) neighbour ids, nearest first, in new-id row order
        (``cloud.py:405``).  With every node as support the N x (N-1) table is
        built on demand, for small clouds only (as ``local_supports``)."""
        if self._support_knn is None:
            ls = self.local_supports
            return torch.as_tensor(np.array([ls[i] for i in range(self.N)], dtype=np.int64).reshape(self.N, -1))
        return torch.as_tensor(self._support_knn)

    # -- dict-valued attributes of the reference, built on demand -------------
    @property
    def renumbering_map(self):
        """{original id: new id} in renumbered order (``cloud.py:150,172``)."""
        return {int(o): int(k) for k, o in enumerate(self._order)}

    @property
    def node_types(self):
        return {int(self._new_id[g]): self._types_orig[g] for g in range(self.N)}

    @property
    def nodes(self):
        return {k: self._nodes_np[k] for k in range(self.N)}

    @property
    def outward_normals(self):
        base = self.Ni + self.Nd
        return {base + k: self._normals_np[k] for k in range(self._normals_np.shape[0])}

    @property
    def global_indices_rev(self):
        I, J = self._ij
        return {int(self._new_id[g]): (int(I[g]), int(J[g])) for g in range(self.N)}

    @property
    def global_indices(self):
        I, J = self._ij
        gi = np.zeros((self.Nx, self.Ny), dtype=np.int64)
        gi[I, J] = self._new_id
        return torch.as_tensor(gi)

    @property
    def local_supports(self):
        """Neighbours of each node, nearest first (``cloud.py:83-112``).  With
        every node as support this is O(N^2): small clouds only."""
        if self._support_knn is not None:
            return {i: [int(j) for j in self._support_knn[i]] for i in range(self.N)}
        if self.N > 4096:
            raise MemoryError("local_supports is an N x (N-1) table; not built for N > 4096")
        x = self._nodes_np
        d = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
        sup = {}
        for i in range(self.N):
            idx = np.argsort(d[i], kind="stable")
            sup[i] = [int(j) for j in idx if j != i]
        return sup

    def print_global_indices(self):
        print(torch.flip(self.global_indices.T, dims=(0,)))
