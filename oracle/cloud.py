"""SquareCloud restatement (oracle, test infrastructure only).

Follows ``updes/cloud.py``:
* periodic facet renaming ``v + str(index)``            (cloud.py:42-49)
* global ids id = i*Ny + j                             (cloud.py:411-422)
* node types with precedence North > South > East > West  (cloud.py:449-490)
* coordinates linspace(0,1,Nx)[i], linspace(0,1,Ny)[j] plus optional
  jitter on nodes whose type is not d/n/r               (cloud.py:425-446)
* outward normals for n/r/p nodes, corners -> North/South (cloud.py:493-510)
* renumbering internal -> d -> n -> r -> periodic groups by sorted key,
  original-id order inside each class                  (cloud.py:115-172)
* sorted_nodes / sorted_outward_normals by new id       (cloud.py:72-81, 403-408)

The reference draws its jitter from ``jax.random``; here the jitter is an
explicit (N, 2) array indexed by the original id, which is what the workloads
use (parity for ``noise_key`` itself is not claimed: parity unpinned).
With ``support_size="max"`` the support table of ``define_local_supports``
(cloud.py:83-112) is not built: it only encodes "every node but itself".  A
smaller ``support_size`` builds ``supports`` (new ids) from a brute-force
nearest-neighbour query (``knn_supports``), or from an explicit table of
neighbours by original id (``supports=``), e.g. one taken from the same
BallTree query the reference makes, for grids whose equidistant neighbours
tie at the cut.
"""
import numpy as np


class SquareCloud:
    def __init__(self, Nx=7, Ny=5, facet_types=None, jitter=None, support_size="max", supports=None):
        if facet_types is None:
            raise ValueError("facet_types is required")
        self.Nx, self.Ny = int(Nx), int(Ny)
        self.N = self.Nx * self.Ny
        self.dim = 2
        self.facet_precedence = {k: i for i, k in enumerate(facet_types)}
        ft = {}
        for i, (k, v) in enumerate(facet_types.items()):
            ft[k] = v + str(i) if v[0] == "p" else v
        self.facet_types = ft

        Nx, Ny, N = self.Nx, self.Ny, self.N
        # global ids
        self.global_indices_rev = {i * Ny + j: (i, j) for i in range(Nx) for j in range(Ny)}

        # node types and facet nodes (original ids)
        self.facet_nodes = {k: [] for k in self.facet_types}
        node_types = {}
        for gid in range(N):
            k, l = self.global_indices_rev[gid]
            if l == Ny - 1:
                f = "North"
            elif l == 0:
                f = "South"
            elif k == Nx - 1:
                f = "East"
            elif k == 0:
                f = "West"
            else:
                f = None
            if f is None:
                node_types[gid] = "i"
            else:
                self.facet_nodes[f].append(gid)
                node_types[gid] = self.facet_types[f]

        self.Nd = sum(len(self.facet_nodes[f]) for f, t in self.facet_types.items() if t == "d")
        self.Nn = sum(len(self.facet_nodes[f]) for f, t in self.facet_types.items() if t == "n")
        self.Nr = sum(len(self.facet_nodes[f]) for f, t in self.facet_types.items() if t == "r")
        npd = {}
        for f, t in self.facet_types.items():
            if t[0] == "p":
                npd[t[:-1]] = npd.get(t[:-1], 0) + len(self.facet_nodes[f])
        self.Np = [npd[k] for k in sorted(npd)]
        self.Ni = N - self.Nd - self.Nn - self.Nr - sum(self.Np)

        # coordinates (original ids)
        xs = np.linspace(0.0, 1.0, Nx)
        ys = np.linspace(0.0, 1.0, Ny)
        coords = np.empty((N, 2))
        for gid in range(N):
            i, j = self.global_indices_rev[gid]
            coords[gid] = (xs[i], ys[j])
            if jitter is not None and node_types[gid] not in ("d", "n", "r"):
                coords[gid] += jitter[gid]

        # outward normals (original ids) for n / r / p nodes
        normals = {}
        for gid in range(N):
            if node_types[gid][0] in ("n", "r", "p"):
                k, l = self.global_indices_rev[gid]
                if l == Ny - 1:
                    normals[gid] = (0.0, 1.0)
                elif l == 0:
                    normals[gid] = (0.0, -1.0)
                elif k == Nx - 1:
                    normals[gid] = (1.0, 0.0)
                elif k == 0:
                    normals[gid] = (-1.0, 0.0)

        # renumbering
        i_nodes, d_nodes, n_nodes, r_nodes, p_nodes = [], [], [], [], {}
        for gid in range(N):
            t = node_types[gid]
            if t[0] == "i":
                i_nodes.append(gid)
            elif t[0] == "d":
                d_nodes.append(gid)
            elif t[0] == "n":
                n_nodes.append(gid)
            elif t[0] == "r":
                r_nodes.append(gid)
            elif t[0] == "p":
                p_nodes.setdefault(t, []).append(gid)
            else:
                raise ValueError("Unknown node type")
        all_p = []
        for key in sorted(p_nodes):
            all_p += p_nodes[key]
        order = i_nodes + d_nodes + n_nodes + r_nodes + all_p
        new_numb = {v: k for k, v in enumerate(order)}
        self.renumbering_map = new_numb
        self.node_types = {new_numb[k]: v for k, v in node_types.items()}
        self.nodes = {new_numb[k]: coords[k] for k in range(N)}
        self.facet_nodes = {f: [new_numb[g] for g in ids] for f, ids in self.facet_nodes.items()}
        self.outward_normals = {new_numb[k]: np.array(v) for k, v in normals.items()}

        self.sorted_nodes = np.array([coords[g] for g in order])
        if normals:
            keys = sorted(self.outward_normals)
            self.sorted_outward_normals = np.array([self.outward_normals[k] for k in keys])
        else:
            self.sorted_outward_normals = np.zeros((0, 2))
        self.original_ids = np.array(order)   # new id -> original id

        # local supports (cloud.py:83-112, 160-163)
        if support_size in ("max", None):
            support_size = N
        if not (0 < int(support_size) <= N):
            raise AssertionError("support size must be in [1, N]")
        self.support_size = int(support_size)
        self.supports = None                  # None: every other node
        if self.support_size < N:
            if supports is None:
                supports = knn_supports(coords, self.support_size)
            supports = np.asarray(supports, dtype=np.int64)
            tab = np.empty_like(supports)
            for g in range(N):
                tab[new_numb[g]] = [new_numb[int(j)] for j in supports[g]]
            self.supports = tab               # (N, support_size-1), new ids, rows by new id


def knn_supports(coords, k):
    """The k-1 nearest other nodes of each node, nearest first (original ids).

    ``cloud.py:108-111`` asks scikit-learn's BallTree for the k nearest
    points of each node and drops the first hit (the node itself).  This is
    the brute-force statement of the same query: squared distances
    dx*dx + dy*dy (the reduced distance BallTree compares), the node itself
    removed, ties between equidistant neighbours broken by lowest original
    id.  BallTree breaks exact ties by its traversal order instead, so the
    two agree wherever the k-1-th and k-th distances differ.
    """
    coords = np.asarray(coords, dtype=np.float64)
    N = coords.shape[0]
    out = np.empty((N, k - 1), dtype=np.int64)
    for i in range(N):
        d2 = (coords[:, 0] - coords[i, 0]) ** 2 + (coords[:, 1] - coords[i, 1]) ** 2
        order = [int(j) for j in np.argsort(d2, kind="stable") if j != i]
        out[i] = order[:k - 1]
    return out


def interpolate_field(field, cloud1, cloud2):
    """``updes/operators.py:457-480``: reorder a field between two numberings."""
    sorted_map1 = sorted(cloud1.renumbering_map.items(), key=lambda kv: kv[0])
    indexer1 = np.array([v for _, v in sorted_map1])
    field_orig = np.asarray(field)[indexer1]
    indexer2 = np.array(list(cloud2.renumbering_map.keys()))
    return field_orig[indexer2]
