"""Shared workload definitions for the tests (small versions of the benchmark
configurations plus coverage cases for every row class and kernel)."""
import functools

import numpy as np

from oracle import rbf as R

# operator descriptors in oracle form
LAPLACIAN = {"rbf": [0, 0, 0, 1, 0, 0, 0], "mon": [0, 0, 0, 1, 0, 0, 0]}
VALUE = {"rbf": [1, 0, 0, 0, 0, 0, 0], "mon": [1, 0, 0, 0, 0, 0, 0]}


def adv_diff_desc(dt=1e-3, nu=0.01, vel=(1.0, 0.5)):
    """phi - dt (nu Lap phi - v . grad phi)."""
    c = [1.0, dt * vel[0], dt * vel[1], -dt * nu, 0, 0, 0]
    return {"rbf": list(c), "mon": list(c)}


def jitter(nx, ny, frac, seed):
    h = min(1.0 / (nx - 1), 1.0 / (ny - 1))
    return np.random.default_rng(seed).uniform(-frac * h, frac * h, size=(nx * ny, 2))


README_FACETS = {"South": "n", "West": "d", "North": "d", "East": "d"}
KINDS = [(R.POLYHARMONIC, 1.0), (R.POLYHARMONIC, 2.0), (R.GAUSSIAN, 5.0), (R.THIN_PLATE, 1.0),
         (R.MULTIQUADRIC, 2.0), (R.INV_MULTIQUADRIC, 2.0)]
FACET_SETS = [
    {"North": "d", "South": "d", "East": "d", "West": "d"},
    {"South": "n", "West": "d", "North": "d", "East": "d"},
    {"South": "r", "West": "n", "North": "d", "East": "r"},
    {"South": "d", "West": "p1", "North": "d", "East": "p1"},
    {"South": "p1", "North": "p1", "West": "p2", "East": "p2"},
    {"South": "n", "North": "r", "West": "p1", "East": "p1"},
]


def torch_rbf(kind, param):
    import updes_amd as u
    f = {R.POLYHARMONIC: (u.polyharmonic, "a"), R.GAUSSIAN: (u.gaussian, "eps"), R.THIN_PLATE: (u.thin_plate, "a"),
         R.MULTIQUADRIC: (u.multiquadric, "eps"), R.INV_MULTIQUADRIC: (u.inverse_multiquadric, "eps")}[kind]
    return functools.partial(f[0], **{f[1]: param})
