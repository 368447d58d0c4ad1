"""Particle initialisation from a triangle mesh: the lattice sampler of
GeometryLoading::LoadMPMPointCloudFromObj (GeometryLoading.cpp:10-143), with
the inside test (generalised winding number >= 1/2) computed on the device
through dmpm_winding_numbers."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def read_obj(path: str):
    """Vertices (nv,3) float64 and triangles (nf,3) int32 of an OBJ file;
    polygons are fanned into triangles, negative indices count from the end."""
    V, F = [], []
    with open(path) as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            if p[0] == "v":
                V.append([float(a) for a in p[1:4]])
            elif p[0] == "f":
                idx = [int(a.split("/")[0]) for a in p[1:]]
                idx = [i - 1 if i > 0 else len(V) + i for i in idx]
                for k in range(1, len(idx) - 1):
                    F.append([idx[0], idx[k], idx[k + 1]])
    return np.asarray(V, np.float64).reshape(-1, 3), np.asarray(F, np.int32).reshape(-1, 3)


def winding_numbers(P, V, F) -> np.ndarray:
    """Generalised winding number of each point of P (n,3) w.r.t. the mesh."""
    P = np.ascontiguousarray(P, np.float64).reshape(-1, 3)
    V = np.ascontiguousarray(V, np.float64).reshape(-1, 3)
    F = np.ascontiguousarray(F, np.int32).reshape(-1, 3)
    w = np.zeros(P.shape[0], np.float64)
    L.check(L.lib().dmpm_winding_numbers(V.ctypes.data, V.shape[0], F.ctypes.data, F.shape[0], P.ctypes.data,
                                         P.shape[0], w.ctypes.data))
    return w


def lattice(lo, hi, h: float) -> np.ndarray:
    """The sample lattice of GeneratePointCloudFromWatertightTriangleMesh
    (GeometryLoading.cpp:19-33) in float32 arithmetic: ceil((hi - lo) / h) + 1
    points per axis, point = lo + h * (i, j, k), i slowest."""
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    h32 = np.float32(h)
    dims = [int(np.ceil(np.float32(hi[a] - lo[a]) / h32)) + 1 for a in range(3)]
    g = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), axis=-1).reshape(-1, 3)
    return lo[None] + h32 * g.astype(np.float32)


def sample_inside(V, F, h: float) -> np.ndarray:
    """Lattice points (float32, lattice order) inside the watertight mesh,
    over the mesh's bounding box."""
    V = np.asarray(V, np.float64)
    P = lattice(V.min(axis=0).astype(np.float32), V.max(axis=0).astype(np.float32), h)
    return P[winding_numbers(P.astype(np.float64), V, F) >= 0.5]
