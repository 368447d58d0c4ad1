"""The oracle's winding number (the checker of the device mesh sampler)
against closed forms, and the oracle sampler against the points the
reference's own LoadMPMPointCloudFromObj generated (tests/golden, written by
tests/golden/make_golden.py mesh)."""
import os

import numpy as np

from conftest import ROOT
from oracle import winding_numbers

GOLDEN = os.path.join(ROOT, "tests", "golden")

TET_V = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
TET_F = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]])  # outward normals


def test_solid_angle_of_an_octant_triangle():
    """The triangle (e1, e2, e3) subtends one eighth of the sphere from the
    origin: w = 1/8 (and -1/8 with the opposite orientation)."""
    V = np.eye(3)
    assert abs(winding_numbers([[0, 0, 0]], V, [[0, 1, 2]])[0] - 0.125) < 1e-15
    assert abs(winding_numbers([[0, 0, 0]], V, [[0, 2, 1]])[0] + 0.125) < 1e-15


def test_tetrahedron_inside_outside_and_orientation():
    inside = [[0.1, 0.1, 0.1], [0.2, 0.3, 0.1], [0.01, 0.01, 0.97]]
    outside = [[1, 1, 1], [-0.1, 0.2, 0.2], [0.5, 0.5, 0.1 - 1e-3 + 0.5], [10, -3, 2]]
    assert np.allclose(winding_numbers(inside, TET_V, TET_F), 1.0, atol=1e-13)
    assert np.allclose(winding_numbers(outside, TET_V, TET_F), 0.0, atol=1e-13)
    assert np.allclose(winding_numbers(inside, TET_V, TET_F[:, ::-1]), -1.0, atol=1e-13)
    # just inside and just outside a face (exactly on it the sum is 0 or 1,
    # depending on the sign of a zero)
    w = winding_numbers([[0.2, 0.2, 1e-9], [0.2, 0.2, -1e-9]], TET_V, TET_F)
    assert abs(w[0] - 1.0) < 1e-12 and abs(w[1]) < 1e-12


def test_oracle_sampler_reproduces_the_reference_points():
    from dmpm.mesh import lattice, read_obj
    d = np.load(os.path.join(GOLDEN, "ref_mesh_torus.npz"))
    V, F = read_obj(os.path.join(GOLDEN, "mesh_torus.obj"))
    h = np.float32(d["grid_dx"]) / np.float32(d["ppc"])
    P = lattice(V.min(axis=0).astype(np.float32), V.max(axis=0).astype(np.float32), float(h))
    inside = P[winding_numbers(P.astype(np.float64), V, F) >= 0.5]
    np.testing.assert_array_equal(inside, d["x"])
    np.testing.assert_array_equal(d["m"], np.full(len(inside), h * h * h * np.float32(d["rho"])))
    np.testing.assert_array_equal(d["vol"], d["m"] / np.float32(d["rho"]))
