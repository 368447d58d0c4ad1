"""Particle initialisation on the device: winding numbers against the
oracle, and load_point_cloud_from_obj against the reference's own sampler."""
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import winding_numbers as oracle_winding

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_winding_numbers_match_the_oracle():
    from dmpm.mesh import read_obj, winding_numbers
    V, F = read_obj(os.path.join(GOLDEN, "mesh_torus.obj"))
    rng = np.random.default_rng(12)
    lo, hi = V.min(axis=0) - 0.05, V.max(axis=0) + 0.05
    P = lo + (hi - lo) * rng.random(size=(20000, 3))
    print(f"[mesh] {P.shape[0]} points, {F.shape[0]} triangles", flush=True)
    w = winding_numbers(P, V, F)
    ref = oracle_winding(P, V, F)
    assert np.abs(w - ref).max() < 1e-12
    assert ((w > 0.5) != (ref > 0.5)).sum() == 0
    # one triangle, many chunks: the octant closed form
    assert abs(winding_numbers([[0, 0, 0]], np.eye(3), [[0, 1, 2]])[0] - 0.125) < 1e-15


def test_load_point_cloud_from_obj_matches_the_reference():
    import diffmpm_bindings as db
    d = np.load(os.path.join(GOLDEN, "ref_mesh_torus.npz"))
    opt = db.OptInput()
    opt.grid_dx = float(d["grid_dx"])
    opt.points_per_cell_cuberoot = int(d["ppc"])
    opt.p_density = float(d["rho"])
    opt.lam, opt.mu = float(d["lam"]), float(d["mu"])
    pc = db.load_point_cloud_from_obj(os.path.join(GOLDEN, "mesh_torus.obj"), opt)
    np.testing.assert_array_equal(np.array(pc.get_positions()), d["x"])
    np.testing.assert_array_equal(np.array(pc.get_masses(), np.float32), d["m"])
    np.testing.assert_array_equal(pc.get_volumes(), d["vol"])


def test_bad_mesh_is_rejected():
    from dmpm._lib import DmpmError
    from dmpm.mesh import winding_numbers
    with pytest.raises(DmpmError, match="vertex index"):
        winding_numbers([[0, 0, 0]], np.eye(3), [[0, 1, 3]])
