"""The reference's own binding test (tests/test.py:157-245) on the meshes it
ships (copied into tests/golden as data): isosphere -> input point cloud,
bunny -> target; grid_dx 1 on a 33^3 grid from -16, 2 points per cell axis,
density 75.  Every step is checked against what the reference computed on
the same inputs (tests/golden/make_golden.py assets): the sampled points
(identical), the volumes, the end-layer mass loss and its dL/dx; then
accumulate_render_grads runs as in the reference test."""
import os

import numpy as np
import pytest

from conftest import ROOT
from helpers import rel_err

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_reference_binding_test_on_its_assets():
    import diffmpm_bindings as db
    d = np.load(os.path.join(GOLDEN, "ref_assets.npz"))
    opt = db.OptInput()
    opt.mpm_input_mesh_path = os.path.join(GOLDEN, "asset_isosphere.obj")
    opt.mpm_target_mesh_path = os.path.join(GOLDEN, "asset_bunny.obj")
    opt.grid_dx = 1
    opt.grid_min_point = [-16.0, -16.0, -16.0]
    opt.grid_max_point = [16.0, 16.0, 16.0]
    opt.points_per_cell_cuberoot = 2
    opt.lam = 38888.89
    opt.mu = 58333.3
    opt.p_density = 75.0
    opt.dt = 0.00833333333
    opt.drag = 0.5
    opt.f_ext = [0.0, -9.8, 0.0]
    opt.smoothing_factor = 0.955
    print("[assets] loading meshes", flush=True)
    input_pc = db.load_point_cloud_from_obj(opt.mpm_input_mesh_path, opt)
    target_pc = db.load_point_cloud_from_obj(opt.mpm_target_mesh_path, opt)
    np.testing.assert_array_equal(np.array(input_pc.get_positions()), d["in_x"])
    np.testing.assert_array_equal(np.array(target_pc.get_positions()), d["tgt_x"])
    np.testing.assert_array_equal(np.array(input_pc.get_masses(), np.float32), d["in_m"])

    dims = [int((opt.grid_max_point[i] - opt.grid_min_point[i]) / opt.grid_dx) + 1 for i in range(3)]
    input_grid = db.Grid(*dims, opt.grid_dx, opt.grid_min_point)
    target_grid = db.Grid(*dims, opt.grid_dx, opt.grid_min_point)
    db.calculate_point_cloud_volumes(input_pc, input_grid)
    db.calculate_point_cloud_volumes(target_pc, target_grid)
    assert rel_err(input_pc.get_volumes(), d["in_vol"]) < 1e-5
    assert rel_err(target_pc.get_volumes(), d["tgt_vol"]) < 1e-5

    cg = db.CompGraph(input_pc, input_grid, target_grid)
    loss = cg.end_layer_mass_loss()
    print(f"[assets] loss {loss:.6f} (reference {float(d['loss']):.6f})", flush=True)
    assert abs(loss - float(d["loss"])) <= 2e-4 * abs(float(d["loss"]))
    gx = cg._sim.get_grad(0, "x", dtype=np.float64)
    assert rel_err(gx, d["dLdx"]) < 1e-3

    n = len(input_pc.get_positions())
    rng = np.random.default_rng(0)
    cg.accumulate_render_grads((rng.normal(size=(n, 3, 3)) * 1e-3).astype(np.float32),
                               (rng.normal(size=(n, 3)) * 1e-3).astype(np.float32))
