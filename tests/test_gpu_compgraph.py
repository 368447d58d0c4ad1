"""The drop-in `diffmpm_bindings` API on the GPU: the reference's own
episode (golden vectors from the compiled reference), the loss, volumes,
Adam, and the API's error behaviour (bind.cpp)."""
import os

import numpy as np
import pytest

from helpers import perturbed_scene, rel_err
from oracle import Oracle, Params, reference_params

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


def _bindings_cg(d, state, precision="f32"):
    import diffmpm_bindings as db
    dims = tuple(int(a) for a in d["dims"])
    pc = db.PointCloud.from_arrays(state["x"], state["m"], state["vol"], state["lam"], state["mu"], v=state["v"],
                                   F=state["F"], C=state["C"], dFc=state.get("dFc"))
    grid = db.Grid(*dims, float(d["dx"]), [0.0, 0.0, 0.0])
    tgt = db.Grid(*dims, float(d["dx"]), [0.0, 0.0, 0.0])
    if "target_m" in d:
        tgt.m = d["target_m"].astype(np.float32)
    cg = db.CompGraph(pc, grid, tgt, precision=precision)
    cg._sim.set_sim_params(float(d["dt"]), float(d["drag"]), float(d["smoothing"]), tuple(d["f_ext"]))
    return cg


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_reference_episode_through_bindings(precision):
    """set_up_comp_graph(4), compute_forward_pass(0), end_layer_mass_loss,
    set_render_gradients, compute_backward_pass(0) -- against the same calls
    on the reference's CompGraph (golden vectors, float32)."""
    d = load("ref_episode.npz")
    s = {k[3:]: d[k] for k in d if k.startswith("in_")}
    print(f"[episode] precision={precision} n={s['x'].shape[0]}", flush=True)
    cg = _bindings_cg(d, s, precision)
    cg.set_up_comp_graph(int(d["num_layers"]))
    cg.compute_forward_pass(0, 0)
    last = cg.get_point_cloud(int(d["num_layers"]) - 1)
    assert rel_err(np.array(last.get_positions()), d["out_x"]) < 2e-4
    assert rel_err(np.array(last.get_def_grads()), d["out_F"]) < 2e-4
    assert rel_err(last.get_velocities(), d["out_v"]) < 2e-4
    loss = cg.end_layer_mass_loss()
    assert abs(loss - float(d["out_loss"])) <= 2e-4 * abs(float(d["out_loss"]))
    cg.set_render_gradients(d["render_gF"], d["render_gx"])
    assert cg.has_render_gradients()
    info = cg.get_render_gradient_info()
    assert info["num_points"] == s["x"].shape[0]
    assert abs(info["grad_F_norm"] - np.linalg.norm(d["render_gF"])) < 1e-4 * np.linalg.norm(d["render_gF"])
    cg.compute_backward_pass(0)
    for k in ("x", "v", "F", "C"):
        e = rel_err(cg._sim.get_grad(0, k, dtype=np.float64), d[f"out_g{k}"])
        assert e < 1e-3, f"d{k}: {e:.3e}"
    cg.clear_render_gradients()
    assert not cg.has_render_gradients()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("order", [2, 3])
def test_mass_loss_matches_oracle(prec, order):
    from dmpm import Physics, Simulator
    sc, s = perturbed_scene(n=3000)
    tsc, _ = perturbed_scene(n=2000, seed=4, axes=(0.2, 0.24, 0.17))
    prm = Params(dims=sc.dims, dx=sc.dx, order=order)
    orc = Oracle(prec)
    tm = orc.p2g_mass(prm, tsc.x, tsc.m)
    loss_o, gx_o, _, _ = orc.mass_loss(prm, s["x"], s["m"], tm)
    sim = Simulator(sc.dims, sc.dx, physics=Physics(order=order), precision=prec)
    sim.set_state(s)
    sim.set_target_grid_mass(tm)
    loss = sim.mass_loss()
    tol = 1e-10 if prec == "f64" else 1e-4
    assert abs(loss - loss_o) <= tol * abs(loss_o)
    assert rel_err(sim.get_grad(0, "x", dtype=np.float64), gx_o) < tol
    for k in ("v", "F", "C"):
        assert np.abs(sim.get_grad(0, k)).max() == 0.0
    sim.close()


def test_volumes_and_p2g_match_reference():
    import diffmpm_bindings as db
    d = load("ref_volumes.npz")
    dims = tuple(int(a) for a in d["dims"])
    n = d["in_x"].shape[0]
    pc = db.PointCloud.from_arrays(d["in_x"], d["in_m"], np.ones(n), 1.0, 1.0)
    grid = db.Grid(*dims, float(d["dx"]), [0, 0, 0])
    db.calculate_point_cloud_volumes(pc, grid)
    assert rel_err(pc.get_volumes(), d["out_vol"]) < 1e-5
    assert rel_err(grid.m, d["out_grid_m"]) < 1e-5
    g2 = db.Grid(*dims, float(d["dx"]), [0, 0, 0])
    db.p2g(pc, g2)
    assert rel_err(g2.m, d["out_grid_m"]) < 1e-5


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_adam_matches_oracle(prec):
    """Descend_Adam on the device against the oracle, two calls in a row, on
    the control gradient of a real backward pass."""
    from dmpm import Physics, Simulator
    sc, s = perturbed_scene(n=1500)
    sim = Simulator(sc.dims, sc.dx, physics=Physics(dt=1e-3), precision=prec)
    sim.set_state(s)
    sim.set_up(3)
    sim.forward(0)
    rng = np.random.default_rng(1)
    sim.set_last_grad(x=rng.normal(size=(sc.n, 3)), F=rng.normal(size=(sc.n, 3, 3)))
    sim.backward(0)
    dt = np.float64 if prec == "f64" else np.float32
    g = sim.get_control_grad(0, dtype=np.float64)
    gn = sim.control_grad_norm()
    assert abs(gn - np.linalg.norm(g)) < 1e-6 * gn
    orc = Oracle(prec)
    z = np.zeros((sc.n, 3, 3))
    st = (s["dFc"], z, z, z)
    for t, alpha in ((1, 0.5), (2, 1e-3)):
        sim.descend_adam(0, alpha, gn, timestep=t)
        st = orc.adam(g, *st, alpha, gn, timestep=t)
        tol = 1e-10 if prec == "f64" else 1e-5
        assert rel_err(sim.get(0, "dFc", dtype=dt), st[0]) < tol
    sim.close()


def test_api_errors_and_semantics():
    import diffmpm_bindings as db
    sc, s = perturbed_scene(n=800, dfc=0.0)
    n = sc.n
    pc = db.PointCloud.from_arrays(s["x"], s["m"], s["vol"], s["lam"], s["mu"], v=s["v"], F=s["F"], C=s["C"])
    g = db.Grid(*sc.dims, sc.dx, [0, 0, 0])
    cg = db.CompGraph(pc, g, db.Grid(*sc.dims, sc.dx, [0, 0, 0]))
    assert cg.get_num_layers() == 1
    cg.set_up_comp_graph(5)
    assert cg.get_num_layers() == 5
    # every layer starts as a copy of layer 0 (CompGraph.cpp:30-40)
    assert np.array_equal(np.array(cg.get_point_cloud(4).get_positions()), np.array(pc.get_positions()))
    with pytest.raises(IndexError):
        cg.get_point_cloud(5)
    with pytest.raises(IndexError):
        cg.get_point_cloud_at_timestep(-1)
    with pytest.raises(RuntimeError, match="dLdF must be"):
        cg.accumulate_render_grads(np.zeros((n, 3)), np.zeros((n, 3)))
    with pytest.raises(RuntimeError, match="Shape mismatch"):
        cg.accumulate_render_grads(np.zeros((n + 1, 3, 3)), np.zeros((n + 1, 3)))
    with pytest.raises(RuntimeError, match="dLdx must be"):
        cg.set_render_gradients(np.zeros((n, 3, 3)), np.zeros((n, 2)))
    cg._sim.set_sim_params(1e-3, 0.0, 0.0)
    cg._sim.set_control(0, 0.01 * np.ones((n, 3, 3)))
    cg._sim.set_control(2, 0.01 * np.ones((n, 3, 3)))
    cg.compute_forward_pass(0, 0)
    # the forward pass zeroes dFc of every layer it writes (ForwardSimulation.cpp:156)
    assert np.abs(cg.get_point_cloud(2).get_def_grads_morph()).max() == 0
    assert np.abs(cg.get_point_cloud(0).get_def_grads_morph() - 0.01).max() < 1e-7
    F0 = cg.get_point_cloud(0).get_def_grads_total()
    assert np.abs(F0 - (s["F"] + 0.01)).max() < 1e-6
    xl = np.array(cg.get_point_cloud(4).get_positions())
    t = cg.get_point_cloud(4).get_positions_torch()
    assert t.dtype.is_floating_point and tuple(t.shape) == (n, 3) and t.device.type == "cpu"
    assert np.array_equal(t.numpy(), xl)
    cg.promote_last_as_initial()
    assert cg.get_num_layers() == 1
    assert np.array_equal(np.array(cg.get_point_cloud(0).get_positions()), xl)
    lam, mu = db.calculate_lame_parameters(1e4, 0.3)
    assert abs(lam - 5769.2307) < 1e-2 and abs(mu - 3846.1538) < 1e-2


def test_optimize_single_timestep_decreases_loss():
    """OptimizeSingleTimestep (CompGraph.cpp:369-465): Adam steps on dFc of
    one layer under a backtracking line search never increase the loss."""
    import diffmpm_bindings as db
    sc, s = perturbed_scene(n=1500, dfc=0.0, fpert=0.0, cpert=0.0)
    tsc, _ = perturbed_scene(n=1500, seed=4, axes=(0.25, 0.18, 0.2))
    pc = db.PointCloud.from_arrays(s["x"], s["m"], s["vol"], s["lam"], s["mu"], v=0 * s["v"])
    g = db.Grid(*sc.dims, sc.dx, [0, 0, 0])
    tg = db.Grid(*sc.dims, sc.dx, [0, 0, 0])
    tpc = db.PointCloud.from_arrays(tsc.x, tsc.m, tsc.vol, 1.0, 1.0)
    db.calculate_point_cloud_volumes(tpc, tg)
    cg = db.CompGraph(pc, g, tg)
    cg._sim.set_sim_params(1e-3, 0.5, 0.1)
    cg.set_up_comp_graph(6)
    cg.compute_forward_pass(0, 0)
    l0 = cg.end_layer_mass_loss()
    cg.optimize_single_timestep(0, max_gd_iters=3, initial_alpha=0.05)
    cg.compute_forward_pass(0, 0)
    l1 = cg.end_layer_mass_loss()
    assert l1 <= l0
    assert np.abs(cg.get_point_cloud(0).get_def_grads_morph()).max() > 0


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_volumes_accumulate_onto_the_grid(prec):
    """calculate_point_cloud_volumes P2Gs onto the masses the Grid already
    holds (ForwardSimulation.cpp:234-237 calls P2G on the Grid it is given):
    with the cloud's own masses already there, the node masses double and
    every volume halves.  The device path (CUDA tensors, on_device = 1)
    gives the host path's results."""
    import ctypes as C
    import torch
    from dmpm import _lib as L
    from dmpm.sim import calculate_volumes, make_config, Physics
    d = load("ref_volumes.npz")
    dims = tuple(int(a) for a in d["dims"])
    dt = np.float64 if prec == "f64" else np.float32
    x, m = d["in_x"].astype(dt), d["in_m"].astype(dt)
    vol1, gm1 = calculate_volumes(dims, float(d["dx"]), (0.0, 0.0, 0.0), x, m, precision=prec)
    vol2, gm2 = calculate_volumes(dims, float(d["dx"]), (0.0, 0.0, 0.0), x, m, precision=prec, grid_m=gm1)
    tol = 1e-12 if prec == "f64" else 1e-6
    assert rel_err(gm2, 2 * gm1) < tol
    assert rel_err(vol2, 0.5 * vol1) < tol
    dev = torch.device("cuda:0")
    tx, tm = torch.tensor(x, device=dev), torch.tensor(m, device=dev)
    tv = torch.zeros(x.shape[0], dtype=tx.dtype, device=dev)
    tg = torch.tensor(gm1, device=dev)
    cfg = make_config(dims, float(d["dx"]), (0.0, 0.0, 0.0), Physics(order=3), prec)
    code = L.DTYPE_F64 if prec == "f64" else L.DTYPE_F32
    torch.cuda.synchronize()
    L.check(L.lib().dmpm_calculate_volumes(C.byref(cfg), x.shape[0], C.c_void_p(tx.data_ptr()),
                                           C.c_void_p(tm.data_ptr()), C.c_void_p(tv.data_ptr()),
                                           C.c_void_p(tg.data_ptr()), code, 1))
    assert rel_err(tg.cpu().numpy(), gm2) < tol
    assert rel_err(tv.cpu().numpy(), vol2) < tol
