"""The drop-in CompGraph with segment storage orders (segment_every): the
reference's episode loop -- set-up, a forward pass from layer 0, then for
control timesteps a forward pass from t and a backward pass to t -- gives
the results of the one-order CompGraph (FP64, up to summation order), with
render gradients injected at the last layer."""
import numpy as np
import pytest

from helpers import perturbed_scene, rel_err

pytestmark = pytest.mark.gpu

T = 12


def _cg(sc, s, segment_every):
    import diffmpm_bindings as db
    pc = db.PointCloud.from_arrays(s["x"], s["m"], s["vol"], s["lam"], s["mu"], v=s["v"], F=s["F"], C=s["C"])
    grid = db.Grid(*sc.dims, sc.dx, list(sc.grid_min))
    tgt = db.Grid(*sc.dims, sc.dx, list(sc.grid_min))
    tgt.m = np.full(int(np.prod(sc.dims)), 1e-4, np.float32)
    cg = db.CompGraph(pc, grid, tgt, precision="f64", segment_every=segment_every)
    cg._sim.set_sim_params(1e-3, 0.0, 0.0, (0.0, 0.0, 0.0))
    return cg


def _loop(cg, n, seed=4):
    rng = np.random.default_rng(seed)
    cg.set_up_comp_graph(T + 1)
    cg.compute_forward_pass(0, 0)
    out = {"loss0": np.array([cg.end_layer_mass_loss()])}
    cg.set_render_gradients(1e-3 * rng.normal(size=(n, 3, 3)), 1e-3 * rng.normal(size=(n, 3)))
    for t in (0, 5, 9):
        cg.compute_forward_pass(t, 0)
        cg.compute_backward_pass(t)
        out[f"g{t}"] = cg._sim.get_control_grad(t, dtype=np.float64) if t > 0 else \
            cg._sim.get_grad(0, "v", dtype=np.float64)
        out[f"x{t}"] = cg._sim.get(T, "x", dtype=np.float64)
    out["F7"] = cg._sim.get(7, "F", dtype=np.float64)
    return out


def test_segment_orders_in_the_reference_loop():
    sc, s = perturbed_scene(dfc=0.0)
    s["v"] = s["v"] + np.array([6.0, -4.0, 2.0])
    print(f"[compgraph_orders] n={sc.n} T={T}", flush=True)
    one = _loop(_cg(sc, s, 0), sc.n)
    seg = _cg(sc, s, 4)
    got = _loop(seg, sc.n)
    assert seg._sim.num_layers == T + 1
    for k in one:
        e = rel_err(got[k], one[k])
        assert e < 1e-10, f"{k}: {e:.3e}"
    # a second episode (set-up again: one order, then re-sorted again)
    got2 = _loop(seg, sc.n)
    for k in one:
        assert rel_err(got2[k], one[k]) < 1e-10, k


def test_segment_every_rejected_with_deterministic_mode():
    import diffmpm_bindings as db
    sc, s = perturbed_scene(n=300, dfc=0.0)
    pc = db.PointCloud.from_arrays(s["x"], s["m"], s["vol"], s["lam"], s["mu"], v=s["v"], F=s["F"], C=s["C"])
    g = db.Grid(*sc.dims, sc.dx, list(sc.grid_min))
    with pytest.raises(ValueError):
        db.CompGraph(pc, g, g, deterministic=True, segment_every=10)
