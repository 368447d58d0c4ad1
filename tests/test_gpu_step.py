"""GPU parity of one step and its adjoint: libdmpm.so (C ABI) vs the oracle.

FP64 GPU must match the FP64 oracle within 1e-10 relative; FP32 GPU must
match the FP32 oracle within 1e-4 (state) and 1e-3 (gradients), relative to
the field's largest magnitude.
"""
import numpy as np
import pytest

from helpers import oracle_params, perturbed_scene, physics_kwargs, random_adjoint, rel_err
from oracle import Oracle

pytestmark = pytest.mark.gpu

DT = 1e-3
TOL = {"f64": (1e-10, 1e-10), "f32": (1e-4, 1e-3)}


def _sim(sc, name, prec, K=1, recompute_grids=False):
    from dmpm import Physics, Simulator
    return Simulator(sc.dims, sc.dx, sc.grid_min, physics=Physics(**physics_kwargs(name, DT)), precision=prec,
                     checkpoint_interval=K, recompute_grids=recompute_grids)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("name", ["plain", "reference", "reference_exact", "quadratic_fcr", "cubic_nh"])
def test_forward_step(name, prec):
    sc, s = perturbed_scene()
    print(f"[forward_step] physics={name} prec={prec} n={sc.n}", flush=True)
    o = Oracle(prec).forward_step(oracle_params(name, sc, DT), s)
    sim = _sim(sc, name, prec)
    sim.set_state(s)
    sim.set_up(2)
    sim.forward(0)
    dt = np.float64 if prec == "f64" else np.float32
    tol = TOL[prec][0]
    for k in ("x", "v", "F", "C"):
        g = sim.get(1, k, dtype=dt)
        e = rel_err(g, o[k])
        assert e < tol, f"{k}: rel err {e:.3e}"
    sim.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("name", ["plain", "reference", "reference_exact", "quadratic_fcr", "cubic_nh"])
def test_backward_step(name, prec):
    sc, s = perturbed_scene()
    adj = random_adjoint(sc.n)
    print(f"[backward_step] physics={name} prec={prec} n={sc.n}", flush=True)
    o = Oracle(prec).backward_step(oracle_params(name, sc, DT), s, adj)
    sim = _sim(sc, name, prec)
    sim.set_state(s)
    sim.set_up(2)
    sim.forward(0)
    sim.set_last_grad(x=adj["x"], v=adj["v"], F=adj["F"], C_=adj["C"])
    sim.backward(0)
    dt = np.float64 if prec == "f64" else np.float32
    tol = TOL[prec][1]
    for k in ("x", "v", "F", "C"):
        g = sim.get_grad(0, k, dtype=dt)
        e = rel_err(g, o[k])
        assert e < tol, f"d{k}: rel err {e:.3e}"
    g = sim.get_control_grad(0, dtype=dt)
    e = rel_err(g, o["dFc"])
    assert e < tol, f"d dFc: rel err {e:.3e}"
    sim.close()


# (K, recompute_grids): every layer with its kept P2G grid; every layer with
# the grid recomputed in the backward pass; checkpoints every 3 layers; a
# short last segment (layers 5-6 of 0-7); one segment holding the whole tape.
TAPES = [(1, False), (1, True), (3, False), (5, False), (7, False)]


@pytest.mark.parametrize("K,recompute", TAPES)
def test_rollout_checkpointing(K, recompute):
    """A 7-step episode with controls at layers 0 and 2: every tape layout
    must give the oracle's states and gradients.  Then a control is set at
    layer 3 and the forward pass re-run from layer 3 only: the layers and
    gradients must follow the new control (kept grids and segments of the
    old control must not be reused)."""
    sc, s = perturbed_scene(dfc=0.0)
    name, prec, T = "plain", "f64", 7
    rng = np.random.default_rng(3)
    ctl = [0.01 * rng.normal(size=(sc.n, 3, 3)) if t in (0, 2, 3) else None for t in range(T)]
    adj = random_adjoint(sc.n, seed=9)
    orc = Oracle(prec)
    prm = oracle_params(name, sc, DT)
    z33 = np.zeros((sc.n, 3, 3))

    def reference(ctl_layers):
        cs = [ctl[t] if t in ctl_layers else z33 for t in range(T)]
        layers = orc.rollout(prm, dict(s, dFc=None), T, cs)
        g0, gctl = orc.rollout_grad(prm, layers, adj)
        return layers, g0, gctl

    print(f"[rollout] K={K} recompute_grids={recompute} T={T} n={sc.n}", flush=True)
    sim = _sim(sc, name, prec, K=K, recompute_grids=recompute)
    sim.set_state(dict(s, dFc=None))
    sim.set_up(T + 1)
    for t in (0, 2):
        sim.set_control(t, ctl[t])
    sim.request_control_grad(2)
    sim.forward(0, keep_controls=True)
    sim.set_last_grad(x=adj["x"], v=adj["v"], F=adj["F"], C_=adj["C"])
    for case, live in (("controls 0, 2", (0, 2)), ("control 3 added, forward from 3", (0, 2, 3))):
        if 3 in live:
            sim.set_control(3, ctl[3])
            sim.request_control_grad(3)
            sim.forward(3, keep_controls=True)
        layers, g0, gctl = reference(live)
        for k in ("x", "v", "F", "C"):
            assert rel_err(sim.get(T, k, dtype=np.float64), layers[T][k]) < 1e-10, (case, k)
            assert rel_err(sim.get(4, k, dtype=np.float64), layers[4][k]) < 1e-10, (case, k)
        sim.backward(0)
        for k in ("x", "v", "F", "C"):
            e = rel_err(sim.get_grad(0, k, dtype=np.float64), g0[k])
            assert e < 1e-10, f"{case}: d{k}: {e:.3e}"
        for t in live:
            assert rel_err(sim.get_control_grad(t, dtype=np.float64), gctl[t]) < 1e-10, (case, t)
    sim.close()


@pytest.mark.parametrize("K", [1, 3])
@pytest.mark.parametrize("name", ["plain", "reference", "reference_exact", "quadratic_fcr"])
def test_adjoint_reads_next_layer(name, K):
    """The adjoint of step t reads G2P's v' and C' from layer t+1 (the
    reference reads layer n+1's C) instead of recomputing them, while layer
    t+1 is the forward step of layer t as it stands.  Gradients must match
    the always-recompute form, also after a control changes without a new
    forward pass (layer t+1 is then stale and must not be read)."""
    from dmpm import Physics, Simulator
    sc, s = perturbed_scene(dfc=0.0)
    T = 6
    rng = np.random.default_rng(21)
    adj = random_adjoint(sc.n, seed=4)
    ctl2 = 0.01 * rng.normal(size=(sc.n, 3, 3))
    print(f"[adjoint_reads_next_layer] physics={name} K={K} T={T} n={sc.n}", flush=True)
    grads = []
    for recompute in (False, True):
        sim = Simulator(sc.dims, sc.dx, sc.grid_min, physics=Physics(**physics_kwargs(name, DT)), precision="f64",
                        checkpoint_interval=K, adjoint_recompute=recompute)
        sim.set_state(dict(s, dFc=None))
        sim.set_up(T + 1)
        sim.request_control_grad(2)
        sim.forward(0, keep_controls=True)
        sim.set_last_grad(x=adj["x"], v=adj["v"], F=adj["F"], C_=adj["C"])
        sim.backward(0)
        g = {k: sim.get_grad(0, k, dtype=np.float64) for k in ("x", "v", "F", "C")}
        g["ctl"] = sim.get_control_grad(2, dtype=np.float64)
        # a control at layer 2 changes; no forward pass: layer 3 is stale
        sim.set_control(2, ctl2)
        sim.backward(0)
        g2 = {k: sim.get_grad(0, k, dtype=np.float64) for k in ("x", "v", "F", "C")}
        g2["ctl"] = sim.get_control_grad(2, dtype=np.float64)
        grads.append((g, g2))
        sim.close()
    for case in (0, 1):
        for k in ("x", "v", "F", "C", "ctl"):
            e = rel_err(grads[0][case][k], grads[1][case][k])
            assert e < 1e-11, f"case {case}: d{k}: {e:.3e}"
    # the stale case differs from the fresh one (the control did act)
    assert rel_err(grads[0][1]["v"], grads[0][0]["v"]) > 1e-8


@pytest.mark.parametrize("name", ["plain", "reference"])
def test_block_table_cuts(name):
    """Thread blocks cut where their stencil box would exceed split_nodes:
    with a tiny limit most blocks are cut short (more, partly empty blocks);
    states and gradients of a 4-step episode must still match the oracle
    within the FP64 bound, and the same as without any cuts."""
    from dmpm import Physics, Simulator
    sc, s = perturbed_scene(dfc=0.0)
    T = 4
    adj = random_adjoint(sc.n, seed=6)
    orc = Oracle("f64")
    prm = oracle_params(name, sc, DT)
    layers = orc.rollout(prm, dict(s, dFc=None), T)
    g0, _ = orc.rollout_grad(prm, layers, adj)
    windows = {}
    for split in (40, -1):
        print(f"[block_table_cuts] physics={name} split_nodes={split} n={sc.n}", flush=True)
        sim = Simulator(sc.dims, sc.dx, sc.grid_min, physics=Physics(**physics_kwargs(name, DT)), precision="f64",
                        split_nodes=split)
        sim.set_state(dict(s, dFc=None))
        sim.set_up(T + 1)
        sim.window_stats(reset=True)
        sim.count_windows(True)
        sim.forward(0)
        sim.set_last_grad(x=adj["x"], v=adj["v"], F=adj["F"], C_=adj["C"])
        sim.backward(0)
        windows[split] = sum(sim.window_stats())
        for k in ("x", "v", "F", "C"):
            assert rel_err(sim.get(T, k, dtype=np.float64), layers[T][k]) < 1e-10, (split, k)
            e = rel_err(sim.get_grad(0, k, dtype=np.float64), g0[k])
            assert e < 1e-10, f"split {split}: d{k}: {e:.3e}"
        sim.close()
    assert windows[40] > 2 * windows[-1], windows


@pytest.mark.parametrize("K", [1, 3])
def test_sort_particles_keeps_results(K):
    """Reordering the device storage by a middle layer's positions (every
    per-particle array permuted, blocks rebuilt) leaves states, gradients and
    control gradients unchanged (FP64, up to summation order), including
    controls and Adam state set before the reorder."""
    from dmpm import Physics, Simulator
    sc, s = perturbed_scene(dfc=0.0)
    # a drift of about a cell over the episode, so that the middle layer's
    # cells, and with them the order, differ from the uploaded ones
    s["v"] = s["v"] + np.array([8.0, 3.0, 0.0])
    T = 6
    rng = np.random.default_rng(8)
    adj = random_adjoint(sc.n, seed=3)
    ctl = 0.01 * rng.normal(size=(sc.n, 3, 3))
    print(f"[sort_particles] K={K} T={T} n={sc.n}", flush=True)
    sim = Simulator(sc.dims, sc.dx, sc.grid_min, physics=Physics(**physics_kwargs("plain", DT)), precision="f64",
                    checkpoint_interval=K)
    sim.set_state(dict(s, dFc=None))
    sim.set_up(T + 1)
    sim.set_control(2, ctl)
    sim.request_control_grad(2)

    def episode():
        sim.forward(0, keep_controls=True)
        sim.set_last_grad(x=adj["x"], v=adj["v"], F=adj["F"], C_=adj["C"])
        sim.backward(0)
        out = {k: sim.get(T, k, dtype=np.float64) for k in ("x", "v", "F", "C")}
        out.update({"g" + k: sim.get_grad(0, k, dtype=np.float64) for k in ("x", "v", "F", "C")})
        out["gctl"] = sim.get_control_grad(2, dtype=np.float64)
        out["ctl"] = sim.get(2, "dFc", dtype=np.float64)
        return out

    before = episode()
    sim.sort_particles(T // 2)
    mid = {k: sim.get(3, k, dtype=np.float64) for k in ("x", "F")}
    after = episode()
    for k in before:
        e = rel_err(after[k], before[k])
        assert e < 1e-10, f"{k}: {e:.3e}"
    assert rel_err(sim.get(3, "x", dtype=np.float64), mid["x"]) < 1e-12
    sim.close()


# (physics, precision, LDS budget, block size): 64-, 128- and 192-thread
# blocks; LDS budgets so small that most windows (1 KiB: every window of more
# than 32 FP64 nodes) or all of them (2 KiB with 64-node cubic stencils and
# seven staged values) fall back to global atomics.
SHAPES = [("plain", "f64", 0, 64), ("plain", "f64", 0, 192), ("plain", "f64", 1024, 256),
          ("reference", "f64", 2048, 128), ("plain", "f32", 1024, 256), ("reference", "f32", 0, 64)]


@pytest.mark.parametrize("name,prec,lds,block", SHAPES)
def test_launch_shapes_match_oracle(name, prec, lds, block):
    """A 3-step episode (the fused forward step and the fused adjoint) under
    other block sizes and LDS budgets, including windows too large for the
    budget, whose particles scatter and gather through global memory:
    states and gradients must match the oracle's."""
    from dmpm import Physics, Simulator
    sc, s = perturbed_scene(dfc=0.0)
    T = 3
    adj = random_adjoint(sc.n, seed=13)
    orc = Oracle(prec)
    prm = oracle_params(name, sc, DT)
    layers = orc.rollout(prm, dict(s, dFc=None), T, [np.zeros((sc.n, 3, 3))] * T)
    g0, _ = orc.rollout_grad(prm, layers, adj)
    print(f"[launch_shapes] physics={name} prec={prec} lds={lds} block={block} T={T} n={sc.n}", flush=True)
    sim = Simulator(sc.dims, sc.dx, sc.grid_min, physics=Physics(**physics_kwargs(name, DT)), precision=prec,
                    lds_bytes=lds, block=block)
    sim.set_state(dict(s, dFc=None))
    sim.set_up(T + 1)
    sim.window_stats(reset=True)
    sim.count_windows(True)
    sim.forward(0)
    sim.set_last_grad(x=adj["x"], v=adj["v"], F=adj["F"], C_=adj["C"])
    sim.backward(0)
    in_lds, fell_back = sim.window_stats()
    if lds:
        assert fell_back > 0, "the small LDS budget should send windows to global atomics"
    else:
        assert in_lds > 0 and fell_back == 0, (in_lds, fell_back)
    dt = np.float64 if prec == "f64" else np.float32
    tol_s, tol_g = (1e-10, 1e-10) if prec == "f64" else (1e-4, 1e-3)
    for k in ("x", "v", "F", "C"):
        e = rel_err(sim.get(T, k, dtype=dt), layers[T][k])
        assert e < tol_s, f"{k}: rel err {e:.3e}"
        e = rel_err(sim.get_grad(0, k, dtype=dt), g0[k])
        assert e < tol_g, f"d{k}: rel err {e:.3e}"
    sim.close()


@pytest.mark.parametrize("K", [1, 3])
@pytest.mark.parametrize("name", ["plain", "reference"])
def test_unfused_kernels_match_oracle(name, K):
    """unfused = 1: separate G2P and P2G launches in the forward pass (and
    no clears folded into them), the same states and gradients as the
    oracle over a 5-step episode with a control."""
    from dmpm import Physics, Simulator
    sc, s = perturbed_scene(dfc=0.0)
    T = 5
    rng = np.random.default_rng(8)
    ctl = 0.01 * rng.normal(size=(sc.n, 3, 3))
    adj = random_adjoint(sc.n, seed=6)
    orc = Oracle("f64")
    prm = oracle_params(name, sc, DT)
    z = np.zeros((sc.n, 3, 3))
    layers = orc.rollout(prm, dict(s, dFc=None), T, [ctl if t == 1 else z for t in range(T)])
    g0, gctl = orc.rollout_grad(prm, layers, adj)
    print(f"[unfused] physics={name} K={K} T={T} n={sc.n}", flush=True)
    sim = Simulator(sc.dims, sc.dx, sc.grid_min, physics=Physics(**physics_kwargs(name, DT)), precision="f64",
                    checkpoint_interval=K, unfused=True)
    sim.set_state(dict(s, dFc=None))
    sim.set_up(T + 1)
    sim.set_control(1, ctl)
    sim.request_control_grad(1)
    sim.forward(0, keep_controls=True)
    sim.set_last_grad(x=adj["x"], v=adj["v"], F=adj["F"], C_=adj["C"])
    sim.backward(0)
    for k in ("x", "v", "F", "C"):
        assert rel_err(sim.get(T, k, dtype=np.float64), layers[T][k]) < 1e-10, k
        e = rel_err(sim.get_grad(0, k, dtype=np.float64), g0[k])
        assert e < 1e-10, f"d{k}: rel err {e:.3e}"
    assert rel_err(sim.get_control_grad(1, dtype=np.float64), gctl[1]) < 1e-10
    sim.close()
