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
            assert e < 1e-9, f"{case}: d{k}: {e:.3e}"
        for t in live:
            assert rel_err(sim.get_control_grad(t, dtype=np.float64), gctl[t]) < 1e-9, (case, t)
    sim.close()
