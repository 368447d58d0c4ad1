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


def _sim(sc, name, prec, K=1):
    from dmpm import Physics, Simulator
    return Simulator(sc.dims, sc.dx, sc.grid_min, physics=Physics(**physics_kwargs(name, DT)), precision=prec,
                     checkpoint_interval=K)


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


@pytest.mark.parametrize("K", [1, 3])
def test_rollout_checkpointing(K):
    """A 7-step episode with controls at layers 0 and 2; the checkpointed
    backward (K=3) must give the oracle's gradients like the stored one."""
    sc, s = perturbed_scene(dfc=0.0)
    name, prec, T = "plain", "f64", 7
    rng = np.random.default_rng(3)
    ctl = [0.01 * rng.normal(size=(sc.n, 3, 3)) if t in (0, 2) else None for t in range(T)]
    adj = random_adjoint(sc.n, seed=9)
    orc = Oracle(prec)
    prm = oracle_params(name, sc, DT)
    layers = orc.rollout(prm, dict(s, dFc=None), T, [c if c is not None else np.zeros((sc.n, 3, 3)) for c in ctl])
    g0, gctl = orc.rollout_grad(prm, layers, adj)
    print(f"[rollout] K={K} T={T} n={sc.n}", flush=True)
    sim = _sim(sc, name, prec, K=K)
    sim.set_state(dict(s, dFc=None))
    sim.set_up(T + 1)
    for t in (0, 2):
        sim.set_control(t, ctl[t])
    sim.request_control_grad(2)
    sim.forward(0, keep_controls=True)
    for k in ("x", "v", "F", "C"):
        assert rel_err(sim.get(T, k, dtype=np.float64), layers[T][k]) < 1e-10, k
        assert rel_err(sim.get(4, k, dtype=np.float64), layers[4][k]) < 1e-10, k
    sim.set_last_grad(x=adj["x"], v=adj["v"], F=adj["F"], C_=adj["C"])
    sim.backward(0)
    for k in ("x", "v", "F", "C"):
        e = rel_err(sim.get_grad(0, k, dtype=np.float64), g0[k])
        assert e < 1e-9, f"d{k}: {e:.3e}"
    assert rel_err(sim.get_control_grad(0, dtype=np.float64), gctl[0]) < 1e-9
    assert rel_err(sim.get_control_grad(2, dtype=np.float64), gctl[2]) < 1e-9
    sim.close()
