"""The deterministic two-pass mode: bitwise reproducible episodes, equal to
the atomic mode within rounding, and to the oracle in FP64."""
import numpy as np
import pytest

from helpers import oracle_params, perturbed_scene, random_adjoint, rel_err
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _episode(sc, s, prec, deterministic, name="plain", T=5, K=2, dt=1e-3):
    from dmpm import Physics, Simulator
    from helpers import physics_kwargs
    sim = Simulator(sc.dims, sc.dx, physics=Physics(**physics_kwargs(name, dt)), precision=prec,
                    checkpoint_interval=K, deterministic=deterministic)
    sim.set_state(s)
    sim.set_up(T + 1)
    sim.forward(0, keep_controls=True)
    adj = random_adjoint(sc.n, seed=3)
    sim.set_last_grad(x=adj["x"], v=adj["v"], F=adj["F"], C_=adj["C"])
    sim.backward(0)
    dt_ = np.float64 if prec == "f64" else np.float32
    out = {k: sim.get(T, k, dtype=dt_) for k in ("x", "v", "F", "C")}
    out.update({"g" + k: sim.get_grad(0, k, dtype=dt_) for k in ("x", "v", "F", "C")})
    out["gdFc"] = sim.get_control_grad(0, dtype=dt_)
    sim.close()
    return out


@pytest.mark.parametrize("name", ["plain", "reference"])
def test_deterministic_is_bitwise_reproducible(name):
    sc, s = perturbed_scene(n=4000)
    print(f"[det] physics={name} n={sc.n}", flush=True)
    a = _episode(sc, s, "f32", True, name)
    b = _episode(sc, s, "f32", True, name)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    c = _episode(sc, s, "f32", False, name)
    for k in a:
        assert rel_err(a[k], c[k]) < (1e-5 if not k.startswith("g") else 1e-4), k


def test_deterministic_fp64_matches_oracle():
    sc, s = perturbed_scene(n=2000)
    T, dt = 5, 1e-3
    a = _episode(sc, s, "f64", True, "plain", T=T, dt=dt)
    orc = Oracle("f64")
    prm = oracle_params("plain", sc, dt)
    # set_up_comp_graph copies layer 0 -- dFc included -- into every layer
    # (CompGraph.cpp:30-40), and the rollout keeps every layer's control
    layers = orc.rollout(prm, s, T, [s["dFc"]] * T)
    adj = random_adjoint(sc.n, seed=3)
    g0, gctl = orc.rollout_grad(prm, layers, adj)
    for k in ("x", "v", "F", "C"):
        assert rel_err(a[k], layers[-1][k]) < 1e-10, k
        assert rel_err(a["g" + k], g0[k]) < 1e-10, k
    assert rel_err(a["gdFc"], gctl[0]) < 1e-10
    # the deterministic path agrees with the atomic one to FP64 rounding
    b = _episode(sc, s, "f64", False, "plain", T=T, dt=dt)
    for k in a:
        assert rel_err(a[k], b[k]) < 1e-11, k
