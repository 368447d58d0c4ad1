"""Batched episodes: every kernel of many sims' episodes as one launch over
all instances (one hipGraph).  Replays must reproduce each instance's eager
episode: the 64 uneven instances of the batched workload in FP64 with a
control and its gradient, the reference physics in FP32 on a subset, the
kept tape (K = 1) in FP32 on a subset, and a mismatched episode structure is
rejected."""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu


def _make(spec, prec, physics, T, K, control_layer):
    from dmpm import Simulator
    from dmpm.scenes import instance_scene
    sc = instance_scene(spec)
    sim = Simulator(sc.dims, sc.dx, physics=physics, precision=prec, checkpoint_interval=K)
    sim.set_state(sc.state())
    sim.set_up(T + 1)
    rng = np.random.default_rng(spec["seed"])
    if control_layer is not None:
        sim.set_control(control_layer, 1e-3 * rng.normal(size=(sc.n, 3, 3)))
        sim.request_control_grad(control_layer)
    sim.set_last_grad(x=rng.normal(size=(sc.n, 3)), F=1e-2 * rng.normal(size=(sc.n, 3, 3)))
    return sim


def _results(sim, T, control_layer):
    out = {"x": sim.get(T, "x", dtype=np.float64), "F": sim.get(T, "F", dtype=np.float64),
           "gv": sim.get_grad(0, "v", dtype=np.float64), "gF": sim.get_grad(0, "F", dtype=np.float64)}
    if control_layer is not None:
        out["gctl"] = sim.get_control_grad(control_layer, dtype=np.float64)
    return out


def _check(specs, prec, physics, T, K, control_layer, tol):
    from dmpm import BatchedEpisode
    sims, eager = [], []
    for spec in specs:
        sim = _make(spec, prec, physics, T, K, control_layer)
        sim.forward(0, keep_controls=True)
        sim.backward(0)
        eager.append(_results(sim, T, control_layer))
        sims.append(sim)
    print(f"[batch] {len(sims)} instances, {sum(e['x'].shape[0] for e in eager)} particles, {prec}", flush=True)
    b = BatchedEpisode(sims, start=0, stop=0)
    for rep in range(2):
        b.launch()
        b.synchronize()
        for k, (sim, e) in enumerate(zip(sims, eager)):
            r = _results(sim, T, control_layer)
            for f in e:
                assert rel_err(r[f], e[f]) < tol, (rep, k, f)
    b.close()
    for sim in sims:
        sim.close()


def test_batched_workload_fp64_matches_eager():
    from dmpm import Physics
    from dmpm.scenes import config4_instances
    _check(config4_instances(), "f64", Physics.plain(dt=2.5e-4), T=6, K=3, control_layer=2, tol=1e-11)


def test_batched_reference_physics_fp32_matches_eager():
    from dmpm import Physics
    from dmpm.scenes import config4_instances
    _check(config4_instances()[:8], "f32", Physics.reference(dt=2e-4, drag=0.0, smoothing=0.0), T=5, K=2,
           control_layer=None, tol=1e-4)


def test_batched_kept_tape_fp32_matches_eager():
    """Every layer's grid kept (K = 1): each fused step of the batched launch
    clears the grid two layers ahead and each fused adjoint launch the next
    dL/dv grid, per instance; two replays must both match the eager
    episodes."""
    from dmpm import Physics
    from dmpm.scenes import config4_instances
    _check(config4_instances()[:16], "f32", Physics.plain(dt=2.5e-4), T=7, K=1, control_layer=3, tol=1e-4)


def test_mismatched_episodes_are_rejected():
    from dmpm import BatchedEpisode, Physics
    from dmpm._lib import DmpmError
    from dmpm.scenes import config4_instances
    specs = config4_instances()[:2]
    a = _make(specs[0], "f32", Physics.plain(dt=2.5e-4), 4, 2, None)
    b = _make(specs[1], "f32", Physics.plain(dt=2.5e-4), 6, 2, None)
    with pytest.raises(DmpmError, match="same structure"):
        BatchedEpisode([a, b])
    c = _make(specs[1], "f64", Physics.plain(dt=2.5e-4), 4, 2, None)
    with pytest.raises(DmpmError, match="share precision"):
        BatchedEpisode([a, c])
    for s in (a, b, c):
        s.close()


def test_batched_instances_against_oracle_and_conservation():
    """All 64 uneven instances of the batched workload, FP64, the full
    100-step episode (dt = 2.5e-4, K = 10) with a control at layer 3, run as
    one batched episode: every instance conserves its mass and linear
    momentum (no wall contact, no external force); the 4 smallest instances'
    final states, layer-0 gradients and control gradients match the oracle
    run on the same inputs (1e-10)."""
    from dmpm import BatchedEpisode, Physics, p2g_mass
    from dmpm.scenes import config4_instances, instance_scene
    from helpers import oracle_params
    from oracle import Oracle
    T, K, ctl_layer, dt = 100, 10, 3, 2.5e-4
    ph = Physics.plain(dt=dt)
    specs = config4_instances()
    sims = [_make(spec, "f64", ph, T, K, ctl_layer) for spec in specs]
    print(f"[batch_oracle] {len(sims)} instances, {sum(s.n for s in sims)} particles", flush=True)
    b = BatchedEpisode(sims, start=0, stop=0)
    b.launch()
    b.synchronize()
    b.close()
    for spec, sim in zip(specs, sims):
        sc = instance_scene(spec)
        m = sc.m.astype(np.float64)
        xT, vT = sim.get(T, "x", dtype=np.float64), sim.get(T, "v", dtype=np.float64)
        gm = p2g_mass(sc.dims, sc.dx, sc.grid_min, xT, m, order=2, precision="f64")
        assert abs(gm.sum() - m.sum()) <= 1e-10 * m.sum(), spec["seed"]
        p0, pT = (m[:, None] * sc.v).sum(0), (m[:, None] * vT).sum(0)
        assert np.abs(pT - p0).max() <= 1e-10 * (m[:, None] * np.abs(sc.v)).sum(), spec["seed"]
    smallest = sorted(range(len(specs)), key=lambda i: specs[i]["n"])[:4]
    orc = Oracle("f64")
    for i in smallest:
        spec, sim = specs[i], sims[i]
        sc = instance_scene(spec)
        rng = np.random.default_rng(spec["seed"])
        ctl = 1e-3 * rng.normal(size=(sc.n, 3, 3))
        gT = dict(x=rng.normal(size=(sc.n, 3)), v=np.zeros((sc.n, 3)), F=1e-2 * rng.normal(size=(sc.n, 3, 3)),
                  C=np.zeros((sc.n, 3, 3)))
        prm = oracle_params("plain", sc, dt)
        controls = [ctl if t == ctl_layer else None for t in range(T)]
        layers = orc.rollout(prm, dict(sc.state(), dFc=None), T, controls)
        g0, gctl = orc.rollout_grad(prm, layers, gT)
        print(f"[batch_oracle] instance {i}: n={sc.n} cells={spec['cells']}", flush=True)
        for k in ("x", "v", "F", "C"):
            assert rel_err(sim.get(T, k, dtype=np.float64), layers[-1][k]) < 1e-10, (i, k)
            assert rel_err(sim.get_grad(0, k, dtype=np.float64), g0[k]) < 1e-10, (i, "d" + k)
        assert rel_err(sim.get_control_grad(ctl_layer, dtype=np.float64), gctl[ctl_layer]) < 1e-10, i
    for sim in sims:
        sim.close()
