"""The named workloads at their own sizes.

* 8,192 particles on 32^3, FP64, 50 steps: the positional loss and its
  gradients w.r.t. initial velocity and F against the oracle (1e-10).
* The 64 uneven batched instances: one hipGraph holding every instance's
  episode must reproduce the per-instance eager episodes.
* 4,194,304 particles on 256^3, the full 500-step episode with checkpoint
  interval 25 and controls at 20 layers: size-independent properties -- mass
  and linear momentum conservation, the control gradient against a central
  difference in FP64, and FP32 gradients against FP64.
"""
import numpy as np
import pytest

from helpers import oracle_params, rel_err
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _target(x0):
    c = np.array([0.5, 0.5, 0.5])
    return c + (x0 - c) * np.array([1.2, 0.8, 1.0])


def test_config1_positional_loss_matches_oracle():
    """8,192 particles (r = 0.2 sphere, 2 per cell per axis), 32^3 cells,
    quadratic B-spline, Neo-Hookean E = 1e4, nu = 0.3, dt = 5e-4, 50 steps,
    sticky walls at 3 cells, no gravity.  L = sum |x_T - x*|^2 with
    x* = c + diag(1.2, 0.8, 1.0)(x_0 - c)."""
    from dmpm import Physics, Simulator
    from dmpm.scenes import config1
    sc = config1()
    s = sc.state()
    T, dt = 50, 5e-4
    print(f"[config1] n={sc.n} T={T}", flush=True)
    xs = _target(s["x"])
    out = {}
    for prec in ("f64", "f32"):
        sim = Simulator(sc.dims, sc.dx, physics=Physics.plain(dt=dt), precision=prec, checkpoint_interval=10)
        sim.set_state(s)
        sim.set_up(T + 1)
        sim.forward(0)
        xT = sim.get(T, "x", dtype=np.float64)
        sim.set_last_grad(x=2 * (xT - xs))
        sim.backward(0)
        out[prec] = dict(xT=xT, v=sim.get_grad(0, "v", dtype=np.float64), F=sim.get_grad(0, "F", dtype=np.float64),
                         L=float(((xT - xs) ** 2).sum()))
        sim.close()

    orc = Oracle("f64")
    prm = oracle_params("plain", sc, dt)
    layers = orc.rollout(prm, dict(s, dFc=None), T)
    xT = layers[-1]["x"]
    z3, z33 = np.zeros((sc.n, 3)), np.zeros((sc.n, 3, 3))
    g0, _ = orc.rollout_grad(prm, layers, dict(x=2 * (xT - xs), v=z3, F=z33, C=z33))
    g = out["f64"]
    print(f"[config1] L = {g['L']:.6e}", flush=True)
    assert rel_err(g["xT"], xT) < 1e-10
    assert rel_err(g["v"], g0["v"]) < 1e-10
    assert rel_err(g["F"], g0["F"]) < 1e-10
    # FP32 against FP64 (the precision claim: <= 1e-3 relative on gradients)
    e_v = np.linalg.norm(out["f32"]["v"] - g["v"]) / np.linalg.norm(g["v"])
    e_F = np.linalg.norm(out["f32"]["F"] - g["F"]) / np.linalg.norm(g["F"])
    print(f"[config1] fp32 vs fp64: dL/dv0 {e_v:.2e}, dL/dF0 {e_F:.2e}", flush=True)
    assert e_v < 1e-3 and e_F < 1e-3


def test_config4_batched_graph_matches_eager():
    """All 64 instances of the batched workload (FP64, 6 steps, K = 3): the
    episode graph with one branch per instance, replayed twice, against each
    instance's eager episode."""
    from dmpm import EpisodeGraph, Physics, Simulator
    from dmpm.scenes import config4_instances, instance_scene
    T = 6
    sims, eager = [], []
    for spec in config4_instances():
        sc = instance_scene(spec)
        sim = Simulator(sc.dims, sc.dx, physics=Physics.plain(dt=2.5e-4), precision="f64", checkpoint_interval=3)
        sim.set_state(sc.state())
        sim.set_up(T + 1)
        rng = np.random.default_rng(spec["seed"])
        sim.set_last_grad(x=rng.normal(size=(sc.n, 3)), F=1e-2 * rng.normal(size=(sc.n, 3, 3)))
        sim.forward(0, keep_controls=True)
        sim.backward(0)
        eager.append({"x": sim.get(T, "x", dtype=np.float64), "gv": sim.get_grad(0, "v", dtype=np.float64),
                      "gF": sim.get_grad(0, "F", dtype=np.float64)})
        sims.append(sim)
    print(f"[config4] {len(sims)} instances, {sum(e['x'].shape[0] for e in eager)} particles", flush=True)
    g = EpisodeGraph(sims, start=0, stop=0)
    for rep in range(2):
        g.launch()
        g.synchronize()
        for sim, e in zip(sims, eager):
            assert rel_err(sim.get(T, "x", dtype=np.float64), e["x"]) < 1e-12, rep
            assert rel_err(sim.get_grad(0, "v", dtype=np.float64), e["gv"]) < 1e-11, rep
            assert rel_err(sim.get_grad(0, "F", dtype=np.float64), e["gF"]) < 1e-11, rep
    g.close()
    for sim in sims:
        sim.close()


def test_config5_full_scale_properties():
    """The full-scale morph episode: 4,194,304 particles (r = 0.3 sphere),
    256^3 cells, dt = 5e-5, 500 steps with checkpoint interval 25, a
    per-particle control at 20 control layers (every 25th), and a seeded
    upstream gradient dL/dx (Gaussian, sigma 1e-3) and dL/dF (sigma 1e-4)
    injected at the last layer -- L = <gx, x_T> + <gF, F_T>.  Properties that
    hold at any size: total mass and linear momentum are conserved over the
    500 steps (no wall contact, no external force); the adjoint's gradient
    w.r.t. the control of layer 250 along a random direction matches a central
    difference (FP64, h = 1e-4); FP32 gradients (control and initial
    velocity) match FP64 within 1e-3."""
    from dmpm import Physics, Simulator, p2g_mass
    from dmpm.scenes import config5
    sc = config5()
    s = sc.state()
    T, K, dt = 500, 25, 5e-5
    layers = list(range(0, T, T // 20))
    lc = 250
    print(f"[config5] n={sc.n} dims={sc.dims} T={T} K={K} control layers {len(layers)}", flush=True)
    rng = np.random.default_rng(55)
    gx = 1e-3 * rng.normal(size=(sc.n, 3))
    gF = 1e-4 * rng.normal(size=(sc.n, 3, 3))
    D = rng.normal(size=(sc.n, 3, 3))

    def make(prec):
        sim = Simulator(sc.dims, sc.dx, physics=Physics.plain(dt=dt), precision=prec, checkpoint_interval=K)
        sim.set_state(s)
        sim.set_up(T + 1)
        for t in layers:
            sim.set_control(t, np.zeros((sc.n, 3, 3)))
            sim.request_control_grad(t)
        return sim

    def loss(sim):
        x = sim.get(T, "x", dtype=np.float64)
        F = sim.get(T, "F", dtype=np.float64)
        return float((gx * x).sum() + (gF * F).sum())

    sim = make("f64")
    sim.forward(0, keep_controls=True)
    xT, vT = sim.get(T, "x", dtype=np.float64), sim.get(T, "v", dtype=np.float64)
    m = s["m"].astype(np.float64)
    gm = p2g_mass(sc.dims, sc.dx, sc.grid_min, xT, m, order=2, precision="f64")
    assert abs(gm.sum() - m.sum()) <= 1e-10 * m.sum()
    p0, pT = (m[:, None] * s["v"]).sum(0), (m[:, None] * vT).sum(0)
    scale = (m[:, None] * np.abs(s["v"])).sum()
    print(f"[config5] momentum drift {np.abs(pT - p0).max() / scale:.2e}", flush=True)
    # exact but for rounding, which accumulates with the steps: 1.6e-10 of
    # the momentum scale after these 500 steps (the 20-step form of this
    # test stayed below 1e-10)
    assert np.abs(pT - p0).max() < 1e-9 * scale
    del xT, vT, gm

    sim.set_last_grad(x=gx, F=gF)
    sim.backward(0)
    gc64 = sim.get_control_grad(lc, dtype=np.float64)
    gv64 = sim.get_grad(0, "v", dtype=np.float64)
    an = float((gc64 * D).sum())
    h = 1e-4
    Ls = []
    for sgn in (1, -1):
        sim.set_control(lc, sgn * h * D)
        sim.forward(0, keep_controls=True)
        Ls.append(loss(sim))
    fd = (Ls[0] - Ls[1]) / (2 * h)
    print(f"[config5] control gradient of layer {lc}: adjoint {an:.12e} central difference {fd:.12e}", flush=True)
    assert abs(an - fd) <= 1e-6 * abs(fd)
    sim.close()

    sim32 = make("f32")
    sim32.forward(0, keep_controls=True)
    sim32.set_last_grad(x=gx, F=gF)
    sim32.backward(0)
    gc32 = sim32.get_control_grad(lc, dtype=np.float64)
    gv32 = sim32.get_grad(0, "v", dtype=np.float64)
    ec = np.linalg.norm(gc32 - gc64) / np.linalg.norm(gc64)
    ev = np.linalg.norm(gv32 - gv64) / np.linalg.norm(gv64)
    print(f"[config5] fp32 vs fp64 (relative L2): control gradient {ec:.2e}, dL/dv0 {ev:.2e}; "
          f"max-relative {rel_err(gc32, gc64):.2e}, {rel_err(gv32, gv64):.2e}", flush=True)
    assert ec < 1e-3 and ev < 1e-3
    sim32.close()


@pytest.mark.parametrize("K", [1, 25])
def test_config3_metric_workload(K):
    """The metric workload itself: 1,048,576 particles (r = 0.3 sphere, sine
    velocity field of amplitude 0.5), 128^3 cells, E = 5e4, nu = 0.45,
    dt = 1e-4, the full 300-step episode, with the whole tape (K = 1) and
    checkpointed (K = 25).  FP64: total mass and linear momentum conserved
    (the body stays clear of the walls, no external force), and the adjoint
    of L = <w, x_T> along a random direction of v_0 against a central
    difference (h = 1e-4).  FP32 (FP64 det/log J/F^-T) against FP64: final
    positions within 1e-4 (the box is the unit cube), layer-0 gradients
    within 1e-3 relative L2."""
    from dmpm import Physics, Simulator, p2g_mass
    from dmpm.scenes import config3
    sc = config3()
    s = sc.state()
    T, dt = 300, 1e-4
    print(f"[config3] n={sc.n} dims={sc.dims} T={T} K={K}", flush=True)
    rng = np.random.default_rng(33)
    w = rng.normal(size=(sc.n, 3))
    dv = rng.normal(size=(sc.n, 3))

    def make(prec):
        return Simulator(sc.dims, sc.dx, physics=Physics.plain(dt=dt), precision=prec, checkpoint_interval=K)

    sim = make("f64")
    sim.set_state(s)
    sim.set_up(T + 1)
    sim.forward(0)
    xT, vT = sim.get(T, "x", dtype=np.float64), sim.get(T, "v", dtype=np.float64)
    m = s["m"].astype(np.float64)
    gm = p2g_mass(sc.dims, sc.dx, sc.grid_min, xT, m, order=2, precision="f64")
    assert abs(gm.sum() - m.sum()) <= 1e-10 * m.sum()
    p0, pT = (m[:, None] * s["v"]).sum(0), (m[:, None] * vT).sum(0)
    scale = (m[:, None] * np.abs(s["v"])).sum()
    print(f"[config3] momentum drift {np.abs(pT - p0).max() / scale:.2e}", flush=True)
    assert np.abs(pT - p0).max() < 1e-10 * scale
    sim.set_last_grad(x=w)
    sim.backward(0)
    g64 = {k: sim.get_grad(0, k, dtype=np.float64) for k in ("x", "v", "F", "C")}
    an = float((g64["v"] * dv).sum())
    # The FP64 scatters sum in an order that varies from run to run (atomics),
    # which leaves ~1e-14 relative noise in x_T after 300 steps; the central
    # difference divides that by 2h, so h is 1e-4 (truncation ~h^2 stays far
    # below) and the bar 1e-6.
    h = 1e-4
    Ls = []
    for sgn in (1, -1):
        sim.set_state(dict(s, v=s["v"] + sgn * h * dv))
        sim.set_up(T + 1)
        sim.forward(0)
        Ls.append(float((w * sim.get(T, "x", dtype=np.float64)).sum()))
    fd = (Ls[0] - Ls[1]) / (2 * h)
    print(f"[config3] adjoint {an:.12e} central difference {fd:.12e}", flush=True)
    assert abs(an - fd) <= 1e-6 * abs(fd)
    sim.close()

    sim32 = make("f32")
    sim32.set_state(s)
    sim32.set_up(T + 1)
    sim32.forward(0)
    ex = np.abs(sim32.get(T, "x", dtype=np.float64) - xT).max()
    sim32.set_last_grad(x=w)
    sim32.backward(0)
    errs = {k: np.linalg.norm(sim32.get_grad(0, k, dtype=np.float64) - g64[k]) / np.linalg.norm(g64[k])
            for k in g64}
    print(f"[config3] fp32 vs fp64: x_T {ex:.2e}, gradients {errs}", flush=True)
    assert ex < 1e-4
    for k, e in errs.items():
        assert e < 1e-3, k
    sim32.close()
