"""The differentiable rollout (torch.autograd.Function over the C ABI):
gradients against the oracle, against central finite differences at the
medium-size configuration, and FP32 against FP64."""
import numpy as np
import pytest

from helpers import oracle_params, perturbed_scene, rel_err
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _torch_state(s, dtype, dev):
    import torch
    t = lambda a: torch.tensor(np.asarray(a), dtype=dtype, device=dev)  # noqa: E731
    return {k: t(s[k]) for k in ("x", "v", "F", "C", "m", "vol", "lam", "mu")}


def test_autograd_rollout_matches_oracle():
    import torch
    from dmpm import Physics, Simulator
    from dmpm.rollout import rollout
    sc, s = perturbed_scene(n=1200, dfc=0.0)
    T, dt = 6, 1e-3
    dev = torch.device("cuda:0")
    st = _torch_state(s, torch.float64, dev)
    for k in ("x", "v", "F", "C"):
        st[k].requires_grad_(True)
    rng = np.random.default_rng(2)
    ctl_np = {1: 0.01 * rng.normal(size=(sc.n, 3, 3)), 4: 0.01 * rng.normal(size=(sc.n, 3, 3))}
    ctl = {t: torch.tensor(c, dtype=torch.float64, device=dev, requires_grad=True) for t, c in ctl_np.items()}
    w = {k: torch.tensor(rng.normal(size=st[k].shape), dtype=torch.float64, device=dev) for k in ("x", "v", "F", "C")}
    sim = Simulator(sc.dims, sc.dx, physics=Physics.plain(dt=dt), precision="f64", checkpoint_interval=2)
    xT, vT, FT, CT = rollout(sim, st, T, ctl)
    L = (xT * w["x"]).sum() + (vT * w["v"]).sum() + (FT * w["F"]).sum() + (CT * w["C"]).sum()
    L.backward()

    orc = Oracle("f64")
    prm = oracle_params("plain", sc, dt)
    controls = [ctl_np.get(t, np.zeros((sc.n, 3, 3))) for t in range(T)]
    layers = orc.rollout(prm, dict(s, dFc=None), T, controls)
    assert rel_err(xT.detach().cpu().numpy(), layers[-1]["x"]) < 1e-10
    g0, gctl = orc.rollout_grad(prm, layers, {k: w[k].cpu().numpy() for k in w})
    for k in ("x", "v", "F", "C"):
        assert rel_err(st[k].grad.cpu().numpy(), g0[k]) < 1e-10, k
    for t in ctl:
        assert rel_err(ctl[t].grad.cpu().numpy(), gctl[t]) < 1e-10, t
    sim.close()


def _config2_loss(sim, sc, T, x0, controls):
    """L = sum |x_T - x*|^2, x* = c + diag(1.2, 0.8, 1.0)(x_0 - c)."""
    sim.set_up(T + 1)
    for t, c in controls.items():
        sim.set_control(t, c)
    sim.forward(0, keep_controls=True)
    xT = sim.get(T, "x", dtype=np.float64)
    c = np.array([0.5, 0.5, 0.5])
    xs = c + (x0 - c) * np.array([1.2, 0.8, 1.0])
    return float(((xT - xs) ** 2).sum()), 2 * (xT - xs)


def test_config2_gradients_vs_finite_differences():
    """Medium-size check: 65,536 particles in an ellipsoid on a 64^3 grid,
    dt = 2.5e-4, 200 steps with checkpoints every 20, per-step controls dFc at
    10 control steps.  dL/dv0 and dL/d dFc along 16 seeded random directions
    against central differences (h = 1e-6, FP64); then the FP32 gradients
    against FP64 (relative L2 <= 1e-3)."""
    from dmpm import Physics, Simulator
    from dmpm.scenes import config2
    sc = config2()
    T, K, dt = 200, 20, 2.5e-4
    s = sc.state()
    rng = np.random.default_rng(22)
    ctl_steps = list(range(0, T, T // 10))
    controls = {t: 1e-3 * rng.normal(size=(sc.n, 3, 3)) for t in ctl_steps}
    print(f"[config2] n={sc.n} T={T} K={K} controls at {ctl_steps}", flush=True)

    def make(prec):
        sim = Simulator(sc.dims, sc.dx, physics=Physics.plain(dt=dt), precision=prec, checkpoint_interval=K)
        return sim

    sim = make("f64")
    sim.set_state(s)
    L0, gT = _config2_loss(sim, sc, T, s["x"], controls)
    for t in ctl_steps:
        sim.request_control_grad(t)
    sim.set_last_grad(x=gT)
    sim.backward(0)
    gv0 = sim.get_grad(0, "v", dtype=np.float64)
    gctl = {t: sim.get_control_grad(t, dtype=np.float64) for t in ctl_steps}

    h = 1e-6
    errs = []
    for d in range(16):
        dv = rng.normal(size=s["v"].shape)
        dc = {t: rng.normal(size=(sc.n, 3, 3)) for t in ctl_steps}
        an = (gv0 * dv).sum() + sum((gctl[t] * dc[t]).sum() for t in ctl_steps)
        Ls = []
        for sgn in (1, -1):
            sim.set_state(dict(s, v=s["v"] + sgn * h * dv))
            Ls.append(_config2_loss(sim, sc, T, s["x"], {t: controls[t] + sgn * h * dc[t] for t in ctl_steps})[0])
        fd = (Ls[0] - Ls[1]) / (2 * h)
        errs.append(abs(fd - an) / max(abs(fd), 1e-12))
        print(f"[config2] direction {d}: fd {fd:.10e} adjoint {an:.10e} rel {errs[-1]:.2e}", flush=True)
    sim.close()
    assert max(errs) < 1e-5, errs

    sim32 = make("f32")
    sim32.set_state(s)
    _, gT32 = _config2_loss(sim32, sc, T, s["x"], controls)
    for t in ctl_steps:
        sim32.request_control_grad(t)
    sim32.set_last_grad(x=gT32)
    sim32.backward(0)
    e_v = np.linalg.norm(sim32.get_grad(0, "v", dtype=np.float64) - gv0) / np.linalg.norm(gv0)
    e_c = max(np.linalg.norm(sim32.get_control_grad(t, dtype=np.float64) - gctl[t]) / np.linalg.norm(gctl[t])
              for t in ctl_steps)
    print(f"[config2] fp32 vs fp64: dL/dv0 rel {e_v:.2e}, dL/d dFc rel (max over steps) {e_c:.2e}", flush=True)
    sim32.close()
    assert e_v < 1e-3 and e_c < 1e-3
