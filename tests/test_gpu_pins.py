"""GPU parity on the scenes that pin the oracle's remaining branches: the
J trust region on a compressed, converging body (both branches on more than
10 % of the particles), degenerate deformation gradients in the
fixed-corotated stress and its Hessian action, the reference's gradient at
the 8,192-particle size, and the reference's optimisation loops.

FP64 GPU against the FP64 oracle within 1e-10 relative; FP32 GPU against the
FP32 oracle within 1e-4 (state) and 1e-3 (gradients); FP32 GPU against the
reference's own float32 outputs (golden fixtures) within the tolerances the
oracle meets against them (the reference's SVD stops at 1024 eps,
ImplicitQRSVD.h:694-711).
"""
import os

import numpy as np
import pytest

from helpers import random_adjoint, rel_err, trust_branches
from oracle import Oracle, Params, reference_params

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = {"f64": (1e-10, 1e-10), "f32": (1e-4, 1e-3)}


def load(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


def fixture_params(d, **kw) -> Params:
    return reference_params(dims=tuple(int(a) for a in d["dims"]), dx=float(d["dx"]), dt=float(d["dt"]),
                            drag=float(d["drag"]), smoothing=float(d["smoothing"]),
                            f_ext=tuple(float(a) for a in d["f_ext"]), **kw)


def sim_for(prm: Params, prec, K=1):
    from dmpm import Physics, Simulator
    ph = Physics(order=prm.order, model=prm.model, trust_region=prm.trust_region, zeta=prm.zeta, jmin=prm.jmin,
                 wall_cells=prm.wall_cells, adjoint=prm.adjoint, dt=prm.dt, drag=prm.drag,
                 smoothing=prm.smoothing, f_ext=tuple(prm.f_ext))
    return Simulator(prm.dims, prm.dx, prm.grid_min, physics=ph, precision=prec, checkpoint_interval=K)


def trust_state(prec):
    d = load("ref_trust_forward.npz")
    dt = np.float64 if prec == "f64" else np.float32
    return d, {k[3:]: np.asarray(d[k], dt) for k in d if k.startswith("in_")}


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("adjoint", [1, 0])
def test_trust_region_step_and_adjoint(adjoint, prec):
    """One step and its adjoint on the compressive body, reference physics
    with the reference's adjoint (1) or the exact one (0, which recomputes
    C' before the clamp), against the oracle."""
    d, s = trust_state(prec)
    prm = fixture_params(d, adjoint=adjoint)
    scaled, clamped = trust_branches(prm, s)
    n = s["x"].shape[0]
    print(f"[trust_step] adjoint={adjoint} prec={prec} n={n} scaled={scaled.sum()} clamped={clamped.sum()}",
          flush=True)
    assert scaled.sum() > 0.1 * n and clamped.sum() > 0.1 * n
    orc = Oracle(prec)
    o = orc.forward_step(prm, s)
    adj = random_adjoint(n, seed=9)
    b = orc.backward_step(prm, s, adj)
    sim = sim_for(prm, prec)
    sim.set_state(s)
    sim.set_up(2)
    sim.request_control_grad(0)
    sim.forward(0)
    dt = np.float64 if prec == "f64" else np.float32
    ts, tg = TOL[prec]
    for k in ("x", "v", "F", "C"):
        e = rel_err(sim.get(1, k, dtype=dt), o[k])
        assert e < ts, f"{k}: {e:.3e}"
    sim.set_last_grad(x=adj["x"], v=adj["v"], F=adj["F"], C_=adj["C"])
    sim.backward(0)
    for k in ("x", "v", "F", "C"):
        e = rel_err(sim.get_grad(0, k, dtype=dt), b[k])
        assert e < tg, f"d{k}: {e:.3e}"
    e = rel_err(sim.get_control_grad(0, dtype=dt), b["dFc"])
    assert e < tg, f"d dFc: {e:.3e}"
    sim.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("K", [1, 2])
@pytest.mark.parametrize("adjoint", [1, 0])
def test_trust_region_rollout(adjoint, K, prec):
    """A 4-step episode of the compressive body with controls at layers 0
    and 2 and a checkpoint interval K, against the oracle."""
    d, s = trust_state(prec)
    prm = fixture_params(d, adjoint=adjoint)
    n = s["x"].shape[0]
    T = 4
    print(f"[trust_rollout] adjoint={adjoint} K={K} prec={prec} n={n}", flush=True)
    rng = np.random.default_rng(3)
    ctl = [s["dFc"], None, (2e-3 * rng.normal(size=(n, 3, 3))).astype(s["x"].dtype), None]
    orc = Oracle(prec)
    layers = orc.rollout(prm, dict(s, dFc=None), T, ctl)
    gT = random_adjoint(n, seed=13)
    g0, gctl = orc.rollout_grad(prm, layers, gT)
    sim = sim_for(prm, prec, K)
    sim.set_state(dict(s, dFc=None))
    sim.set_up(T + 1)
    for t in (0, 2):
        sim.set_control(t, ctl[t])
        sim.request_control_grad(t)
    sim.forward(0, keep_controls=True)
    sim.set_last_grad(x=gT["x"], v=gT["v"], F=gT["F"], C_=gT["C"])
    sim.backward(0)
    dt = np.float64 if prec == "f64" else np.float32
    ts, tg = TOL[prec]
    for k in ("x", "v", "F", "C"):
        e = rel_err(sim.get(T, k, dtype=dt), layers[-1][k])
        assert e < ts, f"{k}: {e:.3e}"
    for k in ("x", "v", "F", "C"):
        e = rel_err(sim.get_grad(0, k, dtype=dt), g0[k])
        assert e < tg, f"d{k}: {e:.3e}"
    for t in (0, 2):
        e = rel_err(sim.get_control_grad(t, dtype=dt), gctl[t])
        assert e < tg, f"d dFc[{t}]: {e:.3e}"
    sim.close()


def test_trust_region_against_reference_outputs():
    """FP32 GPU, reference physics, against the reference's own step,
    adjoint step and 4-layer CompGraph episode on the compressive body
    (float32 golden fixtures)."""
    import diffmpm_bindings as db
    d = load("ref_trust_forward.npz")
    s = {k[3:]: d[k] for k in d if k.startswith("in_")}
    prm = fixture_params(d)
    sim = sim_for(prm, "f32")
    sim.set_state(s)
    sim.set_up(2)
    sim.forward(0)
    tol = {"x": 1e-6, "v": 2e-5, "F": 1e-5, "C": 1e-4}
    for k, t in tol.items():
        e = rel_err(sim.get(1, k), d[f"out_{k}"])
        assert e < t, f"{k}: {e:.3e}"
    b = load("ref_trust_backward.npz")
    sim.set_last_grad(x=b["adj_x"], v=b["adj_v"], F=b["adj_F"], C_=b["adj_C"])
    sim.backward(0)
    for k in ("x", "v", "F", "C"):
        e = rel_err(sim.get_grad(0, k), b[f"out_{k}"])
        assert e < 2e-4, f"d{k}: {e:.3e}"
    sim.close()
    e = load("ref_trust_episode.npz")
    st = {k[3:]: e[k] for k in e if k.startswith("in_")}
    dims = tuple(int(a) for a in e["dims"])
    pc = db.PointCloud.from_arrays(st["x"], st["m"], st["vol"], st["lam"], st["mu"], v=st["v"], F=st["F"], C=st["C"],
                                   dFc=st["dFc"])
    tg = db.Grid(*dims, float(e["dx"]), [0.0, 0.0, 0.0])
    tg.m = e["target_m"].astype(np.float32)
    cg = db.CompGraph(pc, db.Grid(*dims, float(e["dx"]), [0.0, 0.0, 0.0]), tg)
    cg._sim.set_sim_params(float(e["dt"]), float(e["drag"]), float(e["smoothing"]), tuple(e["f_ext"]))
    L = int(e["num_layers"])
    cg.set_up_comp_graph(L)
    cg.compute_forward_pass(0, 0)
    last = cg.get_point_cloud(L - 1)
    assert rel_err(np.array(last.get_positions()), e["out_x"]) < 2e-4
    assert rel_err(np.array(last.get_def_grads()), e["out_F"]) < 2e-4
    loss = cg.end_layer_mass_loss()
    assert abs(loss - float(e["out_loss"])) <= 2e-4 * abs(float(e["out_loss"]))
    cg.set_render_gradients(e["render_gF"], e["render_gx"])
    cg.compute_backward_pass(0)
    for k in ("x", "v", "F", "C"):
        err = rel_err(cg._sim.get_grad(0, k, dtype=np.float64), e[f"out_g{k}"])
        assert err < 1e-3, f"d{k}: {err:.3e}"


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_degenerate_deformation_gradients(prec):
    """Particles whose F are the degenerate matrices of the elasticity
    fixture -- scaled rotations taking the +1e-4 I regularisation of the
    rotation term, repeated singular values, exact diagonal and permutation
    inputs including inverted ones with sigma_2 + sigma_3 = 0 -- through one
    step and its adjoint of the reference physics (the stress P and the
    Hessian action of the GPU's Jacobi SVD), against the oracle."""
    from dmpm.scenes import Body, make_scene
    e = load("ref_elasticity_degenerate.npz")
    Fs = np.concatenate([e["regularised_F"], e["repeated_F"], e["exact_F"]]).astype(np.float64)
    sc = make_scene(Body((0.5, 0.5, 0.5), (0.2, 0.2, 0.2)), 600, 16, 61, 1e4, 0.3, velocity="sine", amplitude=0.3)
    s = sc.state()
    s["F"] = Fs[np.arange(sc.n) % Fs.shape[0]]
    s["C"] = np.zeros((sc.n, 3, 3))
    dt = np.float64 if prec == "f64" else np.float32
    s = {k: np.asarray(v, dt) for k, v in s.items()}
    prm = reference_params(dims=sc.dims, dx=sc.dx, dt=1e-4, drag=0.5, smoothing=0.1, f_ext=(0.0, -9.8, 0.0))
    print(f"[degenerate] prec={prec} n={sc.n} kinds={Fs.shape[0]}", flush=True)
    orc = Oracle(prec)
    o = orc.forward_step(prm, s)
    adj = random_adjoint(sc.n, seed=17)
    b = orc.backward_step(prm, s, adj)
    sim = sim_for(prm, prec)
    sim.set_state(s)
    sim.set_up(2)
    sim.forward(0)
    ts, tg = TOL[prec]
    for k in ("x", "v", "F", "C"):
        err = rel_err(sim.get(1, k, dtype=dt), o[k])
        assert err < ts, f"{k}: {err:.3e}"
    err = rel_err(sim.get(0, "P", dtype=dt), o["P"])
    assert err < ts, f"P: {err:.3e}"
    sim.set_last_grad(x=adj["x"], v=adj["v"], F=adj["F"], C_=adj["C"])
    sim.backward(0)
    for k in ("x", "v", "F", "C"):
        err = rel_err(sim.get_grad(0, k, dtype=dt), b[k])
        assert err < tg, f"d{k}: {err:.3e}"
    sim.close()


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_c1_size_reference_gradient(prec):
    """The reference's physics on the 8,192-particle body of 32^3 cells:
    5 steps and the adjoint of the positional loss.  FP32 against the
    reference's own float32 run (state 2e-4, gradients 1e-3); FP64 against
    the FP64 oracle (1e-10)."""
    d = load("ref_c1_rollout.npz")
    prm = fixture_params(d)
    n = d["in_x"].shape[0]
    T = int(d["steps"])
    dt = np.float64 if prec == "f64" else np.float32
    s = dict(x=d["in_x"], v=d["in_v"], F=np.broadcast_to(np.eye(3), (n, 3, 3)).copy(), C=np.zeros((n, 3, 3)),
             m=d["in_m"], vol=d["in_vol"], lam=d["in_lam"], mu=d["in_mu"])
    s = {k: np.asarray(v, dt) for k, v in s.items()}
    print(f"[c1] prec={prec} n={n} T={T}", flush=True)
    sim = sim_for(prm, prec)
    sim.set_state(s)
    sim.set_up(T + 1)
    sim.forward(0)
    xT = sim.get(T, "x", dtype=np.float64)
    sim.set_last_grad(x=2 * (xT - d["target_x"]))
    sim.backward(0)
    if prec == "f32":
        for k in ("x", "v", "F", "C"):
            e = rel_err(sim.get(T, k), d[f"out_{k}"])
            assert e < 2e-4, f"{k}: {e:.3e}"
        for k in ("x", "v", "F", "C"):
            e = rel_err(sim.get_grad(0, k), d[f"out_g{k}"])
            assert e < 1e-3, f"d{k}: {e:.3e}"
    else:
        orc = Oracle("f64")
        layers = orc.rollout(prm, s, T)
        for k in ("x", "v", "F", "C"):
            e = rel_err(sim.get(T, k, dtype=dt), layers[-1][k])
            assert e < 1e-10, f"{k}: {e:.3e}"
        gT = dict(x=2 * (layers[-1]["x"] - d["target_x"]), v=np.zeros((n, 3)), F=np.zeros((n, 3, 3)),
                  C=np.zeros((n, 3, 3)))
        g0, _ = orc.rollout_grad(prm, layers, gT)
        for k in ("x", "v", "F", "C"):
            e = rel_err(sim.get_grad(0, k, dtype=dt), g0[k])
            assert e < 1e-10, f"d{k}: {e:.3e}"
    sim.close()


def test_optimisation_loops_against_reference():
    """run_optimization (OptimizeDefGradControlSequence, CompGraph.cpp:213-368)
    and then optimize_single_timestep (CompGraph.cpp:369-465) on the same
    graph, through the diffmpm_bindings API in FP32, against the reference's
    own run: the end-layer losses and every layer's dFc after each call, and
    the Adam state of the single-timestep layer.  OptimizeSingleTimestep's
    step counter is process-wide (a function-local static), so it starts
    from 0 as in the reference's fresh process."""
    import ctypes as C
    import diffmpm_bindings as db
    from dmpm import _lib as L
    d = load("ref_optimize.npz")
    s = {k[3:]: d[k] for k in d if k.startswith("in_")}
    dims = tuple(int(a) for a in d["dims"])
    pc = db.PointCloud.from_arrays(s["x"], s["m"], s["vol"], s["lam"], s["mu"], v=s["v"], F=s["F"], C=s["C"])
    tg = db.Grid(*dims, float(d["dx"]), [0.0, 0.0, 0.0])
    tg.m = d["target_m"].astype(np.float32)
    cg = db.CompGraph(pc, db.Grid(*dims, float(d["dx"]), [0.0, 0.0, 0.0]), tg)
    opt = db.OptInput()
    opt.num_timesteps = int(d["arg_num_steps"])
    opt.dt = float(d["dt"])
    opt.drag = float(d["drag"])
    opt.f_ext = list(d["f_ext"])
    opt.control_stride = int(d["arg_control_stride"])
    opt.max_gd_iters = int(d["arg_max_gd_iters"])
    opt.max_ls_iters = int(d["arg_max_ls_iters"])
    opt.initial_alpha = float(d["arg_initial_alpha"])
    opt.gd_tol = float(d["arg_gd_tol"])
    opt.smoothing_factor = float(d["smoothing"])
    print(f"[optimize] n={s['x'].shape[0]} layers={opt.num_timesteps}", flush=True)
    cg.run_optimization(opt)
    loss = cg.end_layer_mass_loss()
    assert abs(loss - float(d["out_loss_run"])) <= 2e-4 * abs(float(d["out_loss_run"]))
    for t in range(opt.num_timesteps):
        got = cg.get_point_cloud(t).get_def_grads_morph()
        want = d["out_dFc_run"][t]
        assert np.abs(got - want).max() <= 2e-4 * max(np.abs(d["out_dFc_run"]).max(), 1e-12), t
    assert np.abs(d["out_dFc_run"]).max() > 0
    L.check(L.lib().dmpm_single_timestep_adam_count(None, 1))
    cg.optimize_single_timestep(int(d["single_t"]), max_gd_iters=int(d["single_iters"]),
                                initial_alpha=float(d["single_alpha"]),
                                max_line_search_iters=int(d["single_ls"]))
    cnt = C.c_int()
    L.check(L.lib().dmpm_single_timestep_adam_count(C.byref(cnt), 0))
    assert cnt.value == int(d["single_iters"])  # every iteration of the fixture accepted a step
    loss = cg.end_layer_mass_loss()
    assert abs(loss - float(d["out_loss_single"])) <= 2e-4 * abs(float(d["out_loss_single"]))
    scale = max(np.abs(d["out_dFc_single"]).max(), 1e-12)
    for t in range(opt.num_timesteps):
        got = cg.get_point_cloud(t).get_def_grads_morph()
        assert np.abs(got - d["out_dFc_single"][t]).max() <= 2e-4 * scale, t
    t = int(d["single_t"])
    for k in ("m", "v", "vmax"):
        got = cg._sim.get(t, "adam_" + k)
        want = d[f"out_adam_{k}"]
        assert np.abs(got - want).max() <= 2e-4 * max(np.abs(want).max(), 1e-30), k
