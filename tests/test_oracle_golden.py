"""The oracle against the reference's own outputs (golden vectors written by
tests/golden/make_golden.py from the compiled reference, float32).

Tolerances are relative to each field's largest magnitude.  The reference's
float32 SVD stops at 1024 eps (ImplicitQRSVD.h:694-711), so its rotation R is
only good to ~1e-4, which bounds how closely anything downstream of the
fixed-corotated stress can agree.
"""
import os

import numpy as np
import pytest

from helpers import rel_err
from oracle import Oracle, reference_params

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


def params(d):
    return reference_params(dims=tuple(int(a) for a in d["dims"]), dx=float(d["dx"]), dt=float(d["dt"]),
                            drag=float(d["drag"]), smoothing=float(d["smoothing"]),
                            f_ext=tuple(float(a) for a in d["f_ext"]))


def state(d):
    return {k[3:]: d[k] for k in d if k.startswith("in_")}


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_forward_step_matches_reference(prec):
    d = load("ref_forward_step.npz")
    o = Oracle(prec).forward_step(params(d), state(d))
    tol = {"x": 1e-6, "v": 2e-5, "F": 2e-6, "C": 1e-4, "P": 5e-4, "grid_m": 1e-6, "grid_v": 1e-4}
    for k, t in tol.items():
        e = rel_err(o[k], d[f"out_{k}"])
        assert e < t, f"{k}: {e:.3e}"


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_backward_step_matches_reference(prec):
    d = load("ref_backward_step.npz")
    adj = {k[4:]: d[k] for k in d if k.startswith("adj_")}
    o = Oracle(prec).backward_step(params(d), state(d), adj)
    for k in ("x", "v", "F", "C", "P"):
        e = rel_err(o[k], d[f"out_{k}"])
        assert e < 2e-4, f"d{k}: {e:.3e}"


def test_mass_loss_matches_reference():
    d = load("ref_mass_loss.npz")
    s = state(d)
    loss, gx, gm, gdm = Oracle("f32").mass_loss(params(d), s["x"], s["m"], d["target_m"])
    # the reference accumulates the loss over every node in float
    # (CompGraph.cpp:72-93); the oracle accumulates in double
    assert abs(loss - float(d["out_loss"])) <= 2e-4 * abs(float(d["out_loss"]))
    assert rel_err(gm, d["out_grid_m"]) < 1e-6
    assert rel_err(gdm, d["out_grid_dLdm"]) < 1e-5
    assert rel_err(gx, d["out_gx"]) < 1e-5


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_episode_matches_reference(prec):
    """SetUpCompGraph + 3 forward steps + EndLayerMassLoss + render gradients
    + 3 Back_Timesteps, against the reference's CompGraph."""
    d = load("ref_episode.npz")
    prm = params(d)
    orc = Oracle(prec)
    s = state(d)
    T = int(d["num_layers"]) - 1
    layers = orc.rollout(prm, dict(s, dFc=None), T)
    last = layers[-1]
    for k in ("x", "v", "F", "C"):
        assert rel_err(last[k], d[f"out_{k}"]) < 2e-4, k
    loss, gx, _, _ = orc.mass_loss(prm, last["x"], s["m"], d["target_m"])
    assert abs(loss - float(d["out_loss"])) <= 1e-4 * abs(float(d["out_loss"]))
    gT = dict(x=gx + d["render_gx"], v=np.zeros_like(gx), F=d["render_gF"].astype(gx.dtype),
              C=np.zeros((gx.shape[0], 3, 3), gx.dtype))
    g0, _ = orc.rollout_grad(prm, layers, gT)
    for k in ("x", "v", "F", "C"):
        assert rel_err(g0[k], d[f"out_g{k}"]) < 1e-3, f"d{k}: {rel_err(g0[k], d[f'out_g{k}']):.3e}"


def test_volumes_match_reference():
    d = load("ref_volumes.npz")
    vol, gm = Oracle("f32").volumes(params(d), d["in_x"], d["in_m"])
    assert rel_err(vol, d["out_vol"]) < 1e-5
    assert rel_err(gm, d["out_grid_m"]) < 1e-6


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_adam_matches_reference(prec):
    d = load("ref_adam.npz")
    orc = Oracle(prec)
    gn = float(d["grad_norm"])
    # 1 - beta2^t cancels in float for small t (~1e-5 relative in the
    # reference's float32 run); the float64 oracle does not round there
    tol = 1e-5 if prec == "f32" else 5e-5
    o1 = orc.adam(d["dLdF"], d["dFc"], d["mom"], d["vec"], d["vmax"], 0.5, gn, timestep=3)
    for a, k in zip(o1, ("dFc", "mom", "vec", "vmax")):
        assert rel_err(a, d[f"out1_{k}"]) < tol, k
    o2 = orc.adam(d["dLdF2"], *o1, 1e-3, gn, timestep=4)
    for a, k in zip(o2, ("dFc", "mom", "vec", "vmax")):
        assert rel_err(a, d[f"out2_{k}"]) < tol, k


def test_elasticity_matches_reference():
    d = load("ref_elasticity.npz")
    orc = Oracle("f64")
    lam, mu = float(d["lam"]), float(d["mu"])
    R, S = orc.polar(d["F"])
    assert rel_err(R, d["R"]) < 5e-4
    assert rel_err(S, d["S"]) < 5e-4
    assert rel_err(orc.pk1(0, d["F"], lam, mu), d["P"]) < 5e-4
    assert rel_err(orc.pk1_hvp(0, d["F"], lam, mu, d["dF"]), d["hvp"]) < 5e-4
    assert rel_err(orc.energy(0, d["F"], lam, mu), d["psi"]) < 5e-4


# --- the J trust region (ForwardSimulation.cpp:136-152) on a compressed,
# converging body (tests/golden/make_golden.py compressive_scene, dt = 2e-3)

def test_trust_fixture_exercises_every_branch():
    """The fixture's step must take both trust-region branches on a good
    share of its particles, and det F must differ from det(F + dFc) in the
    branch taken for some of them (so that J0 taken from the wrong matrix
    fails the comparisons below)."""
    from dataclasses import replace
    from helpers import trust_branches
    d = load("ref_trust_forward.npz")
    s = state(d)
    prm = params(d)
    scaled, clamped = trust_branches(prm, s)
    n = s["x"].shape[0]
    assert scaled.sum() > 0.1 * n and clamped.sum() > 0.1 * n, (scaled.sum(), clamped.sum())
    sF, cF = trust_branches(prm, dict(s, dFc=None))
    assert (sF != scaled).sum() + (cF != clamped).sum() > 0.05 * n


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_trust_forward_step_matches_reference(prec):
    d = load("ref_trust_forward.npz")
    o = Oracle(prec).forward_step(params(d), state(d))
    tol = {"x": 1e-6, "v": 2e-5, "F": 1e-5, "C": 1e-4, "P": 5e-4, "grid_m": 1e-6, "grid_v": 1e-4}
    for k, t in tol.items():
        e = rel_err(o[k], d[f"out_{k}"])
        assert e < t, f"{k}: {e:.3e}"


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_trust_backward_step_matches_reference(prec):
    d = load("ref_trust_backward.npz")
    adj = {k[4:]: d[k] for k in d if k.startswith("adj_")}
    o = Oracle(prec).backward_step(params(d), state(d), adj)
    for k in ("x", "v", "F", "C", "P"):
        e = rel_err(o[k], d[f"out_{k}"])
        assert e < 2e-4, f"d{k}: {e:.3e}"


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_trust_episode_matches_reference(prec):
    """SetUpCompGraph + 3 forward steps + EndLayerMassLoss + render gradients
    + 3 Back_Timesteps on the compressive scene, against the reference."""
    d = load("ref_trust_episode.npz")
    prm = params(d)
    orc = Oracle(prec)
    s = state(d)
    T = int(d["num_layers"]) - 1
    layers = orc.rollout(prm, s, T, [s["dFc"]] + [None] * (T - 1))
    last = layers[-1]
    for k in ("x", "v", "F", "C"):
        assert rel_err(last[k], d[f"out_{k}"]) < 2e-4, k
    loss, gx, _, _ = orc.mass_loss(prm, last["x"], s["m"], d["target_m"])
    assert abs(loss - float(d["out_loss"])) <= 1e-4 * abs(float(d["out_loss"]))
    gT = dict(x=gx + d["render_gx"], v=np.zeros_like(gx), F=d["render_gF"].astype(gx.dtype),
              C=np.zeros((gx.shape[0], 3, 3), gx.dtype))
    g0, _ = orc.rollout_grad(prm, layers, gT)
    for k in ("x", "v", "F", "C"):
        assert rel_err(g0[k], d[f"out_g{k}"]) < 1e-3, f"d{k}: {rel_err(g0[k], d[f'out_g{k}']):.3e}"


# --- degenerate inputs of the fixed-corotated stress and its Hessian action

@pytest.mark.parametrize("group,tol,htol", [("regularised", 1e-6, 2e-6), ("repeated", 2e-6, 2e-6),
                                            ("exact", 1e-6, 2e-6), ("near", 3e-2, 1e-1)])
def test_degenerate_elasticity_matches_reference(group, tol, htol):
    """Polar factors, PK1 and the Hessian action on degenerate F against
    the reference: scaled rotations that take the +1e-4 I regularisation of
    (tr S I - S) (Elasticity.cpp:97-100), repeated singular values, exact
    diagonal and permutation inputs (inverted ones with sigma_2 + sigma_3 = 0
    among them: JIXIE's sign and order choices, ImplicitQRSVD.h:852-875), and
    inverted F with sigma_2 + sigma_3 = +-1e-2, +-2e-2.  Near that
    degeneracy the reference's QR sweeps stop once the off-diagonal is below
    1024 eps max(|B|/2, 1) ~ 1.2e-4 (ImplicitQRSVD.h:694-723), which leaves
    its rotation off by about that over the gap sigma_2 - |sigma_3|, ~1e-2
    here, and the Hessian action's rotation term divides by sigma_2 +
    sigma_3 once more; a wrong sign or order convention would be off by
    O(1)."""
    d = load("ref_elasticity_degenerate.npz")
    orc = Oracle("f64")
    lam, mu = float(d["lam"]), float(d["mu"])
    F, dF = d[f"{group}_F"], d[f"{group}_dF"]
    if group in ("regularised", "exact"):
        # the regularisation must actually be taken for the small ones
        S = d[f"{group}_S"].astype(np.float64)
        detA = np.linalg.det(np.trace(S, axis1=1, axis2=2)[:, None, None] * np.eye(3) - S)
        assert (np.abs(detA) < 1e-6).sum() >= 4
    R, S = orc.polar(F)
    for i in range(F.shape[0]):
        assert rel_err(R[i], d[f"{group}_R"][i]) < tol, ("R", i)
        assert rel_err(S[i], d[f"{group}_S"][i]) < tol, ("S", i)
    P = orc.pk1(0, F, lam, mu)
    H = orc.pk1_hvp(0, F, lam, mu, dF)
    for i in range(F.shape[0]):
        assert rel_err(P[i], d[f"{group}_P"][i]) < max(tol, 1e-6), ("P", i)
        assert rel_err(H[i], d[f"{group}_hvp"][i]) < htol, ("hvp", i)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_c1_rollout_matches_reference(prec):
    """5 steps and 5 adjoint steps of the reference's physics on the
    8,192-particle, 32^3-cell body, seeded with the positional-loss gradient,
    against the reference."""
    d = load("ref_c1_rollout.npz")
    prm = params(d)
    orc = Oracle(prec)
    n = d["in_x"].shape[0]
    s = dict(x=d["in_x"], v=d["in_v"], F=np.broadcast_to(np.eye(3), (n, 3, 3)).copy(), C=np.zeros((n, 3, 3)),
             m=d["in_m"], vol=d["in_vol"], lam=d["in_lam"], mu=d["in_mu"])
    T = int(d["steps"])
    layers = orc.rollout(prm, s, T)
    for k in ("x", "v", "F", "C"):
        assert rel_err(layers[-1][k], d[f"out_{k}"]) < 2e-4, k
    gT = dict(x=2 * (layers[-1]["x"] - d["target_x"]), v=np.zeros((n, 3)), F=np.zeros((n, 3, 3)),
              C=np.zeros((n, 3, 3)))
    g0, _ = orc.rollout_grad(prm, layers, gT)
    for k in ("x", "v", "F", "C"):
        e = rel_err(g0[k], d[f"out_g{k}"])
        assert e < 1e-3, f"d{k}: {e:.3e}"
