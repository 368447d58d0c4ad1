"""Shared test inputs and comparison helpers."""
from __future__ import annotations

import numpy as np

from dmpm.scenes import Body, make_scene
from oracle import Params, plain_params, reference_params


def perturbed_scene(n=1500, cells=16, seed=11, amplitude=0.3, fpert=0.05, cpert=0.5, E=1e4, nu=0.3,
                    axes=(0.22, 0.2, 0.18), dfc=0.01):
    """A small body with a sine velocity field, perturbed F and C, and a
    control dFc, so that every term of the step is exercised."""
    sc = make_scene(Body((0.5, 0.5, 0.5), axes), n, cells, seed, E, nu, velocity="sine", amplitude=amplitude)
    s = sc.state()
    rng = np.random.default_rng(seed + 100)
    s["F"] = s["F"] + fpert * rng.normal(size=s["F"].shape)
    s["C"] = cpert * rng.normal(size=s["C"].shape)
    s["dFc"] = dfc * rng.normal(size=s["F"].shape) if dfc else None
    return sc, s


def random_adjoint(n, seed=5, scale=1.0):
    rng = np.random.default_rng(seed)
    return dict(x=scale * rng.normal(size=(n, 3)), v=scale * rng.normal(size=(n, 3)),
                F=scale * rng.normal(size=(n, 3, 3)), C=scale * rng.normal(size=(n, 3, 3)))


PHYSICS = {
    # name: (oracle params factory, dmpm Physics kwargs)
    "plain": dict(order=2, model=1, trust_region=0, zeta=0.0, jmin=0.6, wall_cells=3, adjoint=0, drag=0.0,
                  smoothing=0.0, f_ext=(0.0, 0.0, 0.0)),
    "reference": dict(order=3, model=0, trust_region=1, zeta=0.25, jmin=0.6, wall_cells=0, adjoint=1, drag=0.5,
                      smoothing=0.1, f_ext=(0.0, -9.8, 0.0)),
    "reference_exact": dict(order=3, model=0, trust_region=1, zeta=0.25, jmin=0.6, wall_cells=0, adjoint=0,
                            drag=0.5, smoothing=0.1, f_ext=(0.0, -9.8, 0.0)),
    "quadratic_fcr": dict(order=2, model=0, trust_region=0, zeta=0.0, jmin=0.6, wall_cells=2, adjoint=0,
                          drag=0.2, smoothing=0.0, f_ext=(0.0, 0.0, 1.0)),
    "cubic_nh": dict(order=3, model=1, trust_region=1, zeta=0.1, jmin=0.9, wall_cells=0, adjoint=0, drag=0.0,
                     smoothing=0.05, f_ext=(0.0, 0.0, 0.0)),
}


def oracle_params(name, sc, dt) -> Params:
    kw = dict(PHYSICS[name])
    return Params(dims=sc.dims, dx=sc.dx, grid_min=sc.grid_min, dt=dt, **kw)


def physics_kwargs(name, dt) -> dict:
    kw = dict(PHYSICS[name])
    kw["dt"] = dt
    return kw


def rel_err(a, b, floor=1e-30):
    """max |a-b| / max |b| over the whole array."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), floor))


def trust_branches(prm: Params, s: dict):
    """Which particles take which branch of the J trust region
    (ForwardSimulation.cpp:136-152) in one step from state s: the step's G2P
    C' is computed with the trust region off (FP64 oracle), J0 = max(det(F +
    dFc), 1e-6), s_J = (J_min / J0 - 1) / min(-1e-6, dt tr C').  Returns
    boolean masks (scaled: tr C' < 0 and 0 < s_J < 1; clamped: tr C' < 0
    and s_J <= 0)."""
    from dataclasses import replace
    from oracle import Oracle
    s64 = {k: (None if v is None else np.asarray(v, np.float64)) for k, v in s.items()}
    raw = Oracle("f64").forward_step(replace(prm, trust_region=0), s64)
    Ft = s64["F"] + (s64["dFc"] if s64.get("dFc") is not None else 0.0)
    J0 = np.maximum(np.linalg.det(Ft), 1e-6)
    trC = np.trace(raw["C"], axis1=1, axis2=2)
    sj = (prm.jmin / J0 - 1.0) / np.minimum(-1e-6, prm.dt * trC)
    neg = trC < 0
    return neg & (sj > 0) & (sj < 1), neg & (sj <= 0)
