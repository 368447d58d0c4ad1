"""Pins the oracle with what the mathematics fixes, for every part the
reference's golden vectors leave open (the quadratic B-spline, Neo-Hookean
stress, walls, the exact adjoint):
  * B-splines: partition of unity, zero first moment, second moment 1/3
    (cubic) or 1/4 (quadratic) -- which fixes D^{-1} = 3/dx^2 or 4/dx^2 --
    and N' against central differences;
  * SVD/polar against LAPACK (numpy);
  * stress: P(R) = 0 for rotations, P(sI) closed forms, frame indifference
    P(QF) = Q P(F), P = dpsi/dF and H = dP/dF by central differences, and
    symmetry of H;
  * P2G mass and momentum conservation, APIC affine reproduction;
  * the exact adjoint against central differences of multi-step rollouts;
  * the reference adjoint's doubled F term (BackPropagation.cpp:27-28, 140).
"""
import numpy as np
import pytest

from dmpm.scenes import Body, make_scene
from helpers import oracle_params, perturbed_scene, random_adjoint, rel_err
from oracle import Oracle, Params

O64 = Oracle("f64")


@pytest.mark.parametrize("order,second", [(3, 1.0 / 3.0), (2, 0.25)])
def test_bspline_moments(order, second):
    rng = np.random.default_rng(0)
    for x in rng.uniform(-0.5, 0.5, size=50):
        k = np.arange(-3, 4, dtype=np.float64)
        u = k - x
        w, _ = O64.bspline(order, u)
        assert abs(w.sum() - 1) < 1e-14
        assert abs((w * u).sum()) < 1e-14
        assert abs((w * u * u).sum() - second) < 1e-14


@pytest.mark.parametrize("order", [2, 3])
def test_bspline_slope_fd(order):
    u = np.linspace(-2.2, 2.2, 997)
    h = 1e-6
    _, dw = O64.bspline(order, u)
    wp, _ = O64.bspline(order, u + h)
    wm, _ = O64.bspline(order, u - h)
    assert np.abs(dw - (wp - wm) / (2 * h)).max() < 1e-7


def test_svd_against_lapack():
    rng = np.random.default_rng(1)
    F = np.eye(3) + 0.5 * rng.normal(size=(300, 3, 3))
    F[:20] = np.diag([1, 1, -1.0]) @ F[:20]  # some inverted
    U, s, V = O64.svd(F)
    sl = np.linalg.svd(F, compute_uv=False)
    assert np.abs(np.abs(s) - sl).max() < 1e-12
    assert np.all(s[:, 0] >= s[:, 1]) and np.all(s[:, 1] >= np.abs(s[:, 2]) - 1e-14)
    assert np.all(s[:, :2] >= 0)
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-12 and np.abs(np.linalg.det(V) - 1).max() < 1e-12
    assert np.abs(np.einsum("nij,nj,nkj->nik", U, s, V) - F).max() < 1e-12
    assert np.all(np.sign(s[:, 2]) == np.sign(np.linalg.det(F)))


def test_polar_is_closest_rotation():
    rng = np.random.default_rng(2)
    F = np.eye(3) + 0.4 * rng.normal(size=(200, 3, 3))
    F = F[np.linalg.det(F) > 0.05]
    R, S = O64.polar(F)
    Ul, _, Vt = np.linalg.svd(F)
    assert np.abs(R - Ul @ Vt).max() < 1e-11
    assert np.abs(R @ S - F).max() < 1e-12
    assert np.abs(S - np.transpose(S, (0, 2, 1))).max() < 1e-12


def _rotations(n, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    Q[np.linalg.det(Q) < 0] *= -1
    return Q


@pytest.mark.parametrize("model", [0, 1])
def test_stress_closed_forms(model):
    lam, mu = 5769.23, 3846.15
    Q = _rotations(20, 3)
    assert np.abs(O64.pk1(model, Q, lam, mu)).max() < 1e-9
    for s in (0.7, 1.0, 1.3):
        P = O64.pk1(model, s * np.eye(3)[None], lam, mu)[0]
        if model == 0:
            expect = 2 * mu * (s - 1) + lam * (s ** 3 - 1) * s ** 3 / s
        else:
            expect = mu * (s - 1 / s) + lam * 3 * np.log(s) / s
        assert np.abs(P - expect * np.eye(3)).max() < 1e-9 * max(1, abs(expect))
    rng = np.random.default_rng(4)
    F = np.eye(3) + 0.2 * rng.normal(size=(20, 3, 3))
    assert rel_err(O64.pk1(model, Q @ F, lam, mu), Q @ O64.pk1(model, F, lam, mu)) < 1e-12


def test_neo_hookean_kirchhoff_closed_form():
    """P F^T = mu F F^T + (lam log J - mu) I for the Neo-Hookean P, and the
    gradient of <G, P F^T> in closed form (the GPU forms both this way)."""
    lam, mu = 5769.23, 3846.15
    rng = np.random.default_rng(9)
    F = np.eye(3) + 0.3 * rng.normal(size=(40, 3, 3))
    F = F[np.linalg.det(F) > 0.05]
    tau = O64.pk1(1, F, lam, mu) @ np.swapaxes(F, 1, 2)
    J = np.linalg.det(F)
    expect = mu * F @ np.swapaxes(F, 1, 2) + (lam * np.log(J) - mu)[:, None, None] * np.eye(3)
    assert rel_err(tau, expect) < 1e-12
    # its gradient: G^T P + dP(F; G F) = mu (G + G^T) F + lam tr(G) F^-T
    G = rng.normal(size=F.shape)
    Gt = np.swapaxes(G, 1, 2)
    vjp = Gt @ O64.pk1(1, F, lam, mu) + O64.pk1_hvp(1, F, lam, mu, G @ F)
    Fit = np.swapaxes(np.linalg.inv(F), 1, 2)
    closed = mu * (G + Gt) @ F + lam * np.trace(G, axis1=1, axis2=2)[:, None, None] * Fit
    assert rel_err(vjp, closed) < 1e-11


@pytest.mark.parametrize("model", [0, 1])
def test_stress_is_energy_gradient_and_hessian(model):
    lam, mu = 5769.23, 3846.15
    rng = np.random.default_rng(5)
    F = np.eye(3) + 0.2 * rng.normal(size=(30, 3, 3))
    F = F[np.linalg.det(F) > 0.2]
    P = O64.pk1(model, F, lam, mu)
    h = 1e-6
    dF = rng.normal(size=F.shape)
    fd = (O64.energy(model, F + h * dF, lam, mu) - O64.energy(model, F - h * dF, lam, mu)) / (2 * h)
    assert np.abs(fd - (P * dF).sum(axis=(1, 2))).max() < 1e-5 * np.abs(fd).max()
    H = O64.pk1_hvp(model, F, lam, mu, dF)
    fdH = (O64.pk1(model, F + h * dF, lam, mu) - O64.pk1(model, F - h * dF, lam, mu)) / (2 * h)
    assert rel_err(H, fdH) < 1e-6
    dG = rng.normal(size=F.shape)
    a = (dG * H).sum(axis=(1, 2))
    b = (dF * O64.pk1_hvp(model, F, lam, mu, dG)).sum(axis=(1, 2))
    assert np.abs(a - b).max() < 1e-9 * np.abs(a).max()


@pytest.mark.parametrize("order", [2, 3])
def test_p2g_conserves_mass_and_momentum(order):
    sc = make_scene(Body((0.5, 0.5, 0.5), (0.2, 0.2, 0.2)), 500, 16, 3, 1e4, 0.3, velocity="sine")
    s = sc.state()
    rng = np.random.default_rng(6)
    s["F"] = s["F"] + 0.05 * rng.normal(size=s["F"].shape)
    s["C"] = rng.normal(size=s["C"].shape)
    prm = Params(dims=sc.dims, dx=sc.dx, dt=1e-3, drag=0.3, zeta=0.25, order=order, model=1, trust_region=0,
                 wall_cells=0, adjoint=0)
    o = O64.forward_step(prm, s)
    gm, gv = o["grid_m"], o["grid_v"]
    assert abs(gm.sum() - s["m"].sum()) < 1e-12 * s["m"].sum()
    p = gm[:, None] * gv
    expect = (s["m"][:, None] * s["v"] * (1 - prm.dt * prm.drag)).sum(axis=0)
    assert np.abs(p.sum(axis=0) - expect).max() < 1e-10 * np.abs(expect).max()


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("model", [0, 1])
def test_apic_affine_reproduction(order, model):
    """With F = I (zero stress) and an affine velocity field v = A x + b,
    C = A, one step returns v' = v and C' = A exactly (up to rounding)."""
    sc = make_scene(Body((0.5, 0.5, 0.5), (0.2, 0.2, 0.2)), 500, 16, 3, 1e4, 0.3)
    s = sc.state()
    rng = np.random.default_rng(7)
    A = rng.normal(size=(3, 3))
    b = rng.normal(size=3)
    s["v"] = s["x"] @ A.T + b
    s["C"] = np.broadcast_to(A, (sc.n, 3, 3)).copy()
    prm = Params(dims=sc.dims, dx=sc.dx, dt=1e-3, drag=0.0, zeta=0.0, order=order, model=model, trust_region=0,
                 wall_cells=0, adjoint=0)
    o = O64.forward_step(prm, s)
    # nodes reached only by a stencil's far tail can hold 0 < m_i < 1e-12,
    # where the grid update leaves v_i = 0 (ForwardSimulation.cpp:84); their
    # weights are ~1e-11, which bounds the reproduction error
    assert np.abs(o["v"] - s["v"]).max() < 1e-10
    assert np.abs(o["C"] - A).max() < 1e-8


def _fd_check(name, steps, dt, keys=("x", "v", "F", "C", "ctl"), tol=1e-5, n=250):
    sc, s0 = perturbed_scene(n=n, dfc=0.0)
    s0 = dict(s0, dFc=None)
    prm = oracle_params(name, sc, dt)
    rng = np.random.default_rng(8)
    ctl = [1e-3 * rng.normal(size=(sc.n, 3, 3)) for _ in range(steps)]
    gT = random_adjoint(sc.n, seed=12)

    def L(s, c):
        lay = O64.rollout(prm, s, steps, c)
        return sum((lay[-1][k] * gT[k]).sum() for k in gT)

    layers = O64.rollout(prm, s0, steps, ctl)
    adj, gctl = O64.rollout_grad(prm, layers, gT)
    h = 1e-6
    for key in keys:
        d = rng.normal(size=(sc.n, 3) if key in ("x", "v") else (sc.n, 3, 3))
        if key == "ctl":
            cp = [c.copy() for c in ctl]
            cm = [c.copy() for c in ctl]
            cp[1] += h * d
            cm[1] -= h * d
            fd = (L(s0, cp) - L(s0, cm)) / (2 * h)
            an = (gctl[1] * d).sum()
        else:
            sp, sm = dict(s0), dict(s0)
            sp[key] = s0[key] + h * d
            sm[key] = s0[key] - h * d
            fd = (L(sp, ctl) - L(sm, ctl)) / (2 * h)
            an = (adj[key] * d).sum()
        assert abs(fd - an) <= tol * max(abs(fd), 1e-8), f"{name} d{key}: fd {fd:.10e} adjoint {an:.10e}"


@pytest.mark.parametrize("name", ["plain", "reference_exact", "quadratic_fcr", "cubic_nh"])
def test_exact_adjoint_matches_finite_differences(name):
    _fd_check(name, steps=3, dt=1e-3)


def test_reference_adjoint_doubles_the_F_term():
    """With s = 0, zeta = 0, drag = 0, f_ext = 0, no trust region and no
    walls, the reference adjoint differs from the exact one only by
    (I + dt C')^T dL/dF' added a second time (BackPropagation.cpp:140)."""
    sc, s = perturbed_scene(n=300)
    base = dict(dims=sc.dims, dx=sc.dx, dt=1e-3, drag=0.0, smoothing=0.0, zeta=0.0, order=3, model=0,
                trust_region=0, wall_cells=0)
    adj = random_adjoint(sc.n, seed=4)
    ex = O64.backward_step(Params(**base, adjoint=0), s, adj)
    rf = O64.backward_step(Params(**base, adjoint=1), s, adj)
    Cn = O64.forward_step(Params(**base, adjoint=0), s)["C"]
    A1 = np.eye(3) + 1e-3 * Cn
    extra = np.einsum("nji,njk->nik", A1, adj["F"])
    assert rel_err(rf["F"] - ex["F"], extra) < 1e-10
    for k in ("x", "v", "C"):
        assert rel_err(rf[k], ex[k]) < 1e-12, k


def test_exact_adjoint_through_the_scaled_trust_region():
    """Central differences of a 3-step FP64 rollout of the compressive body
    (the reference-physics fixture's inputs, dt = 2e-3) under the exact
    adjoint, where more than 10 % of the particles take the scaled branch of
    the J trust region in the first step: the derivative of s_J through
    tr C' and det(F + dFc) (ForwardSimulation.cpp:136-152) is exercised in
    every input direction, the control dFc included."""
    import os
    from helpers import trust_branches
    d = dict(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_trust_forward.npz"),
                     allow_pickle=False))
    s0 = {k[3:]: np.asarray(d[k], np.float64) for k in d if k.startswith("in_")}
    prm = Params(dims=tuple(int(a) for a in d["dims"]), dx=float(d["dx"]), dt=float(d["dt"]),
                 drag=float(d["drag"]), smoothing=float(d["smoothing"]), f_ext=tuple(float(a) for a in d["f_ext"]),
                 order=3, model=0, trust_region=1, zeta=0.25, jmin=0.6, wall_cells=0, adjoint=0)
    scaled, clamped = trust_branches(prm, s0)
    n = s0["x"].shape[0]
    assert scaled.sum() > 0.1 * n and clamped.sum() > 0.1 * n
    steps = 3
    dfc0 = s0.pop("dFc")
    rng = np.random.default_rng(19)
    ctl = [dfc0] + [1e-3 * rng.normal(size=(n, 3, 3)) for _ in range(steps - 1)]
    gT = random_adjoint(n, seed=23)

    def L(s, c):
        lay = O64.rollout(prm, s, steps, c)
        return sum((lay[-1][k] * gT[k]).sum() for k in gT)

    layers = O64.rollout(prm, dict(s0, dFc=None), steps, ctl)
    adj, gctl = O64.rollout_grad(prm, layers, gT)
    h = 1e-7
    for key in ("x", "v", "F", "C", "ctl0", "ctl1"):
        dd = rng.normal(size=(n, 3) if key in ("x", "v") else (n, 3, 3))
        if key.startswith("ctl"):
            t = int(key[3:])
            cp = [c.copy() for c in ctl]
            cm = [c.copy() for c in ctl]
            cp[t] += h * dd
            cm[t] -= h * dd
            fd = (L(s0, cp) - L(s0, cm)) / (2 * h)
            an = (gctl[t] * dd).sum()
        else:
            sp, sm = dict(s0), dict(s0)
            sp[key] = s0[key] + h * dd
            sm[key] = s0[key] - h * dd
            fd = (L(sp, ctl) - L(sm, ctl)) / (2 * h)
            an = (adj[key] * dd).sum()
        assert abs(fd - an) <= 1e-5 * max(abs(fd), 1e-8), f"d{key}: fd {fd:.10e} adjoint {an:.10e}"
