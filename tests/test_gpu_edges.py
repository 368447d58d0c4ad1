"""Edge cases against the oracle: one particle, counts that leave a partial
wave or a partial block, a body inside the sticky-wall band next to a grid
corner, a box-shaped grid with its origin off zero; and an empty particle
set is refused."""
import numpy as np
import pytest

from helpers import oracle_params, physics_kwargs, random_adjoint, rel_err
from oracle import Oracle

pytestmark = pytest.mark.gpu

DT = 1e-3


def _subset(sc, k, seed=7):
    """The first k particles of scene sc, with perturbed F and C."""
    s = {key: np.asarray(val)[:k].copy() for key, val in sc.state().items()}
    rng = np.random.default_rng(seed)
    s["F"] = s["F"] + 0.05 * rng.normal(size=s["F"].shape)
    s["C"] = 0.5 * rng.normal(size=s["C"].shape)
    return s


def _episode_matches(sc, s, name, T=3):
    from dmpm import Physics, Simulator
    n = s["x"].shape[0]
    adj = random_adjoint(n, seed=3)
    orc = Oracle("f64")
    prm = oracle_params(name, sc, DT)
    layers = orc.rollout(prm, dict(s, dFc=None), T, [np.zeros((n, 3, 3))] * T)
    g0, _ = orc.rollout_grad(prm, layers, adj)
    sim = Simulator(sc.dims, sc.dx, sc.grid_min, physics=Physics(**physics_kwargs(name, DT)), precision="f64")
    sim.set_state(dict(s, dFc=None))
    sim.set_up(T + 1)
    sim.forward(0)
    sim.set_last_grad(x=adj["x"], v=adj["v"], F=adj["F"], C_=adj["C"])
    sim.backward(0)
    for k in ("x", "v", "F", "C"):
        e = rel_err(sim.get(T, k, dtype=np.float64), layers[T][k])
        assert e < 1e-10, f"{k}: rel err {e:.3e}"
        e = rel_err(sim.get_grad(0, k, dtype=np.float64), g0[k])
        assert e < 1e-10, f"d{k}: rel err {e:.3e}"
    sim.close()


@pytest.mark.parametrize("n", [1, 65, 257])
@pytest.mark.parametrize("name", ["plain", "reference"])
def test_small_and_ragged_counts(name, n):
    from dmpm.scenes import Body, make_scene
    sc = make_scene(Body((0.5, 0.5, 0.5), (0.2, 0.2, 0.2)), 1500, 16, 11, 1e4, 0.3, velocity="sine", amplitude=0.3)
    print(f"[edges] physics={name} n={n}", flush=True)
    _episode_matches(sc, _subset(sc, n), name)


@pytest.mark.parametrize("name", ["plain", "reference", "quadratic_fcr"])
def test_body_in_the_wall_band(name):
    """A body 0.12-0.28 from three faces of the unit box (16 cells): part of
    it lies within the sticky-wall cells, and its windows reach the grid's
    first nodes."""
    from dmpm.scenes import Body, make_scene
    sc = make_scene(Body((0.2, 0.8, 0.2), (0.08, 0.08, 0.08)), 800, 16, 5, 1e4, 0.3, velocity="sine", amplitude=0.5)
    assert sc.x.min() > 0.1 and sc.x.max() < 0.9
    print(f"[edges] wall band physics={name} n={sc.n}", flush=True)
    _episode_matches(sc, _subset(sc, sc.n), name)


@pytest.mark.parametrize("name", ["plain", "reference"])
def test_box_shaped_grid_with_offset_origin(name):
    """A grid of 25 x 13 x 19 nodes (node index i*ny*nz + j*nz + k with
    unequal extents) whose origin is not at zero: the window strides and the
    node indexing must still agree with the oracle."""
    import dataclasses
    from dmpm.scenes import Body, make_scene
    base = make_scene(Body((0.5, 0.5, 0.5), (0.3, 0.14, 0.2)), 1200, 20, 17, 1e4, 0.3, velocity="sine",
                      amplitude=0.4)
    shift = np.array([-0.7, 0.35, 0.1])
    dx = base.dx
    sc = dataclasses.replace(base, x=base.x + shift, dims=(25, 13, 19),
                             grid_min=tuple(float(c) for c in shift + np.array([0.5 - 12 * dx, 0.5 - 6 * dx, 0.5 - 9 * dx])))
    lo = (sc.x - np.array(sc.grid_min)) / dx
    assert lo.min() > 2.0 and (lo.max(axis=0) < np.array(sc.dims) - 3.0).all()
    print(f"[edges] box grid physics={name} dims={sc.dims} n={sc.n}", flush=True)
    _episode_matches(sc, _subset(sc, sc.n), name)


def test_empty_particle_set_is_refused():
    from dmpm import Physics, Simulator
    from dmpm._lib import DmpmError
    sim = Simulator((17, 17, 17), 1 / 16, (0.0, 0.0, 0.0), physics=Physics.plain(dt=DT), precision="f32")
    z = np.zeros((0, 3))
    with pytest.raises(DmpmError):
        sim.set_state(dict(x=z, v=z, F=np.zeros((0, 3, 3)), C=np.zeros((0, 3, 3)), m=np.zeros(0), vol=np.zeros(0),
                           lam=np.zeros(0), mu=np.zeros(0)))
    sim.close()


@pytest.mark.parametrize("prec,nodes", [("f32", 520), ("f64", 412)])
def test_grid_byte_offsets_beyond_2gib(prec, nodes):
    """A body near the far corner of a grid whose values reach past 2 GiB
    (520³ nodes × 16 B in FP32, 412³ × 32 B in FP64): the kernels' 32-bit
    byte offsets into the grid exceed INT_MAX there.  The same particles on a
    small grid placed around them must give the same episode."""
    from dmpm import Physics, Simulator
    from dmpm.scenes import Body, make_scene
    dx = 1.0 / (nodes - 1)
    c = (nodes - 12) * dx
    base = make_scene(Body((0.5, 0.5, 0.5), (0.2, 0.2, 0.2)), 600, 20, 29, 1e4, 0.3, velocity="sine",
                      amplitude=0.4)
    s = _subset(base, base.n)
    shrink = 4 * dx / 0.2
    s["x"] = c + (s["x"] - 0.5) * shrink  # radius 4 cells, centred 11 cells from the far faces
    s["v"] = s["v"] * shrink              # the same velocity gradients as the unshrunk body
    s["m"] = s["m"] * shrink ** 3         # and the same density
    s["vol"] = s["vol"] * shrink ** 3
    big = ((nodes,) * 3, (0.0, 0.0, 0.0))
    small = ((24,) * 3, tuple(c - 12 * dx for _ in range(3)))
    assert (nodes ** 3) * 4 * (4 if prec == "f32" else 8) > 2 ** 31
    g_last = nodes ** 3 - 1
    assert (np.floor(s["x"] / dx).max() * nodes * nodes) * 4 * (4 if prec == "f32" else 8) > 2 ** 31
    assert g_last * 4 * (4 if prec == "f32" else 8) < 2 ** 32
    T = 3
    adj = random_adjoint(s["x"].shape[0], seed=4)
    dt = np.float64 if prec == "f64" else np.float32
    out = []
    for dims, mn in (big, small):
        print(f"[edges] byte offsets prec={prec} dims={dims}", flush=True)
        # dt within the CFL bound of the fine grid (sqrt(E / rho) dt / dx = 0.16)
        sim = Simulator(dims, dx, mn, physics=Physics(**physics_kwargs("plain", 1e-4)), precision=prec)
        sim.set_state(dict(s, dFc=None))
        sim.set_up(T + 1)
        sim.forward(0)
        sim.set_last_grad(x=adj["x"], v=adj["v"], F=adj["F"], C_=adj["C"])
        sim.backward(0)
        r = {k: sim.get(T, k, dtype=dt) for k in ("x", "v", "F", "C")}
        r.update({"g" + k: sim.get_grad(0, k, dtype=dt) for k in ("x", "v", "F", "C")})
        out.append(r)
        sim.close()
    # relative to each field's largest magnitude (positions: their absolute
    # coordinates, about 1); the two grids round the particles' cell
    # coordinates differently (about 508 against 12 cells), which FP32 shows
    tol_s, tol_g = (1e-10, 1e-10) if prec == "f64" else (1e-4, 1e-3)
    for k in out[1]:
        e = rel_err(out[0][k], out[1][k])
        assert e < (tol_g if k.startswith("g") else tol_s), f"{k}: rel err {e:.3e}"


def test_particle_plane_offsets_beyond_2gib():
    """12,000,000 FP64 particles: a layer's later planes start past 2 GiB, so
    the kernels' plane offsets (particle plane p at p * ld * 8 bytes) exceed
    INT_MAX.  A small cluster far from the bulk of them must step exactly as
    it does alone."""
    from dmpm import Physics, Simulator
    n_bulk, n_small, nodes = 12_000_000, 1_000, 129
    dx = 1.0 / (nodes - 1)
    assert 23 * ((n_bulk + n_small + 63) // 64 * 64) * 8 > 2 ** 31
    rng = np.random.default_rng(41)
    xb = rng.uniform(0.1, 0.6, size=(n_bulk, 3))
    xs = 0.8 + 0.05 * rng.normal(size=(n_small, 3)).clip(-1.0, 1.0)
    def state(x, seed):
        r = np.random.default_rng(seed)
        k = x.shape[0]
        F = np.eye(3) + 0.02 * r.normal(size=(k, 3, 3))
        return dict(x=x, v=0.1 * r.normal(size=(k, 3)), F=F, C=0.5 * r.normal(size=(k, 3, 3)),
                    m=np.full(k, 1e3 * dx ** 3 / 8), vol=np.full(k, dx ** 3 / 8),
                    lam=np.full(k, 5e3), mu=np.full(k, 3e3))
    bulk, small = state(xb, 1), state(xs, 2)
    both = {k: np.concatenate([bulk[k], small[k]]) for k in bulk}
    del bulk, xb
    T = 2
    adj = random_adjoint(n_small, seed=6)
    out = []
    for name, s in (("bulk + cluster", both), ("cluster alone", small)):
        print(f"[edges] plane offsets: {name}, n={s['x'].shape[0]}", flush=True)
        sim = Simulator((nodes,) * 3, dx, (0.0, 0.0, 0.0), physics=Physics(**physics_kwargs("plain", 1e-4)),
                        precision="f64")
        sim.set_state(s)
        sim.set_up(T + 1)
        sim.forward(0)
        n = s["x"].shape[0]
        g = {k: np.zeros((n,) + a.shape[1:]) for k, a in adj.items()}
        for k in g:
            g[k][n - n_small:] = adj[k]
        sim.set_last_grad(x=g["x"], v=g["v"], F=g["F"], C_=g["C"])
        sim.backward(0)
        r = {k: sim.get(T, k, dtype=np.float64)[n - n_small:] for k in ("x", "v", "F", "C")}
        r.update({"g" + k: sim.get_grad(0, k, dtype=np.float64)[n - n_small:] for k in ("x", "v", "F", "C")})
        out.append(r)
        sim.close()
        del g
    for k in out[1]:
        e = rel_err(out[0][k], out[1][k])
        assert e < 1e-10, f"{k}: rel err {e:.3e}"
