"""hipGraph capture of whole episodes (one branch per instance): every replay
must reproduce the eager episode, with the program's own eager work and
synchronisations between replays."""
import numpy as np
import pytest

from dmpm.scenes import Body, make_scene
from helpers import perturbed_scene, physics_kwargs, rel_err

pytestmark = pytest.mark.gpu


def _instance(n, cells, seed, axes, E, nu, prec="f64", T=6, K=2):
    from dmpm import Physics, Simulator
    sc = make_scene(Body((0.5, 0.5, 0.5), axes), n, cells, seed, E, nu, velocity="sine", amplitude=0.3)
    sim = Simulator(sc.dims, sc.dx, physics=Physics.plain(dt=5e-4), precision=prec, checkpoint_interval=K)
    sim.set_state(sc.state())
    sim.set_up(T + 1)
    rng = np.random.default_rng(seed)
    sim.set_control(1, 1e-3 * rng.normal(size=(sc.n, 3, 3)))
    sim.request_control_grad(1)
    sim.set_last_grad(x=rng.normal(size=(sc.n, 3)), F=1e-2 * rng.normal(size=(sc.n, 3, 3)))
    return sim, T


def _results(sim, T):
    out = {k: sim.get(T, k, dtype=np.float64) for k in ("x", "F")}
    out.update({"g" + k: sim.get_grad(0, k, dtype=np.float64) for k in ("x", "v", "F")})
    out["gctl1"] = sim.get_control_grad(1, dtype=np.float64)
    return out


def test_graph_replays_match_eager():
    from dmpm import EpisodeGraph
    a, T = _instance(1500, 16, 1, (0.2, 0.22, 0.18), 1e4, 0.3)
    b, _ = _instance(2600, 24, 2, (0.25, 0.2, 0.15), 3e4, 0.4)
    c, _ = _instance(900, 12, 3, (0.2, 0.2, 0.2), 1e4, 0.3)
    print("[graph] eager episodes", flush=True)
    eager = []
    for s in (a, b):
        s.forward(0, keep_controls=True)
        s.backward(0)
        eager.append(_results(s, T))
    g = EpisodeGraph([a, b], start=0, stop=0)
    for rep in range(3):
        print(f"[graph] replay {rep}", flush=True)
        g.launch()
        g.synchronize()
        for s, e in zip((a, b), eager):
            r = _results(s, T)
            for k in e:
                assert rel_err(r[k], e[k]) < 1e-12, (rep, k)
        # eager work of the program between replays
        c.forward(0, keep_controls=True)
        c.backward(0)
        c.synchronize()
    g.close()
    for s in (a, b, c):
        s.close()


def test_sort_refused_while_a_graph_holds_the_sim():
    """dmpm_sort_particles moves the sim's buffers: it is refused while a
    captured episode graph of the sim is alive, and allowed once it is
    destroyed."""
    from dmpm import EpisodeGraph, Physics, Simulator
    from dmpm._lib import DmpmError
    sc, s = perturbed_scene(n=800, dfc=0.0)
    print(f"[sort_refused] n={sc.n}", flush=True)
    sim = Simulator(sc.dims, sc.dx, sc.grid_min, physics=Physics(**physics_kwargs("plain", 1e-3)), precision="f32")
    sim.set_state(dict(s, dFc=None))
    sim.set_up(4)
    sim.forward(0)
    g = EpisodeGraph([sim], start=0, stop=0)
    with pytest.raises(DmpmError):
        sim.sort_particles(2)
    g.close()
    sim.sort_particles(2)
    sim.close()


def test_graph_inputs_between_replays_and_refusals():
    """Between replays a capture takes new controls (on layers that had one),
    a new last-layer seed and new render gradients, each updated in place; a
    replay then matches an eager episode of a fresh sim on the same inputs.
    Calls that would reallocate the captured buffers or change the captured
    parameters are refused while the graph lives, and a graph whose sim was
    destroyed refuses to launch."""
    from dmpm import EpisodeGraph
    from dmpm._lib import DmpmError
    import ctypes as C
    from dmpm import _lib as L

    a, T = _instance(1500, 16, 1, (0.2, 0.22, 0.18), 1e4, 0.3)
    n = a.n
    rng = np.random.default_rng(77)
    rgF0 = 1e-3 * rng.normal(size=(n, 3, 3))
    rgx0 = 1e-2 * rng.normal(size=(n, 3))

    def set_rg(sim, gF, gx):
        gF = np.ascontiguousarray(gF)
        gx = np.ascontiguousarray(gx)
        L.check(L.lib().dmpm_set_render_gradients(sim.handle, n, gF.ctypes.data_as(C.c_void_p),
                                                   gx.ctypes.data_as(C.c_void_p), L.DTYPE_F64, 0))

    set_rg(a, rgF0, rgx0)
    print("[graph_inputs] capture", flush=True)
    g = EpisodeGraph([a], start=0, stop=0)
    # new inputs, updated in place
    ctl1 = 2e-3 * rng.normal(size=(n, 3, 3))
    seed_x = rng.normal(size=(n, 3))
    rgF1 = 1e-3 * rng.normal(size=(n, 3, 3))
    rgx1 = 1e-2 * rng.normal(size=(n, 3))
    a.set_control(1, ctl1)
    a.set_last_grad(x=seed_x)
    set_rg(a, rgF1, rgx1)
    # the eager episode of a fresh sim with the same inputs
    b, _ = _instance(1500, 16, 1, (0.2, 0.22, 0.18), 1e4, 0.3)
    b.set_control(1, ctl1)
    b.set_last_grad(x=seed_x)
    set_rg(b, rgF1, rgx1)
    b.forward(0, keep_controls=True)
    b.backward(0)
    want = _results(b, T)
    for rep in range(2):
        print(f"[graph_inputs] replay {rep}", flush=True)
        g.launch()
        g.synchronize()
        got = _results(a, T)
        for k in want:
            assert rel_err(got[k], want[k]) < 1e-12, (rep, k)
        # eager work on the captured sim itself between replays
        a.forward(0, keep_controls=True)
        a.backward(0)
        a.set_last_grad(x=seed_x)
        a.synchronize()
    # a control cleared between replays is zeroed in place
    a.set_control(1, None)
    b.set_control(1, None)
    b.set_last_grad(x=seed_x)
    b.forward(0, keep_controls=True)
    b.backward(0)
    want = _results(b, T)
    g.launch()
    g.synchronize()
    got = _results(a, T)
    for k in ("x", "F", "gx", "gv", "gF"):
        assert rel_err(got[k], want[k]) < 1e-12, k
    # refused while the graph lives
    s = a.get(0, "x", dtype=np.float64)
    for call in (lambda: a.set_up(T + 1),
                 lambda: a.set_particles(s, s, np.zeros((n, 3, 3)), np.zeros((n, 3, 3)), np.ones(n), np.ones(n),
                                         np.ones(n), np.ones(n)),
                 lambda: a.set_control(3, ctl1),
                 lambda: a.set_sim_params(1e-3, 0.0, 0.0),
                 lambda: a.sort_particles(1)):
        with pytest.raises(DmpmError):
            call()
    g.close()
    a.set_up(T + 1)  # allowed again
    # a graph whose sim is gone refuses to launch
    c, _ = _instance(900, 12, 3, (0.2, 0.2, 0.2), 1e4, 0.3)
    g2 = EpisodeGraph([c], start=0, stop=0)
    c.close()
    with pytest.raises(DmpmError):
        g2.launch()
    g2.close()
    for sim in (a, b):
        sim.close()
