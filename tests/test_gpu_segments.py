"""Segment storage orders (dmpm_sort_segments): every per-particle array of a
segment's layers stored in an order of its own, the forward pass stepping
into a new order through a G2P, a permutation and a P2G, the adjoint
permuting its carry back.  States, gradients and control gradients must
stay those of the one-order episode (FP64, up to summation order), for the
whole tape and checkpointed, for the physics variants whose adjoint reads
the next layer and the one that recomputes it; a captured graph must replay
the eager episode; calls that need one order return the sim to it."""
import numpy as np
import pytest

from helpers import perturbed_scene, physics_kwargs, random_adjoint, rel_err

pytestmark = pytest.mark.gpu

DT = 1e-3
T = 12
# inside segments and on their first layers (5 and 10 with three segments,
# 6 with K = 3, 4 and 8 with four), where the forward step into the layer
# stages its control in the previous order
CTL_LAYERS = (2, 4, 5, 6, 8, 10, 11)


def _sim(name, K, **kw):
    from dmpm import Physics, Simulator
    sc, s = perturbed_scene(dfc=0.0)
    # a drift of about a cell every few steps, so that the segments' orders
    # differ from each other and from the uploaded one
    s["v"] = s["v"] + np.array([8.0, -5.0, 3.0])
    sim = Simulator(sc.dims, sc.dx, sc.grid_min, physics=Physics(**physics_kwargs(name, DT)), precision="f64",
                    checkpoint_interval=K, **kw)
    sim.set_state(dict(s, dFc=None))
    sim.set_up(T + 1)
    rng = np.random.default_rng(8)
    for t in CTL_LAYERS:
        sim.set_control(t, 0.01 * rng.normal(size=(sc.n, 3, 3)))
        sim.request_control_grad(t)
    return sim, sc


def _episode(sim, sc):
    adj = random_adjoint(sc.n, seed=3)
    sim.forward(0, keep_controls=True)
    sim.set_last_grad(x=adj["x"], v=adj["v"], F=adj["F"], C_=adj["C"])
    sim.backward(0)
    out = {k: sim.get(T, k, dtype=np.float64) for k in ("x", "v", "F", "C")}
    out.update({"g" + k: sim.get_grad(0, k, dtype=np.float64) for k in ("x", "v", "F", "C")})
    for t in CTL_LAYERS:
        out[f"gctl{t}"] = sim.get_control_grad(t, dtype=np.float64)
        out[f"ctl{t}"] = sim.get(t, "dFc", dtype=np.float64)
    for t in (4, 6, 9):
        out[f"x{t}"] = sim.get(t, "x", dtype=np.float64)
        out[f"m{t}"] = sim.get(t, "m", dtype=np.float64)
    out["loss"] = np.array([sim.mass_loss()])
    return out


def _same(a, b, tol, what):
    for k in b:
        e = rel_err(a[k], b[k])
        assert e < tol, f"{what}: {k}: {e:.3e}"


@pytest.mark.parametrize("name,K,kw", [("plain", 1, {}), ("plain", 3, {}), ("reference", 1, {}),
                                       ("reference_exact", 3, {}),
                                       # the unfused kernels (the crossing's G2P, permutation, P2G)
                                       ("plain", 1, {"unfused": True}),
                                       # grids recomputed in the backward pass, not kept
                                       ("plain", 1, {"recompute_grids": True}),
                                       # the adjoint recomputing layer t+1 (nothing permuted for it)
                                       ("reference", 1, {"adjoint_recompute": True}),
                                       # most blocks cut short
                                       ("plain", 3, {"split_nodes": 40})])
def test_segment_orders_keep_results(name, K, kw):
    sim, sc = _sim(name, K, **kw)
    print(f"[segments] {name} K={K} {kw} T={T} n={sc.n}", flush=True)
    sim.set_target_grid_mass(np.full(int(np.prod(sc.dims)), 1e-4))
    one = _episode(sim, sc)
    sim.sort_segments(3)
    # the reorder itself changes no value
    now = {f"x{t}": sim.get(t, "x", dtype=np.float64) for t in (4, 6, 9)}
    now.update({k: sim.get(T, k, dtype=np.float64) for k in ("x", "F")})
    _same(now, {k: one[k] for k in now}, 1e-15, "right after the reorder")
    seg = _episode(sim, sc)
    _same(seg, one, 1e-10, "segment orders")
    seg2 = _episode(sim, sc)
    _same(seg2, one, 1e-10, "segment orders, second episode")
    # one order again (set_up), and the segment orders once more
    sim.set_up(T + 1)
    rng = np.random.default_rng(8)
    for t in CTL_LAYERS:
        sim.set_control(t, 0.01 * rng.normal(size=(sc.n, 3, 3)))
        sim.request_control_grad(t)
    _same(_episode(sim, sc), one, 1e-10, "after set_up")
    sim.sort_segments(2)
    _same(_episode(sim, sc), one, 1e-10, "two segments")
    sim.sort_particles(T // 2)
    _same(_episode(sim, sc), one, 1e-10, "one order by dmpm_sort_particles")
    sim.close()


def test_segment_orders_controls_set_after_the_reorder():
    """Controls, render gradients and the last layer's seed given after the
    reorder land in the order of their own layer."""
    sim, sc = _sim("plain", 1)
    rng = np.random.default_rng(12)
    ctl = {t: 0.02 * rng.normal(size=(sc.n, 3, 3)) for t in CTL_LAYERS}
    rF, rx = 1e-3 * rng.normal(size=(sc.n, 3, 3)), 1e-3 * rng.normal(size=(sc.n, 3))

    def run():
        import ctypes as C
        from dmpm import _lib as L
        for t in CTL_LAYERS:
            sim.set_control(t, ctl[t])
        aF, ax = np.ascontiguousarray(rF), np.ascontiguousarray(rx)
        L.check(L.lib().dmpm_set_render_gradients(sim.handle, sc.n, aF.ctypes.data_as(C.c_void_p),
                                                  ax.ctypes.data_as(C.c_void_p), L.DTYPE_F64, 0))
        return _episode(sim, sc)

    one = run()
    sim.sort_segments(3)
    seg = run()
    _same(seg, one, 1e-10, "controls after the reorder")
    sim.close()


def test_segment_orders_graph_replays_match_eager():
    from dmpm import EpisodeGraph
    sim, sc = _sim("plain", 1)
    adj = random_adjoint(sc.n, seed=3)
    sim.forward(0, keep_controls=True)
    sim.sort_segments(4)
    sim.set_last_grad(x=adj["x"], F=adj["F"])
    sim.forward(0, keep_controls=True)
    sim.backward(0)
    keys = ("x", "v", "F")
    eager = {k: sim.get(T, k, dtype=np.float64) for k in keys}
    eager.update({"g" + k: sim.get_grad(0, k, dtype=np.float64) for k in keys})
    for t in CTL_LAYERS:
        eager[f"gctl{t}"] = sim.get_control_grad(t, dtype=np.float64)
    g = EpisodeGraph([sim], start=0, stop=0)
    for rep in range(3):
        print(f"[segments] graph replay {rep}", flush=True)
        g.launch()
        g.synchronize()
        r = {k: sim.get(T, k, dtype=np.float64) for k in keys}
        r.update({"g" + k: sim.get_grad(0, k, dtype=np.float64) for k in keys})
        for t in CTL_LAYERS:
            r[f"gctl{t}"] = sim.get_control_grad(t, dtype=np.float64)
        _same(r, eager, 1e-10, f"replay {rep}")
        # eager work between replays
        sim.get(5, "x", dtype=np.float64)
    # the captured graph holds the sim's buffers: no reorder meanwhile
    with pytest.raises(Exception):
        sim.sort_segments(2)
    g.close()
    sim.close()


def test_segment_orders_refused_in_deterministic_mode():
    from dmpm._lib import DmpmError
    sim, sc = _sim("plain", 1, deterministic=True)
    sim.forward(0, keep_controls=True)
    with pytest.raises(DmpmError):
        sim.sort_segments(3)
    sim.sort_particles(T // 2)  # one order is fine
    sim.close()


def test_tune_segments_leaves_the_fastest_order_and_the_results():
    from dmpm.autotune import segment_candidates, tune_segments
    sim, sc = _sim("plain", 1)
    one = _episode(sim, sc)
    cands = segment_candidates(T + 1, 1, lengths=(0, 6, 3))
    assert cands == [1, 3, 5]
    info = tune_segments(sim, cands)
    assert info["segments"] in cands and len(info["table"]) == 3
    assert info["segments"] == min(info["table"], key=lambda r: r["ms"])["segments"]
    _same(_episode(sim, sc), one, 1e-10, f"after tuning ({info['segments']} segments)")
    sim.close()
