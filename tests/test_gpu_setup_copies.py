"""Set-up makes every stored layer a copy of layer 0 (CompGraph.cpp:30-40).
The copies are made when a layer is first read, so every path that reads a
layer before the forward pass writes it must see layer 0: the getters, a
forward pass from a later layer, the end-layer loss, the re-sorts and a
captured graph.  Copies compare bitwise; results of the scatters (float
atomics, summed in any order) to 1e-12 in FP64."""
import numpy as np
import pytest

from helpers import perturbed_scene, physics_kwargs, rel_err

pytestmark = pytest.mark.gpu

DT = 1e-3
T = 9


def _sim(K, prec="f64"):
    from dmpm import Physics, Simulator
    sc, s = perturbed_scene(dfc=0.0)
    sim = Simulator(sc.dims, sc.dx, sc.grid_min, physics=Physics(**physics_kwargs("plain", DT)), precision=prec,
                    checkpoint_interval=K)
    sim.set_state(dict(s, dFc=None))
    sim.set_up(T + 1)
    return sim, sc, s


@pytest.mark.parametrize("K", [1, 3])
def test_layers_read_before_the_forward_pass_are_layer_0(K):
    sim, sc, s = _sim(K)
    print(f"[setup_copies] getters K={K} n={sc.n}", flush=True)
    x0, F0, v0 = (sim.get(0, k, dtype=np.float64) for k in ("x", "F", "v"))
    # stored layers (with K = 3 the others are recomputed from the checkpoint below)
    for t in ((T, 3, 6, 1) if K == 1 else (T, 3, 6)):
        np.testing.assert_array_equal(sim.get(t, "x", dtype=np.float64), x0)
        np.testing.assert_array_equal(sim.get(t, "F", dtype=np.float64), F0)
        np.testing.assert_array_equal(sim.get(t, "v", dtype=np.float64), v0)
    sim.close()


@pytest.mark.parametrize("K", [1, 3])
def test_forward_from_a_later_layer_starts_from_layer_0(K):
    """forward(3) right after set-up steps from the copy of layer 0 held by
    layer 3: layer T must equal layer T - 3 of a forward pass from 0."""
    a, sc, s = _sim(K)
    print(f"[setup_copies] forward from 3, K={K}", flush=True)
    a.forward(3)
    b, _, _ = _sim(K)
    b.forward(0)
    for k in ("x", "v", "F", "C"):
        assert rel_err(a.get(T, k, dtype=np.float64), b.get(T - 3, k, dtype=np.float64)) < 1e-12, k
    # the layers below the start keep their copies
    np.testing.assert_array_equal(a.get(1 if K == 1 else 3, "x", dtype=np.float64), a.get(0, "x", dtype=np.float64))
    a.close()
    b.close()


def test_loss_sort_and_capture_before_the_forward_pass():
    from dmpm import EpisodeGraph
    sim, sc, s = _sim(1)
    print("[setup_copies] loss / sorts / capture", flush=True)
    sim.set_target_grid_mass(np.full(int(np.prod(sc.dims)), 1e-4))
    ref, _, _ = _sim(1)
    ref.set_target_grid_mass(np.full(int(np.prod(sc.dims)), 1e-4))
    # the end-layer loss of a set-up sim is layer 0's
    ref0 = ref.mass_loss()
    assert abs(sim.mass_loss() - ref0) <= 1e-12 * abs(ref0)
    # re-sorts keep the copies (one order, then segment orders)
    sim.sort_particles(T // 2)
    np.testing.assert_array_equal(sim.get(5, "x", dtype=np.float64), sim.get(0, "x", dtype=np.float64))
    sim.sort_segments(3)
    np.testing.assert_array_equal(sim.get(7, "F", dtype=np.float64), sim.get(0, "F", dtype=np.float64))
    sim.close()
    # a graph captured right after set-up replays the eager episode
    sim, _, _ = _sim(1)
    rng = np.random.default_rng(3)
    gx = rng.normal(size=(sc.n, 3))
    sim.set_last_grad(x=gx)
    ref.set_last_grad(x=gx)
    g = EpisodeGraph([sim], start=0, stop=0)
    g.launch()
    g.synchronize()
    ref.forward(0, keep_controls=True)
    ref.backward(0)
    for k in ("x", "F"):
        assert rel_err(sim.get(T, k, dtype=np.float64), ref.get(T, k, dtype=np.float64)) < 1e-12, k
    assert rel_err(sim.get_grad(0, "x", dtype=np.float64), ref.get_grad(0, "x", dtype=np.float64)) < 1e-12
    g.close()
    sim.close()
    ref.close()
