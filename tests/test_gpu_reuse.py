"""One Simulator set up again and again: with the same particle count (its
freed device buffers are handed out again), with another count (they are
given back and new ones allocated), with more and with fewer layers.  Every
episode must match the oracle, and setting up again with the same sizes must
not make the sim hold more device memory."""
import numpy as np
import pytest

from helpers import oracle_params, perturbed_scene, random_adjoint, rel_err
from oracle import Oracle

pytestmark = pytest.mark.gpu

DT = 1e-3


def _episode(sim, sc, s, T, adj):
    sim.set_state(dict(s, dFc=None))
    sim.set_up(T + 1)
    sim.forward(0)
    sim.set_last_grad(x=adj["x"], v=adj["v"], F=adj["F"], C_=adj["C"])
    sim.backward(0)
    out = {k: sim.get(T, k, dtype=np.float64) for k in ("x", "F")}
    out.update({"g" + k: sim.get_grad(0, k, dtype=np.float64) for k in ("x", "v", "F", "C")})
    return out


def _reference(sc, s, T, adj):
    orc = Oracle("f64")
    prm = oracle_params("plain", sc, DT)
    layers = orc.rollout(prm, dict(s, dFc=None), T, [np.zeros((sc.n, 3, 3))] * T)
    g0, _ = orc.rollout_grad(prm, layers, adj)
    out = {k: layers[T][k] for k in ("x", "F")}
    out.update({"g" + k: g0[k] for k in ("x", "v", "F", "C")})
    return out


@pytest.mark.parametrize("K", [1, 3])
def test_set_up_again_with_other_sizes(K):
    from dmpm import Physics, Simulator
    from helpers import physics_kwargs
    a, sa = perturbed_scene(n=1500, dfc=0.0)
    b, sb = perturbed_scene(n=997, seed=23, dfc=0.0)
    assert a.dims == b.dims and a.dx == b.dx
    sim = Simulator(a.dims, a.dx, a.grid_min, physics=Physics(**physics_kwargs("plain", DT)), precision="f64",
                    checkpoint_interval=K)
    cases = [("n=1500, 5 steps", a, sa, 5), ("same again", a, sa, 5), ("same again", a, sa, 5),
             ("n=997", b, sb, 5), ("n=1500, 7 steps", a, sa, 7), ("n=1500, 3 steps", a, sa, 3),
             ("same again", a, sa, 3), ("same again", a, sa, 3)]
    held = []
    for name, sc, s, T in cases:
        print(f"[reuse] K={K} {name}", flush=True)
        adj = random_adjoint(sc.n, seed=T)
        got = _episode(sim, sc, s, T, adj)
        want = _reference(sc, s, T, adj)
        for k in want:
            e = rel_err(got[k], want[k])
            assert e < 1e-10, f"{name}: {k}: {e:.3e}"
        held.append(sim.memory()[0])
    # the same sizes set up again take their buffers from what the sim holds
    # (the second time, the particle order's new arrays are allocated while
    # the old ones are still in use; from then on nothing is added)
    assert held[2] == held[1], held
    assert held[7] == held[6], held
    sim.close()
