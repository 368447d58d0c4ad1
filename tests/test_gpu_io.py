"""Layers written to disk and read back into a new sim (both formats), the
stress field P against the oracle's PK1, and the gradient and Adam fields."""
import numpy as np
import pytest

from helpers import perturbed_scene, rel_err
from oracle import Oracle

pytestmark = pytest.mark.gpu

STATE = ("x", "v", "F", "C", "m", "vol", "lam", "mu", "dFc")


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_layer_files_round_trip(tmp_path, prec):
    from dmpm import Physics, Simulator
    from dmpm.io import read_file
    sc, s = perturbed_scene(n=900)
    print(f"[io] precision={prec} n={sc.n}", flush=True)
    mk = lambda: Simulator(sc.dims, sc.dx, physics=Physics.plain(dt=1e-3), precision=prec,  # noqa: E731
                           checkpoint_interval=2)
    sim = mk()
    sim.set_state(s)
    sim.set_up(5)
    rng = np.random.default_rng(4)
    adam = [1e-3 * rng.normal(size=(sc.n, 3, 3)) for _ in range(3)]
    adam[1], adam[2] = np.abs(adam[1]), np.abs(adam[1]) + np.abs(adam[2])
    sim.set_adam_state(0, *adam)
    sim.set_control(3, 1e-3 * rng.normal(size=(sc.n, 3, 3)))
    sim.forward(0, keep_controls=True)
    gx = rng.normal(size=(sc.n, 3))
    sim.set_last_grad(x=gx)
    sim.backward(0)
    exact = prec == "f64"
    for layer in (0, 3, 4):
        for fmt in ("native", "cereal"):
            path = tmp_path / f"l{layer}.{fmt}"
            sim.write_layer(layer, str(path), fmt)
            sim2 = mk()
            sim2.read_particles(str(path), fmt)
            assert sim2.n == sc.n
            sim2.set_up(1)
            tol = 0 if (fmt == "native" or not exact) else 1e-7
            for k in STATE:
                a, b = sim2.get(0, k, dtype=np.float64), sim.get(layer, k, dtype=np.float64)
                assert rel_err(a, b) <= tol, (layer, fmt, k)
            for k, want in zip(("adam_m", "adam_v", "adam_vmax"), adam):
                got = sim2.get(0, k, dtype=np.float64)
                if layer == 0:
                    assert rel_err(got, want) <= (1e-7 if (fmt == "cereal" or not exact) else 1e-15), (fmt, k)
                else:
                    assert not got.any(), (layer, fmt, k)
            sim2.close()
    # the stress field against the oracle's PK1 of F + dFc (Neo-Hookean)
    Ft = sim.get(3, "F_total", dtype=np.float64)
    P_ref = Oracle("f64").pk1(1, Ft, s["lam"], s["mu"])
    assert rel_err(sim.get(3, "P", dtype=np.float64), P_ref) < (1e-12 if exact else 1e-5)
    f = read_file(str(tmp_path / "l3.cereal"), "cereal", dtype=np.float64)
    assert rel_err(f["P"], P_ref) < (1e-6 if exact else 1e-5)
    # gradients: the last layer holds the seed, the control layer of the last
    # backward pass (0) its adjoint; other layers are written as zeros
    f4 = read_file(str(tmp_path / "l4.native"), "native", dtype=np.float64)
    assert rel_err(f4["dLdx"], gx) <= (0 if exact else 1e-7)
    f0 = read_file(str(tmp_path / "l0.native"), "native", dtype=np.float64)
    assert rel_err(f0["dLdF"], sim.get_grad(0, "F", dtype=np.float64)) == 0
    f3 = read_file(str(tmp_path / "l3.native"), "native", dtype=np.float64)
    assert not f3["dLdx"].any() and not f3["dLdF"].any()
    sim.close()
