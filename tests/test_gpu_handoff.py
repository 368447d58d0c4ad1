"""The torch handoff on the device: particles, render gradients and the
layer getters take and give CUDA tensors without a host round trip, with
the same results as the host paths.  Deterministic mode (two-pass scatters
in a fixed order) makes runs bitwise repeatable, so "the same" is bitwise:
the storage order the device computes from device positions must be the
one it computes from host positions."""
import numpy as np
import pytest

from helpers import perturbed_scene, physics_kwargs, random_adjoint

pytestmark = pytest.mark.gpu


def _cuda(a, dtype):
    import torch
    return torch.tensor(np.asarray(a), dtype=dtype, device="cuda:0")


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_particles_from_device_tensors_match_host_arrays(prec):
    import torch
    from dmpm import Physics, Simulator
    sc, s = perturbed_scene(n=3000)
    T = 4
    tdt = torch.float64 if prec == "f64" else torch.float32
    ndt = np.float64 if prec == "f64" else np.float32
    adj = random_adjoint(sc.n, seed=3)
    out = []
    for device_inputs in (False, True):
        print(f"[handoff] prec={prec} device_inputs={device_inputs} n={sc.n}", flush=True)
        sim = Simulator(sc.dims, sc.dx, sc.grid_min, physics=Physics(**physics_kwargs("reference", 1e-3)),
                        precision=prec, deterministic=True)
        st = {k: (None if v is None else (_cuda(v, tdt) if device_inputs else np.asarray(v, ndt)))
              for k, v in s.items()}
        sim.set_state(st)
        sim.set_up(T + 1)
        sim.forward(0)
        sim.set_last_grad(x=adj["x"], F=adj["F"])
        sim.backward(0)
        r = {k: sim.get(T, k, dtype=ndt) for k in ("x", "v", "F", "C")}
        r.update({"g" + k: sim.get_grad(0, k, dtype=ndt) for k in ("x", "v", "F", "C")})
        sim.sort_particles(T // 2)
        sim.forward(0)
        r["x_sorted"] = sim.get(T, "x", dtype=ndt)
        sim.close()
        out.append(r)
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


def test_render_gradients_and_getters_on_the_device():
    """set_render_gradients / accumulate_render_grads with CUDA tensors give
    bitwise the host arrays' norms and backward pass; get_positions_torch and
    get_def_grads_total_torch on the device give bitwise the host getters'
    values (bind.cpp:89-131, 189-279, 359-451)."""
    import torch
    import diffmpm_bindings as db
    sc, s = perturbed_scene(n=2500)
    rng = np.random.default_rng(5)
    gF = (1e-3 * rng.normal(size=(sc.n, 3, 3))).astype(np.float32)
    gx = (1e-2 * rng.normal(size=(sc.n, 3))).astype(np.float32)
    aF = (1e-3 * rng.normal(size=(sc.n, 3, 3))).astype(np.float32)
    ax = (1e-2 * rng.normal(size=(sc.n, 3))).astype(np.float32)
    res = []
    for on_device in (False, True):
        print(f"[handoff] render gradients on_device={on_device}", flush=True)
        pc = db.PointCloud.from_arrays(s["x"], s["m"], s["vol"], s["lam"], s["mu"], v=s["v"], F=s["F"], C=s["C"],
                                       dFc=s["dFc"])
        g = db.Grid(*sc.dims, sc.dx, [0.0, 0.0, 0.0])
        cg = db.CompGraph(pc, g, db.Grid(*sc.dims, sc.dx, [0.0, 0.0, 0.0]), deterministic=True)
        cg._sim.set_sim_params(1e-3, 0.5, 0.1, (0.0, -9.8, 0.0))
        cg.set_up_comp_graph(4)
        cg.compute_forward_pass(0, 0)
        conv = (lambda a: torch.tensor(a, device="cuda:0")) if on_device else (lambda a: a)
        cg.set_render_gradients(conv(gF), conv(gx))
        cg.accumulate_render_grads(conv(aF), conv(ax))
        info = cg.get_render_gradient_info()
        cg.compute_backward_pass(0)
        r = {"info": info}
        r.update({k: cg._sim.get_grad(0, k) for k in ("x", "v", "F", "C")})
        last = cg.get_point_cloud(3)
        hx, hF = last.get_positions_torch(), last.get_def_grads_total_torch()
        dx, dF = last.get_positions_torch(device="cuda:0"), last.get_def_grads_total_torch(device="cuda:0")
        assert dx.is_cuda and dF.is_cuda and not hx.is_cuda
        assert torch.equal(dx.cpu(), hx) and torch.equal(dF.cpu(), hF)
        res.append(r)
    assert res[0]["info"] == res[1]["info"]
    assert abs(res[0]["info"]["grad_F_norm"] - np.linalg.norm(gF.astype(np.float64))) < 1e-6 * np.linalg.norm(gF)
    for k in ("x", "v", "F", "C"):
        assert np.array_equal(res[0][k], res[1][k]), k
