"""Writes the golden vectors in this folder from the reference's own code.

Runs in the build container only: it needs oracle/_ref/libdmpm_ref.so, which
`make -C oracle ref` compiles from the reference's DiffMPMLib3D sources (plus
oracle/ref_capi.cpp, a C-ABI driver that only builds the reference's
PointCloud/Grid/CompGraph objects from arrays and calls its functions).  The
reference runs single-threaded (OMP_NUM_THREADS=1 semantics) so that its
float atomics and critical sections are deterministic.

    python tests/golden/make_golden.py

Every fixture holds the inputs and the reference's outputs (float32).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "shape-morphing-binder_amd"))

from oracle import Ref, reference_params  # noqa: E402
from dmpm.scenes import Body, make_scene  # noqa: E402


def scene(n=400, cells=16, seed=7):
    sc = make_scene(Body((0.5, 0.5, 0.5), (0.2, 0.2, 0.2)), n, cells, seed, 1e4, 0.3, h=1 / 32, velocity="sine",
                    amplitude=0.3)
    s = {k: np.asarray(v, np.float32) for k, v in sc.state().items()}
    rng = np.random.default_rng(seed + 1)
    s["F"] = (s["F"] + 0.05 * rng.normal(size=s["F"].shape)).astype(np.float32)
    s["C"] = (0.5 * rng.normal(size=s["C"].shape)).astype(np.float32)
    s["dFc"] = (0.01 * rng.normal(size=s["F"].shape)).astype(np.float32)
    return sc, s


def main():
    ref = Ref(threads=1)
    sc, s = scene()
    prm = reference_params(dims=sc.dims, dx=sc.dx, dt=1e-3, drag=0.5, smoothing=0.1, f_ext=(0.0, -9.8, 0.0))
    meta = dict(dims=np.array(prm.dims), dx=np.float64(prm.dx), dt=np.float64(prm.dt), drag=np.float64(prm.drag),
                smoothing=np.float64(prm.smoothing), f_ext=np.array(prm.f_ext))
    ins = {f"in_{k}": v for k, v in s.items()}

    # 1. one ForwardTimeStep
    o = ref.forward_step(prm, s)
    np.savez_compressed(os.path.join(HERE, "ref_forward_step.npz"), **meta, **ins,
                        **{f"out_{k}": o[k] for k in ("x", "v", "F", "C", "P", "grid_m", "grid_v")})

    # 2. one Back_Timestep with a seeded adjoint of layer n+1
    rng = np.random.default_rng(21)
    adj = dict(x=rng.normal(size=(sc.n, 3)), v=rng.normal(size=(sc.n, 3)), F=rng.normal(size=(sc.n, 3, 3)),
               C=rng.normal(size=(sc.n, 3, 3)))
    adj = {k: v.astype(np.float32) for k, v in adj.items()}
    b = ref.backward_step(prm, s, adj)
    np.savez_compressed(os.path.join(HERE, "ref_backward_step.npz"), **meta, **ins,
                        **{f"adj_{k}": v for k, v in adj.items()},
                        **{f"out_{k}": b[k] for k in ("x", "v", "F", "C", "P")})

    # 3. EndLayerMassLoss against the node masses of a shifted body
    tsc = make_scene(Body((0.53, 0.5, 0.47), (0.2, 0.18, 0.2)), 400, 16, 8, 1e4, 0.3, h=1 / 32)
    tm = ref.p2g_mass(prm, tsc.x.astype(np.float32), tsc.m.astype(np.float32))
    loss, gx, gm, gdm = ref.mass_loss(prm, s, tm)
    np.savez_compressed(os.path.join(HERE, "ref_mass_loss.npz"), **meta, **ins, target_m=tm,
                        out_loss=np.float64(loss), out_gx=gx, out_grid_m=gm, out_grid_dLdm=gdm)

    # 4. a 4-layer episode: forward, loss, render gradients, backward to layer 0
    s4 = dict(s)
    s4["dFc"] = np.zeros_like(s["F"])
    rgF = (1e-3 * rng.normal(size=(sc.n, 3, 3))).astype(np.float32)
    rgx = (1e-2 * rng.normal(size=(sc.n, 3))).astype(np.float32)
    e = ref.episode(prm, s4, tm, num_layers=4, control_layer=0, render_gF=rgF, render_gx=rgx)
    np.savez_compressed(os.path.join(HERE, "ref_episode.npz"), **meta, **{f"in_{k}": v for k, v in s4.items()},
                        target_m=tm, render_gF=rgF, render_gx=rgx, num_layers=np.int64(4),
                        out_loss=np.float64(e["loss"]),
                        **{f"out_{k}": e[k] for k in ("x", "v", "F", "C", "gx", "gv", "gF", "gC")})

    # 5. CalculatePointCloudVolumes and the dt=0 P2G
    vol, gmv = ref.volumes(prm, s["x"], s["m"])
    np.savez_compressed(os.path.join(HERE, "ref_volumes.npz"), **meta, in_x=s["x"], in_m=s["m"], out_vol=vol,
                        out_grid_m=gmv)

    # 6. Descend_Adam from a non-zero moment state, two calls
    n = 64
    dLdF = rng.normal(size=(n, 3, 3)).astype(np.float32)
    dFc = (0.01 * rng.normal(size=(n, 3, 3))).astype(np.float32)
    mom = (0.1 * rng.normal(size=(n, 3, 3))).astype(np.float32)
    vec = np.abs(0.01 * rng.normal(size=(n, 3, 3))).astype(np.float32)
    vmax = np.maximum(vec, np.abs(0.01 * rng.normal(size=(n, 3, 3)))).astype(np.float32)
    gn = ref.dLdF_norm(dLdF)
    out1 = ref.adam(dLdF, dFc, mom, vec, vmax, 0.5, gn, timestep=3)
    dLdF2 = dLdF.copy()
    dLdF2[::2] *= 1e-4  # near-stationary particles take the damped branch
    out2 = ref.adam(dLdF2, *out1, 1e-3, gn, timestep=4)
    np.savez_compressed(os.path.join(HERE, "ref_adam.npz"), dLdF=dLdF, dLdF2=dLdF2, dFc=dFc, mom=mom, vec=vec, vmax=vmax,
                        grad_norm=np.float64(gn), out1_dFc=out1[0], out1_mom=out1[1], out1_vec=out1[2],
                        out1_vmax=out1[3], out2_dFc=out2[0], out2_mom=out2[1], out2_vec=out2[2], out2_vmax=out2[3])

    # 7. Elasticity primitives on sampled F (including inverted and near-rotation ones)
    F = np.eye(3, dtype=np.float32) + 0.3 * rng.normal(size=(128, 3, 3)).astype(np.float32)
    F[:4] = np.diag([1.0, 1.0, -0.5]).astype(np.float32) @ F[:4]
    lam, mu = ref.lame(1e4, 0.3)
    dF = rng.normal(size=F.shape).astype(np.float32)
    R, S = ref.polar(F)
    np.savez_compressed(os.path.join(HERE, "ref_elasticity.npz"), F=F, dF=dF, lam=np.float64(lam),
                        mu=np.float64(mu), R=R, S=S, P=ref.pk_fcr(F, lam, mu), hvp=ref.pk_fcr_hvp(F, lam, mu, dF),
                        psi=ref.energy_fcr(F, lam, mu))
    print("wrote golden vectors to", HERE)
    io_fixture()
    mesh_fixture()
    assets_fixture()


def io_fixture():
    """8. The reference's cereal PointCloud file (PointCloud::
    WriteEntirePointCloudToBinaryFile, PointCloud.cpp:382-394) of 5 points
    with distinct values in every serialised field, and the inputs."""
    ref = Ref(1)
    rng = np.random.default_rng(808)
    n = 5
    shapes = {3: (n, 3), 9: (n, 3, 3), 1: (n,)}
    width = dict(x=3, v=3, F=9, C=9, P=9, m=1, vol=1, lam=1, mu=1, dFc=9, dLdx=3, dLdv=3, dLdF=9, momentum=9,
                 vector=9, vector_max=9)
    a = {k: rng.normal(size=shapes[width[k]]).astype(np.float32) for k in Ref.CEREAL_FIELDS}
    path = os.path.join(HERE, "ref_pointcloud.cereal")
    ref.write_cereal(path, a)
    np.savez_compressed(os.path.join(HERE, "ref_pointcloud_inputs.npz"), **a)
    print("wrote", path)


def torus_obj(path: str, R=0.3, r=0.12, nu=24, nv=12, noise=0.01, seed=909):
    """A closed, irregular triangle mesh (a torus with seeded vertex noise),
    outward-oriented, centred at (0.5, 0.5, 0.5), written as OBJ text."""
    rng = np.random.default_rng(seed)
    V = []
    for i in range(nu):
        for j in range(nv):
            u, v = 2 * np.pi * i / nu, 2 * np.pi * j / nv
            p = np.array([(R + r * np.cos(v)) * np.cos(u), (R + r * np.cos(v)) * np.sin(u), r * np.sin(v)])
            V.append(p + noise * rng.normal(size=3) + 0.5)
    F = []
    for i in range(nu):
        for j in range(nv):
            a, b = i * nv + j, ((i + 1) % nu) * nv + j
            c, d = ((i + 1) % nu) * nv + (j + 1) % nv, i * nv + (j + 1) % nv
            F += [(a, b, c), (a, c, d)]
    with open(path, "w") as f:
        for p in V:
            f.write("v %.6f %.6f %.6f\n" % tuple(p))
        for t in F:
            f.write("f %d %d %d\n" % (t[0] + 1, t[1] + 1, t[2] + 1))


def mesh_fixture():
    """9. GeometryLoading::LoadMPMPointCloudFromObj (GeometryLoading.cpp:85-143)
    on an irregular torus: the points, masses and volumes the reference
    samples at point_dx = 0.025 (grid_dx 0.05, 2 points per cell)."""
    ref = Ref(1)
    obj = os.path.join(HERE, "mesh_torus.obj")
    torus_obj(obj)
    grid_dx, ppc, rho, E, nu = 0.05, 2, 1000.0, 1e4, 0.3
    lam, mu = ref.lame(E, nu)
    point_dx = float(np.float32(grid_dx) / np.float32(ppc))
    o = ref.load_obj(obj, point_dx, rho, lam, mu)
    np.savez_compressed(os.path.join(HERE, "ref_mesh_torus.npz"), grid_dx=np.float32(grid_dx), ppc=np.int32(ppc),
                        rho=np.float32(rho), lam=np.float32(lam), mu=np.float32(mu), **o)
    print("wrote", obj, "and", o["x"].shape[0], "reference points")


REF_ASSETS = "/root/reference/assets"


def assets_fixture():
    """10. The reference's own binding test (tests/test.py:157-245) on the
    meshes it ships: isosphere.obj -> input, bunny.obj -> target; grid_dx 1,
    grid 33^3 from -16, 2 points per cell axis, density 75, lam 38888.89,
    mu 58333.3.  Copies the two meshes here (data files of the reference's
    test) and records the points, volumes, target grid masses and the
    end-layer mass loss (with its dL/dx) that the reference computes."""
    import shutil
    ref = Ref(1)
    for name in ("isosphere.obj", "bunny.obj"):
        shutil.copyfile(os.path.join(REF_ASSETS, name), os.path.join(HERE, "asset_" + name))
    point_dx = float(np.float32(1.0) / np.float32(2))
    lam, mu, rho = 38888.89, 58333.3, 75.0
    inp = ref.load_obj(os.path.join(HERE, "asset_isosphere.obj"), point_dx, rho, lam, mu)
    tgt = ref.load_obj(os.path.join(HERE, "asset_bunny.obj"), point_dx, rho, lam, mu)
    prm = reference_params(dims=(33, 33, 33), dx=1.0, grid_min=(-16.0, -16.0, -16.0))
    vin, _ = ref.volumes(prm, inp["x"], inp["m"])
    vt, gt = ref.volumes(prm, tgt["x"], tgt["m"])
    n = inp["x"].shape[0]
    s = dict(x=inp["x"], v=np.zeros((n, 3), np.float32), F=np.broadcast_to(np.eye(3, dtype=np.float32), (n, 3, 3)),
             C=np.zeros((n, 3, 3), np.float32), m=inp["m"], vol=vin, lam=np.full(n, lam, np.float32),
             mu=np.full(n, mu, np.float32))
    loss, gx, _, _ = ref.mass_loss(prm, s, gt)
    np.savez_compressed(os.path.join(HERE, "ref_assets.npz"), in_x=inp["x"], in_m=inp["m"], in_vol=vin,
                        tgt_x=tgt["x"], tgt_m=tgt["m"], tgt_vol=vt, target_grid_m=gt, loss=np.float64(loss),
                        dLdx=gx, lam=np.float32(lam), mu=np.float32(mu), rho=np.float32(rho))
    print("assets:", n, "input points,", tgt["x"].shape[0], "target points, loss", loss)


if __name__ == "__main__":
    if sys.argv[1:] == ["assets"]:
        assets_fixture()
    elif sys.argv[1:] == ["io"]:
        io_fixture()
    elif sys.argv[1:] == ["mesh"]:
        mesh_fixture()
    else:
        main()
