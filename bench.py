#!/usr/bin/env python3
"""Benchmark of the differentiable MPM stepper on one MI355X.

Metric: forward+backward particle-steps/s at 1M particles on a 128^3 grid,
% of HBM peak, gradient error versus FP64.

Workload (the precision-stress configuration): 1,048,576 particles in a
sphere of radius 0.3 (jittered lattice), 128^3 cells (129^3 nodes,
dx = 1/128), E = 5e4, nu = 0.45, rho = 1e3, dt = 1e-4, an 8-mode sine
velocity field of amplitude 0.5, 300 steps, plain physics (quadratic
B-spline, Neo-Hookean, sticky walls at 3 cells).  A "step" of this benchmark
is one episode: 300 forward steps plus the 300-step adjoint back to layer 0,
seeded at the last layer with a Gaussian dL/dx (sigma 1e-3) and dL/dF
(sigma 1e-4).  value = particles * 300 * K / wall time of K episodes.

Tape: the whole episode's tape (301 layer states and 300 P2G grids, about
41 GB in FP32) fits in HBM, so by default every layer's state and grid is kept
(checkpoint interval 1) and the backward pass recomputes nothing -- the
reference's CompGraph keeps every layer and its grid too.  The checkpointed
form (checkpoint interval 25: about 5 GB, one extra forward pass of
recomputation) is measured beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import contextlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "shape-morphing-binder_amd"))

METRIC = "fwd+bwd particle-steps/s at 1M pts,128³ grid; % HBM peak; grad err vs FP64"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes of one step's forward pass and one step's adjoint (FP32
# state): per particle and per active grid node (a node with m > 0, i.e. one
# that some particle's stencil touches).  A fused forward launch (k_step) is
# one step's forward pass, a fused adjoint launch (k_adj_fused: the adjoint
# gather of step t with the adjoint scatter of step t-1) one step's adjoint.
#   forward: read x, v, F, C (96 B) + m, vol (8 B), write x, v, F, C (96 B) per
#            particle; per node write m, p (16 B), read m, p + write v (28 B),
#            read v (12 B)
#   adjoint: read layer t's x, v, F, C, m, vol (104 B) + C of layer t+1 (36 B)
#            and the adjoint of layer t+1 (96 B), write the adjoint of layer t
#            (96 B) per particle; per node read m, v (16 B), scatter dL/dv
#            (12 B), read dL/dv + write dL/dp, dL/dm (28 B), read dL/dp, dL/dm
#            (16 B)
# The whole fwd+bwd step is 532 B per particle + 128 B per active node.
ALGO_BYTES = {"step": (200, 56), "adj_fused": (332, 72)}

# The kernels' own byte models, FP32 (4-byte elements): every word each kernel
# reads or writes once, as built (the adjoint kernels read G2P's v' and C' of
# step t from layer t+1 instead of recomputing them).  Grid nodes hold four
# values each: m, p (16 B) in the forward grid, dL/dv and an unused slot
# (16 B) in the adjoint grid.
#   p2g         reads x, v, F, C, m, vol, lam, mu (28 floats); writes m, p per node
#   g2p         reads x, F (12 floats) and m, p per node; writes x, v, F, C (24 floats)
#   adj_scatter reads x, F (12), the adjoint carry (24) and C of layer t+1 (9); writes dL/dv per node
#   adj_gather  reads x, v, F, C, m, vol, lam, mu (28), v and C of layer t+1 (12) and the carry (24),
#               m, p, dL/dv per node; writes the carry (24)
#   adj_fused   adj_gather of step t, then adj_scatter of step t-1: adds x, F of layer t-1 (12 floats)
#               (C of layer t and the carry of layer t are already in registers) and dL/dv of
#               step t-1 per node
#   step        the fused forward step: reads x, F, m, vol, lam, mu (16 floats) and m, p of grid t per node;
#               writes x, v, F, C (24 floats) and m, p of grid t+1 per node
BYTES = {
    "step": (40 * 4, 32),
    "p2g": (28 * 4, 16),
    "g2p": (36 * 4, 16),
    "adj_scatter": (45 * 4, 16),
    "adj_gather": (88 * 4, 32),
    "adj_fused": (100 * 4, 48),
}


def profiled(key: str, kernel: str, name: str = "kernels_1M_128cubed_hbm_traffic.md"):
    """A per-kernel figure of the default workload from a committed rocprofv3
    summary under profiles/ (a line `key kernel value`), or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == key and f[1] == kernel:
                return float(f[2])
    except OSError:
        pass
    return None


def profiled_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the FETCH_SIZE/WRITE_SIZE runs."""
    v = profiled("traffic_bytes_per_launch", kernel)
    return None if v is None else int(v)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads() -> int:
    """Threads for the CPU baseline: OMP_NUM_THREADS when the host sets it
    (the share of the host's cores this process is given), else every core
    this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return max(1, os.cpu_count() or 1)


def host_cpu() -> dict:
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = None
    return {"model": model, "logical_cpus": os.cpu_count(), "cpus_available": aff}


@contextlib.contextmanager
def stdout_to_stderr():
    """Send what native code writes to file descriptor 1 (the reference
    prints its grid size there) to stderr, so that stdout carries only the
    result line."""
    sys.stdout.flush()
    saved = os.dup(1)
    try:
        os.dup2(2, 1)
        yield
    finally:
        sys.stdout.flush()
        os.dup2(saved, 1)
        os.close(saved)


def cpu_baseline(sc, seed_gx, steps: int = 5):
    """The reference's own OpenMP C++ path (compiled from its sources into
    oracle/_ref) on a bounded sample of the same workload: all particles,
    `steps` forward steps and `steps` adjoint steps, its own physics
    (cubic B-spline, fixed-corotated).  Falls back to the oracle's port when
    the reference library was not built."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, Ref, plain_params, reference_params
    threads = cpu_threads()
    s = {k: np.asarray(v, np.float32) for k, v in sc.state().items()}
    if Ref.available():
        ref = Ref(threads)
        prm = reference_params(dims=sc.dims, dx=sc.dx, dt=1e-4, drag=0.0, smoothing=0.0)
        t0 = time.perf_counter()
        with stdout_to_stderr():
            fwd, bwd = ref.time_steps(prm, s, steps, seed_gx.astype(np.float32))
        log(f"[cpu] reference: {steps} steps fwd {fwd:.2f}s bwd {bwd:.2f}s ({time.perf_counter() - t0:.1f}s total)")
        return {"value": sc.n * steps / (fwd + bwd), "unit": "particle-steps/s", "cores": threads,
                "kind": "reference",
                "sample": f"{sc.n} particles, 128^3 grid, {steps} forward + {steps} adjoint steps of the "
                          f"reference's ForwardTimeStep/Back_Timestep (its physics: cubic B-spline, "
                          f"fixed-corotated), {threads} OpenMP threads",
                "fwd_s": fwd, "bwd_s": bwd, "host": host_cpu()}
    n = min(sc.n, 65536)
    sub = {k: v[:n] for k, v in s.items()}
    orc = Oracle("f32")
    prm = plain_params(dims=sc.dims, dx=sc.dx, dt=1e-4)
    t0 = time.perf_counter()
    layers = orc.rollout(prm, dict(sub, dFc=None), steps)
    gT = dict(x=seed_gx[:n], v=np.zeros((n, 3)), F=np.zeros((n, 3, 3)), C=np.zeros((n, 3, 3)))
    orc.rollout_grad(prm, layers, gT)
    dt = time.perf_counter() - t0
    return {"value": n * steps / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"{n} particles, {steps} fwd+bwd steps of the oracle (plain physics, one thread)",
            "host": host_cpu()}


def stream_copy_GBps(mib: int = 1024, reps: int = 5) -> float:
    """Achievable HBM bandwidth: a plain elementwise copy (read + write of
    `mib` MiB each) timed with HIP events on torch's stream."""
    import torch
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2):
        torch.mul(src, 2.0, out=dst)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        torch.mul(src, 2.0, out=dst)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2 * (mib << 20) / (ms * 1e-3) / 1e9


def active_nodes(sc) -> int:
    from dmpm import p2g_mass
    gm = p2g_mass(sc.dims, sc.dx, sc.grid_min, sc.x, sc.m, order=2)
    return int((gm > 0).sum())


def episode_sim(sc, prec, K, T, seed_gx, seed_gF, physics=None, lds_bytes=0, block=0):
    from dmpm import Physics, Simulator
    ph = physics or Physics.plain(dt=1e-4)
    sim = Simulator(sc.dims, sc.dx, sc.grid_min, physics=ph, precision=prec, checkpoint_interval=K,
                    lds_bytes=lds_bytes, block=block)
    sim.set_state(sc.state())
    sim.set_up(T + 1)
    sim.set_last_grad(x=seed_gx, F=seed_gF)
    return sim


def run_episodes(sim, count):
    for _ in range(count):
        sim.forward(0, keep_controls=True)
        sim.backward(0)


def warm_up(sim, count, T, sort=True, segments=0):
    """`count` untimed episodes; after the first, the particle storage is
    reordered per episode segment (dmpm_sort_segments): the particles drift
    from the order they were uploaded in, and each run of layers is stored
    in the Hilbert order of its middle layer.  `segments` = 0 lets the
    autotuner pick the count (one more untimed episode per candidate: the
    whole episode in one order, a segment every 25 or every 10 layers).
    Returns the autotuner's choice and table (None without a sort).  No
    warm-up, no sort."""
    if count <= 0:
        return None
    from dmpm.autotune import tune_segments
    run_episodes(sim, 1 if sort else count)
    info = None
    if sort:
        info = tune_segments(sim, [segments] if segments > 0 else None)
        run_episodes(sim, max(count - 1, 0))
    sim.synchronize()
    return info


def batched_workload(dt: float = 2.5e-4, T: int = 100, K: int = 10, reps: int = 2, lds_bytes: int = 0,
                     block: int = 0) -> dict:
    """64 independent instances of uneven size (particle counts log-uniform
    in [4,096, 131,072], grids of 32/48/64/96 cells, random ellipsoids and
    materials), each a T-step forward + backward episode seeded at its last
    layer: one at a time (eager calls per instance), one hipGraph holding
    every instance's episode as a concurrent branch, and the batched episode
    (each kernel one launch over all instances, the episode one hipGraph).
    Each instance's storage is re-sorted by its middle layer after a first
    episode, as in the main workload."""
    from dmpm import BatchedEpisode, EpisodeGraph, Physics, Simulator
    from dmpm.scenes import config4_instances, instance_scene
    sims, total = [], 0
    for spec in config4_instances():
        sc = instance_scene(spec)
        sim = Simulator(sc.dims, sc.dx, sc.grid_min, physics=Physics.plain(dt=dt), precision="f32",
                        checkpoint_interval=K, lds_bytes=lds_bytes, block=block)
        sim.set_state(sc.state())
        sim.set_up(T + 1)
        rng = np.random.default_rng(spec["seed"])
        sim.set_last_grad(x=1e-3 * rng.normal(size=(sc.n, 3)), F=1e-4 * rng.normal(size=(sc.n, 3, 3)))
        sims.append(sim)
        total += sc.n

    def sequential():
        for s in sims:
            run_episodes(s, 1)
            s.synchronize()

    log(f"[bench] batched: {len(sims)} instances set up, {total} particles")
    for s in sims:
        # one order each: a batched episode runs every instance in one
        warm_up(s, 1, T, segments=1)
    t0 = time.perf_counter()
    for _ in range(reps):
        sequential()
    w_seq = (time.perf_counter() - t0) / reps
    log(f"[bench] batched: one at a time {1e3 * w_seq:.1f} ms")
    t0 = time.perf_counter()
    g = EpisodeGraph(sims, start=0, stop=0)
    log(f"[bench] batched: graph captured and instantiated ({time.perf_counter() - t0:.2f}s)")
    g.launch()
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.launch()
    g.synchronize()
    w_graph = (time.perf_counter() - t0) / reps
    g.close()
    log(f"[bench] batched: graph of branches {1e3 * w_graph:.1f} ms")
    t0 = time.perf_counter()
    b = BatchedEpisode(sims, start=0, stop=0)
    log(f"[bench] batched: one-launch episode recorded and captured ({time.perf_counter() - t0:.2f}s)")
    b.launch()
    b.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.launch()
    b.synchronize()
    w_batch = (time.perf_counter() - t0) / reps
    b.close()
    peak = sum(s.memory()[1] for s in sims)
    for s in sims:
        s.close()
    out = {"instances": len(sims), "particles_total": total, "episode_steps": T, "checkpoint_interval": K,
           "dt": dt, "one_at_a_time": {"value": total * T / w_seq, "ms_per_step": 1e3 * w_seq},
           "graph": {"value": total * T / w_graph, "ms_per_step": 1e3 * w_graph},
           "one_launch": {"value": total * T / w_batch, "ms_per_step": 1e3 * w_batch},
           "speedup_graph": w_seq / w_graph, "speedup_one_launch": w_seq / w_batch,
           "peak_device_bytes": peak, "unit": "particle-steps/s"}
    log(f"[bench] batched: {len(sims)} instances, {total} particles: one at a time {1e3 * w_seq:.1f} ms, "
        f"graph of branches {1e3 * w_graph:.1f} ms ({out['speedup_graph']:.2f}x), one launch per kernel "
        f"{1e3 * w_batch:.1f} ms ({out['speedup_one_launch']:.2f}x)")
    return out


def full_scale_episode(T: int = 500, K: int = 25, controls: int = 20) -> dict:
    """4,194,304 particles in a sphere of radius 0.3 on a 256^3 grid, FP32
    with FP64 det/log J/F^-T, dt = 5e-5, a 500-step episode with checkpoint
    interval 25, Gaussian dL/dx (sigma 1e-3) and dL/dF (sigma 1e-4) injected
    at the last layer and backpropagated to a per-particle control at 20
    control layers (every 25th).  One warm-up episode, then the storage is
    re-sorted per run of layers (whole checkpoint segments), as in the main
    workload."""
    from dmpm import Physics, Simulator
    from dmpm.scenes import config5
    sc = config5()
    sim = Simulator(sc.dims, sc.dx, sc.grid_min, physics=Physics.plain(dt=5e-5), precision="f32",
                    checkpoint_interval=K)
    sim.set_state(sc.state())
    sim.set_up(T + 1)
    zero = np.zeros((sc.n, 3, 3), np.float32)
    layers = list(range(0, T, T // controls))[:controls]
    for t in layers:
        sim.set_control(t, zero)
        sim.request_control_grad(t)
    rng = np.random.default_rng(5)
    sim.set_last_grad(x=1e-3 * rng.normal(size=(sc.n, 3)), F=1e-4 * rng.normal(size=(sc.n, 3, 3)))
    log(f"[bench] full scale: {sc.n} particles set up")
    order = warm_up(sim, 1, T)
    log(f"[bench] full scale: warm-up episode done; storage orders: {order['segments']} segment(s); "
        + ", ".join(f"{r['segments']}: {r['ms']:.0f}ms" for r in order["table"]))
    t0 = time.perf_counter()
    run_episodes(sim, 1)
    sim.synchronize()
    w = time.perf_counter() - t0
    g = sim.get_control_grad(layers[0], dtype=np.float64)
    _, peak = sim.memory()
    sim.close()
    out = {"particles": sc.n, "grid_nodes": list(sc.dims), "episode_steps": T, "checkpoint_interval": K,
           "control_layers": len(layers), "order_segments": order, "value": sc.n * T / w, "unit": "particle-steps/s",
           "ms_per_step": 1e3 * w, "peak_device_bytes": peak,
           "control_grad_finite": bool(np.isfinite(g).all()), "control_grad_rms": float(np.sqrt((g ** 2).mean()))}
    log(f"[bench] full scale: {1e3 * w:.0f} ms/episode, {out['value']:.4g} particle-steps/s, "
        f"peak {peak / 1e9:.1f} GB")
    return out


def rollout_overhead(sc, T: int, lds_bytes: int, block: int) -> dict:
    """What the torch.autograd.Function front end (dmpm.rollout) costs on
    top of the episode it runs, at the metric size: a forward + backward of
    a T-step rollout from CUDA tensors (particle upload and ordering on the
    device, the layer copies of set-up, the final-layer getters, the seed and
    the layer-0 gradient getters) against the same episode on a sim whose
    particles are already set.  The mean of two calls after two untimed ones."""
    import torch
    from dmpm import Physics, Simulator
    from dmpm.rollout import rollout
    dev = torch.device("cuda:0")
    st = {k: torch.tensor(np.asarray(v), dtype=torch.float32, device=dev) for k, v in sc.state().items()}
    for k in ("x", "v", "F", "C"):
        st[k].requires_grad_(True)
    sim = Simulator(sc.dims, sc.dx, sc.grid_min, physics=Physics.plain(dt=1e-4), precision="f32",
                    lds_bytes=lds_bytes, block=block)
    w = torch.randn(sc.n, 3, device=dev)

    def once():
        xT, vT, FT, CT = rollout(sim, st, T)
        (xT * w).sum().backward()
        torch.cuda.synchronize()

    # two untimed calls: the first creates the inputs' .grad, the second is
    # the first to accumulate into them (torch loads those kernels lazily)
    once()
    once()
    t0 = time.perf_counter()
    once()
    once()
    w_roll = (time.perf_counter() - t0) / 2
    t0 = time.perf_counter()
    sim.set_particles(st["x"], st["v"], st["F"], st["C"], st["m"], st["vol"], st["lam"], st["mu"])
    sim.synchronize()
    w_set = time.perf_counter() - t0
    sim.set_up(T + 1)
    sim.set_last_grad(x=w)
    run_episodes(sim, 1)
    sim.synchronize()
    t0 = time.perf_counter()
    run_episodes(sim, 1)
    sim.synchronize()
    w_ep = time.perf_counter() - t0
    sim.close()
    out = {"particles": sc.n, "episode_steps": T, "rollout_fwd_bwd_ms": 1e3 * w_roll, "episode_ms": 1e3 * w_ep,
           "overhead_ms": 1e3 * (w_roll - w_ep), "set_particles_ms": 1e3 * w_set,
           "note": "episodes in the upload order (no re-sort), so both sides run the same kernels"}
    log(f"[bench] rollout front end: {1e3 * w_roll:.1f} ms per fwd+bwd call against {1e3 * w_ep:.1f} ms for the "
        f"episode ({1e3 * (w_roll - w_ep):.1f} ms overhead; set_particles {1e3 * w_set:.1f} ms)")
    return out


def fp32_errors(o32, g32, o64, g64, n) -> dict:
    """Max and RMS of |x32 - x64| (the box is the unit cube) and |F32 - F64|_F
    at the last layer; relative L2 error of each layer-0 gradient."""
    dx = np.linalg.norm(o32["x"] - o64["x"], axis=1)
    dF = np.linalg.norm((o32["F"] - o64["F"]).reshape(n, 9), axis=1)
    return {"x_max": float(dx.max()), "x_rms": float(np.sqrt((dx ** 2).mean())),
            "F_max": float(dF.max()), "F_rms": float(np.sqrt((dF ** 2).mean())),
            **{f"grad_{k}_rel_l2": float(np.linalg.norm(g32[k] - g64[k]) / max(np.linalg.norm(g64[k]), 1e-300))
               for k in g32}}


def implosion_workload(T: int, lds_bytes: int, block: int) -> dict:
    """The 1M-particle sphere with a radial implosion (22/s) on top of its
    sine field, which compresses the body to det F = 0.30 before it rebounds
    (the sine field alone keeps det F above 0.97).  The FP32 path (FP64
    det/log J/F^-T) with the whole tape and checkpointed (K = 25) against a
    pure FP64 run with the whole tape; and, to measure how strongly the
    dynamics amplify small differences, a second FP64 run whose initial
    positions are perturbed by a relative 1e-7 (about one FP32 ulp)."""
    from dmpm.scenes import config3_implosion
    rng = np.random.default_rng(7)
    sc = config3_implosion()
    gx = 1e-3 * rng.normal(size=(sc.n, 3))
    gF = 1e-4 * rng.normal(size=(sc.n, 3, 3))

    def run(prec, K, scene, det_every=0):
        s = episode_sim(scene, prec, K, T, gx, gF, lds_bytes=lds_bytes if prec == "f32" else 0,
                        block=block if prec == "f32" else 0)
        order = warm_up(s, 1, T)
        t0 = time.perf_counter()
        run_episodes(s, 1)
        s.synchronize()
        w = time.perf_counter() - t0
        o = {k: s.get(T, k, dtype=np.float64) for k in ("x", "F")}
        g = {k: s.get_grad(0, k, dtype=np.float64) for k in ("x", "v", "F", "C")}
        jmin = [float(np.linalg.det(s.get(t, "F", dtype=np.float64)).min()) for t in range(0, T + 1, det_every)] \
            if det_every else None
        s.close()
        return w, o, g, jmin, order["segments"]

    w64, o64, g64, jmin, _ = run("f64", 1, sc, det_every=10)
    scp = config3_implosion()
    scp.x = scp.x * (1 + 1e-7 * np.random.default_rng(9).normal(size=scp.x.shape))
    _, op, gp, _, _ = run("f64", 1, scp)
    w1, o1, g1, _, s1 = run("f32", 1, sc)
    w25, o25, g25, _, s25 = run("f32", 25, sc)
    out = {"scene": "sphere r=0.3, 1,048,576 particles, 128^3, sine field (0.5) + radial implosion 22/s",
           "detF_min_over_episode": min(jmin), "detF_min_every_10_layers": jmin,
           "fp64_ms": 1e3 * w64, "fp32_ms": 1e3 * w1, "fp32_checkpointed_ms": 1e3 * w25,
           "fp32_order_segments": s1, "fp32_checkpointed_order_segments": s25,
           "errors_vs_fp64": fp32_errors(o1, g1, o64, g64, sc.n),
           "checkpointed_errors_vs_fp64": fp32_errors(o25, g25, o64, g64, sc.n),
           "fp64_x0_perturbed_1e-7_vs_fp64": fp32_errors(op, gp, o64, g64, sc.n)}
    log(f"[bench] implosion: det F min {min(jmin):.3f}; FP64 {1e3 * w64:.1f} ms, FP32 {1e3 * w1:.1f} ms, "
        f"FP32 K=25 {1e3 * w25:.1f} ms; FP32 errors {out['errors_vs_fp64']}; FP64 x0*(1+1e-7) "
        f"{out['fp64_x0_perturbed_1e-7_vs_fp64']}")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed episodes")
    ap.add_argument("--warmup", type=int, default=1, help="untimed episodes")
    ap.add_argument("--particles", type=int, default=1048576)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--episode-steps", type=int, default=300)
    ap.add_argument("--checkpoint", type=int, default=1,
                    help="checkpoint interval K (1: every layer and its P2G grid kept in HBM)")
    ap.add_argument("--no-checkpointed", action="store_true",
                    help="skip the checkpointed (K = 25) comparison run")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-reference-physics", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="use the default LDS budget and block size")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-instance batched workload")
    ap.add_argument("--no-full-scale", action="store_true", help="skip the 4M-particle, 256^3, 500-step episode")
    ap.add_argument("--no-rollout", action="store_true",
                    help="skip the timing of the torch.autograd front end (dmpm.rollout) at the metric size")
    ap.add_argument("--no-implosion", action="store_true",
                    help="skip the strongly compressed companion scene (FP32 vs FP64 at det F ~ 0.3)")
    ap.add_argument("--no-sort", action="store_true",
                    help="keep the upload order (no re-sort after the first warm-up episode)")
    ap.add_argument("--segments", type=int, default=0,
                    help="storage-order segments of each episode (dmpm_sort_segments); 0 = chosen by timing")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (3 launches): the known-size kernel "
                         "that scales the FETCH_SIZE/WRITE_SIZE counters of a --pmc run")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: the stepper runs on one GPU (sharding across GPUs is designed, not built); "
              f"--gpus {args.gpus} is not supported (value is the one GPU's throughput).", file=sys.stderr)
        sys.exit(2)

    import torch  # (one HIP runtime for torch and libdmpm)
    from dmpm import Physics

    if args.copy_probe:
        src = torch.ones(args.copy_probe << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 2.0, out=dst)  # one elementwise kernel: reads and writes MIB MiB
        torch.cuda.synchronize()
        log(f"[bench] copy probe: 3 launches, {src.numel() * 4} bytes read and written per launch")
        del src, dst
    from dmpm.scenes import config3

    copy_bw = stream_copy_GBps()
    log(f"[bench] plain copy: {copy_bw:.0f} GB/s")
    T, K = args.episode_steps, max(1, args.checkpoint)
    t0 = time.perf_counter()
    sc = config3(n=args.particles, cells=args.cells)
    log(f"[bench] scene: {sc.n} particles, {sc.dims} nodes, h={sc.h:.5f} ({time.perf_counter() - t0:.1f}s)")
    rng = np.random.default_rng(3)
    seed_gx = 1e-3 * rng.normal(size=(sc.n, 3))
    seed_gF = 1e-4 * rng.normal(size=(sc.n, 3, 3))
    gact = active_nodes(sc)

    # ---- launch configuration (LDS window budget, block size) by the autotuner
    tuned = None
    lds_bytes, block = 0, 0
    if not args.no_autotune:
        from dmpm.autotune import tune
        t0 = time.perf_counter()
        tuned = tune(sc, Physics.plain(dt=1e-4), "f32", steps=20, seed_gx=seed_gx)
        lds_bytes, block = tuned["lds_bytes"], tuned["block"]
        log(f"[bench] autotune ({time.perf_counter() - t0:.1f}s): lds_bytes={lds_bytes} block={block}; "
            + ", ".join(f"{r['lds_bytes'] // 1024}K/{r['block']}: {r['ms']:.1f}ms fb {r['fallback_frac']:.3f}"
                        for r in tuned["table"]))

    # ---- FP32 (FP64 det/logJ/F^-T), checkpointed: the measured configuration.
    # Eager episodes (the library queues a whole episode asynchronously, so a
    # replay of the captured episode graph measures the same), then one more
    # episode with the per-kernel HIP-event timers on.
    sim = episode_sim(sc, "f32", K, T, seed_gx, seed_gF, lds_bytes=lds_bytes, block=block)
    order = warm_up(sim, args.warmup, T, sort=not args.no_sort, segments=args.segments)
    if order:
        log(f"[bench] storage orders: {order['segments']} segment(s); "
            + ", ".join(f"{r['segments']}: {r['ms']:.1f}ms" for r in order["table"]))
    t0 = time.perf_counter()
    run_episodes(sim, args.steps)
    sim.synchronize()
    wall = time.perf_counter() - t0
    value = sc.n * T * args.steps / wall
    ms_per_step = 1e3 * wall / args.steps
    sim.kernel_timer(12 * T + 64)
    run_episodes(sim, 1)
    sim.synchronize()
    kstats = sim.kernel_stats()
    sim.kernel_timer(0)
    # window statistics from one more episode (counting adds an atomic per block)
    sim.window_stats(reset=True)
    sim.count_windows(True)
    run_episodes(sim, 1)
    sim.synchronize()
    sim.count_windows(False)
    win_lds, win_fb = sim.window_stats()
    mem_cur, mem_peak = sim.memory()
    log(f"[bench] f32 K={K}: {ms_per_step:.1f} ms/episode, {value:.4g} particle-steps/s, peak {mem_peak / 1e9:.2f} GB")

    kernels = {}
    for name, (ms, cnt) in kstats.items():
        if cnt == 0:
            continue
        avg = ms / cnt
        entry = {"total_ms": ms, "launches": cnt, "avg_us": 1e3 * avg}
        if name in ALGO_BYTES:
            bp, bn = ALGO_BYTES[name]
            b = bp * sc.n + bn * gact
            entry["bytes_per_launch"] = b
            entry["GBps"] = b / (avg * 1e-3) / 1e9
        if name in BYTES:
            bp, bn = BYTES[name]
            b = bp * sc.n + bn * gact
            entry["kernel_model_bytes_per_launch"] = b
            entry["kernel_model_GBps"] = b / (avg * 1e-3) / 1e9
        kernels[name] = entry
    dom = max((k for k in kernels if k in ALGO_BYTES), key=lambda k: kernels[k]["total_ms"])
    d = kernels[dom]
    default_workload = (args.particles, args.cells, T) == (1048576, 128, 300)
    valu = None
    if default_workload:
        mix = "kernels_1M_128cubed_instruction_mix.md"
        vi = profiled("valu_insts_per_launch", "k_" + dom, mix)
        vb = profiled("valu_busy_frac", "k_" + dom, mix)
        if vi is not None:
            valu = {"insts_per_launch": vi, "busy_frac": vb, "source": "profiles/" + mix}
    roofline = {"bound": "hbm", "kernel": "k_" + dom, "achieved": d["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": d["GBps"] / HBM_PEAK_GBS,
                "traffic": profiled_traffic("k_" + dom) if default_workload else None,
                "bytes_per_launch": d["bytes_per_launch"],
                "bytes_model": f"{ALGO_BYTES[dom][0]} B per particle + {ALGO_BYTES[dom][1]} B per active node "
                               f"({sc.n} particles, {gact} active nodes)",
                "kernel_model_bytes_per_launch": d.get("kernel_model_bytes_per_launch"),
                "avg_launch_us": d["avg_us"], "achievable_copy_GBps": copy_bw, "valu": valu}

    # whole-step algorithmic roofline: 532 B per particle-step + 128 B per active node
    step_bytes = 532 * sc.n + 128 * gact
    step_frac = step_bytes * (value / sc.n) / (HBM_PEAK_GBS * 1e9)

    out32 = {k: sim.get(T, k, dtype=np.float64) for k in ("x", "F")}
    g32 = {k: sim.get_grad(0, k, dtype=np.float64) for k in ("x", "v", "F", "C")}
    sim.close()

    extra = {}
    if not args.no_checkpointed and K != 25:
        simc = episode_sim(sc, "f32", 25, T, seed_gx, seed_gF, lds_bytes=lds_bytes, block=block)
        order_c = warm_up(simc, args.warmup, T, sort=not args.no_sort, segments=args.segments)
        t0 = time.perf_counter()
        run_episodes(simc, args.steps)
        simc.synchronize()
        wc = (time.perf_counter() - t0) / args.steps
        _, peak_c = simc.memory()
        simc_out = {k: simc.get(T, k, dtype=np.float64) for k in ("x", "F")}
        simc_out.update({"g" + k: simc.get_grad(0, k, dtype=np.float64) for k in ("x", "v", "F", "C")})
        simc.close()
        extra["checkpointed"] = {"checkpoint_interval": 25, "order_segments": order_c,
                                 "value": sc.n * T / wc, "ms_per_step": 1e3 * wc,
                                 "peak_device_bytes": peak_c}
        out25 = {k: simc_out[k] for k in simc_out}
        log(f"[bench] f32 K=25: {1e3 * wc:.1f} ms/episode, {sc.n * T / wc:.4g} particle-steps/s, "
            f"peak {peak_c / 1e9:.2f} GB")
    if not args.no_fp64:
        sim64 = episode_sim(sc, "f64", K, T, seed_gx, seed_gF)
        warm_up(sim64, 1, T, sort=not args.no_sort, segments=args.segments)
        t0 = time.perf_counter()
        run_episodes(sim64, 1)
        sim64.synchronize()
        w64 = time.perf_counter() - t0
        out64 = {k: sim64.get(T, k, dtype=np.float64) for k in ("x", "F")}
        g64 = {k: sim64.get_grad(0, k, dtype=np.float64) for k in ("x", "v", "F", "C")}
        sim64.close()
        J = np.linalg.det(out64["F"])
        extra["fp64"] = {"value": sc.n * T / w64, "ms_per_step": 1e3 * w64,
                         "speedup_fp32_over_fp64": value / (sc.n * T / w64)}
        extra["errors_vs_fp64"] = {**fp32_errors(out32, g32, out64, g64, sc.n),
                                   "detF_min": float(J.min()), "detF_p01": float(np.percentile(J, 1)),
                                   "detF_p50": float(np.median(J))}
        if "checkpointed" in extra:
            gk = {k: out25["g" + k] for k in ("x", "v", "F", "C")}
            extra["checkpointed"]["errors_vs_fp64"] = fp32_errors(out25, gk, out64, g64, sc.n)
            w = extra["fp64"]["ms_per_step"] / extra["checkpointed"]["ms_per_step"]
            extra["checkpointed"]["speedup_over_fp64"] = w
        log(f"[bench] fp64: {1e3 * w64:.1f} ms/episode; errors {extra['errors_vs_fp64']}")
    if not args.no_implosion and default_workload:
        extra["implosion"] = implosion_workload(T, lds_bytes, block)

    if not args.no_reference_physics:
        simr = episode_sim(sc, "f32", K, T, seed_gx, seed_gF,
                           physics=Physics.reference(dt=1e-4, drag=0.0, smoothing=0.0))
        warm_up(simr, 1, T, sort=not args.no_sort, segments=args.segments)
        t0 = time.perf_counter()
        run_episodes(simr, 1)
        simr.synchronize()
        wr = time.perf_counter() - t0
        simr.close()
        extra["reference_physics"] = {"value": sc.n * T / wr, "ms_per_step": 1e3 * wr,
                                      "physics": "cubic B-spline, fixed-corotated, bulk viscosity 0.25, "
                                                 "trust region 0.6, reference adjoint"}
        log(f"[bench] reference physics: {1e3 * wr:.1f} ms/episode")

    if not args.no_rollout and default_workload:
        extra["rollout"] = rollout_overhead(sc, T, lds_bytes, block)
    if not args.no_batched:
        extra["batched"] = batched_workload(lds_bytes=lds_bytes, block=block)
    if not args.no_full_scale:
        extra["full_scale"] = full_scale_episode()

    cpu = None
    vs_cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sc, seed_gx)
        if "reference_physics" in extra and cpu["kind"] == "reference":
            # like for like: the GPU and the reference's CPU path on the same
            # physics (cubic B-spline, fixed-corotated, reference adjoint)
            vs_cpu = {"value": extra["reference_physics"]["value"] / cpu["value"],
                      "physics": "reference (cubic B-spline, fixed-corotated) on both",
                      "plain_physics_gpu_over_reference_cpu": value / cpu["value"]}

    rec = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "precision-stress sphere: 1,048,576 particles, 128^3 grid, 300-step fwd+bwd episode",
            "particles": sc.n, "grid_nodes": list(sc.dims), "episode_steps": T, "checkpoint_interval": K,
            "tape": ("every layer's state and P2G grid kept in HBM (no recomputation)" if K == 1
                     else f"checkpoints every {K} layers, segments recomputed"),
            "particle_order": ("upload order (Hilbert order of layer 0's cells)" if not order else
                               f"{order['segments']} run(s) of layers, each stored in the Hilbert order of its "
                               f"middle layer's cells (dmpm_sort_segments after the first warm-up episode)"),
            "order_segments": order,
            "physics": "plain: quadratic B-spline, Neo-Hookean, sticky walls (3 cells)",
            "precision": "FP32 state/arithmetic, FP64 det/log J/F^-T", "E": 5e4, "nu": 0.45, "dt": 1e-4,
            "active_nodes": gact, "peak_device_bytes": mem_peak,
            "lds_bytes": lds_bytes or 30 * 1024, "block": block or 256,
        },
        "windows": {"lds_blocks": win_lds, "fallback_blocks": win_fb},
        "autotune": tuned,
        "roofline": roofline,
        "step_roofline": {"bytes_per_particle_step": step_bytes / sc.n, "achieved_GBps":
                          step_bytes * value / sc.n / 1e9, "frac": step_frac},
        "cpu_baseline": cpu,
        "vs_cpu": vs_cpu,
        "kernels": kernels,
        **extra,
    }
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
