"""Launch-configuration autotuner.

The particle kernels stage a per-block box of grid nodes (the window) in LDS;
the LDS budget per block sets the largest window that fits (and how many
blocks share a CU), and the block size sets how many particles share one
window.  `tune` times a short forward+backward episode for each candidate
and returns the fastest configuration with the whole table.
"""
from __future__ import annotations

import time
from typing import Iterable, Optional

import numpy as np

from .sim import Physics, Simulator

DEFAULT_LDS = (16 * 1024, 24 * 1024, 30 * 1024, 32 * 1024, 48 * 1024)
DEFAULT_BLOCKS = (128, 256)


def tune(scene, physics: Physics, precision: str = "f32", steps: int = 20, checkpoint_interval: int = 10,
         lds_options: Iterable[int] = DEFAULT_LDS, block_options: Iterable[int] = DEFAULT_BLOCKS,
         seed_gx: Optional[np.ndarray] = None, reps: int = 3) -> dict:
    """Best (lds_bytes, block) for `scene` under `physics`; each candidate
    runs one warm-up and `reps` timed `steps`-step forward+backward episodes
    and is ranked by the fastest."""
    if seed_gx is None:
        seed_gx = 1e-3 * np.random.default_rng(0).normal(size=(scene.n, 3))
    table = []
    for lds in lds_options:
        for block in block_options:
            sim = Simulator(scene.dims, scene.dx, scene.grid_min, physics=physics, precision=precision,
                            checkpoint_interval=checkpoint_interval, lds_bytes=lds, block=block)
            sim.set_state(scene.state())
            sim.set_up(steps + 1)
            sim.set_last_grad(x=seed_gx)
            times = []
            for rep in range(1 + reps):
                sim.synchronize()
                t0 = time.perf_counter()
                sim.forward(0, keep_controls=True)
                sim.backward(0)
                sim.synchronize()
                times.append(time.perf_counter() - t0)
            dt = min(times[1:])
            sim.window_stats(reset=True)
            sim.count_windows(True)
            sim.forward(0, keep_controls=True)
            lds_blocks, fb_blocks = sim.window_stats()
            sim.close()
            table.append({"lds_bytes": lds, "block": block, "ms": 1e3 * dt,
                          "fallback_frac": fb_blocks / max(1, lds_blocks + fb_blocks)})
    best = min(table, key=lambda r: r["ms"])
    return {"lds_bytes": best["lds_bytes"], "block": best["block"], "table": table}


def segment_candidates(num_layers: int, checkpoint_interval: int = 1, lengths: Iterable[int] = (0, 25, 10)) -> list:
    """Segment counts for `tune_segments`: one per target segment length in
    layers (0 = the whole episode in one order), dropping counts that give
    the same segments once their length is rounded up to a multiple of the
    checkpoint interval (the rule of dmpm_sort_segments)."""
    K = max(1, int(checkpoint_interval))
    out, seen = [], set()
    for length in lengths:
        S = 1 if length <= 0 else max(1, -(-num_layers // length))
        P = -(-num_layers // S)
        P = -(-P // K) * K
        if P not in seen:
            seen.add(P)
            out.append(S)
    return out


def tune_segments(sim: Simulator, candidates: Optional[Iterable[int]] = None, reps: int = 1) -> dict:
    """Storage orders for the sim's episode.  For each candidate segment
    count S: dmpm_sort_segments(S), then `reps` timed forward + backward
    episodes from layer 0 (controls kept); the sim is left sorted with the
    fastest S.  The layers must have been computed (one episode run): each
    segment is sorted by the positions of its middle layer.  The episodes
    recompute the tape and the gradients from the same inputs."""
    if candidates is None:
        candidates = segment_candidates(sim.num_layers, sim.cfg.checkpoint_interval)
    table = []
    for S in candidates:
        sim.sort_segments(S)
        times = []
        for _ in range(max(1, reps)):
            sim.synchronize()
            t0 = time.perf_counter()
            sim.forward(0, keep_controls=True)
            sim.backward(0)
            sim.synchronize()
            times.append(time.perf_counter() - t0)
        table.append({"segments": int(S), "ms": 1e3 * min(times)})
    best = min(table, key=lambda r: r["ms"])
    if best is not table[-1]:
        sim.sort_segments(best["segments"])
    return {"segments": best["segments"], "table": table}
