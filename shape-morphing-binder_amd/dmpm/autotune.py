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
