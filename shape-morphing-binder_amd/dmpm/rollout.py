"""Differentiable rollout as a torch.autograd.Function.

    xT, vT, FT, CT = rollout(sim, state0, steps, controls={t: dFc_t, ...})

runs `steps` MPM steps on the GPU (libdmpm.so) from the layer-0 state
(x, v, F, C tensors plus m, vol, lam, mu) with per-step controls dFc_t added
to F at step t, and returns the final layer.  Its backward pass runs the
library's adjoint from the final layer to layer 0 -- with the simulator's
checkpoint interval K, only every K-th layer is kept and the rest are
recomputed segment by segment -- and returns dL/dx0, dL/dv0, dL/dF0, dL/dC0
and dL/d dFc_t for every control.  This replaces the reference's
store-every-step tape (CompGraph.cpp:30-40, 143-211) for gradient-based use
from PyTorch.
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np
import torch

from .sim import Simulator


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sim: Simulator, steps: int, statics: dict, ctl_layers: tuple, x0, v0, F0, C0, *controls):
        dev = x0.device
        sim.set_particles(x0, v0, F0, C0, statics["m"], statics["vol"], statics["lam"], statics["mu"])
        sim.set_up(steps + 1)
        for t, c in zip(ctl_layers, controls):
            sim.set_control(t, c)
            sim.request_control_grad(t, True)
        sim.forward(0, keep_controls=True)
        dt = np.float64 if sim.precision == "f64" else np.float32
        out = [sim.get(steps, k, dtype=dt, device=dev if dev.type == "cuda" else None) for k in ("x", "v", "F", "C")]
        out = [o if isinstance(o, torch.Tensor) else torch.from_numpy(o) for o in out]
        ctx.sim, ctx.steps, ctx.ctl_layers, ctx.dev = sim, steps, ctl_layers, dev
        ctx.dtype = x0.dtype
        ctx.generation = sim.generation
        return tuple(o.to(dtype=x0.dtype, device=dev) for o in out)

    @staticmethod
    def backward(ctx, gx, gv, gF, gC):
        sim = ctx.sim
        if sim.generation != ctx.generation:
            raise RuntimeError("the simulator ran another rollout (or forward pass) after this one; its tape is "
                               "gone -- use one Simulator per rollout that is still to be differentiated")
        n = sim.n
        zeros = lambda *s: torch.zeros((n,) + s, dtype=ctx.dtype, device=ctx.dev)  # noqa: E731
        gx = zeros(3) if gx is None else gx
        gv = zeros(3) if gv is None else gv
        gF = zeros(3, 3) if gF is None else gF
        gC = zeros(3, 3) if gC is None else gC
        sim.set_last_grad(x=gx, v=gv, F=gF, C_=gC)
        sim.backward(0)
        dt = np.float64 if sim.precision == "f64" else np.float32
        dev = ctx.dev if ctx.dev.type == "cuda" else None

        def g(layer_get):
            o = layer_get
            o = o if isinstance(o, torch.Tensor) else torch.from_numpy(o)
            return o.to(dtype=ctx.dtype, device=ctx.dev)

        g0 = [g(sim.get_grad(0, k, dtype=dt, device=dev)) for k in ("x", "v", "F", "C")]
        gctl = [g(sim.get_control_grad(t, dtype=dt, device=dev)) for t in ctx.ctl_layers]
        return (None, None, None, None, *g0, *gctl)


def rollout(sim: Simulator, state0: Dict[str, torch.Tensor], steps: int,
            controls: Optional[Dict[int, torch.Tensor]] = None):
    """Differentiable `steps`-step rollout; see the module docstring.
    state0 holds x (N,3), v (N,3), F (N,3,3), C (N,3,3) and the statics m,
    vol, lam, mu (N,).  Gradients flow to x, v, F, C and the controls."""
    controls = controls or {}
    ctl_layers = tuple(sorted(int(t) for t in controls))
    for t in ctl_layers:
        if not 0 <= t < steps:
            raise IndexError(f"control step {t} outside [0, {steps})")
    statics = {k: state0[k] for k in ("m", "vol", "lam", "mu")}
    return _Rollout.apply(sim, int(steps), statics, ctl_layers, state0["x"], state0["v"], state0["F"], state0["C"],
                          *[controls[t] for t in ctl_layers])
