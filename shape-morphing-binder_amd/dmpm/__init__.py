"""dmpm: MI355X-native differentiable MPM stepper (HIP kernels behind a C ABI).

    from dmpm import Simulator, Physics
    sim = Simulator(dims, dx, physics=Physics.plain(dt=1e-4), precision="f32")

The drop-in module for the reference's Python API is `diffmpm_bindings`
(next to this package); `dmpm.rollout` holds the torch.autograd.Function.
"""
from . import _lib
from .sim import (BatchedEpisode, EpisodeGraph, Physics, Simulator, calculate_volumes, lame_parameters, make_config,
                  p2g_mass)

__all__ = ["BatchedEpisode", "EpisodeGraph", "Physics", "Simulator", "calculate_volumes", "lame_parameters", "make_config", "p2g_mass", "_lib"]
