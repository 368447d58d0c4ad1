"""Synthetic particle bodies for the benchmark workloads (numpy, seeded).

The reference builds its particles from a watertight mesh on a lattice of
spacing point_dx = grid_dx / points_per_cell_cuberoot, with m = rho h^3,
vol = h^3, F = I (GeometryLoading.cpp:122-140, bind.cpp:531-538).  These
helpers build the same kind of particle set for analytic bodies (spheres and
ellipsoids) so that no mesh is needed: a lattice of spacing h, jittered by a
seeded uniform offset of +-jitter*h per axis, kept inside the body, and
trimmed (seeded) to exactly N points, kept in lattice order.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np


def lame(E: float, nu: float) -> tuple:
    """Lame parameters from Young's modulus and Poisson's ratio, in float32
    arithmetic like CalculateLameParameters (Elasticity.cpp:149-153)."""
    E32, nu32 = np.float32(E), np.float32(nu)
    one, two = np.float32(1), np.float32(2)
    lam = E32 * nu32 / ((one + nu32) * (one - two * nu32))
    mu = E32 / (two + two * nu32)
    return float(lam), float(mu)


@dataclass
class Body:
    center: Sequence[float]
    axes: Sequence[float]  # semi-axes (a sphere has three equal ones)

    def inside(self, p: np.ndarray) -> np.ndarray:
        c = np.asarray(self.center, np.float64)
        a = np.asarray(self.axes, np.float64)
        return (((p - c) / a) ** 2).sum(axis=1) <= 1.0

    @property
    def volume(self) -> float:
        a = np.asarray(self.axes, np.float64)
        return 4.0 / 3.0 * np.pi * float(np.prod(a))


def sample_body(body: Body, n: int, seed: int, h: Optional[float] = None, jitter: float = 0.25):
    """Jittered-lattice sampling of `body` trimmed to exactly `n` points.
    Returns (x (n,3) float64, h).  When h is None it is chosen so that the
    lattice holds slightly more than n points: h = (V/n)^(1/3) shrunk by 1%
    steps until enough points fall inside."""
    rng = np.random.default_rng(seed)
    c = np.asarray(body.center, np.float64)
    a = np.asarray(body.axes, np.float64)
    hh = (body.volume / n) ** (1.0 / 3.0) if h is None else float(h)
    for _ in range(64):
        lo = c - a - hh
        cnt = np.ceil((2 * a + 2 * hh) / hh).astype(int) + 1
        ii = [np.arange(k) for k in cnt]
        g = np.stack(np.meshgrid(*ii, indexing="ij"), axis=-1).reshape(-1, 3)
        pts = lo + (g + 0.5) * hh
        jit = rng.uniform(-jitter * hh, jitter * hh, size=pts.shape)
        pts = pts + jit
        pts = pts[body.inside(pts)]
        if pts.shape[0] >= n:
            keep = np.sort(rng.choice(pts.shape[0], size=n, replace=False))
            return pts[keep], hh
        if h is not None:
            raise ValueError(f"lattice spacing h={h} gives only {pts.shape[0]} < {n} points")
        hh *= 0.99
    raise RuntimeError("could not sample enough points")


def sine_velocity(x: np.ndarray, seed: int, modes: int = 8, amplitude: float = 0.5, kmax: int = 3):
    """Smooth random velocity field: sum of `modes` sine modes
    v(x) = sum_k a_k sin(2 pi n_k . x + phi_k), n_k in {-kmax..kmax}^3 \\ {0},
    a_k random directions, scaled so max |v| over the points is `amplitude`."""
    rng = np.random.default_rng(seed)
    v = np.zeros_like(x)
    for _ in range(modes):
        while True:
            nk = rng.integers(-kmax, kmax + 1, size=3)
            if np.any(nk != 0):
                break
        phi = rng.uniform(0, 2 * np.pi)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        v += np.sin(2 * np.pi * (x @ nk) + phi)[:, None] * d[None, :]
    vmax = np.linalg.norm(v, axis=1).max()
    if vmax > 0:
        v *= amplitude / vmax
    return v


@dataclass
class Scene:
    """A particle set plus the grid it lives on (all float64 numpy)."""

    x: np.ndarray
    v: np.ndarray
    F: np.ndarray
    C: np.ndarray
    m: np.ndarray
    vol: np.ndarray
    lam: np.ndarray
    mu: np.ndarray
    dims: tuple       # node counts (cells + 1 per axis)
    dx: float
    grid_min: tuple
    h: float

    @property
    def n(self) -> int:
        return self.x.shape[0]

    def state(self) -> dict:
        return dict(x=self.x, v=self.v, F=self.F, C=self.C, m=self.m, vol=self.vol,
                    lam=self.lam, mu=self.mu)


def make_scene(body: Body, n: int, cells: int, seed: int, E: float, nu: float, rho: float = 1e3,
               h: Optional[float] = None, velocity: str = "zero", vel_seed: Optional[int] = None,
               amplitude: float = 0.5, domain: float = 1.0) -> Scene:
    """Particles of `body` on a grid of `cells`^3 cells (cells+1 nodes per
    axis) spanning [0, domain]^3."""
    x, hh = sample_body(body, n, seed, h=h)
    dx = domain / cells
    if velocity == "zero":
        v = np.zeros_like(x)
    elif velocity == "sine":
        v = sine_velocity(x, seed if vel_seed is None else vel_seed, amplitude=amplitude)
    else:
        raise ValueError(velocity)
    F = np.broadcast_to(np.eye(3), (n, 3, 3)).copy()
    C = np.zeros((n, 3, 3))
    lam, mu = lame(E, nu)
    return Scene(x=x, v=v, F=F, C=C, m=np.full(n, rho * hh ** 3), vol=np.full(n, hh ** 3),
                 lam=np.full(n, lam), mu=np.full(n, mu), dims=(cells + 1,) * 3, dx=dx,
                 grid_min=(0.0, 0.0, 0.0), h=hh)


# --- the five workloads ------------------------------------------------------

def config1(seed: int = 1) -> Scene:
    """8,192 particles in a sphere (r=0.2, centre of the unit cube), 32^3
    cells, 2 particles per cell per axis (h = dx/2), E=1e4, nu=0.3, rho=1e3."""
    return make_scene(Body((0.5, 0.5, 0.5), (0.2, 0.2, 0.2)), 8192, 32, seed, 1e4, 0.3, h=1.0 / 64)


def config2(seed: int = 2) -> Scene:
    """65,536 particles in an ellipsoid (0.25, 0.2, 0.15), 64^3 cells,
    h = dx/2, E=1e4, nu=0.3."""
    return make_scene(Body((0.5, 0.5, 0.5), (0.25, 0.2, 0.15)), 65536, 64, seed, 1e4, 0.3, h=1.0 / 128)


def config3(seed: int = 3, n: int = 1048576, cells: int = 128) -> Scene:
    """1,048,576 particles in a sphere (r=0.3), 128^3 cells, E=5e4, nu=0.45,
    8-mode sine velocity field of amplitude 0.5."""
    return make_scene(Body((0.5, 0.5, 0.5), (0.3, 0.3, 0.3)), n, cells, seed, 5e4, 0.45,
                      velocity="sine", amplitude=0.5)


def config3_implosion(seed: int = 3, n: int = 1048576, cells: int = 128, rate: float = 22.0) -> Scene:
    """config3 with a radial implosion v -= rate (x - c) on top of its sine
    field.  The sine field of amplitude 0.5 alone leaves det F above 0.97 over
    the 300 steps of dt = 1e-4 (the body's pressure, lam log J with
    lam = 1.55e5, stops it early); converging at 22/s the 1M-particle body
    reaches det F = 0.30 (its minimum over the episode) and then rebounds,
    which is the compression the FP64 det/log J/F^-T of the FP32 path is
    there for."""
    sc = config3(seed, n, cells)
    sc.v = sc.v - rate * (sc.x - 0.5)
    return sc


def config5(seed: int = 5, n: int = 4194304, cells: int = 256) -> Scene:
    """4,194,304 particles in a sphere (r=0.3), 256^3 cells, E=5e4, nu=0.45."""
    return make_scene(Body((0.5, 0.5, 0.5), (0.3, 0.3, 0.3)), n, cells, seed, 5e4, 0.45,
                      velocity="sine", amplitude=0.5)


def config4_instances(seed: int = 4, count: int = 64, nmin: int = 4096, nmax: int = 131072):
    """The batched workload's instance table: particle counts log-uniform in
    [nmin, nmax], grids of 32/48/64/96 cells, ellipsoids with random semi-axes
    in [0.1, 0.3], E log-uniform in [5e3, 5e4], nu uniform in [0.2, 0.45].
    Returns a list of dicts (no particles; build each with instance_scene)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        cells = int(rng.choice([32, 48, 64, 96]))
        axes = tuple(rng.uniform(0.1, 0.3, size=3))
        E = float(np.exp(rng.uniform(np.log(5e3), np.log(5e4))))
        nu = float(rng.uniform(0.2, 0.45))
        out.append(dict(n=n, cells=cells, axes=axes, E=E, nu=nu, seed=1000 + i))
    return out


def instance_scene(spec: dict, amplitude: float = 0.3) -> Scene:
    """One instance of the batched workload, set in motion by a seeded sine
    velocity field (so the episode is not a rest state)."""
    return make_scene(Body((0.5, 0.5, 0.5), spec["axes"]), spec["n"], spec["cells"], spec["seed"],
                      spec["E"], spec["nu"], velocity="sine", amplitude=amplitude)
