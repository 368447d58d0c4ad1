"""ctypes front end for the CPU checkers.  TEST INFRASTRUCTURE ONLY.

Two libraries:
  * ``liboracle.so``       -- the restatement (oracle/dmpm_oracle.cpp), float32
                              and float64;
  * ``_ref/libdmpm_ref.so`` -- the reference's own DiffMPMLib3D sources compiled
                              with ref_capi.cpp (float32 only; present only
                              where it was built from /root/reference).

Only tests/, the smoke check and bench.py's CPU-baseline leg import this
module.  The shipped library never does.

Array conventions: x, v (N,3); F, C, dFc (N,3,3) row-major; grid node
(i,j,k) at flat index i*ny*nz + j*nz + k (Grid.h:89-94).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libdmpm_ref.so")


class _CParams(C.Structure):
    _fields_ = [
        ("dims", C.c_int * 3),
        ("dx", C.c_double),
        ("grid_min", C.c_double * 3),
        ("dt", C.c_double),
        ("drag", C.c_double),
        ("smoothing", C.c_double),
        ("zeta", C.c_double),
        ("jmin", C.c_double),
        ("f_ext", C.c_double * 3),
        ("order", C.c_int),
        ("model", C.c_int),
        ("trust_region", C.c_int),
        ("wall_cells", C.c_int),
        ("adjoint", C.c_int),
    ]


@dataclass
class Params:
    """Step parameters.  Defaults are the reference's own physics
    (ForwardSimulation.cpp:13, Interpolation.h, Elasticity.cpp:45-55)."""

    dims: tuple = (16, 16, 16)
    dx: float = 1.0 / 15
    grid_min: tuple = (0.0, 0.0, 0.0)
    dt: float = 1e-3
    drag: float = 0.0
    smoothing: float = 0.0
    zeta: float = 0.25
    jmin: float = 0.6
    f_ext: tuple = (0.0, 0.0, 0.0)
    order: int = 3
    model: int = 0
    trust_region: int = 1
    wall_cells: int = 0
    adjoint: int = 1

    def c(self) -> _CParams:
        p = _CParams()
        p.dims[:] = [int(d) for d in self.dims]
        p.dx = float(self.dx)
        p.grid_min[:] = [float(a) for a in self.grid_min]
        p.dt, p.drag, p.smoothing = float(self.dt), float(self.drag), float(self.smoothing)
        p.zeta, p.jmin = float(self.zeta), float(self.jmin)
        p.f_ext[:] = [float(a) for a in self.f_ext]
        p.order, p.model = int(self.order), int(self.model)
        p.trust_region, p.wall_cells, p.adjoint = int(self.trust_region), int(self.wall_cells), int(self.adjoint)
        return p

    @property
    def G(self) -> int:
        return int(np.prod(self.dims))


def reference_params(**kw) -> Params:
    """The reference's physics: cubic, fixed-corotated, zeta=0.25, J_min=0.6,
    reference adjoint."""
    base = dict(order=3, model=0, trust_region=1, zeta=0.25, jmin=0.6, wall_cells=0, adjoint=1)
    base.update(kw)
    return Params(**base)


def plain_params(**kw) -> Params:
    """The configs' plain method: quadratic, Neo-Hookean, no bulk viscosity,
    no trust region, no smoothing or drag, sticky walls at 3 cells, exact adjoint."""
    base = dict(order=2, model=1, trust_region=0, zeta=0.0, smoothing=0.0, drag=0.0,
                wall_cells=3, adjoint=0)
    base.update(kw)
    return Params(**base)


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)


_FIELDS = ("x", "v", "F", "C", "dFc", "m", "vol", "lam", "mu")


def _state_args(s: dict, dt):
    out = []
    for k in _FIELDS:
        a = s.get(k)
        if a is None:
            out.append(None)
            continue
        out.append(np.ascontiguousarray(a, dtype=dt))
    return out


class Oracle:
    """The restatement, in float32 ("f32") or float64 ("f64")."""

    _lib = None

    def __init__(self, precision: str = "f64"):
        if Oracle._lib is None:
            if not os.path.exists(ORACLE_SO):
                raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run make -C oracle)")
            Oracle._lib = C.CDLL(ORACLE_SO)
            for s in ("f32", "f64"):
                getattr(Oracle._lib, f"orc_mass_loss_{s}").restype = C.c_double
                getattr(Oracle._lib, f"orc_dLdF_norm_{s}").restype = C.c_double
        self.p = precision
        self.dt = np.float32 if precision == "f32" else np.float64

    def _f(self, name):
        return getattr(self._lib, f"orc_{name}_{self.p}")

    def forward_step(self, prm: Params, s: dict) -> dict:
        n = s["x"].shape[0]
        args = _state_args(s, self.dt)
        o = {k: np.zeros((n, 3), self.dt) for k in ("x", "v")}
        o.update({k: np.zeros((n, 3, 3), self.dt) for k in ("F", "C", "P")})
        gm = np.zeros(prm.G, self.dt)
        gv = np.zeros((prm.G, 3), self.dt)
        cp = prm.c()
        self._f("forward_step")(C.byref(cp), n, *[_ptr(a) for a in args],
                                _ptr(o["x"]), _ptr(o["v"]), _ptr(o["F"]), _ptr(o["C"]), _ptr(o["P"]),
                                _ptr(gm), _ptr(gv))
        o["grid_m"], o["grid_v"] = gm, gv
        return o

    def backward_step(self, prm: Params, s: dict, adj1: dict) -> dict:
        n = s["x"].shape[0]
        args = _state_args(s, self.dt)
        a1 = [np.ascontiguousarray(adj1[k], dtype=self.dt) for k in ("x", "v", "F", "C")]
        o = {k: np.zeros((n, 3), self.dt) for k in ("x", "v")}
        o.update({k: np.zeros((n, 3, 3), self.dt) for k in ("F", "C", "dFc", "P")})
        cp = prm.c()
        self._f("backward_step")(C.byref(cp), n, *[_ptr(a) for a in args], *[_ptr(a) for a in a1],
                                 _ptr(o["x"]), _ptr(o["v"]), _ptr(o["F"]), _ptr(o["C"]), _ptr(o["dFc"]),
                                 _ptr(o["P"]))
        return o

    def pk1(self, model, F, lam, mu):
        F = np.ascontiguousarray(F, self.dt)
        n = F.shape[0]
        lam = np.ascontiguousarray(np.broadcast_to(lam, (n,)), self.dt)
        mu = np.ascontiguousarray(np.broadcast_to(mu, (n,)), self.dt)
        P = np.zeros_like(F)
        self._f("pk1")(int(model), n, _ptr(F), _ptr(lam), _ptr(mu), _ptr(P))
        return P

    def pk1_hvp(self, model, F, lam, mu, dF):
        F = np.ascontiguousarray(F, self.dt)
        dF = np.ascontiguousarray(dF, self.dt)
        n = F.shape[0]
        lam = np.ascontiguousarray(np.broadcast_to(lam, (n,)), self.dt)
        mu = np.ascontiguousarray(np.broadcast_to(mu, (n,)), self.dt)
        out = np.zeros_like(F)
        self._f("pk1_hvp")(int(model), n, _ptr(F), _ptr(lam), _ptr(mu), _ptr(dF), _ptr(out))
        return out

    def energy(self, model, F, lam, mu):
        F = np.ascontiguousarray(F, self.dt)
        n = F.shape[0]
        lam = np.ascontiguousarray(np.broadcast_to(lam, (n,)), self.dt)
        mu = np.ascontiguousarray(np.broadcast_to(mu, (n,)), self.dt)
        psi = np.zeros(n, self.dt)
        self._f("energy")(int(model), n, _ptr(F), _ptr(lam), _ptr(mu), _ptr(psi))
        return psi

    def svd(self, F):
        F = np.ascontiguousarray(F, self.dt)
        n = F.shape[0]
        U, V = np.zeros_like(F), np.zeros_like(F)
        s = np.zeros((n, 3), self.dt)
        self._f("svd")(n, _ptr(F), _ptr(U), _ptr(s), _ptr(V))
        return U, s, V

    def polar(self, F):
        F = np.ascontiguousarray(F, self.dt)
        R, S = np.zeros_like(F), np.zeros_like(F)
        self._f("polar")(F.shape[0], _ptr(F), _ptr(R), _ptr(S))
        return R, S

    def p2g_mass(self, prm: Params, x, m):
        x = np.ascontiguousarray(x, self.dt)
        m = np.ascontiguousarray(m, self.dt)
        gm = np.zeros(prm.G, self.dt)
        cp = prm.c()
        self._f("p2g_mass")(C.byref(cp), x.shape[0], _ptr(x), _ptr(m), _ptr(gm))
        return gm

    def volumes(self, prm: Params, x, m):
        x = np.ascontiguousarray(x, self.dt)
        m = np.ascontiguousarray(m, self.dt)
        vol = np.zeros(x.shape[0], self.dt)
        gm = np.zeros(prm.G, self.dt)
        cp = prm.c()
        self._f("volumes")(C.byref(cp), x.shape[0], _ptr(x), _ptr(m), _ptr(vol), _ptr(gm))
        return vol, gm

    def mass_loss(self, prm: Params, x, m, target_m):
        x = np.ascontiguousarray(x, self.dt)
        m = np.ascontiguousarray(m, self.dt)
        tm = np.ascontiguousarray(target_m, self.dt)
        gx = np.zeros_like(x)
        gm = np.zeros(prm.G, self.dt)
        gdm = np.zeros(prm.G, self.dt)
        cp = prm.c()
        loss = self._f("mass_loss")(C.byref(cp), x.shape[0], _ptr(x), _ptr(m), _ptr(tm), _ptr(gx),
                                    _ptr(gm), _ptr(gdm))
        return loss, gx, gm, gdm

    def adam(self, dLdF, dFc, mom, vec, vmax, alpha, grad_norm, beta1=0.9, beta2=0.999, eps=1e-3,
             timestep=1):
        arrs = [np.array(a, dtype=self.dt, order="C", copy=True) for a in (dFc, mom, vec, vmax)]
        g = np.ascontiguousarray(dLdF, self.dt)
        self._f("adam")(g.shape[0], _ptr(g), *[_ptr(a) for a in arrs], C.c_double(alpha),
                        C.c_double(grad_norm), C.c_double(beta1), C.c_double(beta2), C.c_double(eps),
                        int(timestep))
        return tuple(arrs)

    def dLdF_norm(self, dLdF):
        g = np.ascontiguousarray(dLdF, self.dt)
        return self._f("dLdF_norm")(g.shape[0], _ptr(g))

    def bspline(self, order, u):
        u = np.ascontiguousarray(u, self.dt)
        w, dw = np.zeros_like(u), np.zeros_like(u)
        self._f("bspline")(int(order), u.size, _ptr(u), _ptr(w), _ptr(dw))
        return w, dw

    # -- multi-step helpers (small sizes) -------------------------------------
    def rollout(self, prm: Params, s0: dict, steps: int, controls: Optional[list] = None) -> list:
        """Forward `steps` steps; returns the list of layer states (steps+1).
        controls[t] (N,3,3) is dFc of layer t (zero/None otherwise)."""
        layers = [dict(s0)]
        for t in range(steps):
            s = dict(layers[-1])
            s["dFc"] = None if controls is None else controls[t]
            o = self.forward_step(prm, s)
            nxt = {k: o[k] for k in ("x", "v", "F", "C")}
            for k in ("m", "vol", "lam", "mu"):
                nxt[k] = s0[k]
            layers[-1] = s
            layers.append(nxt)
        return layers

    def rollout_grad(self, prm: Params, layers: list, gT: dict, stop: int = 0) -> tuple:
        """Backpropagate adjoint gT (keys x, v, F, C) of the last layer to layer
        `stop`.  Returns (adjoint of layer stop, list of dFc gradients per layer)."""
        adj = {k: np.asarray(gT[k], self.dt) for k in ("x", "v", "F", "C")}
        gctl = [None] * (len(layers) - 1)
        for t in range(len(layers) - 2, stop - 1, -1):
            o = self.backward_step(prm, layers[t], adj)
            gctl[t] = o["dFc"]
            adj = {k: o[k] for k in ("x", "v", "F", "C")}
        return adj, gctl


def winding_numbers(P, V, F, chunk: int = 2048) -> np.ndarray:
    """Generalised winding number of points P (n,3) with respect to the
    triangle mesh (V, F): the sum of the triangles' signed solid angles over
    4 pi, each from tan(Omega/2) = a.(b x c) / (|a||b||c| + (a.b)|c| + (b.c)|a|
    + (c.a)|b|) (Van Oosterom & Strackee, 1983), in float64.  The reference
    gets the same quantity from libigl (GeometryLoading.cpp:61-77, sign type
    SIGNED_DISTANCE_TYPE_WINDING_NUMBER; inside <=> w >= 1/2)."""
    P = np.asarray(P, np.float64).reshape(-1, 3)
    V = np.asarray(V, np.float64)
    F = np.asarray(F, np.int64)
    out = np.zeros(P.shape[0])
    A, B, Cc = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    for s in range(0, P.shape[0], chunk):
        p = P[s:s + chunk, None, :]
        a, b, c = A[None] - p, B[None] - p, Cc[None] - p
        la, lb, lc = (np.linalg.norm(t, axis=2) for t in (a, b, c))
        num = np.einsum("ijk,ijk->ij", a, np.cross(b, c))
        den = la * lb * lc + np.einsum("ijk,ijk->ij", a, b) * lc + np.einsum("ijk,ijk->ij", b, c) * la + \
            np.einsum("ijk,ijk->ij", c, a) * lb
        out[s:s + chunk] = (2 * np.arctan2(num, den)).sum(axis=1) / (4 * np.pi)
    return out


class Ref:
    """The reference's own code (float32), through oracle/ref_capi.cpp."""

    _lib = None

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_SO)

    def __init__(self, threads: int = 1):
        if Ref._lib is None:
            if not os.path.exists(REF_SO):
                raise RuntimeError(f"reference library missing: {REF_SO} (run make -C oracle ref)")
            Ref._lib = C.CDLL(REF_SO)
            for f in ("ref_episode", "ref_mass_loss", "ref_dLdF_norm"):
                getattr(Ref._lib, f).restype = C.c_float
            Ref._lib.ref_max_threads.restype = C.c_int
        self.set_threads(threads)

    def set_threads(self, n: int):
        self._lib.ref_set_threads(int(n))

    def max_threads(self) -> int:
        return self._lib.ref_max_threads()

    @staticmethod
    def _grid(prm: Params):
        dims = (C.c_int * 3)(*[int(d) for d in prm.dims])
        mn = np.ascontiguousarray(prm.grid_min, np.float32)
        return dims, mn

    @staticmethod
    def _f3(a):
        return np.ascontiguousarray(a, np.float32)

    def _state(self, s):
        n = s["x"].shape[0]
        out = []
        for k in _FIELDS:
            a = s.get(k)
            if a is None and k == "dFc":
                a = np.zeros((n, 3, 3), np.float32)
            out.append(self._f3(a))
        return out

    def forward_step(self, prm: Params, s: dict) -> dict:
        n = s["x"].shape[0]
        dims, mn = self._grid(prm)
        args = self._state(s)
        o = {k: np.zeros((n, 3), np.float32) for k in ("x", "v")}
        o.update({k: np.zeros((n, 3, 3), np.float32) for k in ("F", "C", "P")})
        gm = np.zeros(prm.G, np.float32)
        gv = np.zeros((prm.G, 3), np.float32)
        fe = self._f3(prm.f_ext)
        self._lib.ref_forward_step(n, *[_ptr(a) for a in args], dims, C.c_float(prm.dx), _ptr(mn),
                                   C.c_float(prm.smoothing), C.c_float(prm.dt), C.c_float(prm.drag), _ptr(fe),
                                   _ptr(o["x"]), _ptr(o["v"]), _ptr(o["F"]), _ptr(o["C"]), _ptr(o["P"]),
                                   _ptr(gm), _ptr(gv))
        o["grid_m"], o["grid_v"] = gm, gv
        return o

    def backward_step(self, prm: Params, s: dict, adj1: dict) -> dict:
        n = s["x"].shape[0]
        dims, mn = self._grid(prm)
        args = self._state(s)
        a1 = [self._f3(adj1[k]) for k in ("x", "v", "F", "C")]
        o = {k: np.zeros((n, 3), np.float32) for k in ("x", "v")}
        o.update({k: np.zeros((n, 3, 3), np.float32) for k in ("F", "C", "P")})
        fe = self._f3(prm.f_ext)
        self._lib.ref_backward_step(n, *[_ptr(a) for a in args], dims, C.c_float(prm.dx), _ptr(mn),
                                    C.c_float(prm.smoothing), C.c_float(prm.dt), C.c_float(prm.drag), _ptr(fe),
                                    *[_ptr(a) for a in a1], _ptr(o["x"]), _ptr(o["v"]), _ptr(o["F"]),
                                    _ptr(o["C"]), _ptr(o["P"]))
        return o

    def episode(self, prm: Params, s: dict, target_m, num_layers: int, control_layer: int = 0,
                render_gF=None, render_gx=None) -> dict:
        n = s["x"].shape[0]
        dims, mn = self._grid(prm)
        args = self._state(s)
        tm = self._f3(target_m)
        o = {k: np.zeros((n, 3), np.float32) for k in ("x", "v", "gx", "gv")}
        o.update({k: np.zeros((n, 3, 3), np.float32) for k in ("F", "C", "gF", "gC")})
        fe = self._f3(prm.f_ext)
        rgF = None if render_gF is None else self._f3(render_gF)
        rgx = None if render_gx is None else self._f3(render_gx)
        loss = self._lib.ref_episode(n, *[_ptr(a) for a in args], dims, C.c_float(prm.dx), _ptr(mn), _ptr(tm),
                                     int(num_layers), int(control_layer), C.c_float(prm.smoothing),
                                     C.c_float(prm.dt), C.c_float(prm.drag), _ptr(fe), _ptr(rgF), _ptr(rgx),
                                     _ptr(o["x"]), _ptr(o["v"]), _ptr(o["F"]), _ptr(o["C"]),
                                     _ptr(o["gx"]), _ptr(o["gv"]), _ptr(o["gF"]), _ptr(o["gC"]))
        o["loss"] = float(loss)
        return o

    def rollout(self, prm: Params, s: dict, steps: int, seed: Optional[dict] = None) -> dict:
        """`steps` ForwardTimeStep calls from layer 0 (dFc of layer 0 as
        given, zero after) and, with `seed` (keys x, v, F, C: the last layer's
        adjoint), `steps` Back_Timestep calls.  Returns the last layer's x, v,
        F, C and, with a seed, layer 0's gradients gx, gv, gF, gC."""
        n = s["x"].shape[0]
        dims, mn = self._grid(prm)
        args = self._state(s)
        o = {k: np.zeros((n, 3), np.float32) for k in ("x", "v", "gx", "gv")}
        o.update({k: np.zeros((n, 3, 3), np.float32) for k in ("F", "C", "gF", "gC")})
        sd = [None] * 4 if seed is None else [self._f3(seed[k]) for k in ("x", "v", "F", "C")]
        fe = self._f3(prm.f_ext)
        self._lib.ref_rollout(n, *[_ptr(a) for a in args], dims, C.c_float(prm.dx), _ptr(mn), int(steps),
                              C.c_float(prm.smoothing), C.c_float(prm.dt), C.c_float(prm.drag), _ptr(fe),
                              *[_ptr(a) for a in sd], _ptr(o["x"]), _ptr(o["v"]), _ptr(o["F"]), _ptr(o["C"]),
                              _ptr(o["gx"]), _ptr(o["gv"]), _ptr(o["gF"]), _ptr(o["gC"]))
        if seed is None:
            for k in ("gx", "gv", "gF", "gC"):
                del o[k]
        return o

    def optimize(self, prm: Params, s: dict, target_m, num_steps: int, control_stride: int, max_gd_iters: int,
                 max_ls_iters: int, initial_alpha: float, gd_tol: float, single: Optional[dict] = None) -> dict:
        """CompGraph::OptimizeDefGradControlSequence with prm's dt, drag,
        f_ext and smoothing, then (with `single` = dict(t, iters, alpha, ls))
        CompGraph::OptimizeSingleTimestep on the same graph.  Returns the
        end-layer loss after each, every layer's dFc after each, and the
        single-timestep layer's Adam state."""
        n = s["x"].shape[0]
        dims, mn = self._grid(prm)
        args = self._state(s)
        del args[4]  # no dFc
        fe = self._f3(prm.f_ext)
        loss = np.zeros(2, np.float32)
        d_run = np.zeros((num_steps, n, 3, 3), np.float32)
        d_single = np.zeros((num_steps, n, 3, 3), np.float32)
        adam = np.zeros((3, n, 3, 3), np.float32)
        sg = single or dict(t=0, iters=0, alpha=1.0, ls=1)
        self._lib.ref_optimize(n, *[_ptr(a) for a in args], dims, C.c_float(prm.dx), _ptr(mn),
                               _ptr(self._f3(target_m)), int(num_steps), C.c_float(prm.dt), C.c_float(prm.drag),
                               _ptr(fe), int(control_stride), int(max_gd_iters), int(max_ls_iters),
                               C.c_float(initial_alpha), C.c_float(gd_tol), C.c_float(prm.smoothing), int(sg["t"]),
                               int(sg["iters"]), C.c_float(sg["alpha"]), int(sg["ls"]), _ptr(loss), _ptr(d_run),
                               _ptr(d_single), _ptr(adam))
        return dict(loss_run=float(loss[0]), loss_single=float(loss[1]), dFc_run=d_run, dFc_single=d_single,
                    adam_m=adam[0], adam_v=adam[1], adam_vmax=adam[2])

    def mass_loss(self, prm: Params, s: dict, target_m):
        n = s["x"].shape[0]
        dims, mn = self._grid(prm)
        args = self._state(s)
        del args[4]  # no dFc
        gx = np.zeros((n, 3), np.float32)
        gm = np.zeros(prm.G, np.float32)
        gdm = np.zeros(prm.G, np.float32)
        loss = self._lib.ref_mass_loss(n, *[_ptr(a) for a in args], dims, C.c_float(prm.dx), _ptr(mn),
                                       _ptr(self._f3(target_m)), _ptr(gx), _ptr(gm), _ptr(gdm))
        return float(loss), gx, gm, gdm

    def p2g_mass(self, prm: Params, x, m):
        dims, mn = self._grid(prm)
        gm = np.zeros(prm.G, np.float32)
        x, m = self._f3(x), self._f3(m)
        self._lib.ref_p2g_mass(x.shape[0], _ptr(x), _ptr(m), dims, C.c_float(prm.dx), _ptr(mn), _ptr(gm))
        return gm

    def volumes(self, prm: Params, x, m):
        dims, mn = self._grid(prm)
        x, m = self._f3(x), self._f3(m)
        vol = np.zeros(x.shape[0], np.float32)
        gm = np.zeros(prm.G, np.float32)
        self._lib.ref_volumes(x.shape[0], _ptr(x), _ptr(m), dims, C.c_float(prm.dx), _ptr(mn), _ptr(vol), _ptr(gm))
        return vol, gm

    def adam(self, dLdF, dFc, mom, vec, vmax, alpha, grad_norm, beta1=0.9, beta2=0.999, eps=1e-3, timestep=1):
        arrs = [np.array(a, dtype=np.float32, order="C", copy=True) for a in (dFc, mom, vec, vmax)]
        g = self._f3(dLdF)
        self._lib.ref_adam(g.shape[0], _ptr(g), *[_ptr(a) for a in arrs], C.c_float(alpha), C.c_float(grad_norm),
                           C.c_float(beta1), C.c_float(beta2), C.c_float(eps), int(timestep))
        return tuple(arrs)

    def dLdF_norm(self, dLdF):
        g = self._f3(dLdF)
        return float(self._lib.ref_dLdF_norm(g.shape[0], _ptr(g)))

    def _mat(self, fn, F, lam, mu, extra=None):
        F = self._f3(F)
        n = F.shape[0]
        lam = self._f3(np.broadcast_to(lam, (n,)))
        mu = self._f3(np.broadcast_to(mu, (n,)))
        out = np.zeros_like(F) if fn != "ref_energy_fcr" else np.zeros(n, np.float32)
        a = [n, _ptr(F), _ptr(lam), _ptr(mu)]
        if extra is not None:
            a.append(_ptr(self._f3(extra)))
        getattr(self._lib, fn)(*a, _ptr(out))
        return out

    def pk_fcr(self, F, lam, mu):
        return self._mat("ref_pk_fcr", F, lam, mu)

    def pk_fcr_hvp(self, F, lam, mu, dF):
        return self._mat("ref_pk_fcr_hvp", F, lam, mu, dF)

    def energy_fcr(self, F, lam, mu):
        return self._mat("ref_energy_fcr", F, lam, mu)

    def polar(self, F):
        F = self._f3(F)
        R, S = np.zeros_like(F), np.zeros_like(F)
        self._lib.ref_polar(F.shape[0], _ptr(F), _ptr(R), _ptr(S))
        return R, S

    def time_steps(self, prm: Params, s: dict, steps: int, seed_gx) -> tuple:
        """Seconds spent in `steps` ForwardTimeStep calls and `steps`
        Back_Timestep calls of the reference (its own physics)."""
        n = s["x"].shape[0]
        dims, mn = self._grid(prm)
        args = self._state(s)
        del args[4]  # no dFc
        fe = self._f3(prm.f_ext)
        f, b = C.c_double(), C.c_double()
        self._lib.ref_time_steps(n, *[_ptr(a) for a in args], dims, C.c_float(prm.dx), _ptr(mn), int(steps),
                                 C.c_float(prm.smoothing), C.c_float(prm.dt), C.c_float(prm.drag), _ptr(fe),
                                 _ptr(self._f3(seed_gx)), C.byref(f), C.byref(b))
        return f.value, b.value

    def lame(self, E, nu):
        lam, mu = C.c_float(), C.c_float()
        self._lib.ref_lame(C.c_float(E), C.c_float(nu), C.byref(lam), C.byref(mu))
        return lam.value, mu.value

    def load_obj(self, path: str, point_dx: float, density: float, lam: float, mu: float) -> dict:
        """GeometryLoading::LoadMPMPointCloudFromObj: positions, masses and
        volumes of the points it samples from the OBJ mesh."""
        f = self._lib.ref_load_obj
        f.restype = C.c_int
        args = (str(path).encode(), C.c_float(point_dx), C.c_float(density), C.c_float(lam), C.c_float(mu))
        n = f(*args, 0, None, None, None)
        if n < 0:
            raise RuntimeError(f"the reference could not load {path}")
        o = dict(x=np.zeros((n, 3), np.float32), m=np.zeros(n, np.float32), vol=np.zeros(n, np.float32))
        f(*args, int(n), _ptr(o["x"]), _ptr(o["m"]), _ptr(o["vol"]))
        return o

    CEREAL_FIELDS = ("x", "v", "F", "C", "P", "m", "vol", "lam", "mu", "dFc", "dLdx", "dLdv", "dLdF",
                     "momentum", "vector", "vector_max")

    def write_cereal(self, path: str, a: dict):
        """PointCloud::WriteEntirePointCloudToBinaryFile of a point cloud
        holding the arrays `a` (keys CEREAL_FIELDS)."""
        args = [self._f3(a[k]) for k in self.CEREAL_FIELDS]
        n = args[0].shape[0]
        self._lib.ref_write_cereal(int(n), *[_ptr(x) for x in args], str(path).encode())

    def read_cereal(self, path: str) -> dict:
        """PointCloud::ReadEntirePointCloudFromBinaryFile: x, F, P, m and
        vector_max of the points read."""
        self._lib.ref_read_cereal.restype = C.c_int
        n = self._lib.ref_read_cereal(str(path).encode(), 0, None, None, None, None, None)
        if n < 0:
            raise RuntimeError(f"the reference could not read {path}")
        o = dict(x=np.zeros((n, 3), np.float32), F=np.zeros((n, 3, 3), np.float32),
                 P=np.zeros((n, 3, 3), np.float32), m=np.zeros(n, np.float32),
                 vector_max=np.zeros((n, 3, 3), np.float32))
        self._lib.ref_read_cereal(str(path).encode(), int(n), _ptr(o["x"]), _ptr(o["F"]), _ptr(o["P"]),
                                  _ptr(o["m"]), _ptr(o["vector_max"]))
        return o
