"""Drop-in replacement for the reference's `diffmpm_bindings` module
(bind/bind.cpp:54-559), running the MPM stepper on the GPU through
libdmpm.so.

Same classes, method names, argument names, defaults, return types and error
types as the reference:

    OptInput, PointCloud, Grid, CompGraph,
    load_point_cloud_from_obj, calculate_lame_parameters, p2g,
    calculate_point_cloud_volumes, get_positions_from_pc

Differences a caller can observe:
  * PointClouds returned by CompGraph.get_point_cloud* read the layer from
    the device on each call (they are live views, as in the reference, but
    carry no host copy).
  * Extra keyword arguments (a superset): CompGraph(..., precision="f32",
    physics=None, checkpoint_interval=1, recompute_grids=False,
    deterministic=False);
    PointCloud.from_arrays(...);
    device=... on the *_torch getters; get_velocities/get_volumes.
"""
from __future__ import annotations

import os
import sys
from typing import Optional

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)

from dmpm import _lib as _L  # noqa: E402
from dmpm.sim import Physics, Simulator, calculate_volumes as _volumes, lame_parameters, p2g_mass  # noqa: E402

__all__ = ["OptInput", "PointCloud", "Grid", "CompGraph", "load_point_cloud_from_obj", "calculate_lame_parameters",
           "p2g", "calculate_point_cloud_volumes", "get_positions_from_pc"]


def _is_cuda(a) -> bool:
    return type(a).__module__.startswith("torch") and a.is_cuda


def _host(a):
    """numpy view of a host array or a CPU/CUDA tensor (the latter copied)."""
    if type(a).__module__.startswith("torch"):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def _vec3(v) -> np.ndarray:
    a = np.asarray(v, dtype=np.float32).reshape(3)
    return a.copy()


class OptInput:
    """The 21-field configuration of bind.cpp:30-52 (read/write attributes)."""

    _VEC = ("grid_min_point", "grid_max_point", "f_ext")

    def __init__(self):
        self.mpm_input_mesh_path = ""
        self.mpm_target_mesh_path = ""
        self.grid_dx = 0.0
        self.grid_min_point = np.zeros(3, np.float32)
        self.grid_max_point = np.zeros(3, np.float32)
        self.points_per_cell_cuberoot = 0
        self.lam = 0.0
        self.mu = 0.0
        self.p_density = 0.0
        self.dt = 0.0
        self.drag = 0.0
        self.f_ext = np.zeros(3, np.float32)
        self.num_animations = 0
        self.num_timesteps = 0
        self.control_stride = 0
        self.max_gd_iters = 0
        self.max_ls_iters = 0
        self.initial_alpha = 0.0
        self.gd_tol = 0.0
        self.smoothing_factor = 0.0
        self.current_episodes = 0

    def __setattr__(self, k, v):
        if k in OptInput._VEC:
            v = _vec3(v)
        elif k in ("grid_dx", "lam", "mu", "p_density", "dt", "drag", "initial_alpha", "gd_tol",
                   "smoothing_factor"):
            v = float(np.float32(v))
        elif k in ("points_per_cell_cuberoot", "num_animations", "num_timesteps", "control_stride", "max_gd_iters",
                   "max_ls_iters", "current_episodes"):
            v = int(v)
        elif k in ("mpm_input_mesh_path", "mpm_target_mesh_path"):
            v = str(v)
        else:
            raise AttributeError(f"'OptInput' object has no attribute '{k}'")
        object.__setattr__(self, k, v)


class Grid:
    """Grid(nx, ny, nz, dx, min_point) (bind.cpp:159-160, Grid.cpp:7-26).
    Holds the node masses written by p2g / calculate_point_cloud_volumes."""

    def __init__(self, nx: int, ny: int, nz: int, dx: float, min_point):
        self.dims = (int(nx), int(ny), int(nz))
        self.dx = float(np.float32(dx))
        self.min_point = _vec3(min_point)
        self.m = np.zeros(int(np.prod(self.dims)), np.float32)

    @property
    def num_nodes(self) -> int:
        return int(np.prod(self.dims))


class PointCloud:
    """Particle set.  Either host arrays (a PointCloud loaded or built by the
    caller) or a live view of one CompGraph layer on the device."""

    def __init__(self):
        raise TypeError("diffmpm_bindings.PointCloud: No constructor defined!")

    # -- construction (superset) ---------------------------------------------
    @classmethod
    def _host(cls, x, v, F, C, m, vol, lam, mu, dFc=None) -> "PointCloud":
        pc = object.__new__(cls)
        n = np.asarray(x).shape[0]
        pc._arr = dict(x=np.asarray(x, np.float32).reshape(n, 3).copy(),
                       v=np.asarray(v, np.float32).reshape(n, 3).copy(),
                       F=np.asarray(F, np.float32).reshape(n, 3, 3).copy(),
                       C=np.asarray(C, np.float32).reshape(n, 3, 3).copy(),
                       m=np.asarray(m, np.float32).reshape(n).copy(),
                       vol=np.asarray(vol, np.float32).reshape(n).copy(),
                       lam=np.asarray(lam, np.float32).reshape(n).copy(),
                       mu=np.asarray(mu, np.float32).reshape(n).copy(),
                       dFc=np.zeros((n, 3, 3), np.float32) if dFc is None
                       else np.asarray(dFc, np.float32).reshape(n, 3, 3).copy())
        pc._sim = None
        pc._layer = None
        return pc

    @classmethod
    def from_arrays(cls, x, m, vol, lam, mu, v=None, F=None, C=None, dFc=None) -> "PointCloud":
        """A PointCloud from arrays (the reference can only load meshes)."""
        n = np.asarray(x).shape[0]
        v = np.zeros((n, 3)) if v is None else v
        F = np.broadcast_to(np.eye(3), (n, 3, 3)) if F is None else F
        C = np.zeros((n, 3, 3)) if C is None else C
        b = lambda a: np.broadcast_to(np.asarray(a, np.float32), (n,))  # noqa: E731
        return cls._host(x, v, F, C, b(m), b(vol), b(lam), b(mu), dFc)

    @classmethod
    def _view(cls, cg: "CompGraph", layer: int) -> "PointCloud":
        pc = object.__new__(cls)
        pc._arr = None
        pc._sim = cg
        pc._layer = int(layer)
        return pc

    def _get(self, field: str) -> np.ndarray:
        if self._arr is not None:
            if field == "F_total":
                return self._arr["F"] + self._arr["dFc"]
            return self._arr[field]
        return self._sim._sim.get(self._layer, field, dtype=np.float32)

    @property
    def num_points(self) -> int:
        return self._arr["x"].shape[0] if self._arr is not None else self._sim._sim.n

    def _state(self) -> dict:
        return {k: self._get(k) for k in ("x", "v", "F", "C", "m", "vol", "lam", "mu", "dFc")}

    # -- reference getters (bind.cpp:84-157) ---------------------------------
    def get_positions(self):
        """Positions as a list of (3,) float32 arrays (PointCloud.cpp:433-445)."""
        return list(self._get("x"))

    def get_masses(self):
        return [float(a) for a in self._get("m")]

    def get_def_grads(self):
        return list(self._get("F"))

    def get_def_grads_morph(self) -> np.ndarray:
        return np.array(self._get("dFc"), dtype=np.float32)

    def get_def_grads_total(self) -> np.ndarray:
        return np.array(self._get("F_total"), dtype=np.float32)

    def get_positions_torch(self, requires_grad: bool = False, device=None):
        """(N,3) float32 tensor; a copy on the CPU as in the reference
        (bind.cpp:89-107), or on `device` without a host round trip."""
        import torch
        if device is not None and self._arr is None and torch.device(device).type == "cuda":
            t = self._sim._sim.get(self._layer, "x", dtype=np.float32, device=device)
        else:
            t = torch.from_numpy(np.array(self._get("x"), np.float32))
            if device is not None:
                t = t.to(device)
        return t.requires_grad_(requires_grad)

    def get_def_grads_total_torch(self, requires_grad: bool = False, device=None):
        import torch
        if device is not None and self._arr is None and torch.device(device).type == "cuda":
            t = self._sim._sim.get(self._layer, "F_total", dtype=np.float32, device=device)
        else:
            t = torch.from_numpy(np.array(self._get("F_total"), np.float32))
            if device is not None:
                t = t.to(device)
        return t.requires_grad_(requires_grad)

    # -- on-disk state (superset) --------------------------------------------
    def write_entire_point_cloud_to_binary_file(self, file_path: str):
        """The reference's cereal PointCloud file
        (PointCloud::WriteEntirePointCloudToBinaryFile, PointCloud.cpp:382-394).
        A CompGraph layer is written with its stress P, kept gradients and
        Adam state; a host PointCloud with those fields zero."""
        from dmpm.io import write_file
        if self._arr is None:
            self._sim._sim.write_layer(self._layer, file_path, "cereal")
        else:
            write_file(file_path, dict(self._arr), "cereal")

    @classmethod
    def read_entire_point_cloud_from_binary_file(cls, file_path: str) -> "PointCloud":
        """A host PointCloud from a cereal PointCloud file
        (PointCloud::ReadEntirePointCloudFromBinaryFile, PointCloud.cpp:396-410);
        raises RuntimeError where the reference prints and returns false."""
        from dmpm.io import read_file
        a = read_file(file_path, "cereal")
        return cls._host(a["x"], a["v"], a["F"], a["C"], a["m"], a["vol"], a["lam"], a["mu"], a["dFc"])

    # -- superset ------------------------------------------------------------
    def get_velocities(self) -> np.ndarray:
        return np.array(self._get("v"), np.float32)

    def get_volumes(self) -> np.ndarray:
        return np.array(self._get("vol"), np.float32)


class CompGraph:
    """CompGraph(pc, grid, target_grid) (bind.cpp:163-528, CompGraph.cpp).

    The simulation parameters start at CompGraph's defaults dt = 1/120,
    drag = 0.5, smoothing = 0.1, f_ext = 0 (CompGraph.h:80-83) and are set by
    run_optimization, as in the reference.  `physics` selects the step's
    physics (default: the reference's own); `precision` is "f32" (default,
    like the reference) or "f64".

    `segment_every` (no counterpart in the reference; 0 = off): after the
    first forward pass from layer 0 that follows a set-up, the device storage
    is re-sorted so that every run of about that many layers is stored in
    the Hilbert order of its middle layer (dmpm_sort_segments).  Results are
    unchanged beyond summation order; the later passes of the episode (the
    per-control-timestep forward and backward passes and line searches of
    the reference's loop) run faster, the particles no longer drifting out
    of order.  The re-sort costs about a third of a forward + backward
    episode, so it pays off from a few passes per set-up on."""

    def __init__(self, pc: PointCloud, grid: Grid, target_grid: Grid, precision: str = "f32",
                 physics: Optional[Physics] = None, checkpoint_interval: int = 1, recompute_grids: bool = False,
                 deterministic: bool = False, segment_every: int = 0):
        if pc is None or grid is None or target_grid is None:
            raise TypeError("CompGraph(pc, grid, target_grid): arguments must not be None")
        if segment_every < 0:
            raise ValueError("segment_every must be non-negative")
        if segment_every > 0 and deterministic:
            raise ValueError("segment orders are not used in deterministic mode")
        ph = physics or Physics.reference()
        self._sim = Simulator(grid.dims, grid.dx, tuple(float(a) for a in grid.min_point), physics=ph,
                              precision=precision, checkpoint_interval=checkpoint_interval,
                              recompute_grids=recompute_grids, deterministic=deterministic)
        self._sim.set_state(pc._state())
        self._sim.set_target_grid_mass(target_grid.m)
        self._grid = grid
        self._target = target_grid
        self._segment_every = int(segment_every)
        self._resort_pending = False

    # -- layers ----------------------------------------------------------------
    def get_num_layers(self) -> int:
        return self._sim.num_layers

    def get_point_cloud(self, layer_idx: int) -> PointCloud:
        if layer_idx < 0 or layer_idx >= self.get_num_layers():
            raise IndexError("Layer index out of range.")
        return PointCloud._view(self, layer_idx)

    def get_point_cloud_at_timestep(self, timestep_idx: int) -> PointCloud:
        n = self.get_num_layers()
        if timestep_idx < 0 or timestep_idx >= n:
            raise IndexError(f"Timestep index {timestep_idx} out of range [0, {n})")
        return PointCloud._view(self, timestep_idx)

    def set_up_comp_graph(self, num_layers: int):
        self._sim.set_up(num_layers)
        self._resort_pending = self._segment_every > 0

    def compute_forward_pass(self, start_layer: int, current_episode: int):
        self._sim.forward(start_layer)
        if self._resort_pending and start_layer == 0:
            self._resort_pending = False
            L = self._sim.num_layers
            self._sim.sort_segments(max(1, -(-L // self._segment_every)))

    def compute_backward_pass(self, control_layer: int):
        self._sim.backward(control_layer)

    def end_layer_mass_loss(self) -> float:
        return float(np.float32(self._sim.mass_loss()))

    # -- render gradients (bind.cpp:189-279, 359-528) ------------------------
    def _check_grads(self, dLdF, dLdx, n_expected=None, prefix="dLdF must be (N,3,3), got"):
        """float32, C-contiguous views of the two gradients: numpy arrays for
        host data, CUDA tensors left on their device (passed to the library
        as device pointers, never copied to the host)."""
        if _is_cuda(dLdF) and _is_cuda(dLdx):
            import torch
            F = dLdF.detach().to(torch.float32).contiguous()
            X = dLdx.detach().to(device=F.device, dtype=torch.float32).contiguous()
        else:
            F = np.ascontiguousarray(_host(dLdF), dtype=np.float32)
            X = np.ascontiguousarray(_host(dLdx), dtype=np.float32)
        if F.ndim != 3 or F.shape[1] != 3 or F.shape[2] != 3:
            raise RuntimeError(f"{prefix} ({','.join(str(s) for s in F.shape)})")
        if X.ndim != 2 or X.shape[1] != 3:
            raise RuntimeError(f"dLdx must be (N,3), got ({','.join(str(s) for s in X.shape)})")
        return F, X

    def accumulate_render_grads(self, dLdF_render, dLdx_render):
        if self.get_num_layers() <= 0:
            raise RuntimeError("CompGraph has no layers!")
        F, X = self._check_grads(dLdF_render, dLdx_render)
        n = self._sim.n
        if F.shape[0] != n or X.shape[0] != n:
            raise RuntimeError(f"Shape mismatch: point cloud has {n} points, but got dLdF with {F.shape[0]} "
                               f"and dLdx with {X.shape[0]}")
        self._sim.accumulate_last_grad(F=F, x=X)

    def set_render_gradients(self, dLdF_render, dLdx_render):
        """Stored and added to the last layer on every backward pass
        (bind.cpp:359-451).  CUDA tensors stay on the device: the library
        reads them in place and sums their norms there."""
        if self.get_num_layers() <= 0:
            raise RuntimeError("CompGraph has no layers!")
        F, X = self._check_grads(dLdF_render, dLdx_render, prefix="dLdF must be (N,3,3), got shape")
        if X.shape[0] != F.shape[0]:
            raise RuntimeError(f"Shape mismatch: dLdF has {F.shape[0]} points, but dLdx has {X.shape[0]}")
        import ctypes as C
        if isinstance(F, np.ndarray):
            pF, pX, dev = F.ctypes.data_as(C.c_void_p), X.ctypes.data_as(C.c_void_p), 0
        else:
            import torch
            torch.cuda.current_stream().synchronize()
            pF, pX, dev = C.c_void_p(F.data_ptr()), C.c_void_p(X.data_ptr()), 1
        _L.check(_L.lib().dmpm_set_render_gradients(self._sim.handle, int(F.shape[0]), pF, pX, _L.DTYPE_F32, dev))
        info = self.get_render_gradient_info()
        print(f"[C++] Stored render gradients for {F.shape[0]} points")
        print(f"[C++]   ||dL_render/dF|| = {info['grad_F_norm']}")
        print(f"[C++]   ||dL_render/dx|| = {info['grad_x_norm']}")

    def clear_render_gradients(self):
        _L.check(_L.lib().dmpm_clear_render_gradients(self._sim.handle))
        print("[C++] Cleared render gradients")

    def has_render_gradients(self) -> bool:
        return bool(self.get_render_gradient_info()["has_gradients"])

    def get_render_gradient_info(self) -> dict:
        import ctypes as C
        has, n = C.c_int(), C.c_int()
        nf, nx = C.c_double(), C.c_double()
        _L.check(_L.lib().dmpm_render_gradient_info(self._sim.handle, C.byref(has), C.byref(n), C.byref(nf),
                                                     C.byref(nx)))
        return {"has_gradients": bool(has.value), "num_points": int(n.value),
                "grad_F_norm": float(np.float32(nf.value)), "grad_x_norm": float(np.float32(nx.value))}

    # -- episodes and optimisation -------------------------------------------
    def promote_last_as_initial(self, carry_grid: bool = False):
        _L.check(_L.lib().dmpm_promote_last_as_initial(self._sim.handle, 1 if carry_grid else 0))

    def optimize_single_timestep(self, timestep_idx: int, max_gd_iters: int = 1, current_episode: int = 0,
                                 initial_alpha: float = 1.0, max_line_search_iters: int = 10):
        _L.check(_L.lib().dmpm_optimize_single_timestep(self._sim.handle, int(timestep_idx), int(max_gd_iters),
                                                        int(current_episode), float(initial_alpha),
                                                        int(max_line_search_iters)))

    def run_optimization(self, opt: OptInput):
        import ctypes as C
        fe = (C.c_double * 3)(*[float(a) for a in opt.f_ext])
        out = C.c_double()
        _L.check(_L.lib().dmpm_run_optimization(
            self._sim.handle, int(opt.num_timesteps), float(opt.dt), float(opt.drag), fe, int(opt.control_stride),
            int(opt.max_gd_iters), int(opt.max_ls_iters), float(opt.initial_alpha), float(opt.gd_tol),
            float(opt.smoothing_factor), int(opt.current_episodes), C.byref(out)))
        print(f"Final loss = {out.value}")


# --- free functions (bind.cpp:531-559) ----------------------------------------

def calculate_lame_parameters(young_mod: float, poisson: float):
    lam, mu = lame_parameters(young_mod, poisson)
    return float(np.float32(lam)), float(np.float32(mu))


def p2g(pc: PointCloud, grid: Grid):
    """P2G with dt = 0, drag = 0 into `grid` (node masses accumulate, as the
    reference's P2G adds onto the grid it is given)."""
    if pc is None or grid is None:
        raise RuntimeError("PointCloud or Grid is null.")
    gm = p2g_mass(grid.dims, grid.dx, tuple(float(a) for a in grid.min_point), pc._get("x"), pc._get("m"))
    grid.m = (grid.m + gm).astype(np.float32)


def calculate_point_cloud_volumes(pc: PointCloud, grid: Grid):
    """vol_p = m_p dx^3 / sum_i w_ip m_i after a dt = 0 P2G into `grid`
    (ForwardSimulation.cpp:234-271)."""
    if pc is None or grid is None:
        raise RuntimeError("PointCloud or Grid is null.")
    if pc._arr is None:
        raise RuntimeError("calculate_point_cloud_volumes needs a host PointCloud")
    # the P2G adds onto the masses the grid already holds, as the reference's
    vol, gm = _volumes(grid.dims, grid.dx, tuple(float(a) for a in grid.min_point), pc._get("x"), pc._get("m"),
                       grid_m=grid.m)
    pc._arr["vol"] = vol.astype(np.float32)
    grid.m = gm.astype(np.float32)


def get_positions_from_pc(pc: PointCloud):
    if pc is None:
        raise RuntimeError("PointCloud is null.")
    return pc.get_positions()


def load_point_cloud_from_obj(obj_path: str, opt: OptInput) -> PointCloud:
    """Particles on a lattice of spacing grid_dx / points_per_cell_cuberoot
    over the mesh's bounding box, kept where the generalised winding number
    (computed on the device) is >= 1/2; m = rho h^3, vol = h^3, F = I
    (GeometryLoading.cpp:10-83, 85-143, bind.cpp:531-538)."""
    from dmpm.mesh import read_obj, sample_inside
    if not os.path.exists(obj_path):
        raise RuntimeError("Failed to load PointCloud from: " + obj_path)
    V, F = read_obj(obj_path)
    if V.size == 0 or F.size == 0:
        raise RuntimeError("Failed to load PointCloud from: " + obj_path)
    h = np.float32(opt.grid_dx) / np.float32(opt.points_per_cell_cuberoot)
    x = sample_inside(V, F, float(h))
    n = x.shape[0]
    m = h * h * h * np.float32(opt.p_density)
    return PointCloud._host(x, np.zeros((n, 3)), np.broadcast_to(np.eye(3), (n, 3, 3)), np.zeros((n, 3, 3)),
                            np.full(n, m), np.full(n, m / np.float32(opt.p_density)), np.full(n, opt.lam),
                            np.full(n, opt.mu))
