"""Python front end of a dmpm_sim (the device-resident CompGraph).

Arrays may be numpy arrays (host) or torch tensors (host or device); outputs
are numpy arrays unless a torch device is asked for.  All computation runs in
libdmpm.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, replace
from typing import Optional

import numpy as np

from . import _lib as L


@dataclass
class Physics:
    """Step physics.  `reference()` is the reference's hard-wired physics
    (cubic B-spline, fixed-corotated, bulk viscosity 0.25, J trust region 0.6,
    ForwardSimulation.cpp:13), `plain()` the configs' plain method (quadratic,
    Neo-Hookean, sticky walls at 3 cells, nothing else)."""

    order: int = L.ORDER_CUBIC
    model: int = L.MODEL_FCR
    trust_region: int = 1
    zeta: float = 0.25
    jmin: float = 0.6
    wall_cells: int = 0
    adjoint: int = L.ADJOINT_REFERENCE
    dt: float = 1.0 / 120.0      # CompGraph.h:80-83 defaults
    drag: float = 0.5
    smoothing: float = 0.1
    f_ext: tuple = (0.0, 0.0, 0.0)

    @staticmethod
    def reference(**kw) -> "Physics":
        return replace(Physics(), **kw)

    @staticmethod
    def plain(**kw) -> "Physics":
        base = Physics(order=L.ORDER_QUADRATIC, model=L.MODEL_NEOHOOKEAN, trust_region=0, zeta=0.0,
                       wall_cells=3, adjoint=L.ADJOINT_EXACT, drag=0.0, smoothing=0.0)
        return replace(base, **kw)


def make_config(dims, dx, grid_min=(0.0, 0.0, 0.0), physics: Optional[Physics] = None, precision: str = "f32",
                checkpoint_interval: int = 1, lds_bytes: int = 0, block: int = 0, unfused: bool = False,
                deterministic: bool = False, recompute_grids: bool = False,
                adjoint_recompute: bool = False, split_nodes: int = 0) -> L.Config:
    ph = physics or Physics()
    c = L.Config()
    gm = (C.c_double * 3)(*[float(a) for a in grid_min])
    L.lib().dmpm_config_init(C.byref(c), int(dims[0]), int(dims[1]), int(dims[2]), float(dx), gm)
    c.dt, c.drag, c.smoothing = float(ph.dt), float(ph.drag), float(ph.smoothing)
    c.f_ext[:] = [float(a) for a in ph.f_ext]
    c.zeta, c.jmin = float(ph.zeta), float(ph.jmin)
    c.order, c.model, c.trust_region = int(ph.order), int(ph.model), int(ph.trust_region)
    c.wall_cells, c.adjoint = int(ph.wall_cells), int(ph.adjoint)
    c.precision = L.PREC_F64 if precision in ("f64", "float64", "fp64") else L.PREC_F32
    c.checkpoint_interval = int(checkpoint_interval)
    c.lds_bytes, c.block = int(lds_bytes), int(block)
    c.unfused = 1 if unfused else 0
    c.deterministic = 1 if deterministic else 0
    c.recompute_grids = 1 if recompute_grids else 0
    c.adjoint_recompute = 1 if adjoint_recompute else 0
    c.split_nodes = int(split_nodes)
    return c


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _on_one_device(arrays, code):
    """Borrowed pointers for several arrays that one C call takes with a
    single on_device flag: when any is a CUDA tensor, host arrays are copied
    to that device first; otherwise all stay on the host."""
    dev = None
    for a in arrays:
        if a is not None and _is_torch(a) and a.is_cuda:
            dev = a.device
            break
    if dev is not None:
        import torch
        arrays = [a if a is None or (_is_torch(a) and a.is_cuda) else
                  torch.as_tensor(np.asarray(a) if not _is_torch(a) else a).to(dev) for a in arrays]
    out = [_Arr(a, code) for a in arrays]
    flag = 1 if dev is not None else 0
    if any(a.ptr is not None and a.dev != flag for a in out):
        raise ValueError("arrays are on more than one device")
    return out, flag


class _Arr:
    """A borrowed pointer to caller data for one C call."""

    def __init__(self, a, dtype_code: Optional[int] = None, shape=None):
        self.keep = None
        if a is None:
            self.ptr, self.dev, self.code = None, 0, dtype_code if dtype_code is not None else L.DTYPE_F32
            return
        if _is_torch(a):
            import torch
            t = a.detach()
            if dtype_code is not None:
                t = t.to(torch.float64 if dtype_code == L.DTYPE_F64 else torch.float32)
            elif t.dtype not in (torch.float32, torch.float64):
                t = t.to(torch.float32)
            t = t.contiguous()
            if shape is not None and tuple(t.shape) != tuple(shape):
                t = t.reshape(shape)
            self.keep = t
            self.ptr = C.c_void_p(t.data_ptr())
            self.dev = 1 if t.is_cuda else 0
            self.code = L.DTYPE_F64 if t.dtype == torch.float64 else L.DTYPE_F32
            if self.dev:
                torch.cuda.current_stream().synchronize()
        else:
            arr = np.asarray(a)
            if dtype_code is not None:
                arr = arr.astype(np.float64 if dtype_code == L.DTYPE_F64 else np.float32, copy=False)
            elif arr.dtype not in (np.float32, np.float64):
                arr = arr.astype(np.float32)
            arr = np.ascontiguousarray(arr)
            if shape is not None:
                arr = arr.reshape(shape)
            self.keep = arr
            self.ptr = arr.ctypes.data_as(C.c_void_p)
            self.dev = 0
            self.code = L.DTYPE_F64 if arr.dtype == np.float64 else L.DTYPE_F32


_NCOMP = {L.FIELD_X: 3, L.FIELD_V: 3, L.FIELD_F: 9, L.FIELD_C: 9, L.FIELD_DFC: 9, L.FIELD_FTOTAL: 9,
          L.FIELD_M: 1, L.FIELD_VOL: 1, L.FIELD_LAM: 1, L.FIELD_MU: 1, L.FIELD_P: 9, L.FIELD_ADAM_M: 9,
          L.FIELD_ADAM_V: 9, L.FIELD_ADAM_VMAX: 9}
_SHAPE = {3: (3,), 9: (3, 3), 1: ()}
FIELDS = {"x": L.FIELD_X, "v": L.FIELD_V, "F": L.FIELD_F, "C": L.FIELD_C, "dFc": L.FIELD_DFC,
          "F_total": L.FIELD_FTOTAL, "m": L.FIELD_M, "vol": L.FIELD_VOL, "lam": L.FIELD_LAM, "mu": L.FIELD_MU,
          "P": L.FIELD_P, "adam_m": L.FIELD_ADAM_M, "adam_v": L.FIELD_ADAM_V, "adam_vmax": L.FIELD_ADAM_VMAX}


class Simulator:
    """Owns one dmpm_sim: layers of particle states on the device, the grid
    scratch, adjoint carry, controls and Adam state."""

    def __init__(self, dims, dx, grid_min=(0.0, 0.0, 0.0), physics: Optional[Physics] = None,
                 precision: str = "f32", checkpoint_interval: int = 1, lds_bytes: int = 0, block: int = 0,
                 unfused: bool = False, deterministic: bool = False, recompute_grids: bool = False,
                 adjoint_recompute: bool = False, split_nodes: int = 0):
        self.dims = tuple(int(d) for d in dims)
        self.dx = float(dx)
        self.grid_min = tuple(float(a) for a in grid_min)
        self.physics = physics or Physics()
        self.precision = "f64" if precision in ("f64", "float64", "fp64") else "f32"
        self.cfg = make_config(dims, dx, grid_min, self.physics, self.precision, checkpoint_interval, lds_bytes,
                               block, unfused, deterministic, recompute_grids, adjoint_recompute, split_nodes)
        h = C.c_void_p()
        L.check(L.lib().dmpm_create(C.byref(self.cfg), C.byref(h)))
        self._h = h
        self.n = 0
        # bumped by every call that replaces or changes the tape (particles,
        # layer count, parameters, controls, forward pass), so that a differentiable rollout can tell
        # whether the tape it recorded is still the sim's
        self.generation = 0

    # -- lifetime ------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            L.lib().dmpm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def use_stream(self, stream_ptr: int):
        L.check(L.lib().dmpm_set_stream(self._h, C.c_void_p(int(stream_ptr))))

    def synchronize(self):
        L.check(L.lib().dmpm_synchronize(self._h))

    def sort_particles(self, layer: int):
        """Reorder the device storage by the Hilbert order of `layer`'s
        positions (results unchanged up to summation order)."""
        L.check(L.lib().dmpm_sort_particles(self._h, int(layer)))

    def sort_segments(self, segments: int):
        """Reorder the device storage per episode segment: the layers split
        into `segments` equal runs (multiples of the checkpoint interval),
        each stored in the Hilbert order of its middle layer (results
        unchanged up to summation order; the caller's order everywhere)."""
        L.check(L.lib().dmpm_sort_segments(self._h, int(segments)))

    # -- inputs ----------------------------------------------------------------
    def set_particles(self, x, v, F, C_, m, vol, lam, mu, dFc=None):
        n = int(np.shape(x)[0]) if not _is_torch(x) else int(x.shape[0])
        code = L.DTYPE_F64 if self.precision == "f64" else L.DTYPE_F32
        arrs = [_Arr(a, code) for a in (x, v, F, C_, dFc, m, vol, lam, mu)]
        dev = arrs[0].dev
        if any(a.ptr is not None and a.dev != dev for a in arrs):
            raise ValueError("particle arrays must all be on the host or all on the device")
        L.check(L.lib().dmpm_set_particles(self._h, n, *[a.ptr for a in arrs], code, dev))
        self.n = n
        self.generation += 1

    def set_state(self, state: dict):
        lam = state.get("lam")
        mu = state.get("mu")
        self.set_particles(state["x"], state["v"], state["F"], state["C"], state["m"], state["vol"], lam, mu,
                           state.get("dFc"))

    def set_target_grid_mass(self, m):
        a = _Arr(m)
        L.check(L.lib().dmpm_set_target_grid_mass(self._h, a.ptr, a.code, a.dev))

    def set_sim_params(self, dt, drag, smoothing, f_ext=(0.0, 0.0, 0.0)):
        fe = (C.c_double * 3)(*[float(a) for a in f_ext])
        L.check(L.lib().dmpm_set_sim_params(self._h, float(dt), float(drag), float(smoothing), fe))
        self.generation += 1

    # -- CompGraph -------------------------------------------------------------
    def set_up(self, num_layers: int):
        L.check(L.lib().dmpm_set_up_comp_graph(self._h, int(num_layers)))
        self.generation += 1

    @property
    def num_layers(self) -> int:
        v = C.c_int()
        L.check(L.lib().dmpm_get_num_layers(self._h, C.byref(v)))
        return v.value

    def forward(self, start: int = 0, keep_controls: bool = False):
        """compute_forward_pass(start) -- which zeroes dFc of every later layer,
        as the reference does -- or, with keep_controls, a rollout that keeps
        each layer's control."""
        if keep_controls:
            L.check(L.lib().dmpm_rollout_forward(self._h, int(start)))
        else:
            L.check(L.lib().dmpm_compute_forward_pass(self._h, int(start)))
        self.generation += 1

    def mass_loss(self) -> float:
        v = C.c_double()
        L.check(L.lib().dmpm_end_layer_mass_loss(self._h, C.byref(v)))
        return v.value

    def backward(self, stop: int = 0):
        L.check(L.lib().dmpm_compute_backward_pass(self._h, int(stop)))

    def _out(self, nc, dtype, device):
        shape = (self.n,) + _SHAPE[nc]
        if device is not None:
            import torch
            t = torch.empty(shape, dtype=torch.float64 if dtype == np.float64 else torch.float32, device=device)
            return t, C.c_void_p(t.data_ptr()), 1 if t.is_cuda else 0
        a = np.empty(shape, dtype=dtype)
        return a, a.ctypes.data_as(C.c_void_p), 0

    def get(self, layer: int, field: str, dtype=np.float32, device=None):
        f = FIELDS[field]
        out, ptr, dev = self._out(_NCOMP[f], dtype, device)
        if dev:
            import torch
            torch.cuda.current_stream().synchronize()
        L.check(L.lib().dmpm_get_layer(self._h, int(layer), f, ptr,
                                       L.DTYPE_F64 if dtype == np.float64 else L.DTYPE_F32, dev))
        if dev:
            self.synchronize()
        return out

    def get_grad(self, layer: int, field: str, dtype=np.float32, device=None):
        f = FIELDS[field]
        out, ptr, dev = self._out(_NCOMP[f], dtype, device)
        if dev:
            import torch
            torch.cuda.current_stream().synchronize()
        L.check(L.lib().dmpm_get_layer_grad(self._h, int(layer), f, ptr,
                                            L.DTYPE_F64 if dtype == np.float64 else L.DTYPE_F32, dev))
        if dev:
            self.synchronize()
        return out

    def set_last_grad(self, x=None, v=None, F=None, C_=None):
        code = L.DTYPE_F64 if self.precision == "f64" else L.DTYPE_F32
        a, dev = _on_one_device((x, v, F, C_), code)
        L.check(L.lib().dmpm_set_last_layer_grad(self._h, *[t.ptr for t in a], code, dev))

    def accumulate_last_grad(self, F=None, x=None, v=None, C_=None):
        code = L.DTYPE_F64 if self.precision == "f64" else L.DTYPE_F32
        a, dev = _on_one_device((F, x, v, C_), code)
        L.check(L.lib().dmpm_accumulate_last_layer_grad(self._h, *[t.ptr for t in a], code, dev))

    def set_control(self, layer: int, dFc):
        self.generation += 1
        if dFc is None:
            L.check(L.lib().dmpm_set_control(self._h, int(layer), None, 0, 0))
            return
        code = L.DTYPE_F64 if self.precision == "f64" else L.DTYPE_F32
        a = _Arr(dFc, code)
        L.check(L.lib().dmpm_set_control(self._h, int(layer), a.ptr, a.code, a.dev))

    def set_adam_state(self, layer: int, m, v, vmax):
        code = L.DTYPE_F64 if self.precision == "f64" else L.DTYPE_F32
        a, dev = _on_one_device((m, v, vmax), code)
        L.check(L.lib().dmpm_set_adam_state(self._h, int(layer), a[0].ptr, a[1].ptr, a[2].ptr, code, dev))

    # -- on-disk state -------------------------------------------------------
    def write_layer(self, layer: int, path: str, format: str = "native"):
        """Write one layer (state, control, statics, stress, kept gradients,
        Adam state) to `path`: "native" (this library's checkpoint format) or
        "cereal" (the reference's PointCloud binary file)."""
        from .io import FORMATS
        L.check(L.lib().dmpm_write_layer(self._h, int(layer), str(path).encode(), FORMATS[format]))

    def read_particles(self, path: str, format: str = "native"):
        """Layer-0 particles (and Adam state) from a file written by
        write_layer or by the reference."""
        from .io import FORMATS
        L.check(L.lib().dmpm_read_particles(self._h, str(path).encode(), FORMATS[format]))
        n = C.c_int()
        L.check(L.lib().dmpm_num_particles(self._h, C.byref(n)))
        self.n = n.value

    def request_control_grad(self, layer: int, on: bool = True):
        L.check(L.lib().dmpm_request_control_grad(self._h, int(layer), 1 if on else 0))

    def get_control_grad(self, layer: int, dtype=np.float32, device=None):
        out, ptr, dev = self._out(9, dtype, device)
        if dev:
            import torch
            torch.cuda.current_stream().synchronize()
        L.check(L.lib().dmpm_get_control_grad(self._h, int(layer), ptr,
                                              L.DTYPE_F64 if dtype == np.float64 else L.DTYPE_F32, dev))
        if dev:
            self.synchronize()
        return out

    def control_grad_norm(self) -> float:
        v = C.c_double()
        L.check(L.lib().dmpm_control_grad_norm(self._h, C.byref(v)))
        return v.value

    def descend_adam(self, layer, alpha, grad_norm, beta1=0.9, beta2=0.999, eps=1e-3, timestep=1):
        L.check(L.lib().dmpm_descend_adam(self._h, int(layer), float(alpha), float(grad_norm), float(beta1),
                                          float(beta2), float(eps), int(timestep)))

    def timing(self):
        a, b = C.c_double(), C.c_double()
        L.check(L.lib().dmpm_last_timing(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def kernel_timer(self, capacity: int):
        """Start (capacity > 0 launches) or stop (0) per-kernel event timing."""
        L.check(L.lib().dmpm_kernel_timer(self._h, int(capacity)))

    def kernel_stats(self) -> dict:
        """{kernel name: (total ms, launches)} since kernel_timer was started."""
        out = {}
        for k, name in enumerate(L.KERNEL_NAMES):
            ms, c = C.c_double(), C.c_int()
            L.check(L.lib().dmpm_kernel_stats(self._h, k, C.byref(ms), C.byref(c)))
            out[name] = (ms.value, c.value)
        return out

    def count_windows(self, on: bool = True) -> None:
        """Window counting on or off (off by default: one device-scope atomic
        per block of every scatter launch)."""
        L.check(L.lib().dmpm_window_counters(self._h, 1 if on else 0))

    def window_stats(self, reset: bool = False):
        """(scatter blocks held in LDS, blocks that fell back to global
        atomics) counted while count_windows is on."""
        a, b = C.c_longlong(), C.c_longlong()
        L.check(L.lib().dmpm_window_stats(self._h, C.byref(a), C.byref(b), 1 if reset else 0))
        return a.value, b.value

    def memory(self):
        a, b = C.c_size_t(), C.c_size_t()
        L.check(L.lib().dmpm_memory(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value


def p2g_mass(dims, dx, grid_min, x, m, order=L.ORDER_CUBIC, precision="f32"):
    """Node masses of a dt = 0 P2G on a fresh grid (bind.cpp:546-549)."""
    cfg = make_config(dims, dx, grid_min, Physics(order=order), precision)
    code = L.DTYPE_F64 if precision == "f64" else L.DTYPE_F32
    ax, am = _Arr(x, code), _Arr(m, code)
    out = np.zeros(int(np.prod(dims)), np.float64 if code == L.DTYPE_F64 else np.float32)
    L.check(L.lib().dmpm_p2g_mass(C.byref(cfg), int(np.shape(x)[0]), ax.ptr, am.ptr,
                                  out.ctypes.data_as(C.c_void_p), code, ax.dev))
    return out


def calculate_volumes(dims, dx, grid_min, x, m, order=L.ORDER_CUBIC, precision="f32", grid_m=None):
    """CalculatePointCloudVolumes (ForwardSimulation.cpp:234-271): returns
    (vol per particle, node masses).  grid_m holds the node masses the grid
    had before (the P2G adds to them; zeros when None)."""
    cfg = make_config(dims, dx, grid_min, Physics(order=order), precision)
    code = L.DTYPE_F64 if precision == "f64" else L.DTYPE_F32
    ax, am = _Arr(x, code), _Arr(m, code)
    n = int(np.shape(x)[0])
    dt = np.float64 if code == L.DTYPE_F64 else np.float32
    vol = np.zeros(n, dt)
    gm = np.zeros(int(np.prod(dims)), dt) if grid_m is None else np.array(grid_m, dtype=dt).reshape(-1)
    L.check(L.lib().dmpm_calculate_volumes(C.byref(cfg), n, ax.ptr, am.ptr, vol.ctypes.data_as(C.c_void_p),
                                           gm.ctypes.data_as(C.c_void_p), code, ax.dev))
    return vol, gm


def lame_parameters(E, nu):
    lam, mu = C.c_double(), C.c_double()
    L.check(L.lib().dmpm_lame_parameters(float(E), float(nu), C.byref(lam), C.byref(mu)))
    return lam.value, mu.value


class EpisodeGraph:
    """One hipGraph holding a whole episode -- the forward pass from
    `start` (keeping per-layer controls) and the backward pass to `stop` --
    for each of `sims`, as concurrent branches.  launch() replays it.
    Between launches the controls of layers that had one at capture, the
    last-layer seed and render gradients set before the capture may be
    changed through the usual calls; calls that would reallocate the sims'
    buffers (new particles, a new layer count, new parameters) are refused
    until the graph is closed."""

    def __init__(self, sims, start: int = 0, stop: int = 0):
        arr = (C.c_void_p * len(sims))(*[s.handle for s in sims])
        h = C.c_void_p()
        L.check(L.lib().dmpm_graph_capture(arr, len(sims), int(start), int(stop), C.byref(h)))
        self._h = h
        self.sims = list(sims)

    def launch(self):
        L.check(L.lib().dmpm_graph_launch(self._h))

    def synchronize(self):
        L.check(L.lib().dmpm_graph_synchronize(self._h))

    def close(self):
        if getattr(self, "_h", None):
            L.lib().dmpm_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchedEpisode:
    """Many sims' episodes (forward from `start` keeping per-layer controls,
    backward to `stop`) as one hipGraph of batched launches: each kernel runs
    once over all instances.  The sims must share precision, physics
    templates, launch shape and episode structure.  Synchronise the sims
    before launch() and call synchronize() before reading them."""

    def __init__(self, sims, start: int = 0, stop: int = 0):
        for sm in sims:
            sm.synchronize()
        arr = (C.c_void_p * len(sims))(*[sm.handle for sm in sims])
        h = C.c_void_p()
        L.check(L.lib().dmpm_batch_capture(arr, len(sims), int(start), int(stop), C.byref(h)))
        self._h = h
        self.sims = list(sims)

    def launch(self):
        L.check(L.lib().dmpm_batch_launch(self._h))

    def synchronize(self):
        L.check(L.lib().dmpm_batch_synchronize(self._h))

    def close(self):
        if getattr(self, "_h", None):
            L.lib().dmpm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
