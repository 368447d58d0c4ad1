"""ctypes binding of libdmpm.so (include/dmpm.h).

The library must have been built (``python shape-morphing-binder_amd/dmpm/build.py``
or ``__graft_entry__.build()``); there is no fallback: every entry point of
this package runs on the GPU through this library or fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# DMPM_LIB names another build of the library (a variant built with other
# compile flags, say); by default the in-tree one
LIB_PATH = os.environ.get("DMPM_LIB") or os.path.join(HERE, "libdmpm.so")

DMPM_OK = 0
DTYPE_F32, DTYPE_F64 = 0, 1
# bytes of the device array behind dmpm_calculate_volumes_noalloc (include/dmpm.h)
DMPM_NOALLOC_SCRATCH_BYTES = 16 << 20
PREC_F32, PREC_F64 = 0, 1
ORDER_QUADRATIC, ORDER_CUBIC = 2, 3
MODEL_FCR, MODEL_NEOHOOKEAN = 0, 1
ADJOINT_EXACT, ADJOINT_REFERENCE = 0, 1
(FIELD_X, FIELD_V, FIELD_F, FIELD_C, FIELD_DFC, FIELD_FTOTAL, FIELD_M, FIELD_VOL, FIELD_LAM,
 FIELD_MU, FIELD_P, FIELD_ADAM_M, FIELD_ADAM_V, FIELD_ADAM_VMAX) = range(14)

(KERNEL_P2G, KERNEL_G2P, KERNEL_ADJ_SCATTER, KERNEL_ADJ_GATHER, KERNEL_GRID_CLEAR, KERNEL_STEP,
 KERNEL_ADJ_FUSED) = range(7)
KERNEL_NAMES = ["p2g", "g2p", "adj_scatter", "adj_gather", "grid_clear", "step", "adj_fused"]

ERR_NAMES = {-1: "ERR_ARG", -2: "ERR_INDEX", -3: "ERR_HIP", -4: "ERR_STATE", -5: "ERR_NOMEM"}


class Config(C.Structure):
    _fields_ = [
        ("dims", C.c_int * 3),
        ("dx", C.c_double),
        ("grid_min", C.c_double * 3),
        ("dt", C.c_double),
        ("drag", C.c_double),
        ("smoothing", C.c_double),
        ("f_ext", C.c_double * 3),
        ("zeta", C.c_double),
        ("jmin", C.c_double),
        ("order", C.c_int),
        ("model", C.c_int),
        ("trust_region", C.c_int),
        ("wall_cells", C.c_int),
        ("adjoint", C.c_int),
        ("precision", C.c_int),
        ("checkpoint_interval", C.c_int),
        ("lds_bytes", C.c_int),
        ("block", C.c_int),
        ("unfused", C.c_int),
        ("deterministic", C.c_int),
        ("recompute_grids", C.c_int),
        ("adjoint_recompute", C.c_int),
        ("split_nodes", C.c_int),
    ]


class ParticleArrays(C.Structure):
    """dmpm_particle_arrays: host arrays of a particle file (NULL = absent)."""
    FIELDS = ("x", "v", "F", "C", "P", "m", "vol", "lam", "mu", "dFc", "dLdx", "dLdv", "dLdF",
              "adam_m", "adam_v", "adam_vmax")
    _fields_ = [(f, C.c_void_p) for f in FIELDS]


FILE_NATIVE, FILE_CEREAL = 0, 1


class DmpmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class DmpmIndexError(DmpmError, IndexError):
    pass


_lib = None


def lib() -> C.CDLL:
    """Load libdmpm.so (once)."""
    global _lib
    if _lib is not None:
        return _lib
    # torch bundles its own libamdhip64.so.7; load it first so that this
    # library binds to the same HIP runtime (one runtime per process).
    if os.environ.get("DMPM_NO_TORCH", "0") != "1":
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libdmpm.so is not built ({LIB_PATH}); run `python shape-morphing-binder_amd/dmpm/build.py`")
    L = C.CDLL(LIB_PATH)
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    cfgp = C.POINTER(Config)
    proto = {
        "dmpm_config_init": (None, [cfgp, i, i, i, d, dp]),
        "dmpm_last_error": (C.c_char_p, []),
        "dmpm_version": (C.c_char_p, []),
        "dmpm_create": (i, [cfgp, C.POINTER(vp)]),
        "dmpm_destroy": (i, [vp]),
        "dmpm_set_stream": (i, [vp, vp]),
        "dmpm_synchronize": (i, [vp]),
        "dmpm_set_particles": (i, [vp, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, i, i]),
        "dmpm_num_particles": (i, [vp, ip]),
        "dmpm_set_target_grid_mass": (i, [vp, vp, i, i]),
        "dmpm_set_sim_params": (i, [vp, d, d, d, dp]),
        "dmpm_set_up_comp_graph": (i, [vp, i]),
        "dmpm_get_num_layers": (i, [vp, ip]),
        "dmpm_compute_forward_pass": (i, [vp, i]),
        "dmpm_rollout_forward": (i, [vp, i]),
        "dmpm_end_layer_mass_loss": (i, [vp, dp]),
        "dmpm_compute_backward_pass": (i, [vp, i]),
        "dmpm_get_layer": (i, [vp, i, i, vp, i, i]),
        "dmpm_get_layer_grad": (i, [vp, i, i, vp, i, i]),
        "dmpm_accumulate_last_layer_grad": (i, [vp, vp, vp, vp, vp, i, i]),
        "dmpm_set_last_layer_grad": (i, [vp, vp, vp, vp, vp, i, i]),
        "dmpm_set_render_gradients": (i, [vp, i, vp, vp, i, i]),
        "dmpm_clear_render_gradients": (i, [vp]),
        "dmpm_render_gradient_info": (i, [vp, ip, ip, dp, dp]),
        "dmpm_promote_last_as_initial": (i, [vp, i]),
        "dmpm_sort_particles": (i, [vp, i]),
        "dmpm_sort_segments": (i, [vp, i]),
        "dmpm_set_control": (i, [vp, i, vp, i, i]),
        "dmpm_request_control_grad": (i, [vp, i, i]),
        "dmpm_get_control_grad": (i, [vp, i, vp, i, i]),
        "dmpm_control_grad_norm": (i, [vp, dp]),
        "dmpm_descend_adam": (i, [vp, i, d, d, d, d, d, i]),
        "dmpm_optimize_single_timestep": (i, [vp, i, i, i, d, i]),
        "dmpm_run_optimization": (i, [vp, i, d, d, dp, i, i, i, d, d, d, i, dp]),
        "dmpm_single_timestep_adam_count": (i, [ip, i]),
        "dmpm_p2g_mass": (i, [cfgp, i, vp, vp, vp, i, i]),
        "dmpm_calculate_volumes": (i, [cfgp, i, vp, vp, vp, vp, i, i]),
        "dmpm_calculate_volumes_noalloc": (i, [cfgp, i, vp, vp, vp, vp, i, vp]),
        "dmpm_lame_parameters": (i, [d, d, dp, dp]),
        "dmpm_last_timing": (i, [vp, dp, dp]),
        "dmpm_memory": (i, [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
        "dmpm_kernel_timer": (i, [vp, i]),
        "dmpm_window_stats": (i, [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), i]),
        "dmpm_window_counters": (i, [vp, i]),
        "dmpm_graph_capture": (i, [C.POINTER(vp), i, i, i, C.POINTER(vp)]),
        "dmpm_graph_launch": (i, [vp]),
        "dmpm_graph_synchronize": (i, [vp]),
        "dmpm_graph_destroy": (i, [vp]),
        "dmpm_kernel_stats": (i, [vp, i, dp, ip]),
        "dmpm_get_config": (i, [vp, cfgp]),
        "dmpm_set_adam_state": (i, [vp, i, vp, vp, vp, i, i]),
        "dmpm_file_info": (i, [C.c_char_p, i, ip, ip]),
        "dmpm_write_file": (i, [C.c_char_p, i, i, C.POINTER(ParticleArrays), i]),
        "dmpm_read_file": (i, [C.c_char_p, i, i, C.POINTER(ParticleArrays), i]),
        "dmpm_write_layer": (i, [vp, i, C.c_char_p, i]),
        "dmpm_read_particles": (i, [vp, C.c_char_p, i]),
        "dmpm_winding_numbers": (i, [vp, i, vp, i, vp, i, vp]),
        "dmpm_batch_capture": (i, [C.POINTER(vp), i, i, i, C.POINTER(vp)]),
        "dmpm_batch_launch": (i, [vp]),
        "dmpm_batch_synchronize": (i, [vp]),
        "dmpm_batch_destroy": (i, [vp]),
    }
    for name, (res, args) in proto.items():
        if not hasattr(L, name):
            continue  # an older build; calling it raises AttributeError
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


EXPORTED = [
    "dmpm_config_init", "dmpm_last_error", "dmpm_version", "dmpm_create", "dmpm_destroy", "dmpm_set_stream",
    "dmpm_synchronize", "dmpm_set_particles", "dmpm_num_particles", "dmpm_set_target_grid_mass",
    "dmpm_set_sim_params", "dmpm_set_up_comp_graph", "dmpm_get_num_layers", "dmpm_compute_forward_pass", "dmpm_rollout_forward",
    "dmpm_end_layer_mass_loss", "dmpm_compute_backward_pass", "dmpm_get_layer", "dmpm_get_layer_grad",
    "dmpm_accumulate_last_layer_grad", "dmpm_set_last_layer_grad", "dmpm_set_render_gradients",
    "dmpm_clear_render_gradients", "dmpm_render_gradient_info", "dmpm_promote_last_as_initial", "dmpm_sort_particles", "dmpm_sort_segments",
    "dmpm_set_control", "dmpm_request_control_grad", "dmpm_get_control_grad", "dmpm_control_grad_norm",
    "dmpm_descend_adam", "dmpm_optimize_single_timestep", "dmpm_run_optimization",
    "dmpm_single_timestep_adam_count", "dmpm_p2g_mass",
    "dmpm_calculate_volumes", "dmpm_calculate_volumes_noalloc", "dmpm_lame_parameters", "dmpm_last_timing", "dmpm_memory", "dmpm_kernel_timer",
    "dmpm_kernel_stats", "dmpm_window_stats", "dmpm_window_counters", "dmpm_graph_capture", "dmpm_graph_launch", "dmpm_graph_synchronize", "dmpm_graph_destroy",
    "dmpm_get_config", "dmpm_set_adam_state", "dmpm_file_info", "dmpm_write_file", "dmpm_read_file",
    "dmpm_write_layer", "dmpm_read_particles", "dmpm_winding_numbers",
    "dmpm_batch_capture", "dmpm_batch_launch", "dmpm_batch_synchronize", "dmpm_batch_destroy",
]


def check(rc: int):
    if rc != DMPM_OK:
        msg = lib().dmpm_last_error().decode(errors="replace")
        if rc == -2:
            raise DmpmIndexError(rc, msg)
        raise DmpmError(rc, msg)
