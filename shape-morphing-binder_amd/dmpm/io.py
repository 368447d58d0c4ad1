"""On-disk particle state (host arrays <-> files), over the C ABI.

Formats:
  "native" -- this library's checkpoint format (header, named field table,
              float32 or float64 fields; layout in include/dmpm.h);
  "cereal" -- the reference's PointCloud binary file
              (PointCloud::WriteEntirePointCloudToBinaryFile /
              ReadEntirePointCloudFromBinaryFile, PointCloud.cpp:382-410;
              fields of MaterialPoint.h:55-71), always float32.

Arrays use the names of dmpm_particle_arrays: x, v (N,3); F, C, P, dFc, dLdF,
adam_m, adam_v, adam_vmax (N,3,3); m, vol, lam, mu (N,); dLdx, dLdv (N,3).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

FORMATS = {"native": L.FILE_NATIVE, "cereal": L.FILE_CEREAL}
SHAPES = {"x": (3,), "v": (3,), "F": (3, 3), "C": (3, 3), "P": (3, 3), "m": (), "vol": (), "lam": (), "mu": (),
          "dFc": (3, 3), "dLdx": (3,), "dLdv": (3,), "dLdF": (3, 3), "adam_m": (3, 3), "adam_v": (3, 3),
          "adam_vmax": (3, 3)}
REQUIRED = ("x", "v", "F", "C", "m", "vol", "lam", "mu")


def _fmt(format: str) -> int:
    if format not in FORMATS:
        raise ValueError(f"unknown format {format!r}; one of {sorted(FORMATS)}")
    return FORMATS[format]


def file_info(path: str, format: str = "native"):
    """(particle count, numpy dtype of the stored values)."""
    n, dt = C.c_int(), C.c_int()
    L.check(L.lib().dmpm_file_info(str(path).encode(), _fmt(format), C.byref(n), C.byref(dt)))
    return n.value, (np.float64 if dt.value == L.DTYPE_F64 else np.float32)


def write_file(path: str, arrays: dict, format: str = "native", dtype=np.float32):
    """Write `arrays` (x v F C m vol lam mu required; absent fields are
    written as zeros)."""
    missing = [k for k in REQUIRED if k not in arrays]
    if missing:
        raise ValueError(f"missing fields {missing}")
    unknown = set(arrays) - set(SHAPES)
    if unknown:
        raise ValueError(f"unknown fields {sorted(unknown)}")
    n = np.asarray(arrays["x"]).shape[0]
    keep, pa = {}, L.ParticleArrays()
    for k, a in arrays.items():
        if a is None:
            continue
        keep[k] = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype), (n,) + SHAPES[k]))
        setattr(pa, k, keep[k].ctypes.data)
    code = L.DTYPE_F64 if np.dtype(dtype) == np.float64 else L.DTYPE_F32
    L.check(L.lib().dmpm_write_file(str(path).encode(), _fmt(format), int(n), C.byref(pa), code))


def read_file(path: str, format: str = "native", dtype=None) -> dict:
    """Every field of the file as arrays (absent optional fields as zeros)."""
    n, fdt = file_info(path, format)
    dt = np.dtype(dtype or fdt)
    out, pa = {}, L.ParticleArrays()
    for k, shp in SHAPES.items():
        out[k] = np.zeros((n,) + shp, dt)
        setattr(pa, k, out[k].ctypes.data)
    code = L.DTYPE_F64 if dt == np.float64 else L.DTYPE_F32
    L.check(L.lib().dmpm_read_file(str(path).encode(), _fmt(format), int(n), C.byref(pa), code))
    return out
