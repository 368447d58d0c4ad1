"""dmpm_calculate_volumes_noalloc (the allocation-free variant of
dmpm_calculate_volumes, its particle planes in a static device array)
against dmpm_calculate_volumes with device arrays: two calls in a row, the
second with fewer particles and onto the grid the first one filled; the
GPU's free memory is the same before and after a call; a particle count
beyond the array is refused before any work is queued."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _cfg(dims, dx, prec):
    from dmpm.sim import Physics, make_config
    return make_config(dims, dx, (0.0, 0.0, 0.0), Physics(order=3), prec)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_noalloc_volumes_match_the_allocating_entry_point(prec):
    import torch
    from dmpm import _lib as L
    from dmpm.scenes import Body, make_scene
    lib = L.lib()
    dev = torch.device("cuda:0")
    dt = torch.float64 if prec == "f64" else torch.float32
    code = L.DTYPE_F64 if prec == "f64" else L.DTYPE_F32
    tol = 1e-10 if prec == "f64" else 1e-5
    sc = make_scene(Body((0.5, 0.5, 0.5), (0.25, 0.2, 0.22)), 20000, 32, 71, 1e4, 0.3)
    cfg = _cfg(sc.dims, sc.dx, prec)
    G = int(np.prod(sc.dims))
    grid_a = torch.zeros(G, dtype=dt, device=dev)
    grid_b = torch.zeros(G, dtype=dt, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    for call, n in enumerate((sc.n, 7001)):
        print(f"[noalloc] prec={prec} call={call} n={n}", flush=True)
        x = torch.tensor(sc.x[:n], dtype=dt, device=dev)
        m = torch.tensor(sc.m[:n], dtype=dt, device=dev)
        va = torch.zeros(n, dtype=dt, device=dev)
        vb = torch.full((n,), -1.0, dtype=dt, device=dev)
        torch.cuda.synchronize()
        L.check(lib.dmpm_calculate_volumes(C.byref(cfg), n, _ptr(x), _ptr(m), _ptr(va), _ptr(grid_a), code, 1))
        free0 = torch.cuda.mem_get_info()[0]
        L.check(lib.dmpm_calculate_volumes_noalloc(C.byref(cfg), n, _ptr(x), _ptr(m), _ptr(vb), _ptr(grid_b), code,
                                                   C.c_void_p(stream)))
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        if call > 0:
            assert free1 == free0, (free0, free1)
        assert rel_err(grid_b.cpu().numpy(), grid_a.cpu().numpy()) < tol
        assert rel_err(vb.cpu().numpy(), va.cpu().numpy()) < tol
        assert (vb > 0).all()


def test_noalloc_refuses_more_particles_than_its_array():
    import torch
    from dmpm import _lib as L
    from dmpm._lib import DmpmError
    cfg = _cfg((17, 17, 17), 1.0 / 16, "f64")
    n = L.DMPM_NOALLOC_SCRATCH_BYTES // (4 * 8) + 1
    t = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    with pytest.raises(DmpmError, match="exceed"):
        L.check(L.lib().dmpm_calculate_volumes_noalloc(C.byref(cfg), n, _ptr(t), _ptr(t), _ptr(t), _ptr(t),
                                                       L.DTYPE_F64, None))
