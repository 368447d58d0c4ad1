"""The C ABI library: it loads, exports every function include/dmpm.h
declares, and its host-only entry points behave (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "dmpm.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dmpm_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    from dmpm import _lib
    names = declared()
    assert len(names) >= 30
    assert sorted(_lib.EXPORTED) == names


def test_library_exports_every_declared_symbol():
    from dmpm import _lib
    L = _lib.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_config_init_defaults():
    """dmpm_config_init gives the reference's physics and CompGraph's
    default dt/drag/smoothing (CompGraph.h:80-83)."""
    from dmpm import _lib
    c = _lib.Config()
    gm = (C.c_double * 3)(0.1, 0.2, 0.3)
    _lib.lib().dmpm_config_init(C.byref(c), 33, 34, 35, 1.0 / 32, gm)
    assert list(c.dims) == [33, 34, 35] and list(c.grid_min) == [0.1, 0.2, 0.3]
    assert abs(c.dt - 1 / 120) < 1e-15 and c.drag == 0.5 and c.smoothing == 0.1
    assert c.zeta == 0.25 and c.jmin == 0.6
    assert (c.order, c.model, c.trust_region, c.adjoint, c.checkpoint_interval) == (3, 0, 1, 1, 1)


def test_lame_parameters_match_reference():
    from dmpm import lame_parameters
    d = np.load(os.path.join(ROOT, "tests", "golden", "ref_elasticity.npz"))
    lam, mu = lame_parameters(1e4, 0.3)
    assert abs(lam - float(d["lam"])) < 1e-3 and abs(mu - float(d["mu"])) < 1e-3


def test_version_and_error_string():
    from dmpm import _lib
    L = _lib.lib()
    assert b"gfx950" in L.dmpm_version()
    assert isinstance(L.dmpm_last_error(), bytes)
    # an invalid config is rejected before any device work
    c = _lib.Config()
    L.dmpm_config_init(C.byref(c), 0, 4, 4, 0.1, None)
    h = C.c_void_p()
    rc = L.dmpm_create(C.byref(c), C.byref(h))
    assert rc == -1 and b"dims" in L.dmpm_last_error()
