"""On-disk particle state, host side (no GPU): the reference's cereal
PointCloud file read and written byte for byte, the native checkpoint format
round-tripped, and malformed files rejected with an error."""
import os

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden")
# cereal field names -> dmpm_particle_arrays names
CEREAL_TO_DMPM = {"momentum": "adam_m", "vector": "adam_v", "vector_max": "adam_vmax"}


def _golden():
    d = np.load(os.path.join(GOLDEN, "ref_pointcloud_inputs.npz"))
    return {CEREAL_TO_DMPM.get(k, k): d[k] for k in d.files}


def _random_arrays(n, dtype, seed=3):
    from dmpm.io import SHAPES
    rng = np.random.default_rng(seed)
    return {k: rng.normal(size=(n,) + shp).astype(dtype) for k, shp in SHAPES.items()}


def test_reads_the_reference_cereal_file():
    """A file written by the reference's own WriteEntirePointCloudToBinaryFile
    (tests/golden/make_golden.py io): every field, bit for bit."""
    from dmpm.io import file_info, read_file
    path = os.path.join(GOLDEN, "ref_pointcloud.cereal")
    assert file_info(path, "cereal") == (5, np.float32)
    got = read_file(path, "cereal")
    for k, v in _golden().items():
        np.testing.assert_array_equal(got[k], v, err_msg=k)


def test_writes_the_reference_cereal_bytes(tmp_path):
    from dmpm.io import write_file
    out = tmp_path / "pc.cereal"
    write_file(out, _golden(), "cereal")
    assert out.read_bytes() == open(os.path.join(GOLDEN, "ref_pointcloud.cereal"), "rb").read()


def test_reference_reads_our_cereal_file(tmp_path):
    import sys
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Ref
    if not Ref.available():
        pytest.skip("reference library not built")
    from dmpm.io import write_file
    a = _random_arrays(37, np.float32)
    out = tmp_path / "ours.cereal"
    write_file(out, a, "cereal")
    got = Ref(1).read_cereal(str(out))
    for k in ("x", "F", "P", "m"):
        np.testing.assert_array_equal(got[k], a[k], err_msg=k)
    np.testing.assert_array_equal(got["vector_max"], a["adam_vmax"])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_native_round_trip(tmp_path, dtype):
    from dmpm.io import file_info, read_file, write_file
    a = _random_arrays(101, dtype)
    out = tmp_path / "state.dmpm"
    write_file(out, a, "native", dtype=dtype)
    assert file_info(out, "native") == (101, dtype)
    got = read_file(out, "native")
    for k, v in a.items():
        assert got[k].dtype == dtype
        np.testing.assert_array_equal(got[k], v, err_msg=k)
    # reading into the other precision converts
    other = np.float32 if dtype == np.float64 else np.float64
    got2 = read_file(out, "native", dtype=other)
    np.testing.assert_array_equal(got2["x"], a["x"].astype(other))


def test_absent_optional_fields_are_zeros(tmp_path):
    from dmpm.io import REQUIRED, read_file, write_file
    a = _random_arrays(9, np.float32)
    out = tmp_path / "req.dmpm"
    write_file(out, {k: a[k] for k in REQUIRED}, "native")
    got = read_file(out, "native")
    for k in REQUIRED:
        np.testing.assert_array_equal(got[k], a[k])
    for k in ("P", "dFc", "dLdx", "adam_vmax"):
        assert not got[k].any()


def test_malformed_files_are_rejected(tmp_path):
    from dmpm._lib import DmpmError
    from dmpm.io import read_file, write_file
    a = _random_arrays(8, np.float32)
    good = tmp_path / "good.dmpm"
    write_file(good, a, "native")
    data = good.read_bytes()
    (tmp_path / "trunc.dmpm").write_bytes(data[:-4])
    with pytest.raises(DmpmError, match="size"):
        read_file(tmp_path / "trunc.dmpm", "native")
    (tmp_path / "magic.dmpm").write_bytes(b"XXXXXXXX" + data[8:])
    with pytest.raises(DmpmError, match="not a dmpm state file"):
        read_file(tmp_path / "magic.dmpm", "native")
    cer = tmp_path / "c.cereal"
    write_file(cer, a, "cereal")
    (tmp_path / "ctrunc.cereal").write_bytes(cer.read_bytes()[:-1])
    with pytest.raises(DmpmError, match="432-byte"):
        read_file(tmp_path / "ctrunc.cereal", "cereal")
    bad = bytearray(cer.read_bytes())
    bad[8:12] = (4).to_bytes(4, "little")  # the first point's x row tag
    (tmp_path / "ctag.cereal").write_bytes(bytes(bad))
    with pytest.raises(DmpmError, match="tag"):
        read_file(tmp_path / "ctag.cereal", "cereal")
    with pytest.raises(ValueError, match="missing"):
        write_file(tmp_path / "x.dmpm", {"x": a["x"]}, "native")
    with pytest.raises(DmpmError, match="cannot open"):
        read_file(tmp_path / "nope.dmpm", "native")


def test_bindings_point_cloud_file_round_trip(tmp_path):
    import diffmpm_bindings as db
    a = _random_arrays(12, np.float32)
    pc = db.PointCloud.from_arrays(a["x"], a["m"], a["vol"], a["lam"], a["mu"], v=a["v"], F=a["F"], C=a["C"],
                                   dFc=a["dFc"])
    out = tmp_path / "pc.bin"
    pc.write_entire_point_cloud_to_binary_file(str(out))
    back = db.PointCloud.read_entire_point_cloud_from_binary_file(str(out))
    np.testing.assert_array_equal(np.array(back.get_positions()), a["x"])
    np.testing.assert_array_equal(back.get_def_grads_morph(), a["dFc"])
    np.testing.assert_array_equal(np.array(back.get_masses(), np.float32), a["m"])
