"""The VTR writer (Print_Result_VTK_2D, PrintResult.f90:165-190, through
VTR_mod.f90) and the checkpoint file container, on the host: the files
parse as VTK XML with the reference's extents, array names and order, and
hold the fields as Float32 (i fastest, Velocity interleaved as (u, v, 0));
ASCII and binary forms carry the same values; checkpoint headers reject
foreign, truncated and mismatched files."""
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from simcoflow_amd import cases as C
from simcoflow_amd import output as O


def _fields(g, seed=0):
    rng = np.random.default_rng(seed)
    gh = (g.nj + 2, g.ni + 2)
    ce = (g.nj, g.ni)
    return dict(u=rng.standard_normal(gh), v=rng.standard_normal(gh), p=rng.standard_normal(gh) * 1e4,
                phi=rng.standard_normal(ce), phiS=rng.standard_normal(ce), vof=rng.uniform(size=ce),
                vofS=rng.uniform(size=ce))


@pytest.mark.parametrize("fmt", ["ascii", "binary"])
def test_vtr_layout_and_values(tmp_path, fmt):
    g = C.uniform_grid(7, 5, 1.4, 1.0)
    f = _fields(g)
    path = O.vtr_name(str(tmp_path), 42)
    assert path.endswith("Paraview/FlowField_00000042.vtr")
    O.write_vtr(path, g, fmt=fmt, **f)
    root = ET.parse(path).getroot()
    assert root.tag == "VTKFile" and root.get("type") == "RectilinearGrid" and root.get("version") == "0.1"
    rg = root.find("RectilinearGrid")
    assert rg.get("WholeExtent").split() == ["1", "7", "1", "5", "1", "1"]
    names = [d.get("Name") for d in rg.iter("DataArray")]
    assert names == ["X_COORDINATES", "Y_COORDINATES", "Z_COORDINATES"] + list(O.VTR_FIELDS)
    assert all(d.get("type") == "Float32" for d in rg.iter("DataArray"))
    arr, ext = O.read_vtr(path)
    assert ext == [1, 7, 1, 5, 1, 1]
    f32 = lambda a: np.asarray(a, dtype=np.float32).ravel()
    assert np.array_equal(arr["X_COORDINATES"], f32(g.xc))
    assert np.array_equal(arr["Y_COORDINATES"], f32(g.yc))
    assert np.array_equal(arr["Z_COORDINATES"], np.zeros(1, np.float32))
    inner = lambda a: a[1:-1, 1:-1]
    assert np.array_equal(arr["X_Velocity"], f32(inner(f["u"])))
    assert np.array_equal(arr["Y_Velocity"], f32(inner(f["v"])))
    assert np.array_equal(arr["Pressure"], f32(inner(f["p"])))
    vel = arr["Velocity"]
    assert vel.shape == (35, 3)
    assert np.array_equal(vel[:, 0], arr["X_Velocity"]) and np.array_equal(vel[:, 1], arr["Y_Velocity"])
    assert np.all(vel[:, 2] == 0)
    # i runs fastest: entry 1 is cell (i=2, j=1)
    assert arr["Vof"][1] == np.float32(f["vof"][0, 1]) and arr["Vof"][7] == np.float32(f["vof"][1, 0])
    for k, n in (("phi", "Levelset"), ("phiS", "SolidLvs"), ("vof", "Vof"), ("vofS", "SolidVof")):
        assert np.array_equal(arr[n], f32(f[k]))


def test_vtr_rejects_wrong_shapes(tmp_path):
    g = C.uniform_grid(4, 4, 1.0, 1.0)
    f = _fields(g)
    f["vof"] = f["vof"][:, :3]
    with pytest.raises(ValueError):
        O.write_vtr(str(tmp_path / "a.vtr"), g, **f)


class _FakeContext:
    ni, nj = 8, 4

    def __init__(self, blob=b"state-bytes"):
        self.blob = blob
        self.loaded = None

    def state_save(self):
        return self.blob

    def state_load(self, blob):
        self.loaded = blob


def test_checkpoint_container(tmp_path):
    p = str(tmp_path / "run.ckpt")
    O.save_checkpoint(p, _FakeContext(b"\x00\x01\x02" * 100), 17, 0.125, extra={"case": "bubble"})
    q = _FakeContext()
    meta = O.load_checkpoint(p, q)
    assert meta["itt"] == 17 and meta["time"] == 0.125 and meta["extra"] == {"case": "bubble"}
    assert q.loaded == b"\x00\x01\x02" * 100
    other = _FakeContext()
    other.ni = 9
    with pytest.raises(ValueError, match="grid"):
        O.load_checkpoint(p, other)
    raw = open(p, "rb").read()
    open(p, "wb").write(raw[:-5])
    with pytest.raises(ValueError, match="truncated"):
        O.load_checkpoint(p, _FakeContext())
    open(p, "wb").write(b"NOTACKPT" + raw[8:])
    with pytest.raises(ValueError, match="not a checkpoint"):
        O.load_checkpoint(p, _FakeContext())
    assert struct.unpack("<QQ", raw[8:24])[1] == 300
