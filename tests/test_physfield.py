"""PhysField checkpoint files (the reference's HDF5 dump/restore layout) on
CPU: round trips, the file layout as h5dump reports it, no overwrite, error
paths, and the C ABI of libpsx_h5.so."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from polyswift_plusplus_amd import physfield as pf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "psx_h5.h")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)

pytestmark = pytest.mark.skipif(not os.path.exists(pf.H5LIB_PATH), reason="libpsx_h5.so not built (no HDF5)")


def test_header_symbols_exported():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    names = re.findall(r"^\s*(?:int|void)\s+(psx_h5_\w+)\s*\(", text, flags=re.M)
    assert len(names) == 3
    lib = ctypes.CDLL(pf.H5LIB_PATH)
    assert all(hasattr(lib, n) for n in names)


@pytest.mark.parametrize("shape", [(8, 6, 1), (6, 5, 4), (7, 1, 1)])
def test_round_trip(tmp_path, shape):
    rng = np.random.default_rng(1)
    d, c = rng.standard_normal(shape), rng.standard_normal(shape)
    path = str(tmp_path / pf.dump_name("run", "totA", 40))
    assert path.endswith("run_totA_1000040.h5")
    assert pf.write_physfield(path, "totA", shape, (0.1, 0.2, 0.3), "MonomerDensity", d,
                              "ConjugateDensity", c, 40)
    d2, c2 = pf.read_physfield(path, shape, "MonomerDensity", "ConjugateDensity")
    assert np.array_equal(d, d2) and np.array_equal(c, c2)


def test_existing_file_is_not_overwritten(tmp_path):
    shape = (4, 4, 1)
    path = str(tmp_path / "f.h5")
    assert pf.write_physfield(path, "p", shape, (1, 1, 1), "MonomerDensity", np.zeros(shape),
                              "PressureField", np.ones(shape), 1)
    assert not pf.write_physfield(path, "p", shape, (1, 1, 1), "MonomerDensity", np.ones(shape),
                                  "PressureField", np.zeros(shape), 2)
    _, p = pf.read_physfield(path, shape, "MonomerDensity", "PressureField")
    assert np.all(p == 1.0)


def test_errors(tmp_path):
    shape = (4, 4, 2)
    path = str(tmp_path / "g.h5")
    pf.write_physfield(path, "g", shape, (1, 1, 1), "MonomerDensity", np.zeros(shape),
                       "ConjugateDensity", np.zeros(shape), 0)
    with pytest.raises(OSError, match="no dataset"):
        pf.read_physfield(path, shape, "MonomerDensity", "PressureField")
    with pytest.raises(OSError, match="elements"):
        pf.read_physfield(path, (4, 4, 3), "MonomerDensity", "ConjugateDensity")
    with pytest.raises(OSError, match="cannot open"):
        pf.read_physfield(str(tmp_path / "none.h5"), shape, "MonomerDensity", "ConjugateDensity")


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
@pytest.mark.parametrize("shape,dims,labels,ncell", [
    ((8, 6, 1), "( 8, 6, 1 )", "x,y", "8, 6"),
    ((6, 5, 4), "( 6, 5, 4, 1 )", "x,y,z", "6, 5, 4")])
def test_layout_as_h5dump_sees_it(tmp_path, shape, dims, labels, ncell):
    """Datasets, shapes, VizSchema attributes and groups of
    psphysf/PsPhysFldFuncs.cpp:119-143,253-350."""
    path = str(tmp_path / "x.h5")
    pf.write_physfield(path, "totStyrDens", shape, (0.1, 0.2, 0.3), "MonomerDensity",
                       np.zeros(shape), "ConjugateDensity", np.ones(shape), 100)
    out = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
    for ds in ("MonomerDensity", "ConjugateDensity"):
        assert f'DATASET "{ds}"' in out
    assert out.count(f"SIMPLE {{ {dims} / {dims} }}") == 2
    assert "H5T_IEEE_F64LE" in out
    for grp in ("totStyrDensGlobal", "totStyrDensGlobalLimits", "time"):
        assert f'GROUP "{grp}"' in out
    full = subprocess.run([H5DUMP, "-A", path], capture_output=True, text=True, check=True).stdout
    flat = " ".join(full.split())
    for attr, val in [("vsType", '"variable"'), ("vsMesh", '"totStyrDensGlobal"'),
                      ("vsLimits", '"totStyrDensGlobalLimits"'), ("vsCentering", '"zonal"'),
                      ("vsLabels", f'"{labels}"'), ("vsTimeGroup", '"time"'), ("vsKind", '"uniform"'),
                      ("vsKind", '"Cartesian"'), ("vsType", '"mesh"'), ("vsType", '"limits"')]:
        assert re.search(rf'ATTRIBUTE "{attr}" {{.*?\(0\): {re.escape(val)}', flat), (attr, val)
    assert re.search(rf'ATTRIBUTE "vsNumCells" {{.*?\(0\): {ncell} }}', flat)
    assert re.search(r'ATTRIBUTE "vsStep" {.*?\(0\): 100 }', flat)
    assert re.search(r'ATTRIBUTE "GridSizes" {.*?\(0\): 0.1, 0.2, 0.3 }', flat)
