"""The C ABI: the library loads, exports every entry point include/chinium_xc.h
declares, and reports the reference's failure cases as status codes with the
reference's messages (src/ExchangeCorrelation.cpp:23,35; GridGeneration.cpp:33,114).
No GPU work here."""
import ctypes
import os
import re

import numpy as np
import pytest

import chinium_amd as ca

HEADER = os.path.join(os.path.dirname(ca.PKG_DIR), "include", "chinium_xc.h")
NO_GPU = not os.path.exists("/dev/kfd")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(chx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    names = declared_functions()
    assert len(names) >= 30
    lib = ctypes.CDLL(ca.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_every_declared_function_is_bound_in_python():
    ca.lib()
    src = open(os.path.join(ca.PKG_DIR, "__init__.py")).read()
    for n in declared_functions():
        assert '"%s"' % n in src, n


def _xc_read(root, name):
    h = ctypes.c_void_p()
    rc = ca.lib().chx_xc_read(root.encode(), name.encode(), 1, 0, ctypes.byref(h))
    return rc, ca.lib().chx_last_error().decode(), h


def test_functional_files(tmp_path):
    rc, msg, h = _xc_read(ca.DATA_ROOT, "B3LYP")
    assert rc == 0
    fam, exx, nc = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
    ca.lib().chx_xc_info(h, ctypes.byref(fam), ctypes.byref(exx), ctypes.byref(nc), None)
    assert (fam.value, exx.value, nc.value) == (1, 0.2, 1)
    ca.lib().chx_xc_destroy(h)
    rc, msg, _ = _xc_read(ca.DATA_ROOT, "NOPE")
    assert rc == -2 and msg == "Density functional file is missing!"
    (tmp_path / "DensityFunctionals").mkdir()
    (tmp_path / "DensityFunctionals" / "BAD.df").write_text("101 x3\n")
    rc, msg, _ = _xc_read(str(tmp_path), "BAD")
    assert rc == -2 and msg == "Corrupted density functional file!"
    (tmp_path / "DensityFunctionals" / "MN15.df").write_text("268 269\n")
    rc, msg, _ = _xc_read(str(tmp_path), "MN15")
    assert rc == -5 and "268" in msg


def test_python_mirror_of_exchange_correlation():
    xc = ca.ExchangeCorrelation()
    assert not xc
    xc.Read("PBE")
    assert xc and xc.Codes == [321, 322] and xc.Family == "GGA" and xc.EXX == 0.0
    xc.Read("SVWN5")
    assert xc.Codes == [1, 7] and xc.Family == "LDA"


@pytest.mark.skipif(not NO_GPU, reason="checks the no-device path")
def test_grid_create_without_device_fails_cleanly():
    m = ca.Molecule([8, 1, 1], np.array([[0, 0, 0], [1.4, 0, 1.1], [-1.4, 0, 1.1]]), "cc-pvdz")
    with pytest.raises(ca.ChxError) as e:
        ca.Grid(m, "SG-1")
    assert e.value.code == -4


def test_grid_file_errors_come_before_device_use(tmp_path):
    m = ca.Molecule([8], np.zeros((1, 3)), "cc-pvdz")
    with pytest.raises(ca.ChxError) as e:
        ca.Grid(m, "SG-7")
    assert e.value.code == -2 and str(e.value) == "Missing element grid file in folder!"
    (tmp_path / "Grids" / "X").mkdir(parents=True)
    (tmp_path / "Grids" / "X" / "O.grid").write_text("1 10\nxyz 1.0\n6 10\n")
    with pytest.raises(ca.ChxError, match="Unrecognized radial formula!"):
        ca.Grid(m, "X", data_root=str(tmp_path))


def test_bad_arguments():
    with pytest.raises(ca.ChxError) as e:
        ca.Molecule([], np.zeros((0, 3)))
    assert e.value.code == -1
    m = ca.Molecule([1], np.zeros((1, 3)))
    with pytest.raises(ca.ChxError) as e:
        m.add_shell(0, -4, [1.0], [1.0])
    assert e.value.code == -5
    with pytest.raises(ca.ChxError):
        m.add_shell(3, 0, [1.0], [1.0])
    m.add_shell(0, 0, [1.0], [(2 / np.pi) ** 0.75])
    assert m.nbasis == 1 and abs(m.overlap()[0, 0] - 1) < 1e-15
