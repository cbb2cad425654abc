"""The header-only C++ drop-in (include/chinium/xc.hpp): a caller written
against the reference's Grid / ExchangeCorrelation interface builds, and on the
GPU returns what the Python mirror and the oracle return."""
import os
import subprocess

import numpy as np
import pytest

import chinium_amd as ca
import oracle_lib as ol
from chinium_amd import workloads as wl

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")
BIN = os.path.join(CPP, "xc_step")


def test_cpp_caller_builds():
    subprocess.run(["make", "-C", CPP, "-B", "-s"], check=True)
    assert os.access(BIN, os.X_OK)


@pytest.mark.gpu
def test_cpp_caller_matches_oracle(tmp_path):
    if not os.path.exists(BIN):
        subprocess.run(["make", "-C", CPP, "-s"], check=True)
    w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D, _ = wl.workload_density(mol, w["nocc"], w["seed"])
    geo = tmp_path / "geo.txt"
    geo.write_text("".join("%d %.17g %.17g %.17g\n" % (z, *x) for z, x in zip(w["Z"], w["xyz_bohr"])))
    np.asfortranarray(D).ravel(order="F").tofile(tmp_path / "D.bin")
    env = dict(os.environ, CHINIUM_PATH=ca.DATA_ROOT)
    print("cpp drop-in: running xc_step", flush=True)
    subprocess.run([BIN, w["grid"], w["basis"], w["functional"], str(geo), str(tmp_path / "D.bin"),
                    str(tmp_path / "out.bin")], check=True, env=env, timeout=120)
    out = np.fromfile(tmp_path / "out.bin")
    n = mol.nbasis
    E, N = out[0], out[1]
    V = out[2:2 + n * n].reshape((n, n), order="F")
    F = out[2 + n * n:2 + 2 * n * n].reshape((n, n), order="F")  # response Fock matrix for D^x = D
    k = 2 + 2 * n * n
    na3 = 3 * len(w["Z"])
    Fs = [out[k + m * n * n:k + (m + 1) * n * n].reshape((n, n), order="F") for m in range(na3)]  # getFockU<s_t>
    k += na3 * n * n
    H = out[k:k + na3 * na3].reshape(na3, na3)  # getEnergyHess
    k += na3 * na3
    Fk = [out[k + m * n * n:k + (m + 1) * n * n].reshape((n, n), order="F") for m in range(na3)]  # getFockSkeleton
    k += na3 * n * n
    Eu = out[k]
    Va = out[k + 1:k + 1 + n * n].reshape((n, n), order="F")
    Vb = out[k + 1 + n * n:].reshape((n, n), order="F")
    orc = ol.OracleSystem(w["Z"], w["xyz_bohr"], w["grid"], w["basis"])
    orc.set_functional(w["functional"])
    orc.get_ao()
    Eo, No, Vo = orc.xc(D)
    assert abs(E - Eo) <= 1e-10
    assert abs(N - No) <= 1e-10
    assert np.abs(V - Vo).max() <= 1e-9
    (Fo,) = orc.fock_u(D, [D])
    assert np.abs(F - Fo).max() <= 1e-9
    Fso = orc.fock_s(D)
    assert max(float(np.abs(a - b).max()) for a, b in zip(Fs, Fso)) <= 1e-9
    Ho = orc.xc_hess(D)
    assert np.abs(H - Ho).max() <= 1e-9 * np.abs(Ho).max()
    Fko = orc.fock_skeleton(D)
    assert max(float(np.abs(a - b).max()) for a, b in zip(Fk, Fko)) <= 1e-9
    Euo, _, Vao, Vbo = orc.xc_pol(0.6 * D, 0.4 * D)
    assert abs(Eu - Euo) <= 1e-10
    assert np.abs(Va - Vao).max() <= 1e-9 and np.abs(Vb - Vbo).max() <= 1e-9
