"""The meta-GGA configuration at full size: cyclo[18]carbon / def2-TZVP / SG-3 /
r2SCAN (318k points, 558 AOs). The mixed form against the FP64 form on the same
grid and D (the FP64 form is pinned against the oracle on smaller systems in
test_gpu_functionals.py and test_gpu_mixed.py). Bars: |dE_xc| <= 1e-5 Eh/atom,
max |dV| <= 5e-5, |dN_e| <= 1e-6 N_e."""
import numpy as np
import pytest

import chinium_amd as ca
from chinium_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_c18_r2scan_mixed_vs_fp64():
    w = wl.build_workload("c18_def2tzvp_sg3_r2scan")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D, S = wl.workload_density(mol, w["nocc"], w["seed"])
    xc = ca.ExchangeCorrelation()
    xc.Read(w["functional"])
    out = {}
    for prec in (ca.CHX_PREC_FP64, ca.CHX_PREC_MIXED):
        grid = ca.Grid(mol, w["grid"], 1, 0)
        grid.setPrecision(prec)
        grid.setType(xc.family)
        grid.getAO(0, 0)
        E, V = ca.xc_build(grid, xc, D)
        out[prec] = (E, V, grid.getNumElectrons())
        _say("  C18/def2-TZVP/SG-3/r2SCAN %s: E_xc %.10f N_e %.8f" % ("mixed" if prec else "fp64", E, out[prec][2]))
        del grid
    (E0, V0, N0), (E1, V1, N1) = out[ca.CHX_PREC_FP64], out[ca.CHX_PREC_MIXED]
    _say("  |dE|/atom %.2e  max|dV| %.2e  dN %.2e" % (abs(E1 - E0) / 18, np.abs(V1 - V0).max(), abs(N1 - N0)))
    assert abs(N0 - 108.0) < 1e-3  # SG-3 quadrature: 2 Tr(DS)... = 108 electrons
    assert abs(E1 - E0) / 18 <= 1e-5
    assert np.abs(V1 - V0).max() <= 5e-5
    assert abs(N1 - N0) <= 1e-6 * N0


@pytest.mark.parametrize("func", ["B3LYP", "R2SCAN"])
def test_wide_vxc_tiles_mixed_vs_fp64(func):
    """(H2O)40 / def2-SVP / SG-1: 960 AOs, npad 1024, so the mixed form runs its Vxc over 256-AO
    super-tiles (the large-basis kernel); against the FP64 form on the same grid and D."""
    Z, X = wl.water_cluster_angstrom(40, 10.5, 21)
    mol = ca.Molecule(Z, X * ca.ANGSTROM_TO_BOHR, "def2-svp")
    assert mol.nbasis == 960
    D, S = wl.workload_density(mol, 200, 5)
    xc = ca.ExchangeCorrelation()
    xc.Read(func)
    out = {}
    for prec in (ca.CHX_PREC_FP64, ca.CHX_PREC_MIXED):
        grid = ca.Grid(mol, "SG-1", 1, 0)
        grid.setPrecision(prec)
        grid.setAOCache(1)
        grid.setType(xc.family)
        grid.getAO(0, 0)
        E, V = ca.xc_build(grid, xc, D)
        out[prec] = (E, V)
        del grid
    (E0, V0), (E1, V1) = out[ca.CHX_PREC_FP64], out[ca.CHX_PREC_MIXED]
    _say("  (H2O)40/def2-SVP/SG-1 %s wide Vxc: |dE|/atom %.2e  max|dV| %.2e" % (func, abs(E1 - E0) / 120,
                                                                               np.abs(V1 - V0).max()))
    assert abs(E1 - E0) / 120 <= 1e-5
    assert np.abs(V1 - V0).max() <= 5e-5
    assert np.abs(V1 - V1.T).max() <= 1e-12 * np.abs(V1).max()
