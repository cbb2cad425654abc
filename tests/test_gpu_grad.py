"""GPU XC nuclear gradient (Grid::getDensitySkeleton + Grid::getEnergyGrad,
skeleton form with the grid fixed) against the FP64 oracle, whose gradient is
pinned by finite differences in test_oracle_xc.py. Bar: 1e-10 Eh/bohr."""
import numpy as np
import pytest

import chinium_amd as ca
import oracle_lib as ol
from chinium_amd import workloads as wl

pytestmark = pytest.mark.gpu
G_TOL = 1e-10


def _say(msg):
    print(msg, flush=True)


def _gpu_grad(mol, grid_name, func, D):
    grid = ca.Grid(mol, grid_name, 1, 0)
    grid.setAOCache(1)
    xc = ca.ExchangeCorrelation()
    xc.Read(func)
    grid.setType(xc.family)
    grid.getAO(0, 0)
    E, V = ca.xc_build(grid, xc, D)  # last SCF iteration: density, Evaluate("ev"), energy, Fock
    grid.getAO(1, 0)                 # Restricted/Grad.cpp:73
    grid.getDensitySkeleton([D])
    return grid, grid.getEnergyGrad()


@pytest.mark.parametrize("func", ["PBE", "B3LYP", "SVWN5"])
def test_h2o_gradient(func):
    w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D, S = wl.workload_density(mol, w["nocc"], w["seed"])
    _say("XC gradient h2o/cc-pvdz/SG-1 %s" % func)
    grid, G = _gpu_grad(mol, "SG-1", func, D)
    orc = ol.OracleSystem(w["Z"], w["xyz_bohr"], "SG-1", w["basis"])
    orc.set_functional(func)
    orc.get_ao()
    Go = orc.xc_grad(D)
    _say("  max |G - G_oracle| %.2e  max |G| %.3f" % (np.abs(G - Go).max(), np.abs(G).max()))
    assert np.abs(G - Go).max() <= G_TOL


def test_water_cluster_gradient():
    Z, X = wl.water_cluster_angstrom(4, 5.0, 11)
    xyz = X * ca.ANGSTROM_TO_BOHR
    mol = ca.Molecule(Z, xyz, "def2-svp")
    D, S = wl.workload_density(mol, 20, 3)
    _say("XC gradient (H2O)4/def2-svp/SG-1 B3LYP")
    grid, G = _gpu_grad(mol, "SG-1", "B3LYP", D)
    orc = ol.OracleSystem(Z, xyz, "SG-1", "def2-svp")
    orc.set_functional("B3LYP")
    orc.get_ao()
    Go = orc.xc_grad(D)
    _say("  max |G - G_oracle| %.2e" % np.abs(G - Go).max())
    assert np.abs(G - Go).max() <= G_TOL


def test_gradient_protocol():
    w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D, S = wl.workload_density(mol, w["nocc"], w["seed"])
    grid = ca.Grid(mol, "SG-1", 1, 0)
    grid.setAOCache(1)
    grid.setType(1)
    grid.getAO(0, 0)
    with pytest.raises(ca.ChxError) as e:
        grid.getEnergyGrad()
    assert e.value.code == -3
    grid.getDensitySkeleton([D])
    with pytest.raises(ca.ChxError) as e:  # no Evaluate("ev") yet
        grid.getEnergyGrad()
    assert e.value.code == -3
    m = ca.Grid(mol, "SG-1", 1, 0)
    m.setPrecision(ca.CHX_PREC_MIXED)
    m.setType(1)
    m.getAO(0, 0)
    with pytest.raises(ca.ChxError) as e:
        m.getDensitySkeleton([D])
    assert e.value.code == -5
