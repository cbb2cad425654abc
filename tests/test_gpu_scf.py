"""Kohn-Sham SCF (chinium_amd.scf.RestrictedSCF) with the XC path on the GPU, in both forms,
against the same loop with the CPU oracle in place of the XC path."""
import time

import numpy as np
import pytest

import chinium_amd as ca
import oracle_lib as ol
from chinium_amd import workloads as wl
from chinium_amd.scf import RestrictedSCF

pytestmark = pytest.mark.gpu


class _OracleGrid:
    """getDensity -> getEnergy -> getFock through the oracle (the checker's XC path)."""

    def __init__(self, orc):
        self.orc = orc

    def getDensity(self, Ds):
        self.E, _, self.V = self.orc.xc(Ds[0])

    def getEnergy(self):
        return self.E

    def getFock(self):
        return [self.V]


class _OracleXC:
    def __init__(self, exx):
        self.EXX = exx

    def Evaluate(self, order, grid):
        assert order == "ev"


@pytest.mark.parametrize("system,func", [("h2o", "PBE"), ("h2o", "B3LYP"), ("h2o_dimer", "B3LYP")])
def test_scf_water(system, func):
    if system == "h2o":
        w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
        Z, xyz, basis, nocc = w["Z"], w["xyz_bohr"], w["basis"], 5
    else:
        Z, X = wl.water_cluster_angstrom(2, 4.0, 7)
        xyz, basis, nocc = X * ca.ANGSTROM_TO_BOHR, "cc-pvdz", 10
    mol = ca.Molecule(Z, xyz, basis)
    natoms = len(Z)
    xc = ca.ExchangeCorrelation()
    xc.Read(func)
    print("SCF %s/cc-pVDZ/SG-1/%s (%d AOs, EXX %.2f)" % (system, func, mol.nbasis, xc.EXX), flush=True)
    out = {}
    for prec in (ca.CHX_PREC_FP64, ca.CHX_PREC_MIXED):
        grid = ca.Grid(mol, "SG-1", 1, 0)
        grid.setPrecision(prec)
        grid.setType(xc.family)
        grid.getAO(0, 0)
        t0 = time.perf_counter()
        scf = RestrictedSCF(mol, nocc, xc, grid)
        # the mixed form's V carries the BF16 splitting's rounding (~1e-6), which floors the DIIS
        # error there; the energy error left by a 1e-5 error is second order (~1e-10)
        r = scf.run(e_tol=1e-10, g_tol=1e-7) if prec == ca.CHX_PREC_FP64 else scf.run(e_tol=1e-8, g_tol=1e-5)
        out[prec] = r
        print("  form %d: E %.12f in %d iterations, %.1f s, XC build %.2f ms (median)"
              % (prec, r["E"], r["iterations"], time.perf_counter() - t0, np.median(scf.xc_ms)), flush=True)
    orc = ol.OracleSystem(Z, xyz, "SG-1", basis)
    orc.set_functional(func)
    orc.get_ao()
    ro = RestrictedSCF(mol, nocc, _OracleXC(xc.EXX), _OracleGrid(orc)).run()
    e64, emx = out[ca.CHX_PREC_FP64]["E"], out[ca.CHX_PREC_MIXED]["E"]
    print("  oracle E %.12f; FP64 - oracle %.1e; mixed - FP64 %.1e (%.1e per atom)"
          % (ro["E"], e64 - ro["E"], emx - e64, (emx - e64) / natoms), flush=True)
    assert abs(e64 - ro["E"]) <= 1e-9
    assert abs(emx - e64) / natoms <= 1e-5
    # the converged orbital energies of the FP64 form follow the oracle's
    assert np.abs(out[ca.CHX_PREC_FP64]["eps"] - ro["eps"]).max() <= 1e-7


class _OracleGridPol:
    def __init__(self, orc):
        self.orc = orc

    def getDensity(self, Ds):
        self.E, _, self.Va, self.Vb = self.orc.xc_pol(Ds[0], Ds[1])

    def getEnergy(self):
        return self.E

    def getFock(self):
        return [self.Va, self.Vb]


def test_uks_water_and_oh_radical():
    """Unrestricted Kohn-Sham (PBE, nspin = 2 XC path): closed-shell water equals the restricted
    SCF; the OH radical (5 alpha, 4 beta) in the FP64 form equals the same loop with the oracle's
    polarised XC, and the mixed form lies within 1e-5 Eh/atom of it."""
    from chinium_amd.scf import UnrestrictedSCF
    w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    xr = ca.ExchangeCorrelation()
    xr.Read("PBE")
    xu = ca.ExchangeCorrelation()
    xu.Spin = 2
    xu.Read("PBE")
    g = ca.Grid(mol, "SG-1", 1, 0)
    g.setType(1)
    g.getAO(0, 0)
    er = RestrictedSCF(mol, 5, xr, g).run()["E"]
    eu = UnrestrictedSCF(mol, 5, 5, xu, g).run()["E"]
    print("water PBE: RKS %.12f UKS %.12f diff %.1e" % (er, eu, eu - er), flush=True)
    assert abs(eu - er) <= 1e-9
    Z, xyz = [8, 1], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.834]])
    mol = ca.Molecule(Z, xyz, "cc-pvdz")
    out, grids = {}, {}
    for prec in (ca.CHX_PREC_FP64, ca.CHX_PREC_MIXED):
        g = ca.Grid(mol, "SG-1", 1, 0)
        g.setPrecision(prec)
        g.setType(1)
        g.getAO(0, 0)
        grids[prec] = g
        scf = UnrestrictedSCF(mol, 5, 4, xu, g)
        out[prec] = scf.run() if prec == ca.CHX_PREC_FP64 else scf.run(e_tol=1e-8, g_tol=1e-5)
        print("  OH form %d: E %.12f in %d iterations" % (prec, out[prec]["E"], out[prec]["iterations"]), flush=True)
    # the mixed form's E_xc at the converged FP64 density (the single-point accuracy bar)
    exc = []
    for prec in (ca.CHX_PREC_FP64, ca.CHX_PREC_MIXED):
        grids[prec].getDensity(list(out[ca.CHX_PREC_FP64]["D"]))
        xu.Evaluate("ev", grids[prec])
        exc.append(grids[prec].getEnergy())
    print("  OH E_xc at the FP64 density: mixed - FP64 %.1e Eh (%.1e per atom)"
          % (exc[1] - exc[0], (exc[1] - exc[0]) / 2), flush=True)
    assert abs(exc[1] - exc[0]) / 2 <= 1e-5
    orc = ol.OracleSystem(Z, xyz, "SG-1", "cc-pvdz")
    orc.set_functional("PBE")
    orc.get_ao()
    ro = UnrestrictedSCF(mol, 5, 4, _OracleXC(0.0), _OracleGridPol(orc)).run()
    e64, emx = out[ca.CHX_PREC_FP64]["E"], out[ca.CHX_PREC_MIXED]["E"]
    print("  OH oracle %.12f; FP64 - oracle %.1e; mixed - FP64 %.1e per atom" % (ro["E"], e64 - ro["E"],
                                                                             (emx - e64) / 2), flush=True)
    assert abs(e64 - ro["E"]) <= 1e-9
    # SCF total energy of the mixed form against the FP64 form: the 1e-5 Eh/atom bar
    assert abs(emx - e64) / 2 <= 1e-5
