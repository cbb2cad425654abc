"""The density-fitted Kohn-Sham SCF (chinium_amd.scf.RestrictedDFSCF: J, S, T, V from the device
Coulomb engine, the reference's ADIIS -> CDIIS driver, dense algebra in torch on the GPU) with
r2SCAN, against the exact-Coulomb SCF (RestrictedSCF: host four-centre integrals) on the same
molecule and grid: the difference is the fitting error of def2-SVP-J (bounded here by 1e-4 Eh
for a water molecule; the Coulomb-metric fit's own error is below 1e-5 Eh per water). The mixed
form's SCF against the FP64 form's, both fitted: |dE| <= 1e-5 Eh/atom."""
import numpy as np
import pytest

import chinium_amd as ca
from chinium_amd import workloads as wl
from chinium_amd.scf import RestrictedDFSCF, RestrictedSCF

pytestmark = pytest.mark.gpu


def _grid(mol, prec, family):
    g = ca.Grid(mol, "SG-1", 1, 0)
    g.setPrecision(prec)
    g.setType(family)
    g.getAO(0, 0)
    return g


@pytest.mark.parametrize("nw", [1, 2])
def test_df_scf_r2scan(nw):
    if nw == 1:
        w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
        Z, X = w["Z"], w["xyz_bohr"]
    else:
        Z, Xa = wl.water_cluster_angstrom(2, 4.0, 7)
        X = Xa * ca.ANGSTROM_TO_BOHR
    mol = ca.Molecule(Z, X, "def2-svp")
    nocc, natoms = 5 * nw, len(Z)
    xc = ca.ExchangeCorrelation()
    xc.Read("R2SCAN")
    log = lambda s: print(s, flush=True)
    res = {}
    for prec, tol in ((ca.CHX_PREC_FP64, 1e-6), (ca.CHX_PREC_MIXED, 1e-5)):
        grid = _grid(mol, prec, xc.family)
        scf = RestrictedDFSCF(mol, nocc, xc, grid)
        r = scf.run(tol=tol, log=log)
        res[prec] = r
        print("(H2O)%d r2SCAN %s DF-SCF: E %.10f in %d iterations, max|G| %.1e" % (
            nw, "mixed" if prec else "FP64", r["E"], r["iterations"], r["diis_error"]), flush=True)
    grid = _grid(mol, ca.CHX_PREC_FP64, xc.family)
    ex = RestrictedSCF(mol, nocc, xc, grid).run_reference(tol=1e-6)
    d_fit = res[ca.CHX_PREC_FP64]["E"] - ex["E"]
    d_mx = abs(res[ca.CHX_PREC_MIXED]["E"] - res[ca.CHX_PREC_FP64]["E"]) / natoms
    print("  exact-J SCF %.10f; fitted - exact %.2e Eh; mixed - FP64 %.2e Eh/atom" % (ex["E"], d_fit, d_mx))
    assert abs(d_fit) <= 1e-4 * nw
    assert d_mx <= 1e-5
