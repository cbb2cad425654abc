"""Mixed form: the energy's first-order correction for the split of D.

The contractions see D as its two bf16 terms hi + lo; the remainder R = D - hi - lo changes by
rounding steps as D moves, so the energy of hi + lo alone jumps between nearby densities (about
2^-17 |V . D|). chx_grid_energy adds sum_s V_s . R_s (dE/dD_s = V_s), which makes E smooth in D.
Checked against the FP64 form (smooth by construction) on a line of densities D0 + k eps P: the
corrected mixed energies follow the FP64 ones to a constant offset, the uncorrected ones
(Evaluate("e"), no potential) do not. The correction does not depend on the call order."""
import numpy as np
import pytest

import chinium_amd as ca
from chinium_amd import hip
from chinium_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _grid(mol, name, prec, family):
    g = ca.Grid(mol, name, 1, 0)
    g.setPrecision(prec)
    g.setType(family)
    g.getAO(0, 0)
    return g


def _line(mol, gname, xc, Ds_of_k, K=6):
    """E of both forms ("ev", and "e" alone for the mixed form) at each density of the line."""
    g64, gmx = _grid(mol, gname, ca.CHX_PREC_FP64, xc.family), _grid(mol, gname, ca.CHX_PREC_MIXED, xc.family)
    e64, emx, eonly = [], [], []
    for k in range(K):
        Ds = Ds_of_k(k)
        for g, out in ((g64, e64), (gmx, emx)):
            g.getDensity(Ds)
            xc.Evaluate("ev", g)
            out.append(g.getEnergy())
        gmx.getDensity(Ds)
        xc.Evaluate("e", gmx)
        eonly.append(gmx.getEnergy())
    d = np.array(emx) - np.array(e64)
    dn = np.array(eonly) - np.array(e64)
    return d, dn


def _spread(x):
    return float(x.max() - x.min())


def test_energy_smooth_in_density_closed_shell():
    w = wl.build_workload("h2o16_def2svp_sg2_b3lyp")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D0, _ = wl.workload_density(mol, w["nocc"], w["seed"])
    rng = np.random.default_rng(7)
    P = rng.standard_normal(D0.shape)
    P = 0.5 * (P + P.T) * np.sqrt(np.abs(D0))
    xc = ca.ExchangeCorrelation()
    xc.Read(w["functional"])
    d, dn = _line(mol, w["grid"], xc, lambda k: [D0 + k * 2e-6 * P])
    print("  (H2O)16 B3LYP: mixed - FP64 along the line, corrected %s (spread %.1e); energy only, spread %.1e"
          % (np.array2string(d, precision=3), _spread(d), _spread(dn)), flush=True)
    # measured: 3e-8 corrected, 2.5e-5 without the correction
    assert _spread(d) <= 3e-7
    assert _spread(dn) >= 10 * _spread(d)
    assert np.abs(d).max() / len(w["Z"]) <= 1e-6


def test_energy_smooth_in_density_polarised():
    Z, xyz = [8, 1], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.834]])
    mol = ca.Molecule(Z, xyz, "cc-pvdz")
    S = mol.overlap()
    Da, Db = 0.5 * wl.density_matrix(S, 5, 11), 0.5 * wl.density_matrix(S, 4, 12)
    rng = np.random.default_rng(3)
    Pa, Pb = rng.standard_normal(Da.shape), rng.standard_normal(Db.shape)
    Pa, Pb = 0.5 * (Pa + Pa.T) * np.sqrt(np.abs(Da)), 0.5 * (Pb + Pb.T) * np.sqrt(np.abs(Db))
    for func in ("PBE", "R2SCAN"):
        xc = ca.ExchangeCorrelation()
        xc.Spin = 2
        xc.Read(func)
        d, dn = _line(mol, "SG-2", xc, lambda k: [Da + k * 2e-6 * Pa, Db + k * 2e-6 * Pb])
        print("  OH %s polarised: corrected spread %.1e, energy only %.1e, offset %.1e"
              % (func, _spread(d), _spread(dn), d.mean()), flush=True)
        assert _spread(d) <= 1e-7
        assert np.abs(d).max() / 2 <= 1e-5


def test_correction_independent_of_call_order():
    w = wl.build_workload("h2o16_def2svp_sg2_b3lyp")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D, _ = wl.workload_density(mol, w["nocc"], w["seed"])
    xc = ca.ExchangeCorrelation()
    xc.Read(w["functional"])
    g = _grid(mol, w["grid"], ca.CHX_PREC_MIXED, xc.family)
    n = mol.nbasis
    # the reference's order: energy first (the potential is built for it), then the Fock matrix
    g.getDensity([D])
    xc.Evaluate("ev", g)
    E1 = g.getEnergy()
    V1 = g.getFock()[0]
    V1b = g.getFock()[0]  # a second call returns the same matrix
    E1b = g.getEnergy()
    dV = hip.DeviceArray(n * n)
    g.getFockDevice(dV.ptr, 1, n)
    V1d = dV.download().reshape((n, n), order="F")
    # Fock first (device), then the energy
    g.getDensity([D])
    xc.Evaluate("ev", g)
    g.getFockDevice(dV.ptr, 1, n)
    E2 = g.getEnergy()
    V2 = dV.download().reshape((n, n), order="F")
    # energy only: no potential, no correction
    g.getDensity([D])
    xc.Evaluate("e", g)
    E3 = g.getEnergy()
    print("  E (energy first) %.12f, (Fock first) %.12f; without the correction %.12f" % (E1, E2, E3), flush=True)
    # the later calls return the potential built for the energy, bit for bit; a second build agrees to
    # the order of its atomic FP32 sums
    assert E1 == E1b and abs(E2 - E1) <= 1e-11
    assert np.array_equal(V1, V1b) and np.array_equal(V1, V1d)
    assert np.abs(V2 - V1).max() <= 1e-6 * np.abs(V1).max()
    assert E3 != E1 and abs(E3 - E1) <= 1e-4
