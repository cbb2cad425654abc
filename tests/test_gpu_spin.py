"""Spin-polarised (nspin = 2) XC path on the GPU, LDA/GGA: the device's
polarised functionals against the sympy golden vectors, and the nspin = 2
build (getDensity({Da, Db}) -> Evaluate("ev") -> getEnergy -> getFock, two
matrices) against the FP64 oracle (pinned in test_oracle_spin.py) at
|dE| <= 1e-10 Eh and max|dV| <= 1e-9 (FP64 form, both AO paths); the mixed
form against the FP64 form at <= 1e-5 Eh/atom."""
import json
import os

import numpy as np
import pytest

import chinium_amd as ca
import oracle_lib as ol
from chinium_amd import workloads as wl

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "functionals_pol.json")))


def _say(msg):
    print(msg, flush=True)


def _xc(func=None, codes=None):
    xc = ca.ExchangeCorrelation()
    xc.Spin = 2
    if func:
        xc.Read(func)
    else:
        xc.from_codes(codes)
    return xc


def test_polarised_functionals_match_golden():
    P = np.array(GOLD["points"])
    for code, vals in GOLD["values"].items():
        ref = np.array(vals)
        e, vr, vs = _xc(codes=[int(code)]).eval_points(P[:, :2], P[:, 2:])
        got = np.column_stack([e, vr, vs])
        mask = np.ones_like(got, dtype=bool)
        mask[P[:, 2] <= 1e-39, 3:] = False
        rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
        _say("  polarised code %s: max rel %.1e" % (code, rel[mask].max()))
        assert rel[mask].max() < 1e-10


def _spin_dms(mol, nalpha, nbeta, seed):
    S = mol.overlap()
    Da = wl.density_matrix(S, nalpha, seed) / 2  # C C^T of nalpha orbitals
    Db = wl.density_matrix(S, nbeta, seed + 1) / 2
    return Da, Db


@pytest.mark.parametrize("func,cache", [("SVWN5", 1), ("PBE", 1), ("B3LYP", 1), ("BLYP", 1), ("PBE", 0)])
def test_h2o_cation_against_oracle(func, cache):
    w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    Da, Db = _spin_dms(mol, 5, 4, 7)
    _say("nspin=2 H2O+/cc-pvdz/SG-1 %s, AO cache %d" % (func, cache))
    xc = _xc(func)
    g = ca.Grid(mol, "SG-1", 1, 0)
    g.setAOCache(cache)
    g.setType(xc.family)
    g.getAO(0, 0)
    g.getDensity([Da, Db])
    xc.Evaluate("ev", g)
    E, N = g.getEnergy(), g.getNumElectrons()
    Va, Vb = g.getFock()
    o = ol.OracleSystem(w["Z"], w["xyz_bohr"], "SG-1", w["basis"])
    o.set_functional(func)
    o.get_ao()
    Eo, No, Vao, Vbo = o.xc_pol(Da, Db)
    _say("  E %.12f oracle %.12f  N %.10f  max dVa %.1e dVb %.1e"
         % (E, Eo, N, np.abs(Va - Vao).max(), np.abs(Vb - Vbo).max()))
    assert abs(E - Eo) <= 1e-10 and abs(N - No) <= 1e-10
    assert np.abs(Va - Vao).max() <= 1e-9 and np.abs(Vb - Vbo).max() <= 1e-9
    assert np.abs(Va - Vb).max() > 1e-3  # the two spins really differ


def test_mixed_form_spin_cluster():
    Z, X = wl.water_cluster_angstrom(4, 5.0, 11)
    mol = ca.Molecule(Z, X * ca.ANGSTROM_TO_BOHR, "def2-svp")
    Da, Db = _spin_dms(mol, 20, 19, 3)
    out = {}
    for prec in (ca.CHX_PREC_FP64, ca.CHX_PREC_MIXED):
        xc = _xc("B3LYP")
        g = ca.Grid(mol, "SG-1", 1, 0)
        g.setPrecision(prec)
        g.setType(1)
        g.getAO(0, 0)
        g.getDensity([Da, Db])
        xc.Evaluate("ev", g)
        out[prec] = (g.getEnergy(), g.getFock())
    (E0, V0), (E1, V1) = out[ca.CHX_PREC_FP64], out[ca.CHX_PREC_MIXED]
    dv = max(np.abs(V0[k] - V1[k]).max() for k in range(2))
    _say("  (H2O)4+ B3LYP nspin=2: mixed vs FP64 dE/atom %.2e, max dV %.2e" % (abs(E1 - E0) / len(Z), dv))
    assert abs(E1 - E0) / len(Z) <= 1e-5 and dv <= 5e-5


def test_spin_protocol_errors():
    w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    Da, Db = _spin_dms(mol, 5, 4, 7)
    g = ca.Grid(mol, "SG-1", 1, 0)
    g.setType(1)
    g.getAO(0, 0)
    g.getDensity([Da, Db])
    pbe = ca.ExchangeCorrelation()
    pbe.Read("PBE")  # unpolarised functional on a polarised density
    with pytest.raises(ca.ChxError) as e:
        pbe.Evaluate("ev", g)
    assert e.value.code == -3
    with pytest.raises(ca.ChxError) as e:  # a polarised meta-GGA other than r2SCAN is not built
        _xc(codes=[497, 131])
    assert e.value.code == -5
    r2 = _xc("R2SCAN")
    with pytest.raises(ca.ChxError) as e:  # the GGA grid holds no tau
        r2.Evaluate("ev", g)
    assert e.value.code == -3
    g2 = ca.Grid(mol, "SG-1", 1, 0)
    g2.setType(2)
    g2.getAO(0, 0)
    g2.getDensity([Da, Db])
    with pytest.raises(ca.ChxError) as e:  # f_xc is built for LDA and GGA (the reference's mGGA has no "f")
        r2.Evaluate("f", g2)
    assert e.value.code == -5
    with pytest.raises(ca.ChxError) as e:  # three spin matrices
        g.getDensity([Da, Db, Da])
    assert e.value.code == -1
