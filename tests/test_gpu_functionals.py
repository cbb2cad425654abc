"""Device functionals (chx_xc_eval_points, the libxc-shaped per-point call)
against the golden vectors of the independent sympy/mpmath transcriptions and
against the oracle, and meta-GGA (r2SCAN) XC builds against the oracle."""
import json
import os

import numpy as np
import pytest

import chinium_amd as ca
import oracle_lib as ol
from chinium_amd import workloads as wl

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


@pytest.mark.parametrize("fname,ncol", [("functionals.json", 2), ("r2scan.json", 3)])
def test_device_functionals_match_golden(fname, ncol):
    g = json.load(open(os.path.join(GOLD, fname)))
    P = np.array(g["points"])
    for code, vals in g["values"].items():
        print("device functional %s code %s" % (fname, code), flush=True)
        v = np.array(vals)
        xc = ca.ExchangeCorrelation().from_codes([int(code)])
        e, vr, vs, vt = xc.eval_points(P[:, 0], P[:, 1], P[:, 2] if ncol == 3 else None)
        assert _rel(e, v[:, 0]).max() < 1e-12
        assert _rel(vr, v[:, 1]).max() < 1e-11
        m = P[:, 1] > 0
        assert _rel(vs[m], v[m, 2]).max() < 1e-10
        if ncol == 3:
            assert _rel(vt, v[:, 3]).max() < 1e-9


def test_device_functionals_match_oracle_random_points():
    rng = np.random.default_rng(5)
    n = 4000
    rho = 10 ** rng.uniform(-12, 3, n)
    k2 = (3 * np.pi ** 2) ** (2 / 3)
    s = rng.uniform(0, 3, n)
    sig = 4 * k2 * rho ** (8 / 3) * s * s
    tw = sig / (8 * rho)
    tau = tw + rng.uniform(0, 4, n) * (0.3 * k2 * rho ** (5 / 3) + 0.001 * tw)
    for codes in ([321, 322], [402], [1, 7], [106, 131], [497, 498]):
        xc = ca.ExchangeCorrelation().from_codes(codes)
        d = xc.eval_points(rho, sig, tau)
        o = ol.eval_functional(codes, rho, sig, tau)
        # rounding differences are amplified by cancellation at very low density
        # (rs ~ 1e4 in PW92 / PBE correlation), where the values do not matter,
        # and in the sigma / tau derivatives near s = 0 (measured 3.9e-11)
        for name, a, b in zip(("exc", "vrho", "vsigma", "vtau"), d, o):
            tol = np.where(rho >= 1e-6, 1e-11 if name in ("exc", "vrho") else 1e-10, 1e-7)
            rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
            bad = rel > tol
            if bad.any():
                k = np.argmax(np.where(bad, rel, 0))
                print("  %s %s worst rel %.2e at rho %.3e s %.3f" % (codes, name, rel[k], rho[k], s[k]), flush=True)
            assert not bad.any(), (codes, name)


@pytest.mark.parametrize("cache", [1, 0])
@pytest.mark.parametrize("case", ["h2o", "ch4_tzvp"])
def test_r2scan_parity(case, cache):
    """r2SCAN through the FP64 form against the oracle, with the AO tables in HBM (cache 1) and
    with the AO values recomputed in the kernels (cache 0: tau partials and the tau term of V
    take three more passes over d_r phi)."""
    if case == "h2o":
        w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
        Z, xyz, basis, grid_name, nocc = w["Z"], w["xyz_bohr"], w["basis"], "SG-1", 5
    else:
        Z, X = wl.alkane_angstrom(1, 4)
        Z, xyz, basis, grid_name, nocc = Z, X * ca.ANGSTROM_TO_BOHR, "def2-tzvp", "SG-1", 5
    mol = ca.Molecule(Z, xyz, basis)
    D, S = wl.workload_density(mol, nocc, 11)
    print("r2SCAN parity %s (%d AOs), AO cache %d" % (case, mol.nbasis, cache), flush=True)
    grid = ca.Grid(mol, grid_name, 1, 0)
    grid.setAOCache(cache)
    xc = ca.ExchangeCorrelation()
    xc.Read("R2SCAN")
    assert xc.Family == "mGGA"
    grid.setType(2)
    grid.getAO(0, 0)
    assert grid.aoCached() == bool(cache)
    E, V = ca.xc_build(grid, xc, D)
    orc = ol.OracleSystem(Z, xyz, grid_name, basis)
    orc.set_functional("R2SCAN")
    orc.get_ao()
    Eo, No, Vo = orc.xc(D)
    print("  E gpu %.14f oracle %.14f dE %.2e dV %.2e" % (E, Eo, abs(E - Eo), np.abs(V - Vo).max()), flush=True)
    assert abs(E - Eo) <= 1e-10
    assert np.abs(V - Vo).max() <= 1e-9


def test_device_nonphysical_inputs():
    """The device functionals treat negative / NaN inputs as the oracle does (bit for bit on the
    values that are exactly zero or floored), unpolarised and polarised."""
    nan = float("nan")
    rho = np.array([-1e-3, nan, 0.0, 1e-20, 0.3, 0.3])
    sig = np.array([0.1, 0.1, 0.1, 0.1, nan, -0.5])
    tau = np.array([0.1, 0.1, 0.1, 0.1, nan, -0.5])
    for codes in ([1, 7], [101, 130], [402], [497, 498]):
        xc = ca.ExchangeCorrelation().from_codes(codes)
        d = xc.eval_points(rho, sig, tau)
        o = ol.eval_functional(codes, rho, sig, tau)
        for a, b in zip(d, o):
            assert np.all(np.isfinite(a)), codes
            assert np.array_equal(a[:4], b[:4])
            # floored sigma / tau: rounding-level agreement (derivatives that vanish there agree to 1e-17)
            assert np.all(np.abs(a[4:] - b[4:]) <= 1e-11 * np.abs(b[4:]) + 1e-14)
    xc = ca.ExchangeCorrelation()
    xc.Spin = 2
    xc.from_codes([101, 130])
    R = np.array([[nan, 0.1], [-1.0, 0.2], [0.0, 0.0]])
    S = np.array([[0.1, 0.0, 0.1]] * 3)
    e, vr, vs = xc.eval_points(R, S)
    eo, vro, vso = ol.eval_functional_pol([101, 130], R, S)
    assert np.all(np.isfinite(e)) and np.all(np.isfinite(vr)) and np.all(np.isfinite(vs))
    assert e[2] == 0 and not np.any(vr[2])
    assert np.all(np.abs(e - eo) <= 1e-11 * np.abs(eo))  # an empty channel is a fully polarised point


def test_density_matrix_of_wrong_size():
    w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    g = ca.Grid(mol, "SG-1", 1, 0)
    g.setType(1)
    g.getAO(0, 0)
    with pytest.raises(ca.ChxError) as e:
        g.getDensity([np.zeros((mol.nbasis + 1, mol.nbasis + 1))])
    assert e.value.code == -1
