"""GPU response path (Evaluate("f") -> Grid::getDensityU -> Grid::getFockU<u_t>,
closed shell, LDA/GGA, FP64 form) against the FP64 oracle, whose f_xc and
response matrices are pinned by finite differences in test_oracle_response.py.
Bars: f_xc 1e-11 relative; response matrices 1e-9 elementwise."""
import numpy as np
import pytest

import chinium_amd as ca
import oracle_lib as ol
from chinium_amd import workloads as wl

pytestmark = pytest.mark.gpu
F_TOL = 1e-9


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("func", ["PBE", "B3LYP", "SVWN5", "BLYP"])
def test_fxc_points(func):
    xc = ca.ExchangeCorrelation()
    xc.Read(func)
    rng = np.random.default_rng(9)
    rho = 10.0 ** rng.uniform(-6, 2, 2000)
    sigma = 10.0 ** rng.uniform(-4, 0.5, 2000) * 38.28312000250922 * rho ** (8.0 / 3.0)
    rho[:5] = [0.0, 1e-16, 1e-15, 2e-15, 1e-12]  # threshold neighbourhood
    f0, f1, f2 = xc.eval_points_f(rho, sigma)
    vs, g0, g1, g2 = ol.eval_functional_f(xc.Codes, rho, sigma)
    exc = ol.eval_functional(xc.Codes, rho, sigma)[0]
    e = np.abs(rho * exc)  # energy density: the natural scale of e'' is e / rho^2, e / (rho sigma), e / sigma^2
    with np.errstate(divide="ignore", invalid="ignore"):
        scales = (e / rho ** 2, e / (rho * sigma), e / sigma ** 2)
    worst = 0.0
    for a, b, sc in zip((f0, f1, f2), (g0, g1, g2), scales):
        assert np.all(a[:2] == 0) and np.all(b[:2] == 0)  # below the density threshold
        # different operation orders: rounding-level agreement against the natural scale
        err = np.abs(a[2:] - b[2:]) / (np.abs(b[2:]) + sc[2:])
        worst = max(worst, float(err.max()))
    _say("  f_xc %s: max diff / (|f| + scale) %.1e" % (func, worst))
    assert worst <= 1e-11


def _h2o():
    w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D, S = wl.workload_density(mol, w["nocc"], w["seed"])
    return w, mol, D


def _response(mol, grid_name, func, D, Dxs, cache):
    grid = ca.Grid(mol, grid_name, 1, 0)
    grid.setAOCache(cache)
    xc = ca.ExchangeCorrelation()
    xc.Read(func)
    grid.setType(xc.family)
    grid.getAO(0, 0)
    grid.getDensity([D])
    xc.Evaluate("f", grid)
    grid.getDensityU([Dxs])
    return grid, xc, grid.getFockU()[0]


@pytest.mark.parametrize("func,cache", [("PBE", 1), ("B3LYP", 1), ("SVWN5", 1), ("PBE", 0)])
def test_h2o_response_against_oracle(func, cache):
    w, mol, D = _h2o()
    rng = np.random.default_rng(4)
    Dxs = []
    for _ in range(3):
        A = rng.standard_normal(D.shape)
        Dxs.append(0.1 * (A + A.T))
    _say("response h2o/cc-pvdz/SG-1 %s, AO cache %d, 3 perturbations" % (func, cache))
    grid, xc, Fs = _response(mol, "SG-1", func, D, Dxs, cache)
    orc = ol.OracleSystem(w["Z"], w["xyz_bohr"], "SG-1", w["basis"])
    orc.set_functional(func)
    orc.get_ao()
    Fo = orc.fock_u(D, Dxs)
    for F, G in zip(Fs, Fo):
        _say("  max |F - F_oracle| %.2e  max |F| %.3f" % (np.abs(F - G).max(), np.abs(G).max()))
        assert np.abs(F - G).max() <= F_TOL
        assert np.abs(F - F.T).max() <= 1e-12


def test_water_cluster_response():
    Z, X = wl.water_cluster_angstrom(4, 5.0, 11)
    xyz = X * ca.ANGSTROM_TO_BOHR
    mol = ca.Molecule(Z, xyz, "def2-svp")
    D, S = wl.workload_density(mol, 20, 3)
    rng = np.random.default_rng(8)
    A = rng.standard_normal(D.shape)
    Dx = 0.05 * (A + A.T)
    _say("response (H2O)4/def2-svp/SG-1 B3LYP")
    grid, xc, Fs = _response(mol, "SG-1", "B3LYP", D, [Dx], 1)
    orc = ol.OracleSystem(Z, xyz, "SG-1", "def2-svp")
    orc.set_functional("B3LYP")
    orc.get_ao()
    (Fo,) = orc.fock_u(D, [Dx])
    _say("  max |F - F_oracle| %.2e" % np.abs(Fs[0] - Fo).max())
    assert np.abs(Fs[0] - Fo).max() <= F_TOL
    # the ground-state build on the same grid is unaffected by the response calls
    E, V = ca.xc_build(grid, xc, D)
    Eo, No, Vo = orc.xc(D)
    assert abs(E - Eo) <= 1e-10 and np.abs(V - Vo).max() <= 1e-9


def test_response_protocol_errors():
    w, mol, D = _h2o()
    grid = ca.Grid(mol, "SG-1", 1, 0)
    xc = ca.ExchangeCorrelation()
    xc.Read("PBE")
    grid.setType(1)
    grid.getAO(0, 0)
    grid.getDensity([D])
    grid.getDensityU([[D]])
    with pytest.raises(ca.ChxError) as e:  # Evaluate("f") has not run
        grid.getFockU()
    assert e.value.code == -3
    r2 = ca.ExchangeCorrelation()
    r2.Read("R2SCAN")
    grid.setType(2)
    grid.getAO(0, 0)
    grid.getDensity([D])
    with pytest.raises(ca.ChxError) as e:  # meta-GGA f_xc is not built
        r2.Evaluate("f", grid)
    assert e.value.code == -5
    m = ca.Grid(mol, "SG-1", 1, 0)
    m.setPrecision(ca.CHX_PREC_MIXED)
    m.setType(1)
    m.getAO(0, 0)
    m.getDensity([D])
    with pytest.raises(ca.ChxError) as e:  # FP64 form only
        xc.Evaluate("f", m)
    assert e.value.code == -5


def _skeleton(mol, grid_name, func, D):
    grid = ca.Grid(mol, grid_name, 1, 0)
    xc = ca.ExchangeCorrelation()
    xc.Read(func)
    grid.setType(xc.family)
    grid.getAO(0, 0)
    grid.getDensity([D])
    # the reference's Hessian order (Restricted/Hess.cpp:78-96)
    grid.getAO(2, 1)
    xc.Evaluate("f", grid)
    grid.getDensitySkeleton([D])
    return grid, xc, grid.getFockU(d_t="s")[0]


@pytest.mark.parametrize("func", ["PBE", "B3LYP", "SVWN5"])
def test_h2o_skeleton_response_against_oracle(func):
    """Grid::getFockU<s_t>: one response matrix per nuclear coordinate from the skeleton density
    derivatives, against the oracle (pinned by finite differences of V with one atom's basis moved)."""
    w, mol, D = _h2o()
    _say("skeleton response h2o/cc-pvdz/SG-1 %s, 9 coordinates" % func)
    grid, xc, Fs = _skeleton(mol, "SG-1", func, D)
    orc = ol.OracleSystem(w["Z"], w["xyz_bohr"], "SG-1", w["basis"])
    orc.set_functional(func)
    orc.get_ao()
    Fo = orc.fock_s(D)
    assert len(Fs) == len(Fo) == 9
    worst = max(float(np.abs(F - G).max()) for F, G in zip(Fs, Fo))
    _say("  max |F - F_oracle| %.2e  max |F| %.3f" % (worst, max(float(np.abs(G).max()) for G in Fo)))
    assert worst <= F_TOL
    assert all(np.abs(F - F.T).max() <= 1e-12 for F in Fs)


def test_water_cluster_skeleton_response():
    Z, X = wl.water_cluster_angstrom(4, 5.0, 11)
    xyz = X * ca.ANGSTROM_TO_BOHR
    mol = ca.Molecule(Z, xyz, "def2-svp")
    D, S = wl.workload_density(mol, 20, 3)
    _say("skeleton response (H2O)4/def2-svp/SG-1 B3LYP, 36 coordinates")
    grid, xc, Fs = _skeleton(mol, "SG-1", "B3LYP", D)
    orc = ol.OracleSystem(Z, xyz, "SG-1", "def2-svp")
    orc.set_functional("B3LYP")
    orc.get_ao()
    Fo = orc.fock_s(D)
    worst = max(float(np.abs(F - G).max()) for F, G in zip(Fs, Fo))
    _say("  max |F - F_oracle| %.2e" % worst)
    assert worst <= F_TOL
    # the same grid then still gives the ground-state build and the u_t response
    E, V = ca.xc_build(grid, xc, D)
    Eo, No, Vo = orc.xc(D)
    assert abs(E - Eo) <= 1e-10 and np.abs(V - Vo).max() <= 1e-9


def test_skeleton_response_protocol_errors():
    w, mol, D = _h2o()
    grid = ca.Grid(mol, "SG-1", 1, 0)
    xc = ca.ExchangeCorrelation()
    xc.Read("PBE")
    grid.setType(1)
    grid.getAO(0, 0)
    grid.getDensity([D])
    xc.Evaluate("f", grid)
    with pytest.raises(ca.ChxError) as e:  # getDensitySkeleton has not run
        grid.getFockU(d_t="s")
    assert e.value.code == -3
    with pytest.raises(ca.ChxError) as e:
        grid.getAO(3, 0)
    assert e.value.code == -5


def _skeleton_fock(mol, grid_name, func, D):
    grid = ca.Grid(mol, grid_name, 1, 0)
    xc = ca.ExchangeCorrelation()
    xc.Read(func)
    grid.setType(xc.family)
    grid.getAO(0, 0)
    grid.getDensity([D])
    xc.Evaluate("ev", grid)  # the SCF's last build
    # the reference's Hessian order (Restricted/Hess.cpp:78-96)
    grid.getAO(2, 1)
    xc.Evaluate("f", grid)
    grid.getDensitySkeleton([D])
    Fk = grid.getFockSkeleton()[0]
    Fs = grid.getFockU(d_t="s")[0]
    return grid, xc, Fk, Fs


@pytest.mark.parametrize("func", ["PBE", "B3LYP", "SVWN5"])
def test_h2o_fock_skeleton_against_oracle(func):
    """Grid::getFockSkeleton: the XC Fock matrix's derivative along each nuclear coordinate with the
    potential held fixed, against the oracle (pinned by finite differences of V assembled on one atom's
    moved basis functions); the sum with getFockU<s_t> against the oracle's sum."""
    w, mol, D = _h2o()
    _say("skeleton Fock derivative h2o/cc-pvdz/SG-1 %s, 9 coordinates" % func)
    grid, xc, Fk, Fs = _skeleton_fock(mol, "SG-1", func, D)
    orc = ol.OracleSystem(w["Z"], w["xyz_bohr"], "SG-1", w["basis"])
    orc.set_functional(func)
    orc.get_ao()
    Fo = orc.fock_skeleton(D)
    Fso = orc.fock_s(D)
    assert len(Fk) == len(Fo) == 9
    scale = max(float(np.abs(G).max()) for G in Fo)
    worst = max(float(np.abs(F - G).max()) for F, G in zip(Fk, Fo))
    worst_sum = max(float(np.abs(F + S - G - T).max()) for F, S, G, T in zip(Fk, Fs, Fo, Fso))
    _say("  max |F - F_oracle| %.2e, of the total derivative %.2e, max |F| %.3f" % (worst, worst_sum, scale))
    assert worst <= F_TOL and worst_sum <= F_TOL
    assert all(np.abs(F - F.T).max() <= 1e-12 * scale for F in Fk)


def test_water_cluster_fock_skeleton():
    Z, X = wl.water_cluster_angstrom(4, 5.0, 11)
    xyz = X * ca.ANGSTROM_TO_BOHR
    mol = ca.Molecule(Z, xyz, "def2-svp")
    D, S = wl.workload_density(mol, 20, 3)
    _say("skeleton Fock derivative (H2O)4/def2-svp/SG-1 B3LYP, 36 coordinates")
    grid, xc, Fk, Fs = _skeleton_fock(mol, "SG-1", "B3LYP", D)
    orc = ol.OracleSystem(Z, xyz, "SG-1", "def2-svp")
    orc.set_functional("B3LYP")
    orc.get_ao()
    Fo = orc.fock_skeleton(D)
    worst = max(float(np.abs(F - G).max()) for F, G in zip(Fk, Fo))
    _say("  max |F - F_oracle| %.2e" % worst)
    assert len(Fk) == 36 and worst <= F_TOL


def test_fock_skeleton_protocol_errors():
    w, mol, D = _h2o()
    grid = ca.Grid(mol, "SG-1", 1, 0)
    xc = ca.ExchangeCorrelation()
    xc.Read("PBE")
    grid.setType(1)
    grid.getAO(0, 0)
    grid.getDensity([D])
    with pytest.raises(ca.ChxError) as e:  # no potential yet
        grid.getFockSkeleton()
    assert e.value.code == -3
    m = ca.Grid(mol, "SG-1", 1, 0)
    r2 = ca.ExchangeCorrelation()
    r2.Read("R2SCAN")
    m.setType(2)
    m.getAO(0, 0)
    m.getDensity([D])
    r2.Evaluate("ev", m)
    with pytest.raises(ca.ChxError) as e:  # meta-GGA: not built (the reference's has no tau term)
        m.getFockSkeleton()
    assert e.value.code == -5


def test_water_dimer_energy_hess():
    """The skeleton XC Hessian on (H2O)2 / def2-SVP / SG-1 / PBE (six atoms: atom pairs across
    molecules) against the oracle."""
    Z, X = wl.water_cluster_angstrom(2, 4.0, 5)
    xyz = X * ca.ANGSTROM_TO_BOHR
    mol = ca.Molecule(Z, xyz, "def2-svp")
    D, S = wl.workload_density(mol, 10, 7)
    grid = ca.Grid(mol, "SG-1", 1, 0)
    xc = ca.ExchangeCorrelation()
    xc.Read("PBE")
    grid.setType(1)
    grid.getAO(0, 0)
    grid.getDensity([D])
    xc.Evaluate("ev", grid)
    grid.getAO(2, 1)
    xc.Evaluate("f", grid)
    grid.getDensitySkeleton2([D])
    H = grid.getEnergyHess()
    orc = ol.OracleSystem(Z, xyz, "SG-1", "def2-svp")
    orc.set_functional("PBE")
    orc.get_ao()
    Ho = orc.xc_hess(D)
    err = float(np.abs(H - Ho).max())
    _say("  (H2O)2 PBE: max |H - H_oracle| %.2e, max |H| %.3f" % (err, np.abs(Ho).max()))
    assert H.shape == (18, 18) and err <= 1e-9 * np.abs(Ho).max()


@pytest.mark.parametrize("func", ["PBE", "B3LYP", "SVWN5"])
def test_h2o_energy_hess_against_oracle(func):
    """Grid::getDensitySkeleton2 + Grid::getEnergyHess: the skeleton XC Hessian against the oracle
    (pinned by finite differences of the pinned skeleton gradient); symmetric."""
    w, mol, D = _h2o()
    _say("XC energy Hessian h2o/cc-pvdz/SG-1 %s" % func)
    grid = ca.Grid(mol, "SG-1", 1, 0)
    xc = ca.ExchangeCorrelation()
    xc.Read(func)
    grid.setType(xc.family)
    grid.getAO(0, 0)
    grid.getDensity([D])
    xc.Evaluate("ev", grid)
    grid.getAO(2, 1)
    xc.Evaluate("f", grid)
    grid.getDensitySkeleton2([D])
    H = grid.getEnergyHess()
    orc = ol.OracleSystem(w["Z"], w["xyz_bohr"], "SG-1", w["basis"])
    orc.set_functional(func)
    orc.get_ao()
    Ho = orc.xc_hess(D)
    err = float(np.abs(H - Ho).max())
    _say("  max |H - H_oracle| %.2e, max |H| %.3f, asymmetry %.1e" % (err, np.abs(Ho).max(), np.abs(H - H.T).max()))
    assert H.shape == (9, 9)
    assert err <= 1e-9 * np.abs(Ho).max()
    assert np.abs(H - H.T).max() <= 1e-11 * np.abs(Ho).max()


def test_energy_hess_protocol_errors():
    w, mol, D = _h2o()
    grid = ca.Grid(mol, "SG-1", 1, 0)
    xc = ca.ExchangeCorrelation()
    xc.Read("PBE")
    grid.setType(1)
    grid.getAO(0, 0)
    grid.getDensity([D])
    xc.Evaluate("ev", grid)
    with pytest.raises(ca.ChxError) as e:  # getDensitySkeleton2 has not run
        grid.getEnergyHess()
    assert e.value.code == -3
    grid.getDensitySkeleton2([D])
    with pytest.raises(ca.ChxError) as e:  # Evaluate("f") has not run
        grid.getEnergyHess()
    assert e.value.code == -3
