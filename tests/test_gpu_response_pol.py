"""GPU spin-polarised response path (Evaluate("f") with the functional polarised ->
Grid::getDensityU -> Grid::getFockU<u_t> at nspin = 2, LDA/GGA, FP64 form) against the FP64 oracle,
whose polarised f_xc and response matrices are pinned by finite differences in
test_oracle_response_pol.py. Bars: f_xc 1e-11 against its natural scale (plus the rounding of
1 +- zeta near full polarisation); response matrices 1e-9 elementwise relative to their largest
element."""
import numpy as np
import pytest

import chinium_amd as ca
import oracle_lib as ol
from chinium_amd import workloads as wl

pytestmark = pytest.mark.gpu
F_TOL = 1e-9


def _pol_xc(func):
    xc = ca.ExchangeCorrelation()
    xc.Spin = 2
    xc.Read(func)
    return xc


@pytest.mark.parametrize("func", ["PBE", "B3LYP", "SVWN5", "BLYP"])
def test_fxc_pol_points(func):
    xc = _pol_xc(func)
    rng = np.random.default_rng(19)
    n = 2000
    ra, rb = 10.0 ** rng.uniform(-5, 2, n), 10.0 ** rng.uniform(-5, 2, n)
    saa = 10.0 ** rng.uniform(-4, 0.5, n) * 38.28312000250922 * ra ** (8.0 / 3.0)
    sbb = 10.0 ** rng.uniform(-4, 0.5, n) * 38.28312000250922 * rb ** (8.0 / 3.0)
    sab = rng.uniform(-0.95, 0.95, n) * np.sqrt(saa * sbb)
    rb[:3] = 0.0  # fully polarised points: the empty channel drops out
    rho, sigma = np.stack([ra, rb], 1), np.stack([saa, sab, sbb], 1)
    sigma[:3, 1:] = 0.0
    f0, f1, f2 = xc.eval_points_f(rho, sigma)
    _, g0, g1, g2 = ol.eval_functional_fpol(xc.Codes, rho, sigma)
    e = np.abs((ra + rb) * ol.eval_functional_pol(xc.Codes, rho, sigma)[0])
    r, s = (ra + rb)[:, None], (saa + sbb)[:, None] + 1e-300
    # 1 +- zeta from zeta = (ra - rb) / (ra + rb) carries the rounding of zeta relative to min(1 + z, 1 - z):
    # near full polarisation both sides lose those digits, in their own operation order
    m = np.minimum(2 * ra, 2 * rb) / (ra + rb)
    tol = 1e-11 + 1e-15 / np.where(m > 0, m, 1.0)[:, None]
    worst, where = 0.0, None
    for k, (a, b, sc) in enumerate(zip((f0, f1, f2), (g0, g1, g2),
                                       (e[:, None] / r ** 2, e[:, None] / (r * s), e[:, None] / s ** 2))):
        err = np.abs(a - b) / (np.abs(b) + sc) / tol
        if err.max() > worst:
            i, j = np.unravel_index(np.argmax(err), err.shape)
            worst, where = float(err.max()), (k, int(i), int(j), rho[i], sigma[i])
    print("  polarised f_xc %s: max diff / (|f| + scale) / bar %.2f at %s" % (func, worst, where), flush=True)
    assert worst <= 1.0


def _oh():
    Z, xyz = [8, 1], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.834]])
    mol = ca.Molecule(Z, xyz, "cc-pvdz")
    S = mol.overlap()
    Da, Db = 0.5 * wl.density_matrix(S, 5, 31), 0.5 * wl.density_matrix(S, 4, 32)
    return Z, xyz, mol, Da, Db


def _gpu_response(mol, grid_name, func, Da, Db, Xas, Xbs, cache=1):
    grid = ca.Grid(mol, grid_name, 1, 0)
    grid.setAOCache(cache)
    xc = _pol_xc(func)
    grid.setType(xc.family)
    grid.getAO(0, 0)
    grid.getDensity([Da, Db])
    xc.Evaluate("f", grid)
    grid.getDensityU([Xas, Xbs])
    return grid, grid.getFockU(2)


@pytest.mark.parametrize("func,cache", [("PBE", 1), ("B3LYP", 1), ("SVWN5", 1), ("BLYP", 1), ("PBE", 0)])
def test_oh_polarised_response_against_oracle(func, cache):
    Z, xyz, mol, Da, Db = _oh()
    rng = np.random.default_rng(6)
    Xas, Xbs = [], []
    for _ in range(2):
        A, B = rng.standard_normal(Da.shape), rng.standard_normal(Db.shape)
        Xas.append(0.05 * (A + A.T))
        Xbs.append(0.05 * (B + B.T))
    _, (Fa, Fb) = _gpu_response(mol, "SG-1", func, Da, Db, Xas, Xbs, cache)
    orc = ol.OracleSystem(Z, xyz, "SG-1", "cc-pvdz")
    orc.set_functional(func)
    orc.get_ao()
    Ra, Rb = orc.fock_u_pol(Da, Db, Xas, Xbs)
    worst = 0.0
    for F, R in zip(Fa + Fb, Ra + Rb):
        worst = max(worst, float(np.abs(F - R).max() / np.abs(R).max()))
        assert np.abs(F - F.T).max() <= 1e-12 * np.abs(F).max()
    print("  OH %s (AO tables %s): max |F_gpu - F_oracle| / max |F| %.1e" % (func, "cached" if cache else "recomputed",
                                                                            worst), flush=True)
    assert worst <= F_TOL


def test_polarised_response_closed_shell_limit():
    """Equal spins reproduce the closed-shell response: F_a = F_b = F[D, X] of nspin = 1."""
    w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D, _ = wl.workload_density(mol, w["nocc"], w["seed"])
    rng = np.random.default_rng(2)
    X = rng.standard_normal(D.shape)
    X = X + X.T
    g = ca.Grid(mol, "SG-1", 1, 0)
    xr = ca.ExchangeCorrelation()
    xr.Read("PBE")
    g.setType(1)
    g.getAO(0, 0)
    g.getDensity([D])
    xr.Evaluate("f", g)
    g.getDensityU([[X]])
    (F,), = g.getFockU(1)
    _, ((Fa,), (Fb,)) = _gpu_response(mol, "SG-1", "PBE", D / 2, D / 2, [X / 2], [X / 2])
    print("  closed-shell limit: |F_a - F| %.1e, |F_a - F_b| %.1e (max |F| %.1e)" % (
        np.abs(Fa - F).max(), np.abs(Fa - Fb).max(), np.abs(F).max()), flush=True)
    assert np.abs(Fa - Fb).max() <= 1e-12 * np.abs(F).max()
    assert np.abs(Fa - F).max() <= 1e-10 * np.abs(F).max()


def test_polarised_response_protocol():
    Z, xyz, mol, Da, Db = _oh()
    grid, _ = _gpu_response(mol, "SG-1", "PBE", Da, Db, [Da], [Db])
    with pytest.raises(ca.ChxError):  # spin count of getFockU must match getDensityU / Evaluate("f")
        grid.getFockU(1)
    g = ca.Grid(mol, "SG-1", 1, 0)
    g.setPrecision(ca.CHX_PREC_MIXED)
    g.setType(1)
    g.getAO(0, 0)
    g.getDensity([Da, Db])
    with pytest.raises(ca.ChxError):  # the response path is the FP64 form's
        _pol_xc("PBE").Evaluate("f", g)
    g2 = ca.Grid(mol, "SG-1", 1, 0)
    g2.setType(1)
    g2.getAO(0, 0)
    g2.getDensity([Da, Db])
    with pytest.raises(ca.ChxError):  # a closed-shell functional on a polarised density
        xc1 = ca.ExchangeCorrelation()
        xc1.Read("PBE")
        xc1.Evaluate("f", g2)
