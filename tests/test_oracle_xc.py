"""Oracle XC build (getDensity -> Evaluate -> getEnergy -> getFock) pinned by
identities that hold for any correct implementation:
  * Vxc is the derivative of E_xc: Tr(V dD) = dE_xc[D + t dD]/dt at t = 0;
  * Slater exchange is homogeneous of degree 4/3 in D: sum D V = 4/3 E_x;
  * N_e on the grid equals Tr(D S) to quadrature accuracy;
  * sums over sub-grids are additive (any batching gives the same result)."""
import numpy as np
import pytest

import chinium_amd as ca
import oracle_lib as ol
from chinium_amd import workloads as wl


@pytest.fixture(scope="module")
def h2o():
    w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D, S = wl.workload_density(mol, w["nocc"], w["seed"])
    o = ol.OracleSystem(w["Z"], w["xyz_bohr"], w["grid"], w["basis"])
    return o, D, S


@pytest.mark.parametrize("func", ["PBE", "B3LYP", "SVWN5", "BLYP"])
def test_vxc_is_derivative_of_exc(h2o, func):
    o, D, S = h2o
    o.set_functional(func)
    o.get_ao()
    E, N, V = o.xc(D)
    rng = np.random.default_rng(7)
    dD = rng.standard_normal(D.shape)
    dD = 0.05 * (dD + dD.T)
    f = lambda t: o.xc(D + t * dD)[0]
    # Richardson-extrapolated central difference (error O(h^4))
    h = 1e-3
    d1 = (f(h) - f(-h)) / (2 * h)
    d2 = (f(h / 2) - f(-h / 2)) / h
    fd = (4 * d2 - d1) / 3
    an = np.sum(V * dD)
    assert abs(fd - an) <= 1e-8 * max(1.0, abs(an))


def test_slater_homogeneity(h2o):
    o, D, S = h2o
    o.set_functional("SLATER")
    o.get_ao()
    E, N, V = o.xc(D)
    assert abs(np.sum(D * V) - 4.0 / 3.0 * E) < 1e-12
    # E_x scales as lambda^(4/3) under D -> lambda D
    E2, _, _ = o.xc(2.0 * D)
    assert abs(E2 - 2.0 ** (4.0 / 3.0) * E) < 1e-12


def test_electron_count_and_symmetry(h2o):
    o, D, S = h2o
    o.set_functional("PBE")
    o.get_ao()
    E, N, V = o.xc(D)
    assert abs(np.trace(D @ S) - 10.0) < 1e-12
    assert abs(N - 10.0) < 1e-4           # SG-1 quadrature accuracy
    assert np.abs(V - V.T).max() < 1e-14
    rho, grad, sig, tau, w = o.density()
    assert np.all(rho >= -1e-14)
    assert np.allclose(sig, np.sum(grad ** 2, axis=1), rtol=1e-12, atol=1e-300)


def test_subgrid_additivity(h2o):
    o, D, S = h2o
    o.set_functional("B3LYP")
    o.set_subset(None)
    o.get_ao()
    E, N, V = o.xc(D)
    ids = np.arange(o.nsubgrids)
    a, b = ids[::3], np.setdiff1d(ids, ids[::3])
    o.set_subset(a)
    o.get_ao()
    Ea, Na, Va = o.xc(D)
    o.set_subset(b)
    o.get_ao()
    Eb, Nb, Vb = o.xc(D)
    o.set_subset(None)
    assert abs(Ea + Eb - E) < 1e-12
    assert np.abs(Va + Vb - V).max() < 1e-12


def test_unknown_code_rejected(h2o):
    o, D, S = h2o
    with pytest.raises(RuntimeError):
        o.set_functional(codes=[9999])


def test_r2scan_vxc_is_derivative(h2o):
    o, D, S = h2o
    o.set_functional("R2SCAN")
    assert o.family == 2
    o.get_ao()
    E, N, V = o.xc(D)
    rng = np.random.default_rng(8)
    dD = rng.standard_normal(D.shape)
    dD = 0.05 * (dD + dD.T)
    f = lambda t: o.xc(D + t * dD)[0]
    h = 1e-3
    d1 = (f(h) - f(-h)) / (2 * h)
    d2 = (f(h / 2) - f(-h / 2)) / h
    fd = (4 * d2 - d1) / 3
    an = np.sum(V * dD)
    assert abs(fd - an) <= 1e-7 * max(1.0, abs(an))


@pytest.mark.parametrize("func", ["PBE", "SVWN5", "B3LYP"])
def test_xc_gradient_is_derivative(h2o, func):
    """The skeleton XC gradient (grid fixed, GridEnergy.cpp:15-30) equals the
    derivative of E_xc with respect to moving one atom's basis functions,
    by Richardson-extrapolated central differences (error O(h^4))."""
    o, D, S = h2o
    o.set_functional(func)
    o.get_ao()
    G = o.xc_grad(D)
    h = 1e-4
    Gfd = np.zeros_like(G)
    for A in range(3):
        for t in range(3):
            d = np.zeros(3)
            d[t] = h
            vals = []
            for sgn in (1, -1, 2, -2):
                o.shift_basis(A, sgn * d)
                o.get_ao()
                vals.append(o.xc(D)[0])
                o.shift_basis(A, -sgn * d)
            d1 = (vals[0] - vals[1]) / (2 * h)
            d2 = (vals[2] - vals[3]) / (4 * h)
            Gfd[A, t] = (4 * d1 - d2) / 3
    o.get_ao()
    assert np.abs(G - Gfd).max() <= 1e-9
    assert np.abs(G).max() > 0.1


def test_xc_gradient_rejects_meta_gga(h2o):
    o, D, S = h2o
    o.set_functional("R2SCAN")
    o.get_ao()
    with pytest.raises(RuntimeError):
        o.xc_grad(D)
