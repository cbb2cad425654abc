"""Spin-polarised r2SCAN in the oracle (libxc 497 / 498 with XC_POLARIZED; libxc
itself is absent, so its spin conventions are parity unpinned). Pinned by what the
mathematics fixes:
  * zeta = 0 (rho_a = rho_b, tau_a = tau_b, sigma split evenly) reduces to the
    unpolarised r2SCAN, values and derivatives;
  * exchange obeys the spin-scaling relation E_x[ra, rb] = (E_x[2 ra] + E_x[2 rb]) / 2;
  * a one-electron density (zeta = 1, tau = tau_W) has zero correlation energy
    (r2SCAN is one-electron self-correlation free by construction: alpha = 0,
    f_c(0) = 1, G_c(1) = 0);
  * the uniform gas of either polarisation (sigma = 0, tau_s the Thomas-Fermi
    value of each spin) gives alpha = 1 and hence the polarised PW92 correlation
    and the spin-scaled Slater exchange;
  * v_rho, v_sigma, v_tau are the derivatives of e = rho eps (central differences);
  * the nspin = 2 build with Da = Db = D / 2 equals the closed-shell build, and
    its potentials are the derivatives of E_xc for Da != Db."""
import numpy as np
import pytest

import oracle_lib as ol
from chinium_amd import workloads as wl

R2 = [497, 498]
CTF = 0.3 * (6 * np.pi ** 2) ** (2 / 3)  # tau of a uniform spin channel: CTF rho_s^(5/3)


def _points(n, seed, zeta=None):
    rng = np.random.default_rng(seed)
    rho = 10 ** rng.uniform(-3, 1.5, n)
    z = rng.uniform(-0.9, 0.9, n) if zeta is None else np.full(n, zeta)
    ra, rb = rho * (1 + z) / 2, rho * (1 - z) / 2
    ga = rng.standard_normal((n, 3)) * np.sqrt(10 ** rng.uniform(-2, 0.3, n) * 38.28 * ra ** (8 / 3) / 3)[:, None]
    gb = rng.standard_normal((n, 3)) * np.sqrt(10 ** rng.uniform(-2, 0.3, n) * 38.28 * rb ** (8 / 3) / 3)[:, None]
    saa, sab, sbb = np.sum(ga * ga, 1), np.sum(ga * gb, 1), np.sum(gb * gb, 1)
    ta = saa / (8 * ra) + CTF * ra ** (5 / 3) * rng.uniform(0.1, 3, n)  # above tau_W
    tb = sbb / (8 * rb) + CTF * rb ** (5 / 3) * rng.uniform(0.1, 3, n)
    return np.column_stack([ra, rb]), np.column_stack([saa, sab, sbb]), np.column_stack([ta, tb])


@pytest.mark.parametrize("codes", [[497], [498], R2])
def test_zeta_zero_reduces_to_unpolarised(codes):
    rng = np.random.default_rng(1)
    rho = 10 ** rng.uniform(-3, 2, 60)
    sig = 10 ** rng.uniform(-2, 0.5, 60) * 38.28 * rho ** (8 / 3)
    tau = sig / (8 * rho) + 0.3 * (3 * np.pi ** 2) ** (2 / 3) * rho ** (5 / 3) * rng.uniform(0.05, 4, 60)
    e0, vr0, vs0, vt0 = ol.eval_functional(codes, rho, sig, tau)
    e, vr, vs, vt = ol.eval_functional_mpol(codes, np.column_stack([rho / 2, rho / 2]),
                                            np.column_stack([sig / 4, sig / 4, sig / 4]),
                                            np.column_stack([tau / 2, tau / 2]))
    np.testing.assert_allclose(e, e0, rtol=1e-12)
    np.testing.assert_allclose(vr[:, 0], vr0, rtol=1e-10)
    np.testing.assert_allclose(vr[:, 1], vr0, rtol=1e-10)
    np.testing.assert_allclose(vs.sum(axis=1) / 4, vs0, rtol=1e-10)
    np.testing.assert_allclose(vt[:, 0], vt0, rtol=1e-10)
    np.testing.assert_allclose(vt[:, 1], vt0, rtol=1e-10)


def test_exchange_spin_scaling():
    rho, sig, tau = _points(40, 3)
    e, _, _, _ = ol.eval_functional_mpol([497], rho, sig, tau)
    ea, _, _, _ = ol.eval_functional([497], 2 * rho[:, 0], 4 * sig[:, 0], 2 * tau[:, 0])
    eb, _, _, _ = ol.eval_functional([497], 2 * rho[:, 1], 4 * sig[:, 2], 2 * tau[:, 1])
    tot = rho.sum(axis=1)
    np.testing.assert_allclose(e * tot, 0.5 * (ea * 2 * rho[:, 0] + eb * 2 * rho[:, 1]), rtol=1e-12)


def test_one_electron_density_has_no_correlation():
    rng = np.random.default_rng(4)
    ra = 10 ** rng.uniform(-3, 1, 30)
    saa = 10 ** rng.uniform(-2, 0.5, 30) * 38.28 * ra ** (8 / 3)
    rho = np.column_stack([ra, np.zeros(30)])
    sig = np.column_stack([saa, np.zeros(30), np.zeros(30)])
    tau = np.column_stack([saa / (8 * ra), np.zeros(30)])  # a single orbital: tau = tau_W
    ec, _, _, _ = ol.eval_functional_mpol([498], rho, sig, tau)
    ex, _, _, _ = ol.eval_functional_mpol([497], rho, sig, tau)
    assert np.all(np.abs(ec) <= 1e-12 * np.abs(ex))
    # a partly polarised density or tau above tau_W does correlate
    ec2, _, _, _ = ol.eval_functional_mpol([498], rho, sig, tau * 1.5)
    assert np.all(ec2 < -1e-8)


@pytest.mark.parametrize("zeta", [0.0, 0.3, -0.7, 1.0])
def test_uniform_gas(zeta):
    rho = np.array([1e-3, 0.05, 1.0, 30.0])
    ra, rb = rho * (1 + zeta) / 2, rho * (1 - zeta) / 2
    R = np.column_stack([ra, rb])
    S = np.zeros((4, 3))
    T = np.column_stack([CTF * ra ** (5 / 3), CTF * rb ** (5 / 3)])
    ec, _, _, _ = ol.eval_functional_mpol([498], R, S, T)
    ex, _, _, _ = ol.eval_functional_mpol([497], R, S, T)
    # references: the polarised PW92 inside PBE correlation at sigma = 0 and spin-scaled Slater exchange
    pw, _, _ = ol.eval_functional_pol([322], R, S)
    sl, _, _ = ol.eval_functional_pol([1], R, S)
    np.testing.assert_allclose(ec, pw, rtol=1e-11)
    np.testing.assert_allclose(ex, sl, rtol=1e-11)


@pytest.mark.parametrize("code", [497, 498])
def test_potentials_are_derivatives(code):
    rho, sig, tau = _points(25, 7)
    X = np.column_stack([rho, sig, tau])  # 7 variables per point

    def e_of(X):
        e, _, _, _ = ol.eval_functional_mpol([code], X[:, 0:2], X[:, 2:5], X[:, 5:7])
        return e * X[:, 0:2].sum(axis=1)

    _, vr, vs, vt = ol.eval_functional_mpol([code], rho, sig, tau)
    an = np.column_stack([vr, vs, vt])
    for k in range(7):
        if k == 3:  # sigma_ab: move it inside the Cauchy-Schwarz cap
            h = 1e-5 * np.sqrt(X[:, 2] * X[:, 4])
        else:
            h = 1e-5 * X[:, k]
        def at(t):
            Y = X.copy(); Y[:, k] += t
            return e_of(Y)
        d1 = (at(h) - at(-h)) / (2 * h)
        d2 = (at(h / 2) - at(-h / 2)) / h
        fd = (4 * d2 - d1) / 3
        scale = np.abs(e_of(X)) / np.maximum(np.abs(X[:, k]), h)
        assert np.all(np.abs(fd - an[:, k]) <= 1e-6 * (np.abs(an[:, k]) + 1e-3 * scale)), (code, k)


@pytest.fixture(scope="module")
def water():
    import chinium_amd as ca
    w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
    o = ol.OracleSystem(w["Z"], w["xyz_bohr"], w["grid"], w["basis"])
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D, S = wl.workload_density(mol, w["nocc"], w["seed"])
    o.set_functional("R2SCAN")
    o.get_ao()
    return o, D


def test_equal_spins_is_closed_shell(water):
    o, D = water
    E0, N0, V0 = o.xc(D)
    E, N, Va, Vb = o.xc_pol(D / 2, D / 2)
    assert abs(E - E0) <= 1e-11 * abs(E0) and abs(N - N0) <= 1e-12 * N0
    assert np.abs(Va - V0).max() <= 1e-10 * np.abs(V0).max()
    assert np.abs(Vb - V0).max() <= 1e-10 * np.abs(V0).max()


def test_spin_potentials_are_derivatives(water):
    o, D = water
    rng = np.random.default_rng(2)
    Da, Db = 0.6 * D, 0.4 * D
    A = rng.standard_normal(D.shape); A = 0.02 * (A + A.T)
    B = rng.standard_normal(D.shape); B = 0.02 * (B + B.T)
    E, N, Va, Vb = o.xc_pol(Da, Db)
    assert np.abs(Va - Va.T).max() <= 1e-13 * np.abs(Va).max()
    Ef = lambda t: o.xc_pol(Da + t * A, Db + t * B)[0]
    h = 1e-3
    d1 = (Ef(h) - Ef(-h)) / (2 * h)
    d2 = (Ef(h / 2) - Ef(-h / 2)) / h
    fd = (4 * d2 - d1) / 3
    an = np.sum(Va * A) + np.sum(Vb * B)
    assert abs(an - fd) <= 1e-7 * abs(fd), (an, fd)
