"""Oracle response path (Evaluate("f") -> getDensityU -> getFockU<u_t>) pinned
by what the mathematics fixes, with no libxc:
  * f_xc is the derivative of v_xc: v2rho2 = d v_rho / d rho,
    v2rhosigma = d v_sigma / d rho = d v_rho / d sigma, v2sigma2 = d v_sigma / d sigma
    (central differences of the oracle's exact first derivatives);
  * the response Fock matrix is the linear response of V: F[D^x] = dV[D + t D^x]/dt
    at t = 0 (central differences of the oracle's full XC build);
  * F is linear in D^x;
  * the skeleton response (getFockU<s_t>) is the derivative of V[rho] with the
    density taken from one atom's displaced basis functions (grid and the AO
    values that assemble V held fixed): central differences of V assembled here
    in numpy from the oracle's per-point functional, an assembly checked first
    against the oracle's own build."""
import numpy as np
import pytest

import chinium_amd as ca
import oracle_lib as ol
from chinium_amd import workloads as wl

FUNCS = {"SVWN5": [1, 7], "VWN_RPA": [8], "PBE": [101, 130], "PBE_libxc": [321, 322], "B3LYP": [402],
         "BLYP": [106, 131]}


def _points(n, seed):
    rng = np.random.default_rng(seed)
    rho = 10.0 ** rng.uniform(-3, 1, n)
    s2 = 10.0 ** rng.uniform(-3, 0.5, n)          # reduced gradient squared
    sigma = s2 * 38.28312000250922 * rho ** (8.0 / 3.0)
    return rho, sigma


@pytest.mark.parametrize("name", sorted(FUNCS))
def test_fxc_is_derivative_of_vxc(name):
    codes = FUNCS[name]
    rho, sigma = _points(40, 11)
    vs, frr, frs, fss = ol.eval_functional_f(codes, rho, sigma)
    _, vr0, vs0, _ = ol.eval_functional(codes, rho, sigma)
    np.testing.assert_allclose(vs, vs0, rtol=1e-11, atol=0)  # same derivative, other operation order

    def d(fn, x, h):  # Richardson-extrapolated central difference
        d1 = (fn(x + h) - fn(x - h)) / (2 * h)
        d2 = (fn(x + h / 2) - fn(x - h / 2)) / h
        return (4 * d2 - d1) / 3

    hr, hs = 1e-4 * rho, 1e-4 * sigma
    vr_r = d(lambda r: ol.eval_functional(codes, r, sigma)[1], rho, hr)
    vs_r = d(lambda r: ol.eval_functional(codes, r, sigma)[2], rho, hr)
    vr_s = d(lambda s: ol.eval_functional(codes, rho, s)[1], sigma, hs)
    vs_s = d(lambda s: ol.eval_functional(codes, rho, s)[2], sigma, hs)
    # truncation (1e-7 relative) plus the rounding of the differenced values
    # (|v| eps / h, which dominates where v hardly changes over the step)
    tol = lambda ref, v, h: 1e-7 * np.abs(ref) + 1e-13 * np.abs(v) / h
    assert np.all(np.abs(frr - vr_r) <= tol(frr, vr0, hr))
    if codes[0] not in (1, 7, 8):  # GGA: the mixed derivatives agree both ways
        assert np.all(np.abs(frs - vs_r) <= tol(frs, vs0, hr))
        assert np.all(np.abs(frs - vr_s) <= tol(frs, vr0, hs))
        assert np.all(np.abs(fss - vs_s) <= tol(fss, vs0, hs))
    else:
        assert np.all(frs == 0) and np.all(fss == 0)


def test_fxc_below_threshold_is_zero():
    vs, frr, frs, fss = ol.eval_functional_f([101, 130], np.array([1e-16, 0.0]), np.array([1e-30, 0.0]))
    assert np.all(frr == 0) and np.all(frs == 0) and np.all(fss == 0) and np.all(vs == 0)


@pytest.fixture(scope="module")
def h2o():
    w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D, S = wl.workload_density(mol, w["nocc"], w["seed"])
    o = ol.OracleSystem(w["Z"], w["xyz_bohr"], w["grid"], w["basis"])
    return o, D


@pytest.mark.parametrize("func", ["PBE", "B3LYP", "SVWN5"])
def test_fock_u_is_linear_response_of_v(h2o, func):
    o, D = h2o
    o.set_functional(func)
    o.get_ao()
    rng = np.random.default_rng(3)
    Dx = rng.standard_normal(D.shape)
    Dx = 0.05 * (Dx + Dx.T)
    (F,) = o.fock_u(D, [Dx])
    V = lambda t: o.xc(D + t * Dx)[2]
    h = 1e-3
    d1 = (V(h) - V(-h)) / (2 * h)
    d2 = (V(h / 2) - V(-h / 2)) / h
    fd = (4 * d2 - d1) / 3
    assert np.abs(F - F.T).max() <= 1e-13 * np.abs(F).max()
    assert np.abs(F - fd).max() <= 1e-7 * np.abs(F).max()


def test_fock_u_is_linear_in_dx(h2o):
    o, D = h2o
    o.set_functional("PBE")
    o.get_ao()
    rng = np.random.default_rng(5)
    A = rng.standard_normal(D.shape); A = A + A.T
    B = rng.standard_normal(D.shape); B = B + B.T
    FA, FB, FC = o.fock_u(D, [A, B, 2.0 * A - 3.0 * B])
    assert np.abs(FC - (2.0 * FA - 3.0 * FB)).max() <= 1e-12 * np.abs(FC).max()


def test_fock_u_rejects_meta_gga(h2o):
    o, D = h2o
    o.set_functional("R2SCAN")
    o.get_ao()
    with pytest.raises(RuntimeError):
        o.fock_u(D, [D])


def _df_codes(name):
    with open("%s/DensityFunctionals/%s.df" % (ol.DATA_ROOT, name)) as f:
        return [int(c) for c in f.readline().split()]


@pytest.mark.parametrize("func", ["PBE", "B3LYP", "SVWN5"])
def test_fock_s_is_skeleton_response_of_v(h2o, func):
    o, D = h2o
    o.set_functional(func)
    o.get_ao()
    codes = _df_codes(func)
    gga = func != "SVWN5"
    Fs = o.fock_s(D)
    xyz, _, _, _, w = o.density_at(D)
    ao0, g0, _ = o.ao_values(xyz)  # [mu, g], [r, mu, g]

    def V_of(ao, g):  # V[rho] from the density of D built on (ao, g), assembled on the fixed (ao0, g0)
        rho = np.einsum("mg,mn,ng->g", ao, D, ao)
        grad = 2 * np.einsum("rmg,mn,ng->rg", g, D, ao)
        sigma = np.sum(grad ** 2, axis=0)
        _, vr, vs, _ = ol.eval_functional(codes, rho, sigma if gga else None)
        V = (ao0 * (w * vr)) @ ao0.T
        if gga:
            Y = np.einsum("rg,rmg->mg", 2 * w * vs * grad, g0)
            V += Y @ ao0.T + ao0 @ Y.T
        return V

    _, _, Vref = o.xc(D)
    assert np.abs(V_of(ao0, g0) - Vref).max() <= 1e-12 * np.abs(Vref).max()
    h = 1e-4
    for A, t in ((0, 0), (0, 2), (1, 1), (2, 0)):
        def V_at(eps):
            d = np.zeros(3); d[t] = eps
            o.shift_basis(A, d)
            ao, g, _ = o.ao_values(xyz)
            o.shift_basis(A, -d)
            return V_of(ao, g)
        d1 = (V_at(h) - V_at(-h)) / (2 * h)
        d2 = (V_at(h / 2) - V_at(-h / 2)) / h
        fd = (4 * d2 - d1) / 3
        F = Fs[3 * A + t]
        assert np.abs(F - F.T).max() <= 1e-13 * np.abs(F).max()
        assert np.abs(F - fd).max() <= 1e-8 * np.abs(F).max(), (A, t, np.abs(F - fd).max(), np.abs(F).max())
    assert all(np.abs(Fs[k]).max() > 0 for k in range(len(Fs)))


def test_fock_s_rejects_meta_gga(h2o):
    o, D = h2o
    o.set_functional("R2SCAN")
    o.get_ao()
    with pytest.raises(RuntimeError):
        o.fock_s(D)


@pytest.mark.parametrize("func", ["PBE", "B3LYP", "SVWN5"])
def test_fock_skeleton_is_explicit_derivative_of_v(h2o, func):
    """Grid::getFockSkeleton (GridFock.cpp:51-97) is the derivative of V along a shift of one atom's
    basis functions with the potential (v_rho, v_sigma, grad rho of D) and the grid held fixed; with
    getFockU<s_t> (the response of the potential to the same shift) it makes the total derivative
    of V[D] at fixed D."""
    o, D = h2o
    o.set_functional(func)
    o.get_ao()
    codes = _df_codes(func)
    gga = func != "SVWN5"
    Fk = o.fock_skeleton(D)
    Fs = o.fock_s(D)
    xyz, _, _, _, w = o.density_at(D)
    ao0, g0, _ = o.ao_values(xyz)

    def dens(ao, g):
        rho = np.einsum("mg,mn,ng->g", ao, D, ao)
        grad = 2 * np.einsum("rmg,mn,ng->rg", g, D, ao)
        return rho, grad

    rho0, grad0 = dens(ao0, g0)
    _, vr0, vs0, _ = ol.eval_functional(codes, rho0, np.sum(grad0 ** 2, axis=0) if gga else None)

    def V_on(ao, g, vr, vs, grad):  # V of the potential (vr, vs, grad) assembled on the AOs (ao, g)
        V = (ao * (w * vr)) @ ao.T
        if gga:
            Y = np.einsum("rg,rmg->mg", 2 * w * vs * grad, g)
            V += Y @ ao.T + ao @ Y.T
        return V

    def V_full(ao, g):
        rho, grad = dens(ao, g)
        _, vr, vs, _ = ol.eval_functional(codes, rho, np.sum(grad ** 2, axis=0) if gga else None)
        return V_on(ao, g, vr, vs, grad)

    h = 1e-4
    for A, t in ((0, 0), (0, 2), (1, 1), (2, 0)):
        def at(eps, fn):
            d = np.zeros(3); d[t] = eps
            o.shift_basis(A, d)
            ao, g, _ = o.ao_values(xyz)
            o.shift_basis(A, -d)
            return fn(ao, g)

        def fd(fn):
            d1 = (at(h, fn) - at(-h, fn)) / (2 * h)
            d2 = (at(h / 2, fn) - at(-h / 2, fn)) / h
            return (4 * d2 - d1) / 3

        F = Fk[3 * A + t]
        ref = fd(lambda ao, g: V_on(ao, g, vr0, vs0, grad0))
        assert np.abs(F - F.T).max() <= 1e-13 * np.abs(F).max()
        assert np.abs(F - ref).max() <= 1e-8 * np.abs(F).max(), (A, t, np.abs(F - ref).max(), np.abs(F).max())
        tot = fd(V_full)
        S = F + Fs[3 * A + t]
        assert np.abs(S - tot).max() <= 1e-8 * np.abs(S).max()
    with pytest.raises(RuntimeError):
        o.set_functional("R2SCAN")
        o.get_ao()
        o.fock_skeleton(D)


@pytest.mark.parametrize("func", ["PBE", "B3LYP", "SVWN5"])
def test_xc_hess_is_derivative_of_gradient(h2o, func):
    """getDensitySkeleton2 + getEnergyHess (GridDensity.cpp:132-209, GridEnergy.cpp:32-45): the
    skeleton XC Hessian is the derivative of the skeleton XC gradient (itself pinned as the derivative
    of E_xc) along a shift of one atom's basis functions, D and the grid fixed; it is symmetric."""
    o, D = h2o
    o.set_functional(func)
    o.get_ao()
    H = o.xc_hess(D)
    na = len(o.Z)
    assert H.shape == (3 * na, 3 * na)
    assert np.abs(H - H.T).max() <= 1e-10 * np.abs(H).max()
    h = 1e-4

    def grad_at(B, s, eps):
        d = np.zeros(3); d[s] = eps
        o.shift_basis(B, d)
        o.get_ao()
        G = o.xc_grad(D).ravel()
        o.shift_basis(B, -d)
        o.get_ao()
        return G

    for B, s in ((0, 0), (1, 2), (2, 1)):
        d1 = (grad_at(B, s, h) - grad_at(B, s, -h)) / (2 * h)
        d2 = (grad_at(B, s, h / 2) - grad_at(B, s, -h / 2)) / h
        fd = (4 * d2 - d1) / 3
        col = H[:, 3 * B + s]
        assert np.abs(col - fd).max() <= 1e-7 * np.abs(H).max(), (B, s, np.abs(col - fd).max(), np.abs(H).max())
