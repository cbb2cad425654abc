"""Oracle restatement of the learned functional surrogate (oracle/mlp.cpp),
pinned by what the mathematics fixes:
  * the parameter gradient of the loss and the input gradient of the network
    are derivatives (central differences of the FP64 network);
  * the descriptor derivatives are derivatives (central differences);
  * known answers: uniform gas (sigma = 0, tau = tau_unif) gives s = 0,
    alpha = 1; a one-orbital density (tau = sigma / 8 rho) gives alpha = 0;
  * a network that outputs F = 1 everywhere is Slater exchange, whose eps and
    v_rho the analytic LDA exchange (code 1) gives independently;
  * v_rho, v_sigma, v_tau of a surrogate point are the derivatives of
    rho * eps (central differences), and points outside the box or below
    rho_min return the analytic functional exactly."""
import numpy as np
import pytest

import oracle_lib as ol

R2SCAN = [497, 498]


def _params(seed, scale=0.4):
    rng = np.random.default_rng(seed)
    return rng.uniform(-scale, scale, ol.MLP_PARAMS)


def _fd(fn, x, h):
    return (fn(x + h) - fn(x - h)) / (2 * h)


def test_loss_gradient_is_derivative():
    P = _params(1)
    rng = np.random.default_rng(2)
    U = rng.uniform(-1, 1, (37, 3))
    T = rng.uniform(0.8, 1.6, 37)
    C = rng.uniform(0.1, 2.0, 37)
    L, G = ol.mlp_loss_grad(P, U, T, C)
    assert L > 0
    idx = np.concatenate([np.arange(0, 192, 17), np.arange(192, 256, 9), np.arange(256, 4352, 331),
                          np.arange(4352, 4416, 13), np.arange(4416, 8512, 407), np.arange(8512, 8641, 11), [8640]])
    for k in idx:
        def f(v, k=k):
            Q = P.copy()
            Q[k] = v
            return ol.mlp_loss_grad(Q, U, T, C)[0]
        fd = _fd(f, P[k], 1e-6)
        assert abs(G[k] - fd) <= 1e-7 * max(1.0, abs(fd)), (k, G[k], fd)


def test_input_gradient_is_derivative():
    P = _params(3)
    U = np.random.default_rng(4).uniform(-1, 1, (25, 3))
    F, dU = ol.mlp_forward(P, U)
    for k in range(3):
        e = np.zeros(3)
        e[k] = 1e-6
        fd = (ol.mlp_forward(P, U + e)[0] - ol.mlp_forward(P, U - e)[0]) / 2e-6
        np.testing.assert_allclose(dU[:, k], fd, rtol=1e-7, atol=1e-9)


def test_descriptor_known_answers():
    rho = np.array([0.1, 1.0, 7.0])
    tau_unif = 2.8712340001881915 * rho ** (5.0 / 3.0)
    t, _ = ol.mlp_descriptors(rho, np.zeros(3), tau_unif)
    np.testing.assert_allclose(t[:, 0], np.log(rho), rtol=1e-15)
    assert np.all(np.abs(t[:, 1]) < 1e-12)  # s = 0 (sigma floored at 1e-40)
    np.testing.assert_allclose(t[:, 2], 0.5, rtol=1e-13)  # alpha = 1
    # one-orbital density: tau = sigma / (8 rho) -> alpha = 0; s from its definition
    sigma = np.array([0.01, 0.3, 2.0])
    t, _ = ol.mlp_descriptors(rho, sigma, sigma / (8 * rho))
    assert np.all(np.abs(t[:, 2]) < 1e-12)
    s = np.sqrt(sigma) / (2 * (3 * np.pi ** 2) ** (1 / 3) * rho ** (4 / 3))
    np.testing.assert_allclose(t[:, 1], s / (1 + s), rtol=1e-13)


def test_descriptor_derivatives():
    rng = np.random.default_rng(5)
    rho = 10 ** rng.uniform(-3, 1, 30)
    sigma = 10 ** rng.uniform(-2, 0.5, 30) * 38.28 * rho ** (8 / 3)
    tau = sigma / (8 * rho) + 10 ** rng.uniform(-1, 1, 30) * 2.87 * rho ** (5 / 3)
    t, dt = ol.mlp_descriptors(rho, sigma, tau)
    for c, x in enumerate((rho, sigma, tau)):
        h = 1e-6 * x
        args = [rho, sigma, tau]

        def f(v):
            a = list(args)
            a[c] = v
            return ol.mlp_descriptors(*a)[0]
        fd = (f(x + h) - f(x - h)) / (2 * h)[:, None]
        scale = np.abs(t) / x[:, None] + np.abs(dt[:, :, c])
        assert np.all(np.abs(dt[:, :, c] - fd) <= 1e-6 * scale + 1e-300), c


def test_constant_network_is_slater_exchange():
    P = np.zeros(ol.MLP_PARAMS)
    P[8640] = 1.0  # F = 1
    rng = np.random.default_rng(6)
    rho = 10 ** rng.uniform(-4, 2, 50)
    sigma = 10 ** rng.uniform(-2, 0, 50) * rho ** (8 / 3)
    tau = sigma / (8 * rho) + rho ** (5 / 3)
    lo, hi = np.array([-30.0, 0.0, 0.0]), np.array([10.0, 1.0, 1.0])
    e, vr, vs, vt, used = ol.mlp_point(P, lo, hi, 1e-8, R2SCAN, rho, sigma, tau)
    assert used.all()
    ex, vx, _, _ = ol.eval_functional([1], rho)
    np.testing.assert_allclose(e, ex, rtol=1e-14)
    np.testing.assert_allclose(vr, vx, rtol=1e-14)
    assert np.all(vs == 0) and np.all(vt == 0)


def test_point_potentials_are_derivatives():
    P = _params(7, 0.3)
    rng = np.random.default_rng(8)
    n = 30
    rho = 10 ** rng.uniform(-3, 1, n)
    sigma = 10 ** rng.uniform(-2, 0.5, n) * 38.28 * rho ** (8 / 3)
    tau = sigma / (8 * rho) + 10 ** rng.uniform(-1, 1, n) * 2.87 * rho ** (5 / 3)
    lo, hi = np.array([-20.0, 0.0, 0.0]), np.array([5.0, 1.0, 1.0])
    e, vr, vs, vt, used = ol.mlp_point(P, lo, hi, 1e-8, R2SCAN, rho, sigma, tau)
    assert used.all()
    for v, c, x in ((vr, 0, rho), (vs, 1, sigma), (vt, 2, tau)):
        h = 1e-6 * x
        args = [rho, sigma, tau]

        def E(val):
            a = list(args)
            a[c] = val
            return a[0] * ol.mlp_point(P, lo, hi, 1e-8, R2SCAN, *a)[0]
        fd = (E(x + h) - E(x - h)) / (2 * h)
        scale = np.abs(rho * e) / x
        assert np.all(np.abs(v - fd) <= 1e-6 * (np.abs(fd) + scale)), c


def test_fallback_is_the_analytic_functional():
    P = _params(9)
    rho = np.array([1e-9, 1e-3, 0.5, 3.0])
    sigma = np.array([1e-20, 1e-4, 50.0, 0.2])
    tau = sigma / (8 * rho) + 0.1
    lo, hi = np.array([-8.0, 0.0, 0.0]), np.array([5.0, 0.6, 0.99])
    e, vr, vs, vt, used = ol.mlp_point(P, lo, hi, 1e-8, R2SCAN, rho, sigma, tau)
    ea, vra, vsa, vta = ol.eval_functional(R2SCAN, rho, sigma, tau=tau)
    t, _ = ol.mlp_descriptors(rho, sigma, tau)
    inside = (rho >= 1e-8) & np.all((t >= lo) & (t <= hi), axis=1)
    assert np.array_equal(used, inside) and not used.all()
    for a, b in ((e, ea), (vr, vra), (vs, vsa), (vt, vta)):
        assert np.array_equal(a[~used], b[~used])


def test_zero_network_on_a_base_is_the_base():
    """With an analytic base (Slater + VWN5) the network adds to the base: a network that outputs 0
    returns the base functional's values and derivatives; F = 1 adds Slater exchange once more."""
    rng = np.random.default_rng(12)
    rho = 10 ** rng.uniform(-4, 2, 50)
    sigma = 10 ** rng.uniform(-2, 0, 50) * rho ** (8 / 3)
    tau = sigma / (8 * rho) + rho ** (5 / 3)
    lo, hi = np.array([-30.0, 0.0, 0.0]), np.array([10.0, 1.0, 1.0])
    P = np.zeros(ol.MLP_PARAMS)
    e, vr, vs, vt, used = ol.mlp_point(P, lo, hi, 1e-8, R2SCAN, rho, sigma, tau, base=[1, 7])
    assert used.all()
    eb, vb, _, _ = ol.eval_functional([1, 7], rho)
    np.testing.assert_allclose(e, eb, rtol=1e-14)
    np.testing.assert_allclose(vr, vb, rtol=1e-14)
    assert np.all(vs == 0) and np.all(vt == 0)
    P[8640] = 1.0
    e1, vr1, _, _, _ = ol.mlp_point(P, lo, hi, 1e-8, R2SCAN, rho, sigma, tau, base=[1, 7])
    ex, vx, _, _ = ol.eval_functional([1], rho)
    np.testing.assert_allclose(e1, eb + ex, rtol=1e-13)
    np.testing.assert_allclose(vr1, vb + vx, rtol=1e-13)


def test_point_potentials_are_derivatives_with_base():
    P = _params(13, 0.3)
    rng = np.random.default_rng(14)
    n = 30
    rho = 10 ** rng.uniform(-3, 1, n)
    sigma = 10 ** rng.uniform(-2, 0.5, n) * 38.28 * rho ** (8 / 3)
    tau = sigma / (8 * rho) + 10 ** rng.uniform(-1, 1, n) * 2.87 * rho ** (5 / 3)
    lo, hi = np.array([-20.0, 0.0, 0.0]), np.array([5.0, 1.0, 1.0])
    base = [101, 130]  # a GGA base, so that v_sigma has a base part too
    e, vr, vs, vt, used = ol.mlp_point(P, lo, hi, 1e-8, R2SCAN, rho, sigma, tau, base=base)
    assert used.all()
    for v, c, x in ((vr, 0, rho), (vs, 1, sigma), (vt, 2, tau)):
        h = 1e-6 * x
        args = [rho, sigma, tau]

        def E(val):
            a = list(args)
            a[c] = val
            return a[0] * ol.mlp_point(P, lo, hi, 1e-8, R2SCAN, *a, base=base)[0]
        fd = (E(x + h) - E(x - h)) / (2 * h)
        scale = np.abs(rho * e) / x
        assert np.all(np.abs(v - fd) <= 1e-6 * (np.abs(fd) + scale)), c
