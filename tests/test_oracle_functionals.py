"""Oracle functionals (libxc replacement, src/ExchangeCorrelation.cpp:86-217)
against golden vectors from an independent sympy transcription of the
published forms (tests/golden/gen_functional_vectors.py), plus limits.

libxc itself is absent, so agreement with libxc's exact outputs (thresholds,
the parameter values it ships for IDs 321/322) is not pinned here."""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "functionals.json")))
PTS = np.array(GOLD["points"])


@pytest.mark.parametrize("code", sorted(GOLD["values"], key=int))
def test_against_sympy_vectors(code):
    ref = np.array(GOLD["values"][code])
    e, vr, vs, _ = ol.eval_functional([int(code)], PTS[:, 0], PTS[:, 1])
    rel = lambda a, b: np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    assert rel(e, ref[:, 0]).max() < 1e-12
    assert rel(vr, ref[:, 1]).max() < 1e-12
    has_grad = PTS[:, 1] > 0
    assert rel(vs[has_grad], ref[has_grad, 2]).max() < 1e-12


def test_gga_exchange_reduces_to_slater_at_zero_gradient():
    rho = np.logspace(-6, 3, 40)
    ex, vx, _, _ = ol.eval_functional([1], rho, np.zeros_like(rho))
    for code in (101, 321, 106):
        e, v, _, _ = ol.eval_functional([code], rho, np.zeros_like(rho))
        assert np.abs(e / ex - 1).max() < 1e-12
        assert np.abs(v / vx - 1).max() < 1e-12


def test_slater_closed_form():
    rho = np.array([1e-3, 0.1, 1.0, 10.0])
    e, v, _, _ = ol.eval_functional([1], rho)
    cx = -0.75 * (3 / np.pi) ** (1 / 3)
    assert np.allclose(e, cx * rho ** (1 / 3), rtol=1e-14)
    assert np.allclose(v, 4 / 3 * cx * rho ** (1 / 3), rtol=1e-14)


def test_below_density_threshold_contributes_nothing():
    rho = np.array([0.0, 1e-20, 9e-16])
    for code in (1, 7, 8, 101, 130, 106, 131, 321, 322, 402):
        e, vr, vs, _ = ol.eval_functional([code], rho, np.full(3, 1e-30))
        assert not np.any(e) and not np.any(vr) and not np.any(vs)


def test_codes_sum_like_the_reference():
    """Evaluate() sums eps and potentials over the codes of a .df line (:202-217)."""
    rho = np.array([0.01, 0.2, 3.0])
    sig = np.array([1e-4, 0.05, 2.0])
    a = ol.eval_functional([106, 131], rho, sig)
    b1 = ol.eval_functional([106], rho, sig)
    b2 = ol.eval_functional([131], rho, sig)
    for x, y, z in zip(a[:3], b1[:3], b2[:3]):
        assert np.allclose(x, y + z, rtol=1e-14, atol=0)


def test_derivatives_by_finite_difference():
    """Dual-number derivatives against central differences of rho*eps."""
    rho = np.array([0.003, 0.08, 0.7, 12.0])
    sig = np.array([1e-5, 0.01, 0.4, 30.0])
    for code in (101, 130, 106, 131, 402):
        e, vr, vs, _ = ol.eval_functional([code], rho, sig)
        h = 1e-6
        ep, _, _, _ = ol.eval_functional([code], rho * (1 + h), sig)
        em, _, _, _ = ol.eval_functional([code], rho * (1 - h), sig)
        fd_r = (ep * rho * (1 + h) - em * rho * (1 - h)) / (2 * h * rho)
        assert np.allclose(vr, fd_r, rtol=1e-7)
        ep, _, _, _ = ol.eval_functional([code], rho, sig * (1 + h))
        em, _, _, _ = ol.eval_functional([code], rho, sig * (1 - h))
        fd_s = (ep - em) * rho / (2 * h * sig)
        assert np.allclose(vs, fd_s, rtol=1e-6)


R2 = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "r2scan.json")))


@pytest.mark.parametrize("code", ["497", "498"])
def test_r2scan_against_mpmath_vectors(code):
    P = np.array(R2["points"])
    ref = np.array(R2["values"][code])
    e, vr, vs, vt = ol.eval_functional([int(code)], P[:, 0], P[:, 1], P[:, 2])
    rel = lambda a, b: np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    assert rel(e, ref[:, 0]).max() < 1e-12
    assert rel(vr, ref[:, 1]).max() < 1e-11
    assert rel(vs, ref[:, 2]).max() < 1e-10
    assert rel(vt, ref[:, 3]).max() < 1e-9


def test_r2scan_exact_constraints():
    """Uniform gas (p = 0, alpha = 1): F_x = 1 and eps_c = PW92; one-orbital
    limit (alpha = 0, p = 0): F_x = h_x^0 = 1.174."""
    k2 = (3 * np.pi ** 2) ** (2 / 3)
    rho = np.array([1e-3, 0.05, 0.7, 9.0])
    tu = 0.3 * k2 * rho ** (5 / 3)
    zero = np.full(4, 1e-40)
    ex, _, _, _ = ol.eval_functional([497], rho, zero, tu)
    el, _, _, _ = ol.eval_functional([1], rho)
    assert np.allclose(ex, el, rtol=1e-13)
    ec, _, _, _ = ol.eval_functional([498], rho, zero, tu)
    pbe_c0, _, _, _ = ol.eval_functional([130], rho, zero)   # PBE correlation at zero gradient is PW92
    assert np.allclose(ec, pbe_c0, rtol=1e-12)
    ex0, _, _, _ = ol.eval_functional([497], rho, zero, np.full(4, 1e-30) + 0.0 * rho)
    # tau floored at 1e-20 keeps alpha at ~0 for these densities
    assert np.allclose(ex0 / el, 1.174, rtol=1e-6)


def test_nonphysical_inputs_contribute_nothing():
    """Negative or NaN densities contribute nothing; a NaN or negative sigma / tau is floored
    (sigma at the threshold floor, tau at 1e-20) rather than propagated."""
    nan = float("nan")
    rho = np.array([-1e-3, nan, 0.0, 1e-20, 0.3, 0.3])
    sig = np.array([0.1, 0.1, 0.1, 0.1, nan, -0.5])
    tau = np.array([0.1, 0.1, 0.1, 0.1, nan, -0.5])
    for codes in ([1, 7], [101, 130], [402], [497, 498]):
        e, vr, vs, vt = ol.eval_functional(codes, rho, sig, tau)
        for a in (e, vr, vs, vt):
            assert np.all(np.isfinite(a)), codes
        assert not np.any(e[:4]) and not np.any(vr[:4])
        e0, vr0, _, _ = ol.eval_functional(codes, rho[4:5], np.zeros(1), np.full(1, 1e-20))
        assert e[4] == e0[0] and e[5] == e0[0]  # floored sigma / tau = the zero-gradient values
    e, vr, vs = ol.eval_functional_pol([101, 130], np.array([[nan, 0.1], [-1.0, 0.2], [0.0, 0.0]]),
                                       np.array([[0.1, 0.0, 0.1]] * 3))
    assert np.all(np.isfinite(e)) and np.all(np.isfinite(vr)) and np.all(np.isfinite(vs))
    assert e[2] == 0 and not np.any(vr[2])
