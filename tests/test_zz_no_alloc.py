"""chx_xc_eval_points_noalloc (device staging in a fixed array, no allocation)
against chx_xc_eval_points (device buffers per call) through the C ABI: the
same kernels on the same inputs, so results must agree bit for bit. Two calls
in a row on each functional, the second with fewer points and different
inputs, so that a region the first call left in the shared array would show
(the second GGA call drops sigma, which must read as zeros, not the first
call's sigma). Sizes beyond the fixed maximum are rejected with an error."""
import ctypes

import numpy as np
import pytest

import chinium_amd as ca

pytestmark = pytest.mark.gpu

MAXN = 65536  # CHX_NOALLOC_MAX_POINTS


def _inputs(n, seed, pol=False):
    rng = np.random.default_rng(seed)
    if pol:
        rho = 10.0 ** rng.uniform(-6, 2, size=(n, 2))
        g = rng.normal(size=(n, 2, 3)) * rho[:, :, None] ** (4.0 / 3.0)
        sig = np.stack([np.einsum("ij,ij->i", g[:, 0], g[:, 0]), np.einsum("ij,ij->i", g[:, 0], g[:, 1]),
                        np.einsum("ij,ij->i", g[:, 1], g[:, 1])], axis=1)
        return rho, sig, None
    rho = 10.0 ** rng.uniform(-6, 2, size=n)
    sig = rng.uniform(0, 4, size=n) * rho ** (8.0 / 3.0)
    tau = sig / (8 * rho) * (1 + rng.uniform(0, 3, size=n))
    return rho, sig, tau


@pytest.mark.parametrize("func,spin,use_sigma,use_tau", [
    ("PBE", 1, (True, False), (False, False)),
    ("R2SCAN", 1, (True, True), (True, True)),
    ("B3LYP", 2, (True, True), (False, False)),
])
def test_noalloc_matches_eval_points(func, spin, use_sigma, use_tau):
    xc = ca.ExchangeCorrelation()
    xc.Spin = spin
    xc.Read(func)
    for call, (n, seed) in enumerate(((MAXN, 1), (1000, 2))):
        rho, sig, tau = _inputs(n, seed, pol=spin == 2)
        sig = sig if use_sigma[call] else None
        tau = tau if use_tau[call] else None
        a = xc.eval_points(rho, sig, tau)
        b = xc.eval_points(rho, sig, tau, noalloc=True)
        print("  %s spin %d call %d: %d points, sigma %s, tau %s" % (func, spin, call, n, sig is not None,
                                                                     tau is not None), flush=True)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert np.all(np.isfinite(a[0]))


def test_noalloc_rejects_oversize():
    xc = ca.ExchangeCorrelation()
    xc.Read("PBE")
    n = MAXN + 1
    rho = np.ones(n)
    out = [np.zeros(n) for _ in range(4)]
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = ca.lib().chx_xc_eval_points_noalloc(xc.h, n, p(rho), None, None, *[p(o) for o in out])
    assert rc == -1
    assert "at most 65536" in ca.lib().chx_last_error().decode()
