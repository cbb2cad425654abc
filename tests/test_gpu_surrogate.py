"""Learned functional surrogate on the GPU (chx_mlp_*): forward, input
gradient and the loss / parameter gradient of the training kernels against the
FP64 oracle (oracle/mlp.cpp, pinned by finite differences and known answers in
test_oracle_mlp.py), with and without the analytic base the network corrects;
the out-of-distribution fallback returns the analytic functional bit for bit;
training on the workload grids lowers the loss; on held-out cyclo[18]carbon the
trained correction gains >= 100x in E_xc error over its base alone (what the
network adds to the method without it); Evaluate with the surrogate attached
integrates the same per-point values the oracle gives.

Tolerances: FP32 network arithmetic against FP64 -> 2e-5 relative on F,
1e-4 of the largest component on gradients."""
import numpy as np
import pytest

import chinium_amd as ca
import oracle_lib as ol
from chinium_amd import workloads as wl
from chinium_amd.surrogate import LDA_BASE, Surrogate, workload_samples

pytestmark = pytest.mark.gpu
R2SCAN = [497, 498]


def _say(msg):
    print(msg, flush=True)


def _rand_params(seed, scale=0.3):
    return np.random.default_rng(seed).uniform(-scale, scale, ol.MLP_PARAMS).astype(np.float32)


def _r2scan():
    xc = ca.ExchangeCorrelation()
    xc.Read("R2SCAN")
    return xc


@pytest.mark.parametrize("mode", ["fp32", "bf16", "mfma"])
def test_forward_against_oracle(mode):
    """fp32 / bf16: the per-lane FP32 network on the FP32 / BF16-rounded weights; mfma: the
    inference path (FP32 weights and activations as split-BF16 MFMA operands, transposed products,
    1001 points = 15 full waves and a ragged one) against the FP32 weights. Bars: FP32 arithmetic
    2e-5 / 1e-4; split-BF16 operands (16 significant bits) 1e-4 / 1e-3."""
    s = Surrogate(3)
    s.params = _rand_params(1)
    U = np.random.default_rng(2).uniform(-1.2, 1.2, (1001, 3)).astype(np.float32)
    F, dF = s.forward(U, bf16=mode != "fp32", mfma=mode == "mfma")
    P = s.bf16_params() if mode == "bf16" else s.inference_params()
    Fo, dFo = ol.mlp_forward(P.astype(np.float64), U.astype(np.float64))
    ef = np.abs(F - Fo).max() / (1 + np.abs(Fo).max())
    eg = np.abs(dF - dFo).max() / np.abs(dFo).max()
    _say("  forward (%s): max |dF| %.1e, max |d grad| %.1e" % (mode, ef, eg))
    if mode == "mfma":
        assert ef <= 1e-4 and eg <= 1e-3
    else:
        assert ef <= 2e-5 and eg <= 1e-4


@pytest.mark.parametrize("n,weighted", [(1, False), (64, True), (1000, True), (70001, False)])
def test_loss_gradient_against_oracle(n, weighted):
    s = Surrogate(4)
    s.params = _rand_params(5)
    rng = np.random.default_rng(n)
    U = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    T = rng.uniform(0.8, 1.7, n).astype(np.float32)
    C = rng.uniform(0.1, 3.0, n).astype(np.float32) if weighted else None
    L, G = s.loss_grad(U, T, C)
    Lo, Go = ol.mlp_loss_grad(s.params.astype(np.float64), U.astype(np.float64), T.astype(np.float64),
                              None if C is None else C.astype(np.float64))
    eg = np.abs(G - Go).max() / np.abs(Go).max()
    _say("  loss/grad n=%d: |dL|/L %.1e, max |dG| / max|G| %.1e" % (n, abs(L - Lo) / Lo, eg))
    assert abs(L - Lo) <= 1e-5 * Lo
    assert eg <= 1e-4


@pytest.mark.parametrize("base", [(), (1, 7)])
def test_eval_points_and_fallback(base):
    s = Surrogate(6, base=base)
    s.params = _rand_params(7, 0.2)
    lo, hi = np.array([-12.0, 0.0, 0.02], np.float32), np.array([4.0, 0.8, 0.9], np.float32)
    s.set_box(lo, hi, 1e-8)
    rng = np.random.default_rng(8)
    n = 4000
    rho = 10 ** rng.uniform(-9.5, 2, n)
    sigma = 10 ** rng.uniform(-3, 1, n) * 38.28 * rho ** (8 / 3)
    tau = sigma / (8 * rho) + 10 ** rng.uniform(-1.5, 1.5, n) * 2.87 * rho ** (5 / 3)
    xc = _r2scan()
    e, vr, vs, vt, nf = s.eval_points(xc, rho, sigma, tau)
    eo, vro, vso, vto, used = ol.mlp_point(s.inference_params().astype(np.float64), lo.astype(np.float64),
                                           hi.astype(np.float64), 1e-8, R2SCAN, rho, sigma, tau, base=base)
    _say("  %d points, %d on the network, %d fallback" % (n, used.sum(), nf))
    assert nf == (~used).sum() and 0 < nf < n
    # fallback lanes: the device's analytic functional, bit for bit
    ea, vra, vsa, vta = xc.eval_points(rho, sigma, tau)
    for a, b in ((e, ea), (vr, vra), (vs, vsa), (vt, vta)):
        assert np.array_equal(a[~used], b[~used])
    # network lanes against the FP64 oracle on the same FP32 weights; FP32 network errors are absolute
    # in F, so the scale is the uniform-gas exchange energy density |e_x^unif| (F = 1)
    eu = 0.7385587663820224 * rho ** (4.0 / 3.0)
    for a, b, sc in ((rho * e, rho * eo, eu), (vr, vro, eu / rho), (vs, vso, eu / sigma), (vt, vto, eu / tau)):
        err = np.abs(a - b)[used] / (np.abs(b) + sc)[used]
        assert err.max() <= 1e-4, err.max()


@pytest.fixture(scope="module")
def trained():
    _say("collecting r2SCAN training samples on the workload grids")
    rho, sigma, tau, w = workload_samples(n_alkanes=8)
    _say("  %d samples" % rho.size)
    xc = _r2scan()
    s = Surrogate(11, base=LDA_BASE)
    loss = s.fit(xc, rho, sigma, tau, w, epochs=150, batch=16384, lr=3e-3, lr_final=3e-5)
    _say("  loss: first epoch %.3e, last %.3e" % (loss[0], loss[-1]))
    return s, xc, loss


def _c18():
    w = wl.build_workload("c18_def2tzvp_sg3_r2scan")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D, _ = wl.workload_density(mol, w["nocc"], w["seed"])
    g = ca.Grid(mol, w["grid"], 1, 0)
    g.setAOCache(1)
    g.setType(2)
    g.getAO(0, 0)
    g.getDensity([D])
    return g, len(w["Z"])


def _functional_ms(g, xc, reps=5):
    g.profile(True)
    g.kernel_times(reset=True)
    for _ in range(reps):
        xc.Evaluate("ev", g)
    ms, n = g.kernel_times(reset=True)["functional"]
    g.profile(False)
    return ms / max(1, n)


def test_training_lowers_loss(trained):
    s, xc, loss = trained
    assert np.all(np.isfinite(loss))
    assert loss[-1] < 1e-3 * loss[0]


def test_gain_over_the_method_without_the_network(trained):
    """Held-out cyclo[18]carbon / def2-TZVP / SG-3 (configuration 3; no C18 sample is in the training
    set). The method without the network is its analytic base, Slater + VWN5 (LDA): the trained
    correction must cut the base's E_xc error against r2SCAN by >= 100x and stay within 2e-3 Eh/atom
    (measured 2e-4 to 4e-4 with 2 M samples; runs differ, the gradient sums use float atomics). Also
    recorded, not asserted: the functional stage of each, against the analytic r2SCAN it stands in for
    (the network is slower and less accurate than that)."""
    s, xc, _ = trained
    g, natoms = _c18()
    xc.Evaluate("ev", g)
    Ea = g.getEnergy()
    ta = _functional_ms(g, xc)
    lda = ca.ExchangeCorrelation().from_codes(list(LDA_BASE))
    lda.Evaluate("ev", g)
    Eb = g.getEnergy()
    tb = _functional_ms(g, lda)
    xc.setSurrogate(s)
    try:
        xc.Evaluate("ev", g)
        Es = g.getEnergy()
        nf = g.fallbackCount()
        ts = _functional_ms(g, xc)
    finally:
        xc.setSurrogate(None)
    eb, es = abs(Eb - Ea) / natoms, abs(Es - Ea) / natoms
    _say("  C18 |dE_xc| per atom vs analytic r2SCAN: LDA base alone %.3e, base + network %.3e Eh (gain %.0fx); "
         "%d fallback points; functional stage: r2SCAN %.4f ms, LDA %.4f ms, base + network %.4f ms"
         % (eb, es, eb / es, nf, ta, tb, ts))
    assert es <= 2e-3
    assert eb >= 100 * es


def test_grid_evaluate_matches_points(trained):
    s, xc, _ = trained
    Z, X = wl.water_cluster_angstrom(4, 5.0, 11)
    mol = ca.Molecule(Z, X * ca.ANGSTROM_TO_BOHR, "def2-svp")
    D, _ = wl.workload_density(mol, 20, 3)
    g = ca.Grid(mol, "SG-1", 1, 0)
    g.setAOCache(1)
    g.setType(2)
    g.getAO(0, 0)
    g.getDensity([D])
    xc.setSurrogate(s)
    try:
        xc.Evaluate("ev", g)
        E = g.getEnergy()
        V = g.getFock()[0]
        nf = g.fallbackCount()
    finally:
        xc.setSurrogate(None)
    rho, sigma, tau, w = g.pointDensity()
    lo, hi, rmin = s.box
    eo, _, _, _, used = ol.mlp_point(s.inference_params().astype(np.float64), lo.astype(np.float64),
                                     hi.astype(np.float64), rmin, R2SCAN, rho, sigma, tau, base=s.base)
    live = (w != 0) & (rho >= 1e-15)
    Eo = float(np.sum(w * rho * eo))
    _say("  grid E_xc %.10f, from oracle points %.10f; fallback %d (oracle %d)"
         % (E, Eo, nf, int((~used & live).sum())))
    assert abs(E - Eo) <= 5e-6 * abs(Eo)
    assert nf == int((~used & live).sum())
    assert np.abs(V - V.T).max() <= 1e-12 * np.abs(V).max()
    # detaching restores the analytic build exactly
    xc.Evaluate("ev", g)
    E2 = g.getEnergy()
    xc2 = _r2scan()
    xc2.Evaluate("ev", g)
    assert E2 == g.getEnergy()


def test_surrogate_protocol_errors():
    s = Surrogate(1)
    pbe = ca.ExchangeCorrelation()
    pbe.Read("PBE")
    with pytest.raises(ca.ChxError) as e:  # GGA functional
        pbe.setSurrogate(s)
    assert e.value.code == -5
    xc = _r2scan()
    with pytest.raises(ca.ChxError) as e:  # no descriptor box yet
        xc.setSurrogate(s)
    assert e.value.code == -3
    with pytest.raises(ca.ChxError) as e:
        s.set_box([0, 0, 0], [0, 1, 1])
    assert e.value.code == -1
