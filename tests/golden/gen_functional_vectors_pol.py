"""Golden vectors for the spin-polarised LDA/GGA functionals, from a sympy
transcription of the published spin forms (not of the oracle's or the
library's code), evaluated at 40 digits with mpmath.

  exchange (Slater, PBE, B88)  spin scaling E_x[ra, rb] = (E_x[2 ra] + E_x[2 rb]) / 2
                               (Oliver, Perdew, PRA 20, 397 (1979)) on the unpolarised forms
  VWN5                         Vosko, Wilk, Nusair, Can. J. Phys. 58, 1200 (1980): paramagnetic,
                               ferromagnetic and spin-stiffness fits, interpolation
                               eP + a f(z)(1 - z^4)/f''(0) + (eF - eP) f(z) z^4
  VWN-RPA                      RPA fits, eP + (eF - eP) f(z)
  PW92 (PW_MOD parameters)     Perdew, Wang, PRB 45, 13244 (1992), eqs. 8 and 10
  PBE correlation              Perdew, Burke, Ernzerhof, PRL 77, 3865 (1996), eqs. 3, 7-8, phi(z)
  LYP                          Miehlich, Savin, Stoll, Preuss, CPL 157, 200 (1989), eq. 2
  B3LYP (libxc 402, XC part)   0.08 Slater + 0.72 B88 + 0.19 VWN-RPA + 0.81 LYP

Variables (libxc XC_POLARIZED order): rho_a, rho_b, sigma_aa, sigma_ab, sigma_bb.
Values per point: eps (per particle of rho_a + rho_b) and the five derivatives
of e = rho eps. Run: python tests/golden/gen_functional_vectors_pol.py
(writes functionals_pol.json).
"""
import json
import os
import sys

import mpmath as mp
import sympy as sp

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from gen_functional_vectors import VWN5, VWNRPA, b88, lyp_general, pbe_x, rs_of, slater, vwn_eps  # noqa: E402

ra, rb, saa, sab, sbb = sp.symbols("ra rb saa sab sbb", positive=True)
R = sp.Rational
pi = sp.pi
F = lambda v: sp.Float(v, 40)  # noqa: E731


def fz(z):
    return ((1 + z) ** R(4, 3) + (1 - z) ** R(4, 3) - 2) / (2 ** R(4, 3) - 2)


FPP = 4 / (9 * (2 ** R(1, 3) - 1))


def vwn5_pol(rs, z):
    eP = vwn_eps(rs, *VWN5)
    eF = vwn_eps(rs, F("0.01554535"), F("7.06042"), F("18.0578"), F("-0.32500"))
    ea = vwn_eps(rs, -1 / (6 * pi ** 2), F("1.13107"), F("13.0045"), F("-0.0047584"))
    return eP + ea * fz(z) * (1 - z ** 4) / FPP + (eF - eP) * fz(z) * z ** 4


def vwnrpa_pol(rs, z):
    eP = vwn_eps(rs, *VWNRPA)
    eF = vwn_eps(rs, F("0.01554535"), F("20.1231"), F("101.578"), F("-0.743294"))
    return eP + (eF - eP) * fz(z)


def G(rs, A, a1, b1, b2, b3, b4):
    return -2 * A * (1 + a1 * rs) * sp.log(1 + 1 / (2 * A * (b1 * sp.sqrt(rs) + b2 * rs + b3 * rs ** R(3, 2)
                                                             + b4 * rs ** 2)))


def pw92_pol(rs, z):
    e0 = G(rs, *[F(v) for v in ("0.0310907", "0.21370", "7.5957", "3.5876", "1.6382", "0.49294")])
    e1 = G(rs, *[F(v) for v in ("0.01554535", "0.20548", "14.1189", "6.1977", "3.3662", "0.62517")])
    ac = -G(rs, *[F(v) for v in ("0.0168869", "0.11125", "10.357", "3.6231", "0.88026", "0.49671")])
    return e0 + ac * fz(z) * (1 - z ** 4) / FPP + (e1 - e0) * fz(z) * z ** 4


def pbe_c_pol(r, z, gg, beta):
    gamma = (1 - sp.log(2)) / pi ** 2
    ec = pw92_pol(rs_of(r), z)
    phi = ((1 + z) ** R(2, 3) + (1 - z) ** R(2, 3)) / 2
    kf = (3 * pi ** 2 * r) ** R(1, 3)
    ks = sp.sqrt(4 * kf / pi)
    t2 = gg / (2 * phi * ks * r) ** 2
    A = beta / gamma / (sp.exp(-ec / (gamma * phi ** 3)) - 1)
    H = gamma * phi ** 3 * sp.log(1 + beta / gamma * t2 * (1 + A * t2) / (1 + A * t2 + A ** 2 * t2 ** 2))
    return r * (ec + H)


def xspin(fun):
    return fun(2 * ra, 4 * saa) / 2 + fun(2 * rb, 4 * sbb) / 2


r = ra + rb
z = (ra - rb) / r
gg = saa + 2 * sab + sbb
lyp = lyp_general(ra, rb, saa, sbb, sab)
FUNCS = {
    "1": xspin(lambda x, s: slater(x)),
    "7": r * vwn5_pol(rs_of(r), z),
    "8": r * vwnrpa_pol(rs_of(r), z),
    "101": xspin(lambda x, s: pbe_x(x, s, F("0.804"), F("0.2195149727645171"))),
    "321": xspin(lambda x, s: pbe_x(x, s, F("0.804"), F("0.219510"))),
    "130": pbe_c_pol(r, z, gg, F("0.06672455060314922")),
    "322": pbe_c_pol(r, z, gg, F("0.06672632268006")),
    "106": xspin(b88),
    "131": lyp,
    "402": (R(8, 100) * xspin(lambda x, s: slater(x)) + R(72, 100) * xspin(b88)
            + R(19, 100) * r * vwnrpa_pol(rs_of(r), z) + R(81, 100) * lyp),
}

RHOS = ["1e-3", "0.05", "0.3", "2", "50"]
ZETAS = ["0", "0.3", "-0.6", "0.95"]
SRED = ["0", "0.5", "2"]


def points():
    pts = []
    for rs_ in RHOS:
        for zs in ZETAS:
            rt, zt = mp.mpf(rs_), mp.mpf(zs)
            a, b = rt * (1 + zt) / 2, rt * (1 - zt) / 2
            for s_ in SRED:
                s = mp.mpf(s_)
                if s == 0:
                    pts.append((a, b, mp.mpf("1e-40"), mp.mpf("0"), mp.mpf("1e-40")))
                    continue
                # per-channel reduced gradient s: sigma_ss = 4 (6 pi^2)^(2/3) rho_s^(8/3) s^2, cos = 0.6
                ca = 4 * (6 * mp.pi ** 2) ** (mp.mpf(2) / 3) * a ** (mp.mpf(8) / 3) * s ** 2
                cb = 4 * (6 * mp.pi ** 2) ** (mp.mpf(2) / 3) * b ** (mp.mpf(8) / 3) * (s / 2) ** 2
                pts.append((a, b, ca, mp.mpf("0.6") * mp.sqrt(ca * cb), cb))
    return pts


def main():
    mp.mp.dps = 40
    pts = points()
    out = {"variables": ["rho_a", "rho_b", "sigma_aa", "sigma_ab", "sigma_bb"],
           "points": [[float(v) for v in p] for p in pts], "values": {}}
    var = (ra, rb, saa, sab, sbb)
    for code, e in FUNCS.items():
        fs = [sp.lambdify(var, x, "mpmath") for x in [e] + [sp.diff(e, v) for v in var]]
        vals = []
        for p in pts:
            ev = fs[0](*p)
            vals.append([float(ev / (p[0] + p[1]))] + [float(f(*p)) for f in fs[1:]])
        out["values"][code] = vals
        print("code", code, "done", flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "functionals_pol.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
