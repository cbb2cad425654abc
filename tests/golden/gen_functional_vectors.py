"""Golden vectors for the XC functionals, from an independent sympy transcription
of the published forms (not of the oracle's code), evaluated at 40 digits.

  Slater/Dirac exchange      e = -(3/4)(3/pi)^(1/3) rho^(4/3)
  VWN (5 and RPA fits)       Vosko, Wilk, Nusair, Can. J. Phys. 58, 1200 (1980), eq. 4.4
  PW92 (libxc PW_MOD params) Perdew, Wang, PRB 45, 13244 (1992), eq. 10
  PBE exchange/correlation   Perdew, Burke, Ernzerhof, PRL 77, 3865 (1996), eqs. 7-8, 14
  B88 exchange               Becke, PRA 38, 3098 (1988), eq. 8, per spin channel
  LYP correlation            Miehlich, Savin, Stoll, Preuss, CPL 157, 200 (1989), eq. 2,
                             in its general spin-resolved form, then rho_a = rho_b.
  B3LYP (libxc 402, XC part) 0.08 Slater + 0.72 B88 + 0.19 VWN-RPA + 0.81 LYP

Derivatives are sympy's, of the energy density e = rho * eps.
Run: python tests/golden/gen_functional_vectors.py  (writes functionals.json)
"""
import json
import os

import sympy as sp

rho, sigma = sp.symbols("rho sigma", positive=True)
pi = sp.pi
R = sp.Rational


def slater(r):
    return -R(3, 4) * (3 / pi) ** R(1, 3) * r ** R(4, 3)


def rs_of(r):
    return (3 / (4 * pi * r)) ** R(1, 3)


def vwn_eps(rs, A, b, c, x0):
    x = sp.sqrt(rs)
    X = lambda t: t ** 2 + b * t + c
    Q = sp.sqrt(4 * c - b ** 2)
    at = sp.atan(Q / (2 * x + b))
    return A * (sp.log(x ** 2 / X(x)) + 2 * b / Q * at
                - b * x0 / X(x0) * (sp.log((x - x0) ** 2 / X(x)) + 2 * (b + 2 * x0) / Q * at))


VWN5 = (sp.Float("0.0310907", 40), sp.Float("3.72744", 40), sp.Float("12.9352", 40), sp.Float("-0.10498", 40))
VWNRPA = (sp.Float("0.0310907", 40), sp.Float("13.0720", 40), sp.Float("42.7198", 40), sp.Float("-0.409286", 40))


def pw92_eps(rs):
    A, a1, b1, b2, b3, b4 = [sp.Float(v, 40) for v in ("0.0310907", "0.21370", "7.5957", "3.5876", "1.6382", "0.49294")]
    return -2 * A * (1 + a1 * rs) * sp.log(1 + 1 / (2 * A * (b1 * sp.sqrt(rs) + b2 * rs + b3 * rs ** R(3, 2) + b4 * rs ** 2)))


def pbe_x(r, s2g, kappa, mu):
    kf = (3 * pi ** 2 * r) ** R(1, 3)
    s2 = s2g / (2 * kf * r) ** 2
    F = 1 + kappa - kappa / (1 + mu * s2 / kappa)
    return slater(r) * F


def pbe_c(r, s2g, beta):
    gamma = (1 - sp.log(2)) / pi ** 2
    ec = pw92_eps(rs_of(r))
    kf = (3 * pi ** 2 * r) ** R(1, 3)
    ks = sp.sqrt(4 * kf / pi)
    t2 = s2g / (2 * ks * r) ** 2
    A = beta / gamma / (sp.exp(-ec / gamma) - 1)
    H = gamma * sp.log(1 + beta / gamma * t2 * (1 + A * t2) / (1 + A * t2 + A ** 2 * t2 ** 2))
    return r * (ec + H)


def b88(r, s2g):
    beta = sp.Float("0.0042", 40)
    ra = r / 2
    ga = sp.sqrt(s2g) / 2
    x = ga / ra ** R(4, 3)
    per = -R(3, 2) * (3 / (4 * pi)) ** R(1, 3) * ra ** R(4, 3) - beta * ra ** R(4, 3) * x ** 2 / (1 + 6 * beta * x * sp.asinh(x))
    return 2 * per


def lyp_general(ra, rb, gaa, gbb, gab):
    """Miehlich et al. eq. 2; gaa = |grad rho_a|^2, gab = grad rho_a . grad rho_b."""
    a, b, c, d = [sp.Float(v, 40) for v in ("0.04918", "0.132", "0.2533", "0.349")]
    r = ra + rb
    CF = R(3, 10) * (3 * pi ** 2) ** R(2, 3)
    rm = r ** R(-1, 3)
    omega = sp.exp(-c * rm) / (1 + d * rm) * r ** R(-11, 3)
    delta = c * rm + d * rm / (1 + d * rm)
    g = gaa + gbb + 2 * gab  # |grad rho|^2
    t1 = -a * 4 / (1 + d * rm) * ra * rb / r
    inner = (ra * rb * (2 ** R(11, 3) * CF * (ra ** R(8, 3) + rb ** R(8, 3))
                        + (R(47, 18) - R(7, 18) * delta) * g
                        - (R(5, 2) - delta / 18) * (gaa + gbb)
                        - (delta - 11) / 9 * (ra / r * gaa + rb / r * gbb))
             - R(2, 3) * r ** 2 * g + (R(2, 3) * r ** 2 - ra ** 2) * gbb + (R(2, 3) * r ** 2 - rb ** 2) * gaa)
    return t1 - a * b * omega * inner


def lyp(r, s2g):
    return lyp_general(r / 2, r / 2, s2g / 4, s2g / 4, s2g / 4)


def b3lyp(r, s2g):
    return (R(8, 100) * slater(r) + R(72, 100) * b88(r, s2g) + R(19, 100) * r * vwn_eps(rs_of(r), *VWNRPA)
            + R(81, 100) * lyp(r, s2g))


FUNCS = {
    "1": lambda r, s: slater(r),
    "7": lambda r, s: r * vwn_eps(rs_of(r), *VWN5),
    "8": lambda r, s: r * vwn_eps(rs_of(r), *VWNRPA),
    "101": lambda r, s: pbe_x(r, s, sp.Float("0.804", 40), sp.Float("0.2195149727645171", 40)),
    "321": lambda r, s: pbe_x(r, s, sp.Float("0.804", 40), sp.Float("0.219510", 40)),
    "130": lambda r, s: pbe_c(r, s, sp.Float("0.06672455060314922", 40)),
    "322": lambda r, s: pbe_c(r, s, sp.Float("0.06672632268006", 40)),
    "106": b88,
    "131": lyp,
    "402": b3lyp,
}

RHOS = ["1e-6", "1e-3", "0.05", "0.3", "2", "50", "500"]
S_RED = ["0", "0.3", "1", "3"]  # reduced gradient s; sigma = (2 kF rho s)^2


def main():
    out = {"points": [], "values": {}}
    pts = []
    for rs_ in RHOS:
        r = sp.Float(rs_, 40)
        kf = (3 * pi ** 2 * r) ** R(1, 3)
        for s_ in S_RED:
            sg = (2 * kf * r * sp.Float(s_, 40)) ** 2
            pts.append((r, sp.N(sg, 40)))
    out["points"] = [[float(r), float(s)] for r, s in pts]
    for code, f in FUNCS.items():
        e = f(rho, sigma)
        de_r = sp.diff(e, rho)
        de_s = sp.diff(e, sigma)
        vals = []
        for r, s in pts:
            if s == 0:
                s = sp.Float("1e-40", 40)  # the functionals' sigma floor
            sub = {rho: r, sigma: s}
            ev = sp.N(e.subs(sub), 30)
            vals.append([float(ev / r), float(sp.N(de_r.subs(sub), 30)), float(sp.N(de_s.subs(sub), 30))])
        out["values"][code] = vals
        print("code", code, "done", flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "functionals.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
