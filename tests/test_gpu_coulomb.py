"""The Coulomb engine (chx_jfit_*): one-electron integrals and density-fitted J
on the device.

Pins:
* S, T, V against the host McMurchie-Davidson integrals (chx_mol_int1e, pinned
  by the published STO-3G water RHF energy and closed forms in test_scf.py),
  to 1e-12;
* density fitting against an independent route: the three-centre integrals
  (mn|P) and the metric (P|Q) taken from the host four-centre ERI code as
  (mn|P s0) and (P s0|Q s0), with s0 an s function of exponent 0 on the
  auxiliary function's atom (exactly 1, so the four-centre integral is the
  three-centre one). The fitted energy 1/2 gamma^T (P|Q)^-1 gamma and the
  fitted J depend only on the span of the auxiliary basis, so the engine's
  own harmonic polynomials and the host's pure components must give the same
  numbers: 1e-11 relative in the energy, 1e-10 of max|J| elementwise;
* the Coulomb-metric fit as a variational lower bound of the exact Coulomb
  energy 1/2 D (mn|ls) D from the full ERI tensor, with a small gap for the
  full def2-SVP-J basis (which carries the g shells the host check cannot).
"""
import math
import os

import numpy as np
import pytest

import chinium_amd as ca
from chinium_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _water(basis, n=1):
    if n == 1:
        w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
        Z, X = w["Z"], w["xyz_bohr"]
    else:
        w = wl.build_workload("h2o16_def2svp_sg2_b3lyp")
        Z, X = w["Z"][: 3 * n], w["xyz_bohr"][: 3 * n]
    return ca.Molecule(Z, X, basis), np.asarray(Z), np.asarray(X)


def _density(mol, nocc, seed):
    D, _ = wl.workload_density(mol, nocc, seed)
    return D


@pytest.mark.parametrize("basis,nw", [("cc-pvdz", 1), ("def2-svp", 1), ("def2-svp", 2)])
def test_int1e_matches_host(basis, nw):
    mol, _, _ = _water(basis, nw)
    S, T, V = mol.int1e()
    jf = ca.CoulombFit(mol)
    S2, T2, V2 = jf.int1e()
    d = [float(np.abs(a - b).max()) for a, b in ((S, S2), (T, T2), (V, V2))]
    _say("  %s (H2O)%d: max |dS| %.1e |dT| %.1e |dV| %.1e (pairs %s)" % (basis, nw, *d, jf.info()))
    assert d[0] <= 1e-13 and d[1] <= 1e-12 and d[2] <= 1e-11
    assert np.abs(S2 - S2.T).max() <= 1e-15  # both triangles are written by FP64 atomics


def _norm_prim(l, e):
    """Primitive factor of the library's normalisation (libint2 convention)."""
    df = 1.0
    for k in range(2 * l - 1, 1, -2):
        df *= k
    return math.sqrt(2 ** l * (2 * e) ** (l + 1.5) / (math.pi ** 1.5 * df))


# A small fitting basis with one contracted shell and every l the host code
# can write (s, p, d, f): exponents from def2-SVP-J.
AUX_SMALL = {
    8: [(0, [24.3864983, 10.6622662], [2.3737865, 0.0489414]), (0, [1.0447795], [1.0]),
        (1, [1.2019554], [1.0]), (2, [0.9955759], [1.0]), (3, [0.7345613], [1.0])],
    1: [(0, [0.4726794], [1.0]), (1, [0.5358730], [1.0])],
}
SYM = {1: "H", 8: "O"}
LETTER = "SPDF"


def _write_aux(root):
    os.makedirs(os.path.join(root, "BasisSets"), exist_ok=True)
    with open(os.path.join(root, "BasisSets", "fitsmall.gbs"), "w") as f:
        f.write("! test fitting basis\n****\n")
        for z, shells in AUX_SMALL.items():
            f.write("-%s    0\n" % SYM[z])
            for l, e, c in shells:
                f.write("%s  %d  1.00\n" % (LETTER[l], len(e)))
                for a, b in zip(e, c):
                    f.write("   %.10f   %.10f\n" % (a, b))
            f.write("****\n")


def test_fit_matches_four_centre_route(tmp_path):
    mol, Z, X = _water("cc-pvdz")
    n = mol.nbasis
    D = _density(mol, 5, 0)
    root = str(tmp_path)
    _write_aux(root)
    jf = ca.CoulombFit(mol, "fitsmall", data_root=root)
    g = jf.gamma(D)
    M = jf.metric()
    c = np.linalg.solve(M, g)
    Jd = jf.coulomb(c)
    E = 0.5 * g @ c
    # host route: orbital shells, then the auxiliary shells, then one s0 per atom
    m2 = ca.Molecule(Z, X, "cc-pvdz")
    aux_idx, s0_of = [], []
    nb = n
    atom_s0 = {}
    for a, z in enumerate(Z):
        for l, e, cc in AUX_SMALL[int(z)]:
            coef = [ci * _norm_prim(l, ei) for ei, ci in zip(e, cc)]
            m2.add_shell(a, {0: 0, 1: 1, 2: -2, 3: -3}[l], e, coef)
            k = 3 if l == 1 else 2 * l + 1
            aux_idx += list(range(nb, nb + k))
            s0_of += [a] * k
            nb += k
    for a in range(len(Z)):
        m2.add_shell(a, 0, [0.0], [1.0])
        atom_s0[a] = nb
        nb += 1
    eri = m2.eri()
    P = np.array(aux_idx)
    S0 = np.array([atom_s0[a] for a in s0_of])
    B = eri[:n, :n][:, :, P, S0]           # (mn|P s0)
    V = eri[P, S0][:, P, S0]                # (P s0|Q s0)
    gh = np.einsum("mnp,mn->p", B, D)
    ch = np.linalg.solve(V, gh)
    Jh = np.einsum("mnp,p->mn", B, ch)
    Eh = 0.5 * gh @ ch
    dJ = float(np.abs(Jd - Jh).max())
    _say("  fitted E_J engine %.14f host %.14f rel %.1e; max|dJ| %.1e of %.1e" % (E, Eh, abs(E - Eh) / Eh, dJ,
                                                                                np.abs(Jh).max()))
    assert abs(E - Eh) <= 1e-11 * abs(Eh)
    assert dJ <= 1e-10 * np.abs(Jh).max()
    assert np.abs(Jd - Jd.T).max() <= 1e-13 * np.abs(Jd).max()


@pytest.mark.parametrize("basis", ["cc-pvdz", "def2-svp"])
def test_fit_is_a_lower_bound_of_exact_coulomb(basis):
    """At a physical density (the RHF density: fitting bases are made for those, a random density
    matrix has structure they do not carry) the fit's deficit is small and never negative."""
    from chinium_amd.scf import RestrictedSCF
    mol, _, _ = _water(basis)
    C = RestrictedSCF(mol, 5).run()["C"]
    D = 2.0 * C[:, :5] @ C[:, :5].T
    eri = mol.eri()
    Jx = np.einsum("mnls,ls->mn", eri, D)
    Ex = 0.5 * np.sum(D * Jx)
    jf = ca.CoulombFit(mol)
    g = jf.gamma(D)
    c = np.linalg.solve(jf.metric(), g)
    Jd = jf.coulomb(c)
    Ef = 0.5 * g @ c
    _say("  %s: exact E_J %.10f fitted %.10f gap %.2e; max|J - J_fit| %.2e (naux %d)" % (
        basis, Ex, Ef, Ex - Ef, np.abs(Jx - Jd).max(), jf.naux))
    assert -1e-10 <= Ex - Ef <= 2e-4
    assert np.abs(Jx - Jd).max() <= 1e-3
