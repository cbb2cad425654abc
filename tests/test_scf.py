"""Integrals and the restricted SCF loop (chx_mol_int1e / chx_mol_eri, chinium_amd.scf) on the
CPU: Hartree-Fock energies and closed-form or quadrature checks of the integrals."""
import numpy as np
import pytest

import chinium_amd as ca
import oracle_lib as ol
from chinium_amd import workloads as wl
from chinium_amd.scf import RestrictedSCF

# Water geometry (bohr) of T. D. Crawford's programming project 3, whose published STO-3G RHF
# energy is -74.942079928192 Eh with nuclear repulsion 8.002367061810450 Eh.
H2O_XYZ = np.array([[0.0, -0.143225816552, 0.0], [1.638036840407, 1.136548822547, 0.0],
                    [-1.638036840407, 1.136548822547, 0.0]])


def _rotation(a, b, c):
    Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    Rx = np.array([[1, 0, 0], [0, np.cos(b), -np.sin(b)], [0, np.sin(b), np.cos(b)]])
    Ry = np.array([[np.cos(c), 0, np.sin(c)], [0, 1, 0], [-np.sin(c), 0, np.cos(c)]])
    return Rz @ Rx @ Ry


def test_reference_diis_driver():
    """ADIIS to 1e-1, then CDIIS (taking over ADIIS's history) to 1e-6, as SP.cpp:73-79 runs them:
    the published STO-3G water energy, and the same fixed point as the plain CDIIS loop for
    cc-pVDZ, from the core guess."""
    mol = ca.Molecule([8, 1, 1], H2O_XYZ, "sto-3g")
    r = RestrictedSCF(mol, 5).run_reference()
    phases = [h[0] for h in r["history"]]
    print("STO-3G water RHF (ADIIS -> CDIIS) %.12f in %d iterations (%d ADIIS)" % (
        r["E"], r["iterations"], phases.count("ADIIS")))
    assert abs(r["E"] - (-74.942079928192)) <= 1e-9
    assert r["diis_error"] < 1e-6
    assert phases[0] == "ADIIS" and phases[-1] == "CDIIS"
    adiis = [h[2] for h in r["history"] if h[0] == "ADIIS"]
    assert all(e >= 1e-1 for e in adiis[:-1]) and adiis[-1] < 1e-1
    mol = ca.Molecule([8, 1, 1], H2O_XYZ, "cc-pvdz")
    a = RestrictedSCF(mol, 5).run_reference()
    b = RestrictedSCF(mol, 5).run()
    print("cc-pVDZ water RHF: reference driver %.12f (%d it), plain CDIIS %.12f (%d it)" % (
        a["E"], a["iterations"], b["E"], b["iterations"]))
    assert abs(a["E"] - b["E"]) <= 1e-9


def test_adiis_weights_on_the_simplex():
    """ADIIS weights lie on the simplex and minimise the reference's quadratic model: no vertex
    or random simplex point does better."""
    from chinium_amd.scf import DIIS
    rng = np.random.default_rng(3)
    d = DIIS("ADIIS")
    for _ in range(5):
        A = rng.normal(size=(6, 6))
        d.append(A + A.T, rng.normal(size=(6, 6)), rng.normal(size=(6, 6)), 0.0)
    dot = lambda a, b: float(np.sum(a * b))
    w = d.weights(dot)
    assert np.all(w >= 0) and abs(w.sum() - 1) <= 1e-12
    n = len(d.F)
    Dn, Fn = d.D[-1], d.F[-1]
    A = np.array([dot(d.D[i] - Dn, Fn) for i in range(n)])
    B = np.array([[dot(d.D[i] - Dn, d.F[j] - Fn) for j in range(n)] for i in range(n)])
    f = lambda x: x @ A + 0.5 * x @ B @ x
    trial = [np.eye(n)[i] for i in range(n)] + [rng.dirichlet(np.ones(n)) for _ in range(200)]
    assert f(w) <= min(f(x) for x in trial) + 1e-9


def test_sto3g_water_rhf():
    mol = ca.Molecule([8, 1, 1], H2O_XYZ, "sto-3g")
    assert mol.nbasis == 7
    assert abs(mol.nuclear_repulsion() - 8.002367061810450) <= 1e-12
    r = RestrictedSCF(mol, 5).run()
    print("STO-3G water RHF %.12f in %d iterations" % (r["E"], r["iterations"]))
    assert abs(r["E"] - (-74.942079928192)) <= 1e-9


def test_h2_sto3g_closed_form():
    """H2 / STO-3G at R = 1.4 bohr: two AOs, so the RHF orbital is fixed by symmetry,
    sigma_g = (1 + 2)/sqrt(2 + 2 S12), and E follows from the integrals in closed form
    (Szabo, Ostlund, ch. 3: E = -1.1167 Eh)."""
    mol = ca.Molecule([1, 1], np.array([[0, 0, 0], [0, 0, 1.4]]), "sto-3g")
    S, T, V = mol.int1e()
    g = mol.eri()
    c = np.array([1.0, 1.0]) / np.sqrt(2 + 2 * S[0, 1])
    h = c @ (T + V) @ c
    J = np.einsum("i,j,k,l,ijkl->", c, c, c, c, g)
    E = 2 * h + J + 1 / 1.4
    r = RestrictedSCF(mol, 1).run()
    print("H2 STO-3G: closed form %.10f, SCF %.10f, S12 %.6f" % (E, r["E"], S[0, 1]))
    assert abs(r["E"] - E) <= 1e-10
    assert abs(S[0, 1] - 0.6593) <= 1e-4  # Szabo-Ostlund eq. 3.229
    assert abs(E - (-1.1167)) <= 1e-4


@pytest.mark.parametrize("case", ["h2o_ccpvdz", "ch4_tzvp"])
def test_one_electron_integrals(case):
    """S against the analytic overlap; T and V against quadrature of 1/2 grad phi . grad phi and
    -sum_C Z_C phi phi / |r - C| on an SG-3 grid (pure d and f functions included)."""
    if case == "h2o_ccpvdz":
        w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
        Z, xyz, basis = w["Z"], w["xyz_bohr"], w["basis"]
    else:
        Z, X = wl.alkane_angstrom(1, 4)
        Z, xyz, basis = Z, X * ca.ANGSTROM_TO_BOHR, "def2-tzvp"
    mol = ca.Molecule(Z, xyz, basis)
    S, T, V = mol.int1e()
    assert np.abs(S - mol.overlap()).max() <= 1e-13
    orc = ol.OracleSystem(Z, xyz, "SG-3", basis)
    P = orc.points()
    ao, g, _ = orc.ao_values(P[:, :3])
    wq = P[:, 3]
    Tq = 0.5 * sum((g[r] * wq) @ g[r].T for r in range(3))
    vn = np.zeros(len(P))
    for zc, c in zip(Z, np.asarray(xyz)):
        vn -= zc / np.linalg.norm(P[:, :3] - c, axis=1)
    Vq = (ao * (wq * vn)) @ ao.T
    dT, dV = np.abs(T - Tq).max(), np.abs(V - Vq).max()
    print("%s: %d AOs, max|T - T_quad| %.1e, max|V - V_quad| %.1e" % (case, mol.nbasis, dT, dV))
    assert dT <= 2e-4 * np.abs(T).max()
    assert dV <= 2e-4 * np.abs(V).max()


def test_rotation_invariance():
    """The RHF energy (cc-pVDZ: pure d) and the core-Hamiltonian levels (def2-TZVP: pure f) do not
    change when the molecule is rotated and translated."""
    R, t = _rotation(0.7, 1.1, -0.4), np.array([0.3, -1.2, 2.0])
    w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
    E = [RestrictedSCF(ca.Molecule(w["Z"], x, w["basis"]), 5).run()["E"]
         for x in (w["xyz_bohr"], w["xyz_bohr"] @ R.T + t)]
    print("H2O/cc-pVDZ RHF %.10f, rotated %+.1e" % (E[0], E[1] - E[0]))
    assert abs(E[1] - E[0]) <= 1e-10
    Z, X = wl.alkane_angstrom(1, 4)
    lev = []
    for x in (X * ca.ANGSTROM_TO_BOHR, X * ca.ANGSTROM_TO_BOHR @ R.T + t):
        S, T, V = ca.Molecule(Z, x, "def2-tzvp").int1e()
        s, U = np.linalg.eigh(S)
        Zm = U @ np.diag(s ** -0.5) @ U.T
        lev.append(np.linalg.eigvalsh(Zm @ (T + V) @ Zm))
    assert np.abs(lev[1] - lev[0]).max() <= 1e-9


def test_scf_argument_errors():
    mol = ca.Molecule([8, 1, 1], H2O_XYZ, "sto-3g")
    with pytest.raises(ValueError):
        RestrictedSCF(mol, 0)
    with pytest.raises(ValueError):
        RestrictedSCF(mol, 8)
    xc = ca.ExchangeCorrelation()
    xc.Read("PBE")
    with pytest.raises(ValueError):
        RestrictedSCF(mol, 5, xc)  # Kohn-Sham without a grid


def test_eri_limits():
    Z, X = wl.water_cluster_angstrom(5, 6.0, 3)
    mol = ca.Molecule(Z, X * ca.ANGSTROM_TO_BOHR, "cc-pvdz")
    assert mol.nbasis > 96
    with pytest.raises(ca.ChxError) as e:
        mol.eri()
    assert e.value.code == -5


def test_uhf_h_atom_and_closed_shell():
    """UHF: one electron has no two-electron energy (J and K cancel), so E is the lowest level of
    the core Hamiltonian (H / STO-3G: -0.4665819 Eh); for closed-shell water UHF = RHF."""
    from chinium_amd.scf import UnrestrictedSCF
    mol = ca.Molecule([1], np.zeros((1, 3)), "sto-3g")
    S, T, V = mol.int1e()
    r = UnrestrictedSCF(mol, 1, 0).run()
    print("H atom UHF %.10f, Hcore %.10f" % (r["E"], (T + V)[0, 0] / S[0, 0]))
    assert abs(r["E"] - (T + V)[0, 0] / S[0, 0]) <= 1e-12
    assert abs(r["E"] - (-0.4665819)) <= 1e-6
    mol = ca.Molecule([8, 1, 1], H2O_XYZ, "sto-3g")
    ru = UnrestrictedSCF(mol, 5, 5).run()
    assert abs(ru["E"] - (-74.942079928192)) <= 1e-9
    with pytest.raises(ValueError):
        UnrestrictedSCF(mol, 4, 5)
