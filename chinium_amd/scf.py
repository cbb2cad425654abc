"""Restricted SCF around the XC path, for small molecules.

Mirrors RestrictedDIIS (src/HartreeFockKohnSham/Restricted/SP.cpp:20-81), the
caller of the XC path. Each iteration:

* diagonalises F in an orthonormal basis;
* forms D = C_occ C_occ^T and Ghf = J - K from the integrals (J = 2 (mu nu|la si) D_la,si,
  K = EXX (mu la|nu si) D_la,si, the factors of Int4C2E::ContractInts as SP.cpp:43-46 uses them);
* runs the XC sequence grid.getDensity({2 D}) -> xc.Evaluate("ev") -> grid.getEnergy() ->
  grid.getFock() on the GPU (SP.cpp:49-53);
* sets E = sum D (Hcore + Fhf) + E_xc (SP.cpp:55) and F = Hcore + Ghf + Vxc;
* extrapolates F by DIIS on G = 2 (F D S - S D F) (SP.cpp:58).

The reference runs ADIIS to 1e-1 and then CDIIS to 1e-6 on the DIIS error
(SP.cpp:73-77); this loop uses CDIIS alone (up to 8 vectors) from the
core-Hamiltonian guess. The integrals come from chx_mol_int1e / chx_mol_eri
(McMurchie-Davidson on the host, full ERI tensor), so the loop is meant for
small bases (n <= 96); the XC part runs wherever the grid is.
"""
import time

import numpy as np


class RestrictedSCF:
    """nocc doubly occupied orbitals; xc = None gives Hartree-Fock (EXX = 1).

    With xc set, grid must be a Grid of the molecule, typed and planned
    (setType(xc.family), getAO) in either precision form."""

    def __init__(self, mol, nocc, xc=None, grid=None):
        if xc is not None and grid is None:
            raise ValueError("a Kohn-Sham run needs a grid")
        if not 0 < int(nocc) <= mol.nbasis:
            raise ValueError("nocc must be in 1..nbasis")
        self.mol, self.nocc, self.xc, self.grid = mol, int(nocc), xc, grid
        self.S, T, V = mol.int1e()
        self.Hcore = T + V
        self.eri = mol.eri()
        self.enuc = mol.nuclear_repulsion()
        s, U = np.linalg.eigh(self.S)
        self.Z = U @ np.diag(s ** -0.5) @ U.T  # S^{-1/2}
        self.exx = 1.0 if xc is None else float(xc.EXX)
        self.xc_ms = []  # wall time of each XC build (getDensity -> getFock)

    def update(self, F):
        """One reference update: orbitals of F, then the new Fock matrix, energy and DIIS error."""
        eps, Cp = np.linalg.eigh(self.Z.T @ F @ self.Z)
        C = self.Z @ Cp
        D = C[:, :self.nocc] @ C[:, :self.nocc].T
        J = 2.0 * np.einsum("ijkl,kl->ij", self.eri, D, optimize=True)
        K = self.exx * np.einsum("ikjl,kl->ij", self.eri, D, optimize=True)
        Fhf = self.Hcore + J - K
        Exc, Vxc = 0.0, 0.0
        if self.xc is not None:
            t0 = time.perf_counter()
            self.grid.getDensity([2.0 * D])
            self.xc.Evaluate("ev", self.grid)
            Exc = self.grid.getEnergy()
            Vxc = self.grid.getFock()[0]
            self.xc_ms.append(1e3 * (time.perf_counter() - t0))
        Fnew = Fhf + Vxc
        E = float(np.sum(D * (self.Hcore + Fhf))) + Exc
        G = 2.0 * (Fnew @ D @ self.S - self.S @ D @ Fnew)
        return Fnew, E, G, eps, C, Exc

    def run(self, max_iter=100, e_tol=1e-10, g_tol=1e-7, ndiis=8):
        """Returns dict(E total, E_xc, iterations, orbital energies, C). Raises RuntimeError
        ("Convergence failed!", as the reference) when max_iter is reached."""
        F = self.Hcore.copy()
        Fs, Gs = [], []
        E_old = None
        for it in range(1, max_iter + 1):
            Fnew, E, G, eps, C, Exc = self.update(F)
            err = np.abs(G).max()
            if E_old is not None and abs(E - E_old) < e_tol and err < g_tol:
                return dict(E=E + self.enuc, E_elec=E, E_xc=Exc, iterations=it, eps=eps, C=C,
                            D=C[:, :self.nocc] @ C[:, :self.nocc].T, diis_error=err)
            E_old = E
            Fs.append(Fnew)
            Gs.append(G)
            Fs, Gs = Fs[-ndiis:], Gs[-ndiis:]
            m = len(Fs)
            B = -np.ones((m + 1, m + 1))
            B[m, m] = 0.0
            for a in range(m):
                for b in range(m):
                    B[a, b] = np.sum(Gs[a] * Gs[b])
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            try:
                w = np.linalg.solve(B, rhs)[:m]
                F = sum(wi * Fi for wi, Fi in zip(w, Fs))
            except np.linalg.LinAlgError:
                F = Fnew
        raise RuntimeError("Convergence failed!")


class UnrestrictedSCF:
    """na alpha and nb beta electrons; xc = None gives UHF (EXX = 1). With xc set (Spin = 2 before
    Read, LDA/GGA), grid must be a Grid of the molecule, typed and planned.

    Mirrors the unrestricted update (src/HartreeFockKohnSham/Unrestricted/SP.cpp:42-80): orbitals of
    F_a and F_b, D_s = C_s,occ C_s,occ^T, J from D_a + D_b and K_s = EXX (mu la|nu si) D_s,la,si per
    spin, the XC sequence grid.getDensity({D_a, D_b}) -> Evaluate("ev") -> getEnergy -> getFock
    ({V_a, V_b}), E = 1/2 sum_s D_s (Hcore + Fhf_s) + E_xc, and DIIS on the errors
    S^{-1/2} (F_s D_s S - S D_s F_s) S^{-1/2} of both spins together."""

    def __init__(self, mol, na, nb, xc=None, grid=None):
        if xc is not None and grid is None:
            raise ValueError("a Kohn-Sham run needs a grid")
        if not (0 < int(na) <= mol.nbasis and 0 <= int(nb) <= int(na)):
            raise ValueError("need 0 <= nb <= na <= nbasis and na >= 1")
        self.mol, self.na, self.nb, self.xc, self.grid = mol, int(na), int(nb), xc, grid
        self.S, T, V = mol.int1e()
        self.Hcore = T + V
        self.eri = mol.eri()
        self.enuc = mol.nuclear_repulsion()
        s, U = np.linalg.eigh(self.S)
        self.Z = U @ np.diag(s ** -0.5) @ U.T
        self.exx = 1.0 if xc is None else float(xc.EXX)

    def _density(self, F, nocc):
        eps, Cp = np.linalg.eigh(self.Z.T @ F @ self.Z)
        C = self.Z @ Cp
        return eps, C, C[:, :nocc] @ C[:, :nocc].T

    def update(self, Fa, Fb):
        epsa, Ca, Da = self._density(Fa, self.na)
        epsb, Cb, Db = self._density(Fb, self.nb)
        J = np.einsum("ijkl,kl->ij", self.eri, Da + Db, optimize=True)
        Ka = self.exx * np.einsum("ikjl,kl->ij", self.eri, Da, optimize=True)
        Kb = self.exx * np.einsum("ikjl,kl->ij", self.eri, Db, optimize=True)
        Fhfa, Fhfb = self.Hcore + J - Ka, self.Hcore + J - Kb
        Exc, Va, Vb = 0.0, 0.0, 0.0
        if self.xc is not None:
            self.grid.getDensity([Da, Db])
            self.xc.Evaluate("ev", self.grid)
            Exc = self.grid.getEnergy()
            Va, Vb = self.grid.getFock()
        Fna, Fnb = Fhfa + Va, Fhfb + Vb
        E = 0.5 * float(np.sum(Da * (self.Hcore + Fhfa)) + np.sum(Db * (self.Hcore + Fhfb))) + Exc
        Ga = self.Z @ (Fna @ Da @ self.S - self.S @ Da @ Fna) @ self.Z
        Gb = self.Z @ (Fnb @ Db @ self.S - self.S @ Db @ Fnb) @ self.Z
        return Fna, Fnb, E, Ga, Gb, (epsa, epsb), (Da, Db), Exc

    def run(self, max_iter=150, e_tol=1e-10, g_tol=1e-7, ndiis=8):
        """Returns dict(E total, E_xc, iterations, eps (alpha, beta), D (alpha, beta))."""
        Fa, Fb = self.Hcore.copy(), self.Hcore.copy()
        Fs, Gs = [], []
        E_old = None
        n = self.mol.nbasis
        for it in range(1, max_iter + 1):
            Fna, Fnb, E, Ga, Gb, eps, D, Exc = self.update(Fa, Fb)
            err = max(np.abs(Ga).max(), np.abs(Gb).max())
            if E_old is not None and abs(E - E_old) < e_tol and err < g_tol:
                return dict(E=E + self.enuc, E_elec=E, E_xc=Exc, iterations=it, eps=eps, D=D, diis_error=err)
            E_old = E
            Fs.append(np.hstack([Fna, Fnb]))
            Gs.append(np.hstack([Ga, Gb]))
            Fs, Gs = Fs[-ndiis:], Gs[-ndiis:]
            m = len(Fs)
            B = -np.ones((m + 1, m + 1))
            B[m, m] = 0.0
            for a in range(m):
                for b in range(m):
                    B[a, b] = np.sum(Gs[a] * Gs[b])
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            try:
                w = np.linalg.solve(B, rhs)[:m]
                F = sum(wi * Fi for wi, Fi in zip(w, Fs))
            except np.linalg.LinAlgError:
                F = np.hstack([Fna, Fnb])
            Fa, Fb = F[:, :n], F[:, n:]
        raise RuntimeError("Convergence failed!")
