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


    def run_reference(self, adiis_tol=1e-1, tol=1e-6, max_iter=300, log=None):
        """The reference's driver on this update: ADIIS to max|G| < adiis_tol, then CDIIS to
        max|G| < tol (SP.cpp:73-79), from the core guess."""
        def update(F):
            Fnew, E, G, eps, C, Exc = self.update(F)
            return Fnew, E, G, C[:, :self.nocc] @ C[:, :self.nocc].T, eps, C, Exc

        last, hist = run_reference_diis(update, self.Hcore.copy(), lambda a, b: float(np.sum(a * b)),
                                        lambda G: float(np.abs(G).max()), adiis_tol, tol, max_iter, log, self.enuc)
        Fnew, E, G, D, eps, C, Exc = last
        return dict(E=E + self.enuc, E_elec=E, E_xc=Exc, iterations=len(hist), eps=eps, C=C,
                    diis_error=float(np.abs(G).max()), history=hist)


class UnrestrictedSCF:
    """na alpha and nb beta electrons; xc = None gives UHF (EXX = 1). With xc set (Spin = 2 before
    Read; LDA/GGA or r2SCAN), grid must be a Grid of the molecule, typed and planned.

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


# ---------------------------------------------------------------------------
# DIIS as the reference runs it (src/DIIS/DIIS.cpp:88-188, CDIIS.cpp, AEDIIS.cpp)

class DIIS:
    """History of (update F, residual G, density D, energy E) and one extrapolation rule:
    "ADIIS" (AEDIIS.cpp: weights on the simplex minimising x.A + x.B x / 2 with
    A_i = <D_i - D_n, F_n>, B_ij = <D_i - D_n, F_j - F_n>) or "CDIIS" (CDIIS.cpp:
    the Pulay equations on <G_i, G_j> with a ridge factor that grows by 0.001 each time
    the space is cut for weights above 5 with a near-singular matrix). Fewer than two
    entries: the plain update (DIIS.cpp:151-163, no damping). The inner products are
    cached per pair of entries, so an iteration computes only the new ones."""

    def __init__(self, kind, max_size=20, tol=1e-6, max_iter=300):
        self.kind, self.max_size, self.tol, self.max_iter = kind, max_size, tol, max_iter
        self.F, self.G, self.D, self.E, self.ids = [], [], [], [], []
        self._next = 0
        self._dots = {}

    def steal(self, other):
        """DIIS::Steal: take the other's history, dropping the oldest beyond max_size."""
        k = max(0, len(other.F) - self.max_size)
        self.F, self.G, self.D, self.E = other.F[k:], other.G[k:], other.D[k:], other.E[k:]
        self.ids = other.ids[k:]
        self._next, self._dots = other._next, dict(other._dots)

    def append(self, F, G, D, E):
        if len(self.F) == self.max_size:
            self.F.pop(0); self.G.pop(0); self.D.pop(0); self.E.pop(0); self.ids.pop(0)
        self.F.append(F); self.G.append(G); self.D.append(D); self.E.append(E)
        self.ids.append(self._next)
        self._next += 1
        live = set(self.ids)
        self._dots = {k: v for k, v in self._dots.items() if k[1] in live and k[2] in live}

    def _dot(self, dot, kind, i, j):
        key = (kind, self.ids[i], self.ids[j])
        if key not in self._dots:
            a = self.D[i] if kind == "DF" else self.G[i]
            b = self.F[j] if kind == "DF" else self.G[j]
            self._dots[key] = dot(a, b)
        return self._dots[key]

    def weights(self, dot):
        n = len(self.F)
        if self.kind == "CDIIS":
            ridge = 1.0
            for size in range(n, 1, -1):
                H = np.zeros((size + 1, size + 1))
                for i in range(size):
                    H[i, size] = H[size, i] = 1.0
                    for j in range(i, size):
                        H[i, j] = H[j, i] = self._dot(dot, "GG", n - size + i, n - size + j)
                H[np.diag_indices(size + 1)] *= ridge
                b = np.zeros(size + 1)
                b[size] = 1.0
                x = np.linalg.lstsq(H, b, rcond=None)[0][:size]
                w = np.zeros(n)
                w[n - size:] = x
                if np.abs(x).max() < 5 or np.linalg.det(H) > 1e-12:
                    return w
                ridge += 0.001
            return w
        # ADIIS: minimise on the simplex, x = t^2 / sum t^2
        from scipy.optimize import minimize
        DF = np.array([[self._dot(dot, "DF", i, j) for j in range(n)] for i in range(n)])  # <D_i, F_j>
        A = DF[:, n - 1] - DF[n - 1, n - 1]
        B = DF - DF[:, [n - 1]] - DF[[n - 1], :] + DF[n - 1, n - 1]
        Bs = 0.5 * (B + B.T)

        def f(t):
            s = t @ t
            x = t * t / s
            val = x @ A + 0.5 * x @ Bs @ x
            gx = A + Bs @ x
            gt = 2 * t / s * (gx - x @ gx)
            return val, gt

        x0 = np.full(n, 0.1 / max(1, n - 1))
        x0[-1] = 0.9
        r = minimize(f, np.sqrt(x0), jac=True, method="L-BFGS-B", options=dict(maxiter=1000, gtol=1e-10))
        t = r.x
        return t * t / (t @ t)


class RestrictedDFSCF:
    """Restricted Kohn-Sham SCF (RestrictedDIIS, src/HartreeFockKohnSham/Restricted/SP.cpp:20-81) with
    S, T, V and J from the device Coulomb engine (CoulombFit: J fitted in an auxiliary basis with the
    Coulomb metric), the XC part from the grid, and the O(n^3) algebra (diagonalisation, products)
    on the device through rocBLAS / rocSOLVER (chinium_amd.linalg). For functionals without exact
    exchange (EXX = 0, e.g. r2SCAN), which is what the engine serves.

    Update (SP.cpp:36-66): orbitals of F in the S^{-1/2} basis, D = C_occ C_occ^T,
    J = J[2D] (fitted), XC sequence grid.getDensity({2D}) -> Evaluate("ev") -> getEnergy ->
    getFock, E = sum D (Hcore + Fhf) + E_xc, G = 2 (F D S - S D F). Then ADIIS until
    max|G| < adiis_tol, CDIIS (taking over ADIIS's history) until max|G| < tol
    (SP.cpp:73-79)."""

    def __init__(self, mol, nocc, xc, grid, coulomb=None):
        from . import CoulombFit
        from . import linalg as la

        if float(xc.EXX) != 0.0:
            raise ValueError("the density-fitted SCF serves functionals without exact exchange")
        if not 0 < int(nocc) <= mol.nbasis:
            raise ValueError("nocc must be in 1..nbasis")
        self.la = la
        self.mol, self.nocc, self.xc, self.grid = mol, int(nocc), xc, grid
        n = self.n = mol.nbasis
        t0 = time.perf_counter()
        self.jf = coulomb if coulomb is not None else CoulombFit(mol)
        S, T, V = self.jf.int1e()
        M = self.jf.metric()
        self.S, self.Hcore = S, T + V
        self.info = la.DMat(1)
        self.Lm = la.DMat.of(M)
        la.potrf(self.Lm, self.info)
        self.gam = la.DMat(M.shape[0])
        # S^{-1/2} from the eigenpairs of S
        self.w, self.work = la.DMat(n), la.DMat(n)
        U = la.DMat.of(S)
        la.syevd(U, self.w, self.work, self.info)
        s = self.w.get()
        Us = U.get() * s ** -0.25
        self.Z = la.DMat.of(Us @ Us.T)
        self.Sd = la.DMat.of(S)
        self.A, self.B, self.C = la.DMat(n, n), la.DMat(n, n), la.DMat(n, n)
        self.Dd = la.DMat(n, n)
        self.t_setup = time.perf_counter() - t0
        self.enuc = mol.nuclear_repulsion()
        self.timing = dict(j=[], xc=[], linalg=[])

    def coulomb(self, P):
        """Fitted J[P] and the fitted Coulomb energy 1/2 gamma^T c."""
        g = self.jf.gamma(P)
        self.gam.set(g)
        self.la.potrs(self.Lm, self.gam)
        c = self.gam.get()
        return self.jf.coulomb(c), 0.5 * float(g @ c)

    def update(self, F):
        la = self.la
        t0 = time.perf_counter()
        self.A.set(F)
        la.gemm(self.Z, self.A, self.B, ta=True)           # Z^T F
        la.gemm(self.B, self.Z, self.A)                    # Z^T F Z
        la.syevd(self.A, self.w, self.work, self.info)     # eigenvectors C' over A
        la.gemm(self.Z, self.A, self.C)                    # C = Z C'
        la.gemm(self.C, self.C, self.Dd, tb=True, k=self.nocc)  # D = C_occ C_occ^T
        D = self.Dd.get()
        eps = self.w.get()
        t1 = time.perf_counter()
        J, _ = self.coulomb(2.0 * D)
        t2 = time.perf_counter()
        self.grid.getDensity([2.0 * D])
        self.xc.Evaluate("ev", self.grid)
        Exc = self.grid.getEnergy()
        Vxc = self.grid.getFock()[0]
        t3 = time.perf_counter()
        Fhf = self.Hcore + J
        Fnew = Fhf + Vxc
        E = float(np.sum(D * (self.Hcore + Fhf))) + Exc
        self.A.set(Fnew)
        la.gemm(self.A, self.Dd, self.B)                   # F D
        la.gemm(self.B, self.Sd, self.A)                   # F D S (its transpose is S D F)
        X = self.A.get()
        G = 2.0 * (X - X.T)
        t4 = time.perf_counter()
        self.timing["linalg"].append((t1 - t0) + (t4 - t3))
        self.timing["j"].append(t2 - t1)
        self.timing["xc"].append(t3 - t2)
        return Fnew, E, G, D, eps, Exc

    def run(self, F0=None, adiis_tol=1e-1, tol=1e-6, max_iter=300, log=None):
        """Returns dict(E total, E_elec, E_xc, iterations, max|G|, history, F, eps). Raises
        RuntimeError("Convergence failed!") as the reference when a phase runs out of iterations.
        F0: starting Fock matrix (default Hcore, the core guess)."""
        F = self.Hcore.copy() if F0 is None else np.asarray(F0, dtype=np.float64)
        last, hist = run_reference_diis(self.update, F, lambda a, b: float(np.vdot(a, b)),
                                        lambda G: float(np.abs(G).max()), adiis_tol, tol, max_iter, log, self.enuc)
        Fnew, E, G, D, eps, Exc = last
        return dict(E=E + self.enuc, E_elec=E, E_xc=Exc, iterations=len(hist), diis_error=float(np.abs(G).max()),
                    history=hist, F=Fnew, eps=eps)


def run_reference_diis(update, F, dot, maxabs, adiis_tol=1e-1, tol=1e-6, max_iter=300, log=None, enuc=0.0):
    """The reference's SCF driver (SP.cpp:73-79): ADIIS until max|G| < adiis_tol, then CDIIS,
    which takes over ADIIS's history (CDIIS::Steal) and restarts hot from its last entry, until
    max|G| < tol. update(F) -> (Fnew, E, G, D, ...). Returns (last update's tuple, history of
    (phase, E, max|G|)). Raises RuntimeError("Convergence failed!") when a phase runs out."""
    hist = []
    last = None

    def phase(d, hot):
        nonlocal F, last
        for it in range(d.max_iter):
            if not (it == 0 and hot and d.F):  # a hot restart skips the first computation (DIIS.cpp:116-122)
                out = update(F)
                last = out
                Fnew, E, G, D = out[:4]
                d.append(Fnew, G, D, E)
                hist.append((d.kind, E, maxabs(G)))
                if log:
                    log("  %s %3d  E %.10f  max|G| %.3e" % (d.kind, len(hist), E + enuc, hist[-1][2]))
            if maxabs(d.G[-1]) < d.tol:
                return True
            if len(d.F) < 2:
                F = d.F[-1]
            else:
                w = d.weights(dot)
                F = sum(float(wi) * Fi for wi, Fi in zip(w, d.F))
        return False

    a = DIIS("ADIIS", 20, adiis_tol, max_iter)
    if not phase(a, False):
        raise RuntimeError("Convergence failed!")
    c = DIIS("CDIIS", 20, tol, max_iter)
    c.steal(a)
    if not phase(c, True):
        raise RuntimeError("Convergence failed!")
    return last, hist
