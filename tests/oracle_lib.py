"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure: only the tests, the smoke check and the benchmark's CPU
baseline use this module. It is the checker, never the thing shipped.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
DATA_ROOT = os.path.join(ROOT, "chinium_amd", "data")

_c_int_p = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
    L = ctypes.CDLL(ORACLE_SO)
    sig = {
        "ox_create": (_vp, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, _vp, _vp, ctypes.c_int]),
        "ox_destroy": (None, [_vp]),
        "ox_last_error": (ctypes.c_char_p, []),
        "ox_nbasis": (ctypes.c_int, [_vp]),
        "ox_npoints": (ctypes.c_long, [_vp]),
        "ox_nsubgrids": (ctypes.c_int, [_vp]),
        "ox_nshells": (ctypes.c_int, [_vp]),
        "ox_get_points": (ctypes.c_int, [_vp, _vp]),
        "ox_get_raw_weights": (ctypes.c_int, [_vp, _vp]),
        "ox_get_subgrid_sizes": (ctypes.c_int, [_vp, _vp]),
        "ox_get_shells": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
        "ox_lebedev": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _vp]),
        "ox_set_functional_codes": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
        "ox_set_functional_file": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
        "ox_family": (ctypes.c_int, [_vp]),
        "ox_set_subgrid_subset": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
        "ox_get_ao": (ctypes.c_int, [_vp]),
        "ox_xc": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
        "ox_xc_grad": (ctypes.c_int, [_vp, _vp, _vp]),
        "ox_shift_basis": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "ox_density": (ctypes.c_int, [_vp, _vp]),
        "ox_evaluate": (ctypes.c_int, [_vp]),
        "ox_energy": (ctypes.c_int, [_vp, _vp]),
        "ox_num_electrons": (ctypes.c_int, [_vp, _vp]),
        "ox_fock": (ctypes.c_int, [_vp, _vp]),
        "ox_eval_functional": (ctypes.c_int, [ctypes.c_int, _vp, ctypes.c_long] + [_vp] * 9),
        "ox_ao_values": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp]),
        "ox_active_points": (ctypes.c_long, [_vp]),
        "ox_get_density": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
        "ox_get_active_xyz": (ctypes.c_int, [_vp, _vp]),
        "ox_get_active_raw_weights": (ctypes.c_int, [_vp, _vp]),
        "ox_eval_functional_f": (ctypes.c_int, [ctypes.c_int, _vp, ctypes.c_long] + [_vp] * 6),
        "ox_fock_u": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
        "ox_fock_s": (ctypes.c_int, [_vp, _vp, _vp]),
        "ox_fock_skeleton": (ctypes.c_int, [_vp, _vp, _vp]),
        "ox_xc_hess": (ctypes.c_int, [_vp, _vp, _vp]),
        "ox_xc_pol": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "ox_eval_functional_pol": (ctypes.c_int, [ctypes.c_int, _vp, ctypes.c_long, _vp, _vp, _vp, _vp, _vp]),
        "ox_eval_functional_fpol": (ctypes.c_int, [ctypes.c_int, _vp, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, _vp]),
        "ox_fock_u_pol": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp]),
        "ox_eval_functional_mpol": (ctypes.c_int, [ctypes.c_int, _vp, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, _vp,
                                                   _vp]),
        "ox_mlp_forward": (ctypes.c_int, [ctypes.c_long, _vp, _vp, _vp, _vp]),
        "ox_mlp_loss_grad": (ctypes.c_int, [ctypes.c_long, _vp, _vp, _vp, ctypes.c_double, _vp, _vp, _vp]),
        "ox_mlp_descriptors": (ctypes.c_int, [ctypes.c_long, _vp, _vp, _vp, _vp, _vp]),
        "ox_mlp_point": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, ctypes.c_int, _vp, ctypes.c_int, _vp,
                                        ctypes.c_long] + [_vp] * 8),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().ox_last_error().decode())


class OracleSystem:
    """One molecule + basis + grid, evaluated the reference's way on the CPU."""

    def __init__(self, Z, xyz_bohr, grid, basis, nthreads=0, data_root=DATA_ROOT):
        L = lib()
        self.Z = np.ascontiguousarray(Z, dtype=np.int32)
        self.xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
        self.data_root = data_root
        h = L.ox_create(data_root.encode(), grid.encode(), basis.encode(), len(self.Z),
                        _ptr(self.Z), _ptr(self.xyz), int(nthreads))
        if not h:
            raise RuntimeError(L.ox_last_error().decode())
        self.h = h
        self.nbasis = L.ox_nbasis(h)
        self.npoints = L.ox_npoints(h)
        self.nsubgrids = L.ox_nsubgrids(h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ox_destroy(self.h)
            self.h = None

    def points(self):
        P = np.zeros((self.npoints, 4))
        lib().ox_get_points(self.h, _ptr(P))
        return P

    def raw_weights(self):
        w = np.zeros(self.npoints)
        lib().ox_get_raw_weights(self.h, _ptr(w))
        return w

    def subgrid_sizes(self):
        s = np.zeros(self.nsubgrids, dtype=np.int32)
        lib().ox_get_subgrid_sizes(self.h, _ptr(s))
        return s

    def shells(self):
        L = lib()
        ns = L.ox_nshells(self.h)
        t = np.zeros(ns, np.int32); npr = np.zeros(ns, np.int32); c = np.zeros(ns, np.int32)
        e = np.zeros(4096); k = np.zeros(4096)
        L.ox_get_shells(self.h, _ptr(t), _ptr(npr), _ptr(c), _ptr(e), _ptr(k))
        n = int(npr.sum())
        return t, npr, c, e[:n].copy(), k[:n].copy()

    def set_functional(self, name=None, codes=None):
        L = lib()
        if codes is not None:
            cc = np.ascontiguousarray(codes, dtype=np.int32)
            _check(L.ox_set_functional_codes(self.h, len(cc), _ptr(cc)))
        else:
            _check(L.ox_set_functional_file(self.h, self.data_root.encode(), name.encode()))
        self.family = L.ox_family(self.h)

    def set_subset(self, ids):
        ids = np.ascontiguousarray(ids if ids is not None else [], dtype=np.int32)
        _check(lib().ox_set_subgrid_subset(self.h, len(ids), _ptr(ids)))

    def get_ao(self):
        _check(lib().ox_get_ao(self.h))

    def xc(self, D):
        n = self.nbasis
        Df = np.asfortranarray(D, dtype=np.float64)
        V = np.zeros((n, n), order="F")
        E = ctypes.c_double()
        N = ctypes.c_double()
        _check(lib().ox_xc(self.h, _ptr(Df), _ptr(V), ctypes.addressof(E), ctypes.addressof(N)))
        return E.value, N.value, np.array(V)

    def xc_pol(self, Da, Db):
        """Spin-polarised build (LDA/GGA): E, N, Va, Vb."""
        n = self.nbasis
        A = np.asfortranarray(Da, dtype=np.float64)
        B = np.asfortranarray(Db, dtype=np.float64)
        Va, Vb = np.zeros((n, n), order="F"), np.zeros((n, n), order="F")
        E, N = ctypes.c_double(), ctypes.c_double()
        _check(lib().ox_xc_pol(self.h, _ptr(A), _ptr(B), _ptr(Va), _ptr(Vb), ctypes.addressof(E), ctypes.addressof(N)))
        return E.value, N.value, np.array(Va), np.array(Vb)

    def fock_u(self, D, Dxs):
        """Response Fock matrices for the perturbed density matrices Dxs at the
        reference density D (Evaluate("f"), getDensityU, getFockU<u_t>)."""
        n = self.nbasis
        Df = np.asfortranarray(D, dtype=np.float64)
        _check(lib().ox_density(self.h, _ptr(Df)))
        X = np.stack([np.asfortranarray(d, dtype=np.float64).ravel(order="F") for d in Dxs])
        F = np.zeros_like(X)
        _check(lib().ox_fock_u(self.h, len(Dxs), _ptr(X), _ptr(F)))
        return [F[m].reshape((n, n), order="F") for m in range(len(Dxs))]

    def fock_u_pol(self, Da, Db, Dxa, Dxb):
        """Spin-polarised response matrices (Fa[m], Fb[m]) for the perturbed pairs (Dxa[m], Dxb[m])
        at the ground state (Da, Db): Evaluate("f") polarised, getDensityU, getFockU<u_t>."""
        n, m = self.nbasis, len(Dxa)
        A = np.asfortranarray(Da, dtype=np.float64)
        B = np.asfortranarray(Db, dtype=np.float64)
        XA = np.stack([np.asfortranarray(d, dtype=np.float64).ravel(order="F") for d in Dxa])
        XB = np.stack([np.asfortranarray(d, dtype=np.float64).ravel(order="F") for d in Dxb])
        FA, FB = np.zeros_like(XA), np.zeros_like(XB)
        _check(lib().ox_fock_u_pol(self.h, _ptr(A), _ptr(B), m, _ptr(XA), _ptr(XB), _ptr(FA), _ptr(FB)))
        return ([FA[k].reshape((n, n), order="F") for k in range(m)],
                [FB[k].reshape((n, n), order="F") for k in range(m)])

    def fock_s(self, D):
        """Skeleton XC response matrices, one per nuclear coordinate [3 atom + t], at the
        ground-state density D (getDensitySkeleton, Evaluate("f"), getFockU<s_t>)."""
        n = self.nbasis
        Df = np.asfortranarray(D, dtype=np.float64)
        F = np.zeros((3 * len(self.Z), n * n))
        _check(lib().ox_fock_s(self.h, _ptr(Df), _ptr(F)))
        return [F[m].reshape((n, n), order="F") for m in range(F.shape[0])]

    def fock_skeleton(self, D):
        """Skeleton derivative of the XC Fock matrix with the potential held fixed, one matrix per
        nuclear coordinate [3 atom + t] (Grid::getFockSkeleton)."""
        n = self.nbasis
        Df = np.asfortranarray(D, dtype=np.float64)
        F = np.zeros((3 * len(self.Z), n * n))
        _check(lib().ox_fock_skeleton(self.h, _ptr(Df), _ptr(F)))
        return [F[m].reshape((n, n), order="F") for m in range(F.shape[0])]

    def xc_hess(self, D):
        """XC energy Hessian (3 natoms x 3 natoms), skeleton form with the grid fixed."""
        Df = np.asfortranarray(D, dtype=np.float64)
        na = len(self.Z)
        H = np.zeros(9 * na * na)
        _check(lib().ox_xc_hess(self.h, _ptr(Df), _ptr(H)))
        return H.reshape(3 * na, 3 * na)

    def xc_grad(self, D):
        """XC nuclear gradient (natoms x 3), skeleton form with the grid fixed."""
        Df = np.asfortranarray(D, dtype=np.float64)
        G = np.zeros(3 * len(self.Z))
        _check(lib().ox_xc_grad(self.h, _ptr(Df), _ptr(G)))
        return G.reshape(-1, 3)

    def shift_basis(self, atom, d):
        """Moves one atom's basis functions (not its grid); call get_ao afterwards."""
        _check(lib().ox_shift_basis(self.h, int(atom), float(d[0]), float(d[1]), float(d[2])))

    def density(self):
        L = lib()
        n = L.ox_active_points(self.h)
        rho = np.zeros(n); grad = np.zeros((n, 3)); sig = np.zeros(n); tau = np.zeros(n); w = np.zeros(n)
        L.ox_get_density(self.h, _ptr(rho), _ptr(grad), _ptr(sig), _ptr(tau), _ptr(w))
        return rho, grad, sig, tau, w

    def density_at(self, D):
        """Density of D on the active sub-grids (after get_ao): xyz, rho, sigma, tau, w per unpadded point."""
        n = self.nbasis
        Df = np.asfortranarray(D, dtype=np.float64)
        _check(lib().ox_density(self.h, _ptr(Df)))
        rho, grad, sig, tau, w = self.density()
        xyz = np.zeros((len(rho), 3))
        _check(lib().ox_get_active_xyz(self.h, _ptr(xyz)))
        return xyz, rho, sig, tau, w

    def active_raw_weights(self):
        """Radial x angular weights (before the Becke partition) of the active sub-grids' points."""
        w = np.zeros(lib().ox_active_points(self.h))
        _check(lib().ox_get_active_raw_weights(self.h, _ptr(w)))
        return w

    def ao_values(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        n = len(xyz)
        ao = np.zeros((self.nbasis, n)); g = np.zeros((3, self.nbasis, n)); l = np.zeros((self.nbasis, n))
        _check(lib().ox_ao_values(self.h, n, _ptr(xyz), _ptr(ao), _ptr(g), _ptr(l)))
        return ao, g, l


def lebedev(n, data_root=DATA_ROOT):
    out = np.zeros((n, 4))
    _check(lib().ox_lebedev(data_root.encode(), int(n), _ptr(out)))
    return out


def eval_functional(codes, rho, sigma=None, tau=None):
    """Per-point e (per particle), v_rho, v_sigma, v_tau (libxc conventions)."""
    L = lib()
    cc = np.ascontiguousarray(codes, dtype=np.int32)
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    n = len(rho)
    sigma = None if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64)
    tau = None if tau is None else np.ascontiguousarray(tau, dtype=np.float64)
    e = np.zeros(n); vr = np.zeros(n); vs = np.zeros(n); vl = np.zeros(n); vt = np.zeros(n)
    _check(L.ox_eval_functional(len(cc), _ptr(cc), n, _ptr(rho), _ptr(sigma), None, _ptr(tau),
                                _ptr(e), _ptr(vr), _ptr(vs), _ptr(vl), _ptr(vt)))
    return e, vr, vs, vt


def eval_functional_f(codes, rho, sigma=None):
    """Per-point v_sigma, v2rho2, v2rhosigma, v2sigma2 (libxc xc_*_fxc conventions)."""
    L = lib()
    cc = np.ascontiguousarray(codes, dtype=np.int32)
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    n = len(rho)
    sigma = None if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64)
    vs, frr, frs, fss = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    _check(L.ox_eval_functional_f(len(cc), _ptr(cc), n, _ptr(rho), _ptr(sigma), _ptr(vs), _ptr(frr), _ptr(frs),
                                  _ptr(fss)))
    return vs, frr, frs, fss


MLP_PARAMS = 8641


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def mlp_forward(P, U):
    """FP64 network value F and input gradient dF/dU at normalised inputs U (n x 3)."""
    P, U = _d(P), _d(U).reshape(-1, 3)
    n = U.shape[0]
    F, dU = np.zeros(n), np.zeros((n, 3))
    _check(lib().ox_mlp_forward(n, _ptr(U), _ptr(P), _ptr(F), _ptr(dU)))
    return F, dU


def mlp_loss_grad(P, U, T, C=None):
    """L = sum C_i (y_i - T_i)^2 / sum C_i and dL/dP (FP64)."""
    P, U, T = _d(P), _d(U).reshape(-1, 3), _d(T)
    n = U.shape[0]
    Cd = None if C is None else _d(C)
    cs = 1.0 / (n if C is None else float(Cd.sum()))
    G, L = np.zeros(MLP_PARAMS), ctypes.c_double()
    _check(lib().ox_mlp_loss_grad(n, _ptr(U), _ptr(T), _ptr(Cd), cs, _ptr(P), _ptr(G), ctypes.byref(L)))
    return L.value, G


def mlp_descriptors(rho, sigma, tau):
    """t = (ln rho, s/(1+s), alpha/(1+alpha)) (n x 3) and dt/d(rho, sigma, tau) (n x 3 x 3)."""
    rho, sigma, tau = _d(rho), _d(sigma), _d(tau)
    n = rho.size
    t, dt = np.zeros((n, 3)), np.zeros((n, 3, 3))
    _check(lib().ox_mlp_descriptors(n, _ptr(rho), _ptr(sigma), _ptr(tau), _ptr(t), _ptr(dt)))
    return t, dt


def mlp_point(P, lo, hi, rho_min, codes, rho, sigma, tau, base=()):
    """Surrogate eps, v_rho, v_sigma, v_tau with the analytic fallback (base: the analytic codes the
    network's part is added to); used_net flags."""
    P, lo, hi = _d(P), _d(lo), _d(hi)
    rho, sigma, tau = _d(rho), _d(sigma), _d(tau)
    c = np.ascontiguousarray(codes, dtype=np.int32)
    b = np.ascontiguousarray(list(base) or [0], dtype=np.int32)
    n = rho.size
    out = [np.zeros(n) for _ in range(4)]
    used = np.zeros(n, dtype=np.int32)
    _check(lib().ox_mlp_point(_ptr(P), _ptr(lo), _ptr(hi), rho_min, c.size, _ptr(c), len(base), _ptr(b), n,
                              _ptr(rho), _ptr(sigma), _ptr(tau), *[_ptr(o) for o in out], _ptr(used)))
    return out[0], out[1], out[2], out[3], used.astype(bool)


def eval_functional_pol(codes, rho, sigma):
    """Spin-polarised eps, vrho (n x 2), vsigma (n x 3); rho (n x 2: a, b), sigma (n x 3: aa, ab, bb)."""
    cc = np.ascontiguousarray(codes, dtype=np.int32)
    rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1, 2)
    sigma = np.ascontiguousarray(sigma, dtype=np.float64).reshape(-1, 3)
    n = rho.shape[0]
    e, vr, vs = np.zeros(n), np.zeros((n, 2)), np.zeros((n, 3))
    _check(lib().ox_eval_functional_pol(len(cc), _ptr(cc), n, _ptr(rho), _ptr(sigma), _ptr(e), _ptr(vr), _ptr(vs)))
    return e, vr, vs


def eval_functional_fpol(codes, rho, sigma):
    """Spin-polarised f_xc: vsigma (n x 3), v2rho2 (n x 3), v2rhosigma (n x 6), v2sigma2 (n x 6) in libxc's
    XC_POLARIZED order; rho (n x 2), sigma (n x 3)."""
    cc = np.ascontiguousarray(codes, dtype=np.int32)
    rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1, 2)
    sigma = np.ascontiguousarray(sigma, dtype=np.float64).reshape(-1, 3)
    n = rho.shape[0]
    vs, frr, frs, fss = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 6)), np.zeros((n, 6))
    _check(lib().ox_eval_functional_fpol(len(cc), _ptr(cc), n, _ptr(rho), _ptr(sigma), _ptr(vs), _ptr(frr), _ptr(frs),
                                         _ptr(fss)))
    return vs, frr, frs, fss


def eval_functional_mpol(codes, rho, sigma, tau):
    """Spin-polarised r2SCAN: eps, vrho (n x 2), vsigma (n x 3), vtau (n x 2); rho, tau (n x 2), sigma (n x 3)."""
    cc = np.ascontiguousarray(codes, dtype=np.int32)
    rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1, 2)
    sigma = np.ascontiguousarray(sigma, dtype=np.float64).reshape(-1, 3)
    tau = np.ascontiguousarray(tau, dtype=np.float64).reshape(-1, 2)
    n = rho.shape[0]
    e, vr, vs, vt = np.zeros(n), np.zeros((n, 2)), np.zeros((n, 3)), np.zeros((n, 2))
    _check(lib().ox_eval_functional_mpol(len(cc), _ptr(cc), n, _ptr(rho), _ptr(sigma), _ptr(tau), _ptr(e), _ptr(vr),
                                         _ptr(vs), _ptr(vt)))
    return e, vr, vs, vt
