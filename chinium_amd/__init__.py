"""MI355X-native Kohn-Sham exchange-correlation path (Chinium's Grid + ExchangeCorrelation).

Python mirror of the reference's C++ interface over the C ABI in
include/chinium_xc.h (libchinium_xc.so, built in-tree for gfx950):

    Grid(mol, "SG-1", nthreads, output)   src/Grid/Grid.h:90-116
    grid.setType(1); grid.getAO(0, 1)     src/Grid/Grid.cpp:14-33
    grid.getDensity([D])                  src/Grid/Grid.cpp:35-49
    xc.Read("PBE", output); xc.Evaluate("ev", grid)   src/ExchangeCorrelation.cpp:18-248
    grid.getEnergy(); grid.getFock()      src/Grid/Grid.cpp:112-171

The call order per SCF iteration is the reference's (Restricted/SP.cpp:52-55):
getDensity -> Evaluate("ev") -> getEnergy -> getFock. There is no CPU
fallback: without the built library every call raises.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libchinium_xc.so")
DATA_ROOT = os.path.join(PKG_DIR, "data")
ANGSTROM_TO_BOHR = 1.8897259886  # src/Macro/Unit.h:1

CHX_PREC_FP64 = 0
CHX_PREC_MIXED = 1  # split-BF16 MFMA contractions, FP32 accumulation
# status codes of the C ABI (chinium_xc.h)
CHX_ERR_ARG, CHX_ERR_FILE, CHX_ERR_STATE, CHX_ERR_DEVICE, CHX_ERR_UNSUPPORTED = -1, -2, -3, -4, -5

_vp = ctypes.c_void_p
_lib = None


class ChxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    """Load libchinium_xc.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libchinium_xc.so is not built; run __graft_entry__.build() "
                          "(make -C chinium_amd/csrc)")
    L = ctypes.CDLL(LIB_PATH)
    i, d, cp, l = ctypes.c_int, ctypes.c_double, ctypes.c_char_p, ctypes.c_long
    P = ctypes.POINTER
    sig = {
        "chx_last_error": (cp, []),
        "chx_device_count": (i, [P(i)]),
        "chx_mol_create": (i, [i, _vp, _vp, P(_vp)]),
        "chx_mol_set_basis_file": (i, [_vp, cp, cp]),
        "chx_mol_add_shell": (i, [_vp, i, i, i, _vp, _vp]),
        "chx_mol_nbasis": (i, [_vp, P(i)]),
        "chx_mol_nshells": (i, [_vp, P(i)]),
        "chx_mol_get_shells": (i, [_vp, _vp, _vp, _vp, _vp, _vp]),
        "chx_mol_overlap": (i, [_vp, _vp]),
        "chx_mol_int1e": (i, [_vp, _vp, _vp, _vp]),
        "chx_mol_eri": (i, [_vp, _vp]),
        "chx_jfit_create": (i, [_vp, cp, cp, P(_vp)]),
        "chx_jfit_info": (i, [_vp, P(i), P(l), P(l), P(i)]),
        "chx_jfit_int1e": (i, [_vp, _vp, _vp, _vp]),
        "chx_jfit_metric": (i, [_vp, _vp]),
        "chx_jfit_gamma": (i, [_vp, _vp, _vp]),
        "chx_jfit_coulomb": (i, [_vp, _vp, _vp]),
        "chx_jfit_destroy": (None, [_vp]),
        "chx_mol_nuclear_repulsion": (i, [_vp, P(ctypes.c_double)]),
        "chx_mol_destroy": (None, [_vp]),
        "chx_grid_create": (i, [_vp, cp, cp, i, i, P(_vp)]),
        "chx_grid_copy": (i, [_vp, P(_vp)]),
        "chx_grid_destroy": (None, [_vp]),
        "chx_grid_num_threads": (i, [_vp, P(i)]),
        "chx_grid_num_points": (i, [_vp, P(l)]),
        "chx_grid_get_points": (i, [_vp, _vp]),
        "chx_grid_set_type": (i, [_vp, i]),
        "chx_grid_get_ao": (i, [_vp, i, i]),
        "chx_grid_get_density": (i, [_vp, _vp, i, i]),
        "chx_grid_get_density_device": (i, [_vp, _vp, i, i]),
        "chx_grid_num_electrons": (i, [_vp, P(d)]),
        "chx_grid_energy": (i, [_vp, P(d)]),
        "chx_grid_fock": (i, [_vp, _vp, i, i]),
        "chx_grid_fock_device": (i, [_vp, _vp, i, i]),
        "chx_grid_set_vrho": (i, [_vp, _vp, l]),
        "chx_grid_stream": (i, [_vp, P(_vp)]),
        "chx_grid_set_precision": (i, [_vp, i]),
        "chx_grid_set_core_radius": (i, [_vp, d]),
        "chx_grid_core_batches": (i, [_vp, P(l)]),
        "chx_grid_set_screening": (i, [_vp, d]),
        "chx_grid_stats": (i, [_vp, P(l), P(l), P(d), P(d)]),
        "chx_grid_profile": (i, [_vp, i]),
        "chx_grid_set_ao_cache": (i, [_vp, i]),
        "chx_grid_ao_cached": (i, [_vp, P(i)]),
        "chx_grid_kernel_times": (i, [_vp, i, _vp, _vp]),
        "chx_xc_read": (i, [cp, cp, i, i, P(_vp)]),
        "chx_xc_create": (i, [i, _vp, i, P(_vp)]),
        "chx_xc_destroy": (None, [_vp]),
        "chx_xc_info": (i, [_vp, P(i), P(d), P(i), _vp]),
        "chx_xc_evaluate": (i, [_vp, _vp, cp]),
        "chx_selftest": (i, [P(d)]),
        "chx_lebedev_rule": (i, [i, _vp]),
        "chx_xc_eval_points": (i, [_vp, l, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "chx_xc_eval_points_noalloc": (i, [_vp, l, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "chx_grid_get_density_skeleton": (i, [_vp, _vp, i, i]),
        "chx_grid_autotune": (i, [_vp, _vp, _vp, _vp, i, P(i), _vp]),
        "chx_grid_autotune_tiles": (i, [_vp, _vp, _vp, _vp, i, P(i), P(i), _vp]),
        "chx_grid_set_vxc_shape": (i, [_vp, i]),
        "chx_grid_set_vxc_chunk": (i, [_vp, i]),
        "chx_grid_core_flops": (i, [_vp, P(d)]),
        "chx_grid_vxc_shape": (i, [_vp, P(i)]),
        "chx_pack_vxc_wide_from": (i, [_vp, P(i)]),
        "chx_grid_energy_grad": (i, [_vp, _vp]),
        "chx_grid_get_density_u": (i, [_vp, _vp, i, i, i]),
        "chx_grid_fock_u": (i, [_vp, _vp, i, i, i]),
        "chx_grid_fock_s": (i, [_vp, _vp, i, i]),
        "chx_grid_fock_skeleton": (i, [_vp, _vp, i, i]),
        "chx_grid_get_density_skeleton2": (i, [_vp, _vp, i, i]),
        "chx_grid_energy_hess": (i, [_vp, _vp]),
        "chx_xc_eval_points_f": (i, [_vp, l, _vp, _vp, _vp, _vp, _vp]),
        "chx_pack_create": (i, [_vp, i, P(_vp)]),
        "chx_pack_xc": (i, [_vp, _vp, _vp, _vp, _vp, _vp]),
        "chx_pack_set_graph": (i, [_vp, i]),
        "chx_pack_set_lanes": (i, [_vp, i]),
        "chx_pack_lanes": (i, [_vp, P(i)]),
        "chx_pack_set_form": (i, [_vp, i]),
        "chx_pack_form": (i, [_vp, P(i)]),
        "chx_pack_autotune": (i, [_vp, _vp, _vp, _vp, i, P(i), _vp]),
        "chx_pack_stream": (i, [_vp, P(_vp)]),
        "chx_pack_destroy": (None, [_vp]),
        "chx_mlp_create": (i, [ctypes.c_ulonglong, P(_vp)]),
        "chx_mlp_destroy": (None, [_vp]),
        "chx_mlp_get_params": (i, [_vp, _vp]),
        "chx_mlp_set_params": (i, [_vp, _vp]),
        "chx_mlp_set_box": (i, [_vp, _vp, _vp, d]),
        "chx_mlp_set_base": (i, [_vp, i, _vp]),
        "chx_mlp_get_base": (i, [_vp, P(i), _vp]),
        "chx_mlp_get_box": (i, [_vp, _vp, _vp, P(d)]),
        "chx_mlp_features": (i, [_vp, l, _vp, _vp, _vp, _vp, _vp]),
        "chx_mlp_set_data": (i, [_vp, l, _vp, _vp, _vp]),
        "chx_mlp_train": (i, [_vp, i, l, d, d, _vp]),
        "chx_mlp_loss_grad": (i, [_vp, l, _vp, _vp, _vp, _vp, P(d)]),
        "chx_mlp_forward": (i, [_vp, l, _vp, i, _vp, _vp]),
        "chx_mlp_eval_points": (i, [_vp, _vp, l, _vp, _vp, _vp, _vp, _vp, _vp, _vp, P(l)]),
        "chx_xc_set_surrogate": (i, [_vp, _vp]),
        "chx_grid_point_density": (i, [_vp, _vp, _vp, _vp, _vp, P(l)]),
        "chx_grid_fallback_count": (i, [_vp, P(l)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise ChxError(rc, lib().chx_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data


def device_count():
    n = ctypes.c_int()
    _check(lib().chx_device_count(ctypes.byref(n)))
    return n.value


def selftest():
    """Max deviation of one FP64 MFMA tile from the scalar product (0 = layout holds)."""
    e = ctypes.c_double()
    _check(lib().chx_selftest(ctypes.byref(e)))
    return e.value


def lebedev_rule(n):
    """The library's Lebedev-Laikov rule with n points: rows (x, y, z, w), sum w = 1."""
    out = np.zeros((n, 4))
    _check(lib().chx_lebedev_rule(int(n), _ptr(out)))
    return out


class Molecule:
    """Atoms (Z, bohr) and a normalised basis: the Mwfn fields the XC path reads."""

    def __init__(self, Z, xyz_bohr, basis=None, data_root=DATA_ROOT):
        L = lib()
        self.Z = np.ascontiguousarray(Z, dtype=np.int32)
        self.xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
        h = _vp()
        _check(L.chx_mol_create(len(self.Z), _ptr(self.Z), _ptr(self.xyz), ctypes.byref(h)))
        self.h = h
        if basis is not None:
            _check(L.chx_mol_set_basis_file(self.h, data_root.encode(), basis.encode()))

    def add_shell(self, center, type_, exps, norm_coefs):
        e = np.ascontiguousarray(exps, dtype=np.float64)
        c = np.ascontiguousarray(norm_coefs, dtype=np.float64)
        _check(lib().chx_mol_add_shell(self.h, int(center), int(type_), len(e), _ptr(e), _ptr(c)))

    @property
    def nbasis(self):
        n = ctypes.c_int()
        _check(lib().chx_mol_nbasis(self.h, ctypes.byref(n)))
        return n.value

    def shells(self):
        L = lib()
        ns = ctypes.c_int()
        _check(L.chx_mol_nshells(self.h, ctypes.byref(ns)))
        ns = ns.value
        t = np.zeros(ns, np.int32); npr = np.zeros(ns, np.int32); c = np.zeros(ns, np.int32)
        _check(L.chx_mol_get_shells(self.h, _ptr(t), _ptr(npr), _ptr(c), None, None))
        e = np.zeros(int(npr.sum())); k = np.zeros(int(npr.sum()))
        _check(L.chx_mol_get_shells(self.h, None, None, None, _ptr(e), _ptr(k)))
        return t, npr, c, e, k

    def overlap(self):
        n = self.nbasis
        S = np.zeros((n, n), order="F")
        _check(lib().chx_mol_overlap(self.h, _ptr(S)))
        return np.array(S)

    def int1e(self):
        """(S, T, V): overlap, kinetic and nuclear-attraction matrices (chx_mol_int1e)."""
        n = self.nbasis
        S, T, V = (np.zeros((n, n), order="F") for _ in range(3))
        _check(lib().chx_mol_int1e(self.h, _ptr(S), _ptr(T), _ptr(V)))
        return np.array(S), np.array(T), np.array(V)

    def eri(self):
        """Electron-repulsion integrals as an n^4 array, eri[mu, nu, la, si] = (mu nu|la si)."""
        n = self.nbasis
        g = np.zeros(n ** 4)
        _check(lib().chx_mol_eri(self.h, _ptr(g)))
        return g.reshape((n, n, n, n), order="F")

    def nuclear_repulsion(self):
        e = ctypes.c_double()
        _check(lib().chx_mol_nuclear_repulsion(self.h, ctypes.byref(e)))
        return e.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.chx_mol_destroy(self.h)
            self.h = None


class Grid:
    """Device-resident molecular grid; mirrors class Grid (src/Grid/Grid.h:90-116)."""

    def __init__(self, mol=None, grid="", nthreads=1, output=0, data_root=DATA_ROOT, _handle=None):
        self.h = None
        if _handle is not None:
            self.h = _handle
        else:
            h = _vp()
            _check(lib().chx_grid_create(mol.h, grid.encode(), data_root.encode(), int(nthreads), int(output),
                                         ctypes.byref(h)))
            self.h = h
        n = ctypes.c_long()
        _check(lib().chx_grid_num_points(self.h, ctypes.byref(n)))
        self.npoints = n.value
        self.nbasis = mol.nbasis if mol is not None else None
        self.natoms = len(mol.Z) if mol is not None else None

    def copy(self):
        h = _vp()
        _check(lib().chx_grid_copy(self.h, ctypes.byref(h)))
        g = Grid(_handle=h)
        g.nbasis = self.nbasis
        g.natoms = self.natoms
        return g

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.chx_grid_destroy(self.h)
            self.h = None

    def getNumThreads(self):
        n = ctypes.c_int()
        _check(lib().chx_grid_num_threads(self.h, ctypes.byref(n)))
        return n.value

    def points(self):
        P = np.zeros((self.npoints, 4))
        _check(lib().chx_grid_get_points(self.h, _ptr(P)))
        return P

    def setType(self, t):
        _check(lib().chx_grid_set_type(self.h, int(t)))

    def setScreening(self, thr):
        _check(lib().chx_grid_set_screening(self.h, float(thr)))

    def setPrecision(self, p):
        _check(lib().chx_grid_set_precision(self.h, int(p)))

    def setCoreRadius(self, r):
        """Mixed form: batches within r bohr of a heavy nucleus take three-term operands (0: off)."""
        _check(lib().chx_grid_set_core_radius(self.h, float(r)))

    def coreFlops(self):
        v = ctypes.c_double()
        _check(lib().chx_grid_core_flops(self.h, ctypes.byref(v)))
        return v.value

    def coreBatches(self):
        n = ctypes.c_long()
        _check(lib().chx_grid_core_batches(self.h, ctypes.byref(n)))
        return n.value

    def getAO(self, derivative=0, output=0):
        _check(lib().chx_grid_get_ao(self.h, int(derivative), int(output)))

    def _stack(self, Ds):
        """One or two nbasis x nbasis matrices, back to back in column-major order (the C ABI's layout).
        Wrong counts or shapes raise here, before the library copies nspin * n * n doubles."""
        Ds = list(Ds)
        if len(Ds) not in (1, 2):
            raise ChxError(CHX_ERR_ARG, "Wrong number of spin types!")
        n = self.nbasis
        for D in Ds:
            if np.shape(D) != (n, n):
                raise ChxError(CHX_ERR_ARG, "density matrix size does not match the basis")
        if len(Ds) == 1:  # no copy when D is already a column-major float64 array
            return np.asfortranarray(Ds[0], dtype=np.float64), 1
        return np.concatenate([np.asfortranarray(D, dtype=np.float64).ravel(order="F") for D in Ds]), len(Ds)

    def getDensity(self, Ds):
        buf, ns = self._stack(Ds)
        _check(lib().chx_grid_get_density(self.h, _ptr(buf), ns, self.nbasis))
        self._nspin = ns

    def getDensitySkeleton(self, Ds):
        """SubGrid::getDensitySkeleton (GridDensity.cpp:79-130), closed shell."""
        buf, ns = self._stack(Ds)
        _check(lib().chx_grid_get_density_skeleton(self.h, _ptr(buf), ns, self.nbasis))

    def getEnergyGrad(self):
        """Grid::getEnergyGrad: XC nuclear gradient (natoms x 3), grid held fixed."""
        G = np.zeros(3 * self.natoms)
        _check(lib().chx_grid_energy_grad(self.h, _ptr(G)))
        return G.reshape(-1, 3)

    def autotune(self, xc, D_ptr, V_ptr, reps=3):
        """Times each Vxc work-item length, keeps the fastest; returns (chunk, {chunk: ms})."""
        c = ctypes.c_int()
        ms = (ctypes.c_float * 4)()
        _check(lib().chx_grid_autotune(self.h, xc.h, D_ptr, V_ptr, int(reps), ctypes.byref(c), ms))
        return c.value, dict(zip((8, 16, 32, 128), list(ms)))

    def autotuneTiles(self, xc, D_ptr, V_ptr, reps=3):
        """Mixed form: times both Vxc block shapes (0: 128-AO super-tiles, 1: 256-AO union groups) at
        each work-item length, keeps the fastest; returns (shape, chunk, {(shape, chunk): ms})."""
        sh, c = ctypes.c_int(), ctypes.c_int()
        ms = (ctypes.c_float * 8)()
        _check(lib().chx_grid_autotune_tiles(self.h, xc.h, D_ptr, V_ptr, int(reps), ctypes.byref(sh),
                                             ctypes.byref(c), ms))
        keys = [(a, b) for a in (0, 1) for b in (8, 16, 32, 128)]
        return sh.value, c.value, dict(zip(keys, list(ms)))

    def setVxcChunk(self, chunk):
        _check(lib().chx_grid_set_vxc_chunk(self.h, int(chunk)))

    def setVxcShape(self, shape):
        _check(lib().chx_grid_set_vxc_shape(self.h, int(shape)))

    def vxcShape(self):
        v = ctypes.c_int()
        _check(lib().chx_grid_vxc_shape(self.h, ctypes.byref(v)))
        return v.value

    def getDensityU(self, Ds):
        """Grid::getDensityU (src/Grid/Grid.cpp:63-78): Ds[spin][mat], closed shell."""
        Ds = [list(x) for x in Ds]
        n = self.nbasis
        if len(Ds) not in (1, 2) or any(len(x) != len(Ds[0]) for x in Ds):
            raise ChxError(CHX_ERR_ARG, "getDensityU takes [spin][mat] lists with equal matrix counts per spin")
        if any(np.shape(D) != (n, n) for x in Ds for D in x):
            raise ChxError(CHX_ERR_ARG, "density matrix size does not match the basis")
        buf = np.concatenate([np.asfortranarray(D, dtype=np.float64).ravel(order="F") for x in Ds for D in x])
        self._nmats_u = len(Ds[0])
        _check(lib().chx_grid_get_density_u(self.h, _ptr(buf), len(Ds[0]), len(Ds), n))

    def getDensitySkeleton2(self, Ds):
        """Grid::getDensitySkeleton2 (src/Grid/Grid.cpp:96-111): D for the XC Hessian (closed shell)."""
        buf, ns = self._stack(Ds)
        _check(lib().chx_grid_get_density_skeleton2(self.h, _ptr(buf), ns, self.nbasis))

    def getEnergyHess(self):
        """Grid::getEnergyHess (src/Grid/Grid.cpp:139-154): the skeleton XC energy Hessian,
        (3 natoms) x (3 natoms), [3 A + t][3 B + s]."""
        m = 3 * self.natoms
        H = np.zeros(m * m)
        _check(lib().chx_grid_energy_hess(self.h, _ptr(H)))
        return H.reshape(m, m)

    def getFockSkeleton(self):
        """Grid::getFockSkeleton (src/Grid/Grid.cpp:173-191): the XC Fock matrix's derivative along each
        nuclear coordinate with the potential held fixed, [spin][3 atom + t] (closed shell)."""
        n, m = self.nbasis, 3 * self.natoms
        F = np.zeros(m * n * n)
        _check(lib().chx_grid_fock_skeleton(self.h, _ptr(F), 1, n))
        return [[F[j * n * n:(j + 1) * n * n].reshape((n, n), order="F").copy() for j in range(m)]]

    def getFockU(self, nspin=1, d_t="u"):
        """Grid::getFockU<u_t> (src/Grid/Grid.cpp:193-216): response matrices [spin][mat]; d_t="s":
        Grid::getFockU<s_t>, one matrix per nuclear coordinate [3 atom + t] from the skeleton density."""
        n = self.nbasis
        if d_t == "s":
            m = 3 * self.natoms
            F = np.zeros(nspin * m * n * n)
            _check(lib().chx_grid_fock_s(self.h, _ptr(F), nspin, n))
        else:
            m = getattr(self, "_nmats_u", 0)
            F = np.zeros(nspin * m * n * n)
            _check(lib().chx_grid_fock_u(self.h, _ptr(F), m, nspin, n))
        return [[F[(k * m + j) * n * n:(k * m + j + 1) * n * n].reshape((n, n), order="F").copy() for j in range(m)]
                for k in range(nspin)]

    def pointDensity(self):
        """rho, sigma, tau, weight per slot of the last getDensity (padding slots have weight 0)."""
        n = ctypes.c_long()
        _check(lib().chx_grid_point_density(self.h, None, None, None, None, ctypes.byref(n)))
        out = [np.zeros(n.value) for _ in range(4)]
        _check(lib().chx_grid_point_density(self.h, *[_ptr(o) for o in out], ctypes.byref(n)))
        return tuple(out)

    def fallbackCount(self):
        """Points of the last Evaluate that took the surrogate's analytic fallback."""
        n = ctypes.c_long()
        _check(lib().chx_grid_fallback_count(self.h, ctypes.byref(n)))
        return n.value

    def getDensityDevice(self, ptr, nspin, nbasis):
        _check(lib().chx_grid_get_density_device(self.h, ctypes.c_void_p(ptr), int(nspin), int(nbasis)))
        self._nspin = int(nspin)

    def getNumElectrons(self):
        v = ctypes.c_double()
        _check(lib().chx_grid_num_electrons(self.h, ctypes.byref(v)))
        return v.value

    def getEnergy(self):
        v = ctypes.c_double()
        _check(lib().chx_grid_energy(self.h, ctypes.byref(v)))
        return v.value

    def getFock(self, nspin=None):
        """Grid::getFock (src/Grid/Grid.cpp:156-171): one matrix per spin of the last getDensity."""
        if nspin is None:
            nspin = getattr(self, "_nspin", 1)
        n = self.nbasis
        F = np.empty(nspin * n * n)
        _check(lib().chx_grid_fock(self.h, _ptr(F), nspin, n))
        return [F[k * n * n:(k + 1) * n * n].reshape((n, n), order="F") for k in range(nspin)]

    def getFockDevice(self, ptr, nspin, nbasis):
        _check(lib().chx_grid_fock_device(self.h, ctypes.c_void_p(ptr), int(nspin), int(nbasis)))

    def setVrho(self, vrho):
        v = np.ascontiguousarray(vrho, dtype=np.float64)
        _check(lib().chx_grid_set_vrho(self.h, _ptr(v), len(v)))

    def stream(self):
        s = _vp()
        _check(lib().chx_grid_stream(self.h, ctypes.byref(s)))
        return s.value

    STAGES = ("permute_d", "density", "functional", "vxc", "finalize_v", "make_b", "density_core")

    def setAOCache(self, mode):
        """-1 automatic (default), 0 recompute AO values in the kernels, 1 cache them in HBM."""
        _check(lib().chx_grid_set_ao_cache(self.h, int(mode)))

    def aoCached(self):
        c = ctypes.c_int()
        _check(lib().chx_grid_ao_cached(self.h, ctypes.byref(c)))
        return bool(c.value)

    def profile(self, enable=True):
        _check(lib().chx_grid_profile(self.h, int(bool(enable))))

    def kernel_times(self, reset=True):
        """Accumulated HIP-event milliseconds and launch counts per stage."""
        ms = np.zeros(7)
        n = np.zeros(7, dtype=np.int64)
        _check(lib().chx_grid_kernel_times(self.h, int(bool(reset)), _ptr(ms), _ptr(n)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(self.STAGES)}

    def vxcKernel(self):
        """Name of the kernel the mixed form runs for the Vxc contraction (the block shape in use)."""
        return "k_vxc_mx_wide" if self.vxcShape() == 1 else "k_vxc_mx"

    def stats(self):
        nb, npnt, mt, fl = ctypes.c_long(), ctypes.c_long(), ctypes.c_double(), ctypes.c_double()
        _check(lib().chx_grid_stats(self.h, ctypes.byref(nb), ctypes.byref(npnt), ctypes.byref(mt), ctypes.byref(fl)))
        return {"batches": nb.value, "points": npnt.value, "mean_tiles": mt.value, "flops_per_build": fl.value}


class CoulombFit:
    """Density-fitted Coulomb matrix and one-electron integrals on the device (chx_jfit_*), in an
    auxiliary basis from $CHINIUM_PATH/BasisSets (default def2-svp-j): gamma = (P|mn) D_mn,
    (P|Q) c = gamma, J_mn = (mn|P) c_P. Matrices are n x n in the reference AO order."""

    def __init__(self, mol, aux_basis="def2-svp-j", data_root=DATA_ROOT):
        self.mol = mol  # keeps the molecule alive
        self.h = None
        h = _vp()
        _check(lib().chx_jfit_create(mol.h, data_root.encode(), aux_basis.encode(), ctypes.byref(h)))
        self.h = h
        self.n = mol.nbasis

    def info(self):
        """dict(naux, pairs kept, pairs before screening, auxiliary primitives)."""
        na, npr, nraw, nap = ctypes.c_int(), ctypes.c_long(), ctypes.c_long(), ctypes.c_int()
        _check(lib().chx_jfit_info(self.h, ctypes.byref(na), ctypes.byref(npr), ctypes.byref(nraw), ctypes.byref(nap)))
        return dict(naux=na.value, pairs=npr.value, pairs_unscreened=nraw.value, aux_prims=nap.value)

    @property
    def naux(self):
        return self.info()["naux"]

    def int1e(self):
        """(S, T, V), computed once and kept."""
        if getattr(self, "_int1e", None) is None:
            n = self.n
            S, T, V = (np.zeros((n, n), order="F") for _ in range(3))
            _check(lib().chx_jfit_int1e(self.h, _ptr(S), _ptr(T), _ptr(V)))
            self._int1e = (np.array(S), np.array(T), np.array(V))
        return self._int1e

    def metric(self):
        """(P|Q), computed once and kept."""
        if getattr(self, "_metric", None) is None:
            na = self.naux
            M = np.zeros((na, na), order="F")
            _check(lib().chx_jfit_metric(self.h, _ptr(M)))
            self._metric = np.array(M)
        return self._metric

    def gamma(self, D):
        Df = np.asfortranarray(D, dtype=np.float64)
        if Df.shape != (self.n, self.n):
            raise ValueError("D must be %d x %d" % (self.n, self.n))
        g = np.zeros(self.naux)
        _check(lib().chx_jfit_gamma(self.h, _ptr(Df), _ptr(g)))
        return g

    def coulomb(self, c):
        c = np.ascontiguousarray(c, dtype=np.float64)
        if c.shape != (self.naux,):
            raise ValueError("c must have %d entries" % self.naux)
        J = np.zeros((self.n, self.n), order="F")
        _check(lib().chx_jfit_coulomb(self.h, _ptr(c), _ptr(J)))
        return np.array(J)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.chx_jfit_destroy(self.h)
            self.h = None


class ExchangeCorrelation:
    """Mirrors class ExchangeCorrelation (src/ExchangeCorrelation.h:8-17)."""

    def __init__(self):
        self.h = None
        self.Codes = []
        self.EXX = 1.0
        self.Spin = 1
        self.Family = ""

    def _load_info(self):
        L = lib()
        fam, exx, nc = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
        _check(L.chx_xc_info(self.h, ctypes.byref(fam), ctypes.byref(exx), ctypes.byref(nc), None))
        codes = np.zeros(nc.value, np.int32)
        _check(L.chx_xc_info(self.h, None, None, None, _ptr(codes)))
        self.Codes = [int(c) for c in codes]
        self.EXX = exx.value
        self.Family = ["LDA", "GGA", "mGGA"][fam.value]
        self.family = fam.value

    def Read(self, df, output=0, data_root=DATA_ROOT):
        h = _vp()
        _check(lib().chx_xc_read(data_root.encode(), df.encode(), int(self.Spin), int(output), ctypes.byref(h)))
        self._set(h)

    def from_codes(self, codes):
        c = np.ascontiguousarray(codes, dtype=np.int32)
        h = _vp()
        _check(lib().chx_xc_create(len(c), _ptr(c), int(self.Spin), ctypes.byref(h)))
        self._set(h)
        return self

    def _set(self, h):
        if self.h is not None:
            lib().chx_xc_destroy(self.h)
        self.h = h
        self._load_info()

    def eval_points(self, rho, sigma=None, tau=None, noalloc=False):
        """exc, v_rho, v_sigma, v_tau at the given points, evaluated on the device. Spin = 2:
        rho (n x 2), sigma (n x 3), tau (n x 2, meta-GGA) in libxc's polarised layout; returns exc,
        v_rho (n x 2), v_sigma (n x 3) and, for a meta-GGA functional, v_tau (n x 2).
        noalloc: through chx_xc_eval_points_noalloc (fixed device staging, at most CHX_NOALLOC_MAX_POINTS)."""
        f = lib().chx_xc_eval_points_noalloc if noalloc else lib().chx_xc_eval_points
        if self.Spin == 2:
            rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1, 2)
            n = rho.shape[0]
            sg = None if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64).reshape(-1, 3)
            ta = None if tau is None else np.ascontiguousarray(tau, dtype=np.float64).reshape(-1, 2)
            e, vr, vs, vt = np.zeros(n), np.zeros((n, 2)), np.zeros((n, 3)), np.zeros((n, 2))
            _check(f(self.h, n, _ptr(rho), _ptr(sg), _ptr(ta), _ptr(e), _ptr(vr), _ptr(vs), _ptr(vt)))
            return (e, vr, vs, vt) if self.family >= 2 else (e, vr, vs)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        n = len(rho)
        sg = None if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64)
        ta = None if tau is None else np.ascontiguousarray(tau, dtype=np.float64)
        e, vr, vs, vt = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
        _check(f(self.h, n, _ptr(rho), _ptr(sg), _ptr(ta), _ptr(e), _ptr(vr), _ptr(vs), _ptr(vt)))
        return e, vr, vs, vt

    def eval_points_f(self, rho, sigma=None):
        """v2rho2, v2rhosigma, v2sigma2 at the given points (LDA/GGA), on the device. Spin = 2: rho (n x 2),
        sigma (n x 3) in libxc's polarised layout; returns (n x 3), (n x 6), (n x 6)."""
        pol = self.Spin == 2
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape((-1, 2) if pol else (-1,))
        n = rho.shape[0]
        sg = None if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64).reshape((-1, 3) if pol else (-1,))
        shp = ((n, 3), (n, 6), (n, 6)) if pol else ((n,), (n,), (n,))
        f0, f1, f2 = (np.zeros(x) for x in shp)
        _check(lib().chx_xc_eval_points_f(self.h, n, _ptr(rho), _ptr(sg), _ptr(f0), _ptr(f1), _ptr(f2)))
        return f0, f1, f2

    def Evaluate(self, order, grid):
        _check(lib().chx_xc_evaluate(self.h, grid.h, order.encode()))

    def setSurrogate(self, surrogate):
        """Attach a trained Surrogate (meta-GGA functionals; None detaches):
        Evaluate then runs the network, with the analytic fallback."""
        _check(lib().chx_xc_set_surrogate(self.h, None if surrogate is None else surrogate.h))
        self._surrogate = surrogate  # keeps the network alive while attached

    def __bool__(self):
        return len(self.Codes) > 0

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.chx_xc_destroy(self.h)
            self.h = None


class Pack:
    """Many molecules per XC build (chx_pack_*): every grid's kernels on
    staggered lanes (streams), replayed as one hipGraph. Grids must be in the
    mixed-precision form with getAO done; they stay referenced by the pack."""

    def __init__(self, grids):
        self.grids = list(grids)
        n = len(self.grids)
        arr = (ctypes.c_void_p * n)(*[g.h for g in self.grids])
        h = ctypes.c_void_p()
        _check(lib().chx_pack_create(arr, n, ctypes.byref(h)))
        self.h = h

    def _ptrs(self, ptrs):
        assert len(ptrs) == len(self.grids)
        return (ctypes.c_void_p * len(ptrs))(*ptrs)

    def xc(self, xc, D_ptrs, V_ptrs):
        """One XC build of every molecule: D_ptrs / V_ptrs are device pointers
        (column-major nbasis^2 each, reference order). Returns (E_xc, N_e) arrays."""
        n = len(self.grids)
        E, N = np.zeros(n), np.zeros(n)
        _check(lib().chx_pack_xc(self.h, xc.h, self._ptrs(D_ptrs), self._ptrs(V_ptrs), _ptr(E), _ptr(N)))
        return E, N

    def set_graph(self, on):
        _check(lib().chx_pack_set_graph(self.h, 1 if on else 0))

    def set_lanes(self, lanes):
        """Splits the molecules into `lanes` groups on streams of their own, staggered by a stage."""
        _check(lib().chx_pack_set_lanes(self.h, int(lanes)))

    def lanes(self):
        v = ctypes.c_int()
        _check(lib().chx_pack_lanes(self.h, ctypes.byref(v)))
        return v.value

    FORMS = {0: "lanes", 1: "persistent", 2: "lane stages"}

    def set_form(self, form):
        """0: per-molecule kernels on lanes; 1: one persistent launch per stage over every molecule's
        items; 2: one launch per stage per lane (1 and 2: LDA/GGA grids)."""
        _check(lib().chx_pack_set_form(self.h, int(form)))

    def form(self):
        v = ctypes.c_int()
        _check(lib().chx_pack_form(self.h, ctypes.byref(v)))
        return v.value

    def autotune(self, xc, D_ptrs, V_ptrs, reps=3):
        """Times each Vxc work-item length, keeps the fastest, then picks the lane count;
        returns (chunk, {chunk: ms})."""
        c = ctypes.c_int()
        ms = (ctypes.c_float * 4)()
        _check(lib().chx_pack_autotune(self.h, xc.h, self._ptrs(D_ptrs), self._ptrs(V_ptrs), int(reps),
                                       ctypes.byref(c), ms))
        return c.value, dict(zip((8, 16, 32, 128), list(ms)))

    def vxcWideFrom(self):
        """The autotuned Vxc shape threshold: 256-AO union groups for npad >= it (-1: none)."""
        v = ctypes.c_int()
        _check(lib().chx_pack_vxc_wide_from(self.h, ctypes.byref(v)))
        return v.value

    def stream(self):
        p = ctypes.c_void_p()
        _check(lib().chx_pack_stream(self.h, ctypes.byref(p)))
        return p.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.chx_pack_destroy(self.h)
            self.h = None


def xc_build(grid, xc, D):
    """One XC build in the reference's order; returns (E_xc, V)."""
    grid.getDensity([D])
    xc.Evaluate("ev", grid)
    E = grid.getEnergy()
    V = grid.getFock()[0]
    return E, V


from .surrogate import Surrogate  # noqa: E402  (uses lib/_check above)
