"""Learned functional surrogate over the C ABI (chx_mlp_*, include/chinium_xc.h).

A 3 -> 64 -> 64 -> 64 -> 1 tanh network on the spin-restricted descriptors
t = (ln rho, s/(1+s), alpha/(1+alpha)) predicts F = e_xc / e_x^unif, or, with
an analytic base (e.g. LDA_BASE: Slater + VWN5), the correction
F - F_base that the base leaves; attached to a meta-GGA ExchangeCorrelation
(setSurrogate) it stands in for the analytic functional in Evaluate, with the
analytic functional as the fallback outside the training box and below
rho_min. Training runs on the GPU (forward, backward and Adam kernels); the
samples come from the analytic functional on the densities of real grids. No
reference counterpart: the reference calls libxc
(src/ExchangeCorrelation.cpp:86-163).
"""
import ctypes

import numpy as np

from . import _check, _ptr, lib

MLP_PARAMS = 8641
CX_UNIF = -0.7385587663820224  # e_x^unif = CX_UNIF rho^(4/3)
LDA_BASE = (1, 7)  # Slater exchange + VWN5 correlation


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Surrogate:
    def __init__(self, seed=1, base=()):
        h = ctypes.c_void_p()
        _check(lib().chx_mlp_create(int(seed), ctypes.byref(h)))
        self.h = h
        self.set_base(base)

    def set_base(self, codes):
        """Analytic base (0 to 2 LDA / GGA codes): the network then predicts F - F_base."""
        c = np.ascontiguousarray(list(codes) or [0], dtype=np.int32)
        _check(lib().chx_mlp_set_base(self.h, len(codes), _ptr(c)))
        self.base = tuple(int(x) for x in codes)

    def __del__(self):
        if getattr(self, "h", None):
            lib().chx_mlp_destroy(self.h)
            self.h = None

    # parameters and box
    @property
    def params(self):
        P = np.zeros(MLP_PARAMS, np.float32)
        _check(lib().chx_mlp_get_params(self.h, _ptr(P)))
        return P

    @params.setter
    def params(self, P):
        P = _f32(P)
        assert P.size == MLP_PARAMS
        _check(lib().chx_mlp_set_params(self.h, _ptr(P)))

    def set_box(self, lo, hi, rho_min=1e-8):
        _check(lib().chx_mlp_set_box(self.h, _ptr(_f32(lo)), _ptr(_f32(hi)), float(rho_min)))

    @property
    def box(self):
        lo, hi, rm = np.zeros(3, np.float32), np.zeros(3, np.float32), ctypes.c_double()
        _check(lib().chx_mlp_get_box(self.h, _ptr(lo), _ptr(hi), ctypes.byref(rm)))
        return lo, hi, rm.value

    def normalise(self, t):
        lo, hi, _ = self.box
        lo, hi = lo.astype(np.float64), hi.astype(np.float64)
        return (np.asarray(t, np.float64) - 0.5 * (lo + hi)) * (2.0 / (hi - lo))

    # samples
    @staticmethod
    def samples_from_grid(grid, rho_min=1e-8):
        """(rho, sigma, tau, w) of the grid's live points after getDensity (meta-GGA type)."""
        rho, sigma, tau, w = grid.pointDensity()
        keep = (w != 0) & (rho >= rho_min)
        return rho[keep], sigma[keep], tau[keep], w[keep]

    @staticmethod
    def features(xc, rho, sigma, tau):
        """Descriptors t (n x 3) and targets F = e_xc / e_x^unif of xc's analytic functional (GPU)."""
        rho, sigma, tau = _f64(rho), _f64(sigma), _f64(tau)
        n = rho.size
        t, F = np.zeros((n, 3), np.float32), np.zeros(n, np.float32)
        _check(lib().chx_mlp_features(xc.h, n, _ptr(rho), _ptr(sigma), _ptr(tau), _ptr(t), _ptr(F)))
        return t, F

    def targets(self, xc, rho, sigma, tau):
        """Descriptors t and the network's targets: F of xc, less F of the base when one is set (GPU, FP64
        functionals, the difference taken in FP64 before rounding to FP32)."""
        from . import ExchangeCorrelation

        rho, sigma, tau = _f64(rho), _f64(sigma), _f64(tau)
        t, _ = self.features(xc, rho, sigma, tau)
        e = xc.eval_points(rho, sigma, tau)[0]
        F = rho * e / (CX_UNIF * rho ** (4.0 / 3.0))
        if self.base:
            b = ExchangeCorrelation().from_codes(list(self.base))
            F = F - rho * b.eval_points(rho, sigma, tau)[0] / (CX_UNIF * rho ** (4.0 / 3.0))
        return t, F.astype(np.float32)

    # training
    def set_data(self, U, T, C=None):
        U, T = _f32(U).reshape(-1, 3), _f32(T)
        Cc = None if C is None else _f32(C)
        _check(lib().chx_mlp_set_data(self.h, U.shape[0], _ptr(U), _ptr(T), _ptr(Cc)))
        self.ndata = int(U.shape[0])

    def train(self, epochs, batch=16384, lr=3e-3, lr_final=3e-5):
        loss = np.zeros(max(1, epochs))
        _check(lib().chx_mlp_train(self.h, int(epochs), int(batch), float(lr), float(lr_final), _ptr(loss)))
        return loss[:epochs]

    def fit(self, xc, rho, sigma, tau, w, epochs=60, batch=16384, lr=3e-3, lr_final=3e-5, rho_min=1e-8,
            energy_weight=0.5, seed=0):
        """Samples -> box, normalised inputs, sample weights -> GPU training.

        The box is the descriptor range of the samples (the out-of-distribution
        test). Targets: F of xc, less F of the base (set_base). Sample weights
        mix uniform weight with the point's share of |sum w e_x^unif|
        (energy_weight), so that the loss follows the energy error. Returns the
        per-epoch loss."""
        t, F = self.targets(xc, rho, sigma, tau)
        lo, hi = t.min(axis=0), t.max(axis=0)
        self.set_box(lo, hi, rho_min)
        U = self.normalise(t)
        n = rho.size
        ew = np.abs(np.asarray(w) * CX_UNIF * np.asarray(rho) ** (4.0 / 3.0))
        C = (1.0 - energy_weight) + energy_weight * ew * (n / ew.sum())  # mean 1
        perm = np.random.default_rng(seed).permutation(n)
        self.set_data(U[perm], F[perm], C[perm])
        return self.train(epochs, batch, lr, lr_final)

    def loss_grad(self, U, T, C=None):
        """One loss and parameter-gradient evaluation on the GPU (FP32)."""
        U, T = _f32(U).reshape(-1, 3), _f32(T)
        Cc = None if C is None else _f32(C)
        G, L = np.zeros(MLP_PARAMS, np.float32), ctypes.c_double()
        _check(lib().chx_mlp_loss_grad(self.h, U.shape[0], _ptr(U), _ptr(T), _ptr(Cc), _ptr(G), ctypes.byref(L)))
        return L.value, G

    # inference
    def forward(self, U, bf16=True, mfma=False):
        """Network value and dF/dU at normalised inputs (GPU): the FP32 weights, the BF16-rounded
        weights (bf16) on the per-lane network, or the inference path itself (mfma: FP32 weights and
        activations as split-BF16 MFMA operands)."""
        U = _f32(U).reshape(-1, 3)
        n = U.shape[0]
        F, dF = np.zeros(n, np.float32), np.zeros((n, 3), np.float32)
        mode = 2 if mfma else int(bool(bf16))
        _check(lib().chx_mlp_forward(self.h, n, _ptr(U), mode, _ptr(F), _ptr(dF)))
        return F, dF

    def eval_points(self, xc, rho, sigma, tau):
        """eps, v_rho, v_sigma, v_tau of the surrogate (analytic fallback outside the box) and the fallback count."""
        rho, sigma, tau = _f64(rho), _f64(sigma), _f64(tau)
        n = rho.size
        out = [np.zeros(n) for _ in range(4)]
        nf = ctypes.c_long()
        _check(lib().chx_mlp_eval_points(self.h, xc.h, n, _ptr(rho), _ptr(sigma), _ptr(tau), *[_ptr(o) for o in out],
                                         ctypes.byref(nf)))
        return out[0], out[1], out[2], out[3], nf.value

    def inference_params(self):
        """Parameters of the inference path: the FP32 parameters (its MFMA products split them in two
        BF16 terms)."""
        return self.params.astype(np.float32)

    def bf16_params(self):
        """Canonical parameters rounded to BF16 (round to nearest even), the forward(bf16=True) network."""
        P = self.params.astype(np.float32)
        b = P.view(np.uint32).astype(np.uint64)
        r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
        return r.astype(np.uint32).view(np.float32)


def grid_samples(mol, grid_name, D, rho_min=1e-8):
    """(rho, sigma, tau, w) at the live points of a meta-GGA grid of mol with density matrix D."""
    from . import Grid

    g = Grid(mol, grid_name, 1, 0)
    g.setAOCache(1)
    g.setType(2)
    g.getAO(0, 0)
    g.getDensity([D])
    return Surrogate.samples_from_grid(g, rho_min)


def workload_samples(n_alkanes=8, rho_min=1e-8, max_atoms=None):
    """Training samples from the grids of the H2O / cc-pVDZ / SG-1, (H2O)16 / def2-SVP / SG-2
    and alkane / def2-SVP / SG-1 workloads (the first n_alkanes molecules of the 64-alkane
    set), with their seeded synthetic density matrices. Returns concatenated (rho, sigma, tau, w)."""
    from . import ANGSTROM_TO_BOHR, Molecule
    from . import workloads as wl

    parts = []
    for name in ("h2o_ccpvdz_sg1_pbe", "h2o16_def2svp_sg2_b3lyp"):
        w = wl.build_workload(name)
        mol = Molecule(w["Z"], w["xyz_bohr"], w["basis"])
        D, _ = wl.workload_density(mol, w["nocc"], w["seed"])
        parts.append(grid_samples(mol, w["grid"], D, rho_min))
    for k, (n, Z, X) in enumerate(wl.alkane_set(64, 2, 40)[:n_alkanes]):
        if max_atoms is not None and len(Z) > max_atoms:
            continue
        mol = Molecule(Z, X * ANGSTROM_TO_BOHR, "def2-svp")
        D, _ = wl.workload_density(mol, 4 * n + 1, 100 + k)
        parts.append(grid_samples(mol, "SG-1", D, rho_min))
    return tuple(np.concatenate([p[i] for p in parts]) for i in range(4))
