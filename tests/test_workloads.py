"""Seeded synthetic inputs: SplitMix64 known answers and the geometric and
density-matrix constraints each workload promises."""
import numpy as np

import chinium_amd as ca
from chinium_amd import workloads as wl


def test_splitmix64_known_answers():
    # reference outputs of SplitMix64 from seed 0
    r = wl.SplitMix64(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_uniform_and_normal_moments():
    r = wl.SplitMix64(42)
    u = np.array([r.uniform() for _ in range(20000)])
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.01
    z = wl.SplitMix64(1).normal(20001)
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1) < 0.03


def test_water_cluster_constraints():
    Z, X = wl.water_cluster_angstrom(16, 7.5, 1)
    assert len(Z) == 48 and list(Z[:3]) == [8, 1, 1]
    O = X[Z == 8]
    d = np.linalg.norm(O[:, None] - O[None], axis=-1) + np.eye(16) * 99
    assert d.min() >= 2.7
    assert O.min() >= 0 and O.max() <= 7.5
    for k in range(16):
        oh = np.linalg.norm(X[3 * k + 1:3 * k + 3] - X[3 * k], axis=1)
        assert np.allclose(oh, 0.9572, atol=1e-12)
        v1, v2 = X[3 * k + 1] - X[3 * k], X[3 * k + 2] - X[3 * k]
        ang = np.degrees(np.arccos(v1 @ v2 / (np.linalg.norm(v1) * np.linalg.norm(v2))))
        assert abs(ang - 104.52) < 1e-9
    # same seed, same cluster
    Z2, X2 = wl.water_cluster_angstrom(16, 7.5, 1)
    assert np.array_equal(X, X2)


def test_alkane_geometry():
    for n in (1, 2, 7, 40):
        Z, X = wl.alkane_angstrom(n, 2)
        assert (Z == 6).sum() == n and (Z == 1).sum() == 2 * n + 2
        C, H = X[Z == 6], X[Z == 1]
        if n > 1:
            cc = np.linalg.norm(np.diff(C, axis=0), axis=1)
            assert np.all(np.abs(cc - 1.53) <= 0.02 + 1e-12)
        ch = np.linalg.norm(H[:, None] - C[None], axis=-1).min(axis=1)
        assert np.all(np.abs(ch - 1.09) <= 0.02 + 1e-12)
        hh = np.linalg.norm(H[:, None] - H[None], axis=-1) + np.eye(len(H)) * 99
        assert hh.min() > 1.6


def test_cyclocarbon_geometry():
    Z, X = wl.cyclocarbon_angstrom()
    assert len(Z) == 18
    b = np.linalg.norm(np.roll(X, -1, axis=0) - X, axis=1)
    assert np.allclose(b[0::2], 1.23, atol=1e-12) and np.allclose(b[1::2], 1.35, atol=1e-12)
    assert np.allclose(np.linalg.norm(X, axis=1), 3.7145, atol=1e-4)


def test_density_matrix_properties():
    w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D, S = wl.workload_density(mol, w["nocc"], w["seed"])
    assert D.shape == (24, 24)
    assert np.abs(D - D.T).max() < 1e-14
    assert abs(np.trace(D @ S) - 10) < 1e-12
    DS = D @ S
    assert np.abs(DS @ DS - 2 * DS).max() < 1e-11   # idempotent (closed shell, D = 2 C C^T)
