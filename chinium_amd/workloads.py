"""Seeded synthetic workloads: molecules, basis/grid/functional choices and
S-orthonormal density matrices D = 2 C_occ C_occ^T.

Random streams come from a project-written SplitMix64 so that the same seed
gives the same inputs everywhere. Geometries:
  water_monomer     O-H 0.9572 A, H-O-H 104.52 deg, O at the origin.
  water_cluster     monomers at uniform random O sites in a cube (rejection
                    until all O-O >= 2.7 A), random orientation (unit quaternion).
  alkane            all-trans zig-zag C_nH_{2n+2}, C-C 1.53 A, C-H 1.09 A,
                    tetrahedral angles, each bond length jittered by U(-0.02, 0.02) A.
  cyclocarbon       C_18 ring with alternating 1.23/1.35 A bonds (R = 3.7145 A
                    closes the ring with these bond lengths).
"""
import math

import numpy as np

from . import ANGSTROM_TO_BOHR, Molecule

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        """Uniform in [0, 1) with 53 random bits."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def normal(self, n):
        """n standard normals (Box-Muller)."""
        out = np.empty(n)
        i = 0
        while i < n:
            u1 = self.uniform()
            u2 = self.uniform()
            r = math.sqrt(-2.0 * math.log(1.0 - u1))
            out[i] = r * math.cos(2.0 * math.pi * u2)
            if i + 1 < n:
                out[i + 1] = r * math.sin(2.0 * math.pi * u2)
            i += 2
        return out


def water_monomer_angstrom():
    half = math.radians(104.52) / 2
    r = 0.9572
    return np.array([8, 1, 1]), np.array([[0.0, 0.0, 0.0],
                                          [r * math.sin(half), 0.0, r * math.cos(half)],
                                          [-r * math.sin(half), 0.0, r * math.cos(half)]])


def _quat_rotation(rng):
    # uniform random rotation from a uniform unit quaternion (Shoemake)
    u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
    q = (math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
         math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3))
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster_angstrom(n, box, seed, min_oo=2.7):
    rng = SplitMix64(seed)
    Zm, Xm = water_monomer_angstrom()
    sites = []
    tries = 0
    while len(sites) < n:
        tries += 1
        if tries > 1000000:
            raise RuntimeError("could not place waters")
        p = np.array([rng.uniform() * box, rng.uniform() * box, rng.uniform() * box])
        if all(np.linalg.norm(p - s) >= min_oo for s in sites):
            sites.append(p)
    Z, X = [], []
    for s in sites:
        R = _quat_rotation(rng)
        Z.extend(Zm)
        X.extend((Xm @ R.T) + s)
    return np.array(Z), np.array(X)


def alkane_angstrom(nc, seed):
    """All-trans C_nH_{2n+2}: zig-zag carbons in the xy plane, sp3 hydrogens."""
    rng = SplitMix64(seed)
    jit = lambda: (rng.uniform() * 2 - 1) * 0.02
    tet = math.radians(109.4712206)
    half = tet / 2
    if nc == 1:
        dirs = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) / math.sqrt(3)
        return np.array([6, 1, 1, 1, 1]), np.vstack([np.zeros(3)] + [(1.09 + jit()) * d for d in dirs])
    # unit bond vectors of the zig-zag: alternating +/- y
    bond = lambda i: np.array([math.sin(half), math.cos(half) * (1 if i % 2 == 0 else -1), 0.0])
    C = [np.zeros(3)]
    for i in range(nc - 1):
        C.append(C[-1] + (1.53 + jit()) * bond(i))
    Z, X = [6] * nc, list(C)
    for i in range(nc):
        # neighbour directions; a terminal carbon gets the virtual continuation bond
        u1 = -bond(i - 1) if i > 0 else -bond(-1)
        u2 = bond(i) if i < nc - 1 else bond(i)
        b = -(u1 + u2)
        b /= np.linalg.norm(b)
        n = np.cross(u1, u2)
        n /= np.linalg.norm(n)
        hd = [b * math.cos(half) + n * math.sin(half), b * math.cos(half) - n * math.sin(half)]
        if i == 0:
            hd.append(u1)
        if i == nc - 1:
            hd.append(u2)
        for d in hd:
            Z.append(1)
            X.append(C[i] + (1.09 + jit()) * d)
    return np.array(Z), np.array(X)


def alkane_set(count=64, seed=2, nmax=40):
    """The many-molecule workload: `count` linear alkanes C_nH_{2n+2} with n
    drawn seeded-uniform in 1..nmax. Returns [(n, Z, xyz_angstrom)]; molecule k
    takes jitter seed 1000 * seed + k. Closed shell with 4n + 1 occupied orbitals."""
    rng = SplitMix64(seed)
    out = []
    for k in range(count):
        n = min(nmax, 1 + int(rng.uniform() * nmax))
        Z, X = alkane_angstrom(n, 1000 * seed + k)
        out.append((n, Z, X))
    return out


def cyclocarbon_angstrom(n=18, b1=1.23, b2=1.35):
    # chord lengths b = 2 R sin(theta/2); alternate angles theta1, theta2 with
    # n/2 (theta1 + theta2) = 2 pi; solve for R by bisection.
    def total(R):
        return (n // 2) * (2 * math.asin(b1 / (2 * R)) + 2 * math.asin(b2 / (2 * R))) - 2 * math.pi
    lo, hi = max(b1, b2) / 2 + 1e-9, 50.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if total(mid) > 0:
            lo = mid
        else:
            hi = mid
    R = 0.5 * (lo + hi)
    t1, t2 = 2 * math.asin(b1 / (2 * R)), 2 * math.asin(b2 / (2 * R))
    ang = 0.0
    X = []
    for i in range(n):
        X.append([R * math.cos(ang), R * math.sin(ang), 0.0])
        ang += t1 if i % 2 == 0 else t2
    return np.full(n, 6), np.array(X)


def density_matrix(S, nocc, seed):
    """D = 2 C C^T with C (n x nocc) seeded N(0,1), Loewdin-orthonormalised in S."""
    n = S.shape[0]
    rng = SplitMix64(seed)
    C = rng.normal(n * nocc).reshape(n, nocc)
    M = C.T @ S @ C
    w, U = np.linalg.eigh(M)
    C = C @ (U @ np.diag(w ** -0.5) @ U.T)
    return 2.0 * C @ C.T


WORKLOADS = {
    # name: (builder, basis, grid, functional, nocc, seed)
    "h2o_ccpvdz_sg1_pbe": ("water", "cc-pvdz", "SG-1", "PBE", 5, 0),
    "h2o16_def2svp_sg2_b3lyp": ("water16", "def2-svp", "SG-2", "B3LYP", 80, 1),
    # cyclo[18]carbon: bonds fixed at 1.23 / 1.35 A, radius solved for (3.7145 A)
    "c18_def2tzvp_sg3_r2scan": ("c18", "def2-tzvp", "SG-3", "R2SCAN", 54, 3),
    # (H2O)128 in a 15 A cube, O-O >= 2.7 A: the full-size cluster (3072 AOs)
    "h2o128_def2svp_sg2_b3lyp": ("water128", "def2-svp", "SG-2", "B3LYP", 640, 5),
    "h2o128_def2svp_sg2_r2scan": ("water128", "def2-svp", "SG-2", "R2SCAN", 640, 5),
}

DESCRIPTIONS = {
    "h2o_ccpvdz_sg1_pbe": "H2O / cc-pVDZ / SG-1 / PBE",
    "h2o16_def2svp_sg2_b3lyp": "(H2O)16 / def2-SVP / SG-2 / B3LYP",
    "c18_def2tzvp_sg3_r2scan": "cyclo[18]carbon / def2-TZVP / SG-3 / r2SCAN",
    "h2o128_def2svp_sg2_b3lyp": "(H2O)128 / def2-SVP / SG-2 / B3LYP",
    "h2o128_def2svp_sg2_r2scan": "(H2O)128 / def2-SVP / SG-2 / r2SCAN",
}


def build_workload(name):
    """Returns dict(Z, xyz_bohr, basis, grid, functional, nocc, seed)."""
    kind, basis, grid, func, nocc, seed = WORKLOADS[name]
    if kind == "water":
        Z, X = water_monomer_angstrom()
    elif kind == "water16":
        Z, X = water_cluster_angstrom(16, 7.5, seed)
    elif kind == "water128":
        Z, X = water_cluster_angstrom(128, 15.0, seed)
    elif kind == "c18":
        Z, X = cyclocarbon_angstrom(18)
    else:
        raise KeyError(name)
    return dict(Z=np.asarray(Z, dtype=np.int32), xyz_bohr=np.asarray(X) * ANGSTROM_TO_BOHR, basis=basis,
                grid=grid, functional=func, nocc=nocc, seed=seed)


def workload_density(mol, nocc, seed):
    S = mol.overlap()
    return density_matrix(S, nocc, seed), S
