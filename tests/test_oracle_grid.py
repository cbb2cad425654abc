"""Oracle grid (Grid::Grid, GridGeneration.cpp:27-310): pinned by the data files'
point counts, closed-form integrals and the Becke partition of unity."""
import numpy as np
import pytest

import chinium_amd as ca
import oracle_lib as ol
from chinium_amd import workloads as wl

# points per element, computed from Grids/<name>/<El>.grid as sum(npoints * nshells)
COUNTS = {("SG-1", 1): 3752, ("SG-1", 6): 3816, ("SG-1", 8): 3816,
          ("SG-2", 1): 7094, ("SG-2", 6): 7790, ("SG-2", 8): 8574,
          ("SG-3", 1): 16710, ("SG-3", 6): 17674, ("SG-3", 8): 18946}


@pytest.mark.parametrize("grid,Z", sorted(COUNTS))
def test_single_atom_counts_and_closed_form(grid, Z):
    o = ol.OracleSystem([Z], np.zeros((1, 3)), grid, "def2-svp")
    assert o.npoints == COUNTS[(grid, Z)]
    P = o.points()
    # one atom: Becke weight is 1, so the weights are the raw radial x angular weights
    assert np.array_equal(P[:, 3], o.raw_weights())
    r2 = np.sum(P[:, :3] ** 2, axis=1)
    # integral of exp(-a r^2) over space is (pi/a)^(3/2); of r^2 exp(-r^2) is 3/2 pi^(3/2)
    for a in (0.5, 1.0, 3.0):
        assert abs(np.sum(P[:, 3] * np.exp(-a * r2)) / (np.pi / a) ** 1.5 - 1) < 2e-6
    assert abs(np.sum(P[:, 3] * r2 * np.exp(-r2)) / (1.5 * np.pi ** 1.5) - 1) < 2e-6
    # angular exactness on every shell: integral of x^2 exp(-r^2) = pi^(3/2)/2
    assert abs(np.sum(P[:, 3] * P[:, 0] ** 2 * np.exp(-r2)) / (0.5 * np.pi ** 1.5) - 1) < 2e-6


def test_becke_partition_of_unity():
    """Sum of one normalised Gaussian per atom integrates to natoms only if the
    Becke fractions of each point sum to one across atoms."""
    Z, X = wl.water_cluster_angstrom(3, 4.5, 5)
    xyz = X * ca.ANGSTROM_TO_BOHR
    o = ol.OracleSystem(Z, xyz, "SG-2", "def2-svp")
    P = o.points()
    f = np.zeros(len(P))
    for c in xyz:
        f += (2.0 / np.pi) ** 1.5 * np.exp(-2.0 * np.sum((P[:, :3] - c) ** 2, axis=1))
    assert abs(np.sum(P[:, 3] * f) - len(Z)) < 1e-5
    assert np.all(P[:, 3] >= 0)
    # each point's weight is at most its raw weight (fraction in [0, 1])
    assert np.all(P[:, 3] <= o.raw_weights() * (1 + 1e-14))


def test_subgrid_halving_matches_reference_sizes():
    """H2O / SG-1: 11320 points in 128 sub-grids of 88-89 points (recursive
    halving, GridGeneration.cpp:203-210)."""
    Z, X = wl.water_monomer_angstrom()
    o = ol.OracleSystem(Z, X * ca.ANGSTROM_TO_BOHR, "SG-1", "cc-pvdz")
    s = o.subgrid_sizes()
    assert o.npoints == 11320
    assert len(s) == 128 and s.sum() == 11320 and set(s) == {88, 89}


def test_missing_grid_file_message():
    with pytest.raises(RuntimeError, match="Missing element grid file in folder!"):
        ol.OracleSystem([8], np.zeros((1, 3)), "SG-9", "def2-svp")
