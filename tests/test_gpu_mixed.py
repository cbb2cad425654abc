"""GPU parity of the mixed-precision form (CHX_PREC_MIXED: split-BF16 MFMA
contractions with FP32 accumulation, FP64 functional and sums) against the
FP64 oracle, through the C ABI.

Bar: |E_xc - E_oracle| <= 1e-5 Eh per atom; max |V - V_oracle| is reported
and held below 5e-5 (a split-BF16 product carries ~16 significant bits, so
V elements of order 1 agree to a few 1e-6).
"""
import numpy as np
import pytest

import chinium_amd as ca
import oracle_lib as ol
from chinium_amd import workloads as wl

pytestmark = pytest.mark.gpu

E_TOL_PER_ATOM = 1e-5
V_TOL = 5e-5


def _say(msg):
    print(msg, flush=True)


def _mixed_grid(mol, grid_name, gtype):
    grid = ca.Grid(mol, grid_name, 1, 0)
    grid.setPrecision(ca.CHX_PREC_MIXED)
    grid.setType(gtype)
    grid.getAO(0, 0)
    assert grid.aoCached()
    return grid


def _compare(tag, Z, xyz, basis, grid_name, func, D, gtype=1):
    mol = ca.Molecule(Z, xyz, basis)
    _say("mixed %s: %d AOs, %s, %s" % (tag, mol.nbasis, grid_name, func))
    grid = _mixed_grid(mol, grid_name, gtype)
    xc = ca.ExchangeCorrelation()
    xc.Read(func)
    E, V = ca.xc_build(grid, xc, D)
    N = grid.getNumElectrons()
    orc = ol.OracleSystem(Z, xyz, grid_name, basis)
    orc.set_functional(func)
    orc.get_ao()
    Eo, No, Vo = orc.xc(D)
    na = len(Z)
    dE, dV, dN = abs(E - Eo), np.abs(V - Vo).max(), abs(N - No)
    _say("  E mixed %.12f oracle %.12f  dE/atom %.2e  max dV %.2e  dN %.2e" % (E, Eo, dE / na, dV, dN))
    assert dE <= E_TOL_PER_ATOM * na
    assert dV <= V_TOL
    assert np.abs(V - V.T).max() <= 1e-12
    return grid, orc


@pytest.mark.parametrize("func", ["PBE", "B3LYP", "SVWN5"])
def test_mixed_h2o(func):
    w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D, S = wl.workload_density(mol, w["nocc"], w["seed"])
    _compare("h2o", w["Z"], w["xyz_bohr"], w["basis"], w["grid"], func, D)


def test_mixed_water_cluster_and_vrho():
    Z, X = wl.water_cluster_angstrom(4, 5.0, 11)
    xyz = X * ca.ANGSTROM_TO_BOHR
    mol = ca.Molecule(Z, xyz, "def2-svp")
    D, S = wl.workload_density(mol, 20, 3)
    grid, orc = _compare("(H2O)4/def2-svp", Z, xyz, "def2-svp", "SG-2", "B3LYP", D)
    # supplied-potential path (SAP guess) through the mixed Vxc kernels
    P = grid.points()
    v = -np.exp(-0.1 * np.sum(P[:, :3] ** 2, axis=1))
    grid.setVrho(v)
    V2 = grid.getFock()[0]
    ao, _, _ = orc.ao_values(P[:, :3])
    dV = np.abs(V2 - (ao * (P[:, 3] * v)) @ ao.T).max()
    _say("  set_vrho max dV %.2e" % dV)
    assert dV <= V_TOL


def test_mixed_r2scan():
    Z, X = wl.alkane_angstrom(1, 4)
    xyz = X * ca.ANGSTROM_TO_BOHR
    mol = ca.Molecule(Z, xyz, "def2-tzvp")
    D, S = wl.workload_density(mol, 5, 11)
    _compare("CH4/def2-tzvp", Z, xyz, "def2-tzvp", "SG-1", "R2SCAN", D, gtype=2)


def test_mixed_cluster16_b3lyp():
    """The mixed-precision workload: (H2O)16 / def2-SVP / SG-2 / B3LYP."""
    w = wl.build_workload("h2o16_def2svp_sg2_b3lyp")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D, S = wl.workload_density(mol, w["nocc"], w["seed"])
    _compare("(H2O)16/def2-svp", w["Z"], w["xyz_bohr"], w["basis"], w["grid"], "B3LYP", D)


def test_precision_switch_needs_getao():
    w = wl.build_workload("h2o_ccpvdz_sg1_pbe")
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D, S = wl.workload_density(mol, w["nocc"], w["seed"])
    xc = ca.ExchangeCorrelation()
    xc.Read("PBE")
    grid = ca.Grid(mol, w["grid"], 1, 0)
    grid.setType(1)
    grid.getAO(0, 0)
    E64, V64 = ca.xc_build(grid, xc, D)
    grid.setPrecision(ca.CHX_PREC_MIXED)
    with pytest.raises(ca.ChxError) as e:
        grid.getDensity([D])
    assert e.value.code == -3
    grid.getAO(0, 0)
    Emx, Vmx = ca.xc_build(grid, xc, D)
    grid.setPrecision(ca.CHX_PREC_FP64)
    grid.getAO(0, 0)
    E2, V2 = ca.xc_build(grid, xc, D)
    _say("  fp64 %.14f mixed %.14f back %.14f" % (E64, Emx, E2))
    assert abs(E2 - E64) <= 1e-13 and np.abs(V2 - V64).max() <= 1e-13
    assert 0 < abs(Emx - E64) <= 3 * E_TOL_PER_ATOM
    with pytest.raises(ca.ChxError):
        grid.setPrecision(7)


def test_mixed_meta_grid_after_gga_potentials():
    """A meta-GGA grid (type 2) in the mixed form that serves a meta-GGA build, then a potential with one
    component (SAP-style v_rho, GuessSCF.cpp:37-60), then a GGA build, and the meta-GGA build again. A Vxc
    item reads the all-zero B^T slot for tiles that are not alive in a batch; it sits at the end of each
    build's B^T block, whose size follows the potential's component count. (H2O)4 / def2-SVP has three
    AO tiles, so dead tiles occur at the grid's outer points."""
    Z, X = wl.water_cluster_angstrom(4, 4.0, 11)
    xyz = X * ca.ANGSTROM_TO_BOHR
    mol = ca.Molecule(Z, xyz, "def2-svp")
    D, S = wl.workload_density(mol, 20, 3)
    grid = _mixed_grid(mol, "SG-1", 2)
    oracles = {}
    for func in ("R2SCAN", "PBE"):
        o = ol.OracleSystem(Z, xyz, "SG-1", "def2-svp")
        o.set_functional(func)
        o.get_ao()
        oracles[func] = o.xc(D)
    na = len(Z)

    def check(func):
        xc = ca.ExchangeCorrelation()
        xc.Read(func)
        E, V = ca.xc_build(grid, xc, D)
        Eo, No, Vo = oracles[func]
        _say("  type-2 grid, %s: dE/atom %.2e max dV %.2e" % (func, abs(E - Eo) / na, np.abs(V - Vo).max()))
        assert abs(E - Eo) <= E_TOL_PER_ATOM * na
        assert np.abs(V - Vo).max() <= V_TOL

    check("R2SCAN")
    P = grid.points()
    r = np.linalg.norm(P[:, :3] - xyz.mean(axis=0), axis=1)
    v = -1.0 / np.sqrt(r * r + 0.5)
    grid.setVrho(v)
    V = grid.getFock()[0]
    ao, _, _ = ol.OracleSystem(Z, xyz, "SG-1", "def2-svp").ao_values(P[:, :3])
    Vref = (ao * (P[:, 3] * v)) @ ao.T
    _say("  type-2 grid, v_rho only: max dV %.2e" % np.abs(V - Vref).max())
    assert np.abs(V - Vref).max() <= V_TOL
    check("PBE")
    check("R2SCAN")


@pytest.mark.parametrize("name", ["h2o16_def2svp_sg2_b3lyp", "c18_def2tzvp_sg3_r2scan"])
def test_vxc_block_shapes_agree(name):
    """Both Vxc block shapes of the mixed form (128-AO super-tile pairs; 256-AO groups of each batch
    chunk's alive-tile union) on bases below the default switch (npad 384 and 640): the same E_xc (the
    shape only changes Vxc and, through it, E's first-order correction for the split of D,
    sum V . (D - hi - lo), whose size is about 2^-17 |V . D|: the shapes' V differ by their rounding,
    which moves E by far less than 1e-9 Eh), V within the mixed form's rounding of each other and of
    the FP64 form; the tile autotuner keeps one of them."""
    from chinium_amd import hip
    w = wl.build_workload(name)
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    D, _ = wl.workload_density(mol, w["nocc"], w["seed"])
    xc = ca.ExchangeCorrelation()
    xc.Read(w["functional"])
    g64 = ca.Grid(mol, w["grid"], 1, 0)
    g64.setType(xc.family)
    g64.getAO(0, 0)
    E64, V64 = ca.xc_build(g64, xc, D)
    g = ca.Grid(mol, w["grid"], 1, 0)
    g.setPrecision(ca.CHX_PREC_MIXED)
    g.setType(xc.family)
    g.getAO(0, 0)
    out = {}
    for shape in (0, 1):
        g.setVxcShape(shape)
        assert g.vxcShape() == shape
        out[shape] = ca.xc_build(g, xc, D)
    (E0, V0), (E1, V1) = out[0], out[1]
    print("  %s: shape 0 vs 1: dE %.1e, max dV %.1e; vs FP64: %.1e, %.1e" % (
        name, E1 - E0, np.abs(V1 - V0).max(), np.abs(V0 - V64).max(), np.abs(V1 - V64).max()), flush=True)
    assert abs(E1 - E0) <= 1e-9
    assert np.abs(V1 - V0).max() <= 5e-6
    assert np.abs(V0 - V64).max() <= 5e-5 and np.abs(V1 - V64).max() <= 5e-5
    n = mol.nbasis
    dD, dV = hip.DeviceArray(n * n), hip.DeviceArray(n * n)
    dD.upload(np.asfortranarray(D).ravel(order="F"))
    shape, chunk, ms = g.autotuneTiles(xc, dD.ptr, dV.ptr, reps=1)
    print("  autotuned shape %d, length %d: %s" % (shape, chunk, {k: round(v, 3) for k, v in ms.items()}),
          flush=True)
    assert shape in (0, 1) and g.vxcShape() == shape and min(ms.values()) == ms[(shape, chunk)]
    E2, V2 = ca.xc_build(g, xc, D)
    assert abs(E2 - E0) <= 1e-9 and np.abs(V2 - V64).max() <= 5e-5
