"""Packs (chx_pack_*): many molecules' XC builds on staggered lanes, replayed
as one hipGraph, against the same grids built one at a time and against the
FP64 oracle."""
import numpy as np
import pytest

import chinium_amd as ca
import oracle_lib as ol
from chinium_amd import hip
from chinium_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def alkanes():
    mols = []
    for k, n in enumerate((1, 3, 2, 6)):
        Z, X = wl.alkane_angstrom(n, 70 + k)
        xyz = X * ca.ANGSTROM_TO_BOHR
        mol = ca.Molecule(Z, xyz, "def2-svp")
        D, S = wl.workload_density(mol, 4 * n + 1, 30 + k)
        grid = ca.Grid(mol, "SG-1", 1, 0)
        grid.setPrecision(ca.CHX_PREC_MIXED)
        grid.setType(1)
        grid.getAO(0, 0)
        mols.append(dict(n=n, Z=Z, xyz=xyz, mol=mol, D=D, grid=grid))
    xc = ca.ExchangeCorrelation()
    xc.Read("PBE")
    # one-at-a-time builds first (the grids are used on their own before packing)
    for m in mols:
        m["E1"], m["V1"] = ca.xc_build(m["grid"], xc, m["D"])
        m["N1"] = m["grid"].getNumElectrons()
    dD, dV = [], []
    for m in mols:
        nb = m["mol"].nbasis
        a = hip.DeviceArray(nb * nb)
        a.upload(np.asfortranarray(m["D"]).ravel(order="F"))
        dD.append(a)
        dV.append(hip.DeviceArray(nb * nb))
    _say("pack of %d alkanes (n = %s)" % (len(mols), [m["n"] for m in mols]))
    pack = ca.Pack([m["grid"] for m in mols])
    return dict(mols=mols, xc=xc, pack=pack, dD=dD, dV=dV)


def _run(a):
    E, N = a["pack"].xc(a["xc"], [d.ptr for d in a["dD"]], [v.ptr for v in a["dV"]])
    Vs = [v.download().reshape((m["mol"].nbasis,) * 2, order="F") for v, m in zip(a["dV"], a["mols"])]
    return E, N, Vs


def test_pack_matches_single_grids(alkanes):
    E, N, Vs = _run(alkanes)
    for k, m in enumerate(alkanes["mols"]):
        dE, dV = abs(E[k] - m["E1"]), np.abs(Vs[k] - m["V1"]).max()
        _say("  C%d: E pack %.12f single %.12f dE %.1e dV %.1e" % (m["n"], E[k], m["E1"], dE, dV))
        # same kernels and partial sums; V differs only by the order of FP64 atomics
        assert dE <= 1e-12 * abs(m["E1"])
        assert abs(N[k] - m["N1"]) <= 1e-12 * m["N1"]
        assert dV <= 1e-12


def test_pack_graph_replays(alkanes):
    """Replays of the captured graph, with eager work between them, agree with eager launches."""
    pack = alkanes["pack"]
    pack.set_graph(False)
    E0, N0, V0 = _run(alkanes)
    pack.set_graph(True)
    for rep in range(3):
        E, N, Vs = _run(alkanes)
        m = alkanes["mols"][rep % len(alkanes["mols"])]
        ca.xc_build(m["grid"], alkanes["xc"], m["D"])  # eager work on a grid stream between replays
        assert np.all(np.abs(E - E0) <= 1e-12 * np.abs(E0))
        for a, b in zip(Vs, V0):
            assert np.abs(a - b).max() <= 1e-12


def test_pack_against_oracle(alkanes):
    E, N, Vs = _run(alkanes)
    m = alkanes["mols"][3]
    orc = ol.OracleSystem(m["Z"], m["xyz"], "SG-1", "def2-svp")
    orc.set_functional("PBE")
    orc.get_ao()
    Eo, No, Vo = orc.xc(m["D"])
    dE = abs(E[3] - Eo) / len(m["Z"])
    _say("  C%d pack vs oracle: dE/atom %.2e  max dV %.2e" % (m["n"], dE, np.abs(Vs[3] - Vo).max()))
    assert dE <= 1e-5
    assert np.abs(Vs[3] - Vo).max() <= 5e-5


def test_pack_autotune(alkanes):
    xc, pack = alkanes["xc"], alkanes["pack"]
    E0, _, V0 = _run(alkanes)
    chunk, ms = pack.autotune(xc, [d.ptr for d in alkanes["dD"]], [v.ptr for v in alkanes["dV"]], reps=2)
    _say("  autotune: chunk %d, lanes %d, form %d, ms %s" % (chunk, pack.lanes(), pack.form(),
                                                                   {k: round(v, 4) for k, v in ms.items()}))
    assert chunk in (8, 16, 32, 128) and all(v > 0 for v in ms.values())
    assert 1 <= pack.lanes() <= len(alkanes["mols"])
    E, _, Vs = _run(alkanes)
    # the density side does not depend on the Vxc item length; M is summed in
    # FP32 within an item, so another length changes V at the FP32 level only
    assert np.all(np.abs(E - E0) <= 1e-12 * np.abs(E0))
    for a, b in zip(Vs, V0):
        assert np.abs(a - b).max() <= 2e-6


def test_pack_lanes(alkanes):
    """Molecules split over lanes (streams forked and joined inside the graph)
    give the same results for any lane count, replayed and eager; the default
    (more lanes than molecules) clamps to one molecule per lane."""
    pack = alkanes["pack"]
    E0, N0, V0 = _run(alkanes)
    try:
        for lanes, graph in ((1, True), (1, False), (2, True), (3, True), (3, False), (8, True)):
            pack.set_lanes(lanes)
            pack.set_graph(graph)
            for rep in range(2):
                E, N, Vs = _run(alkanes)
                dV = max(np.abs(a - b).max() for a, b in zip(Vs, V0))
                _say("  lanes %d graph %d rep %d: max dE %.1e max dV %.1e"
                     % (lanes, graph, rep, np.abs(E - E0).max(), dV))
                assert np.all(np.abs(E - E0) <= 1e-12 * np.abs(E0))
                assert np.all(np.abs(N - N0) <= 1e-12 * N0)
                assert dV <= 1e-12
        with pytest.raises(ca.ChxError) as e:
            pack.set_lanes(0)
        assert e.value.code == -1
    finally:
        pack.set_graph(True)
        pack.set_lanes(16)


def test_pack_forms(alkanes):
    """Persistent stages (one launch per stage over every molecule's items) and per-lane stage
    launches give the per-molecule kernels' results, eager and replayed; setting lanes leaves the
    persistent form."""
    pack = alkanes["pack"]
    pack.set_lanes(16)
    pack.set_form(0)
    E0, N0, V0 = _run(alkanes)
    try:
        for form, lanes in ((1, 16), (2, 2), (2, 4)):
            pack.set_lanes(lanes)
            pack.set_form(form)
            assert pack.form() == form
            for graph in (False, True, True):
                pack.set_graph(graph)
                E, N, Vs = _run(alkanes)
                dV = max(np.abs(a - b).max() for a, b in zip(Vs, V0))
                _say("  form %d lanes %d graph %d: max dE %.1e max dN %.1e max dV %.1e"
                     % (form, lanes, graph, np.abs(E - E0).max(), np.abs(N - N0).max(), dV))
                assert np.all(np.abs(E - E0) <= 1e-12 * np.abs(E0))
                assert np.all(np.abs(N - N0) <= 1e-12 * N0)
                assert dV <= 1e-12
        pack.set_form(1)
        pack.set_lanes(16)
        assert pack.form() == 0
    finally:
        pack.set_graph(True)
        pack.set_form(0)
        pack.set_lanes(16)


def test_pack_with_wide_member(alkanes):
    """A pack member large enough for the 256-AO Vxc super-tiles ((H2O)40 /
    def2-SVP, 960 AOs, npad 1024) keeps its own plan inside the pack: the pack
    gives its single-grid result, beside a small alkane."""
    Z, X = wl.water_cluster_angstrom(40, 10.5, 21)
    mol = ca.Molecule(Z, X * ca.ANGSTROM_TO_BOHR, "def2-svp")
    D, S = wl.workload_density(mol, 200, 5)
    grid = ca.Grid(mol, "SG-1", 1, 0)
    grid.setPrecision(ca.CHX_PREC_MIXED)
    grid.setType(1)
    grid.getAO(0, 0)
    xc = alkanes["xc"]
    E1, V1 = ca.xc_build(grid, xc, D)
    small = alkanes["mols"][1]
    Es1, Vs1 = ca.xc_build(small["grid"], xc, small["D"])  # at the Vxc item length the autotuner left
    nb = mol.nbasis
    dD, dV = hip.DeviceArray(nb * nb), hip.DeviceArray(nb * nb)
    dD.upload(np.asfortranarray(D).ravel(order="F"))
    pack = ca.Pack([grid, small["grid"]])
    E, N = pack.xc(xc, [dD.ptr, alkanes["dD"][1].ptr], [dV.ptr, alkanes["dV"][1].ptr])
    V = dV.download().reshape((nb, nb), order="F")
    Vs = alkanes["dV"][1].download().reshape((small["mol"].nbasis,) * 2, order="F")
    _say("  (H2O)40 in a pack: dE %.1e dV %.1e; C%d dE %.1e dV %.1e"
         % (abs(E[0] - E1), np.abs(V - V1).max(), small["n"], abs(E[1] - Es1), np.abs(Vs - Vs1).max()))
    assert abs(E[0] - E1) <= 1e-12 * abs(E1)
    assert np.abs(V - V1).max() <= 1e-12
    assert abs(E[1] - Es1) <= 1e-12 * abs(Es1)
    assert np.abs(Vs - Vs1).max() <= 1e-12
    for form in (1, 2):  # the wide member's items go to the 256-AO stage launch
        pack.set_form(form)
        E, N = pack.xc(xc, [dD.ptr, alkanes["dD"][1].ptr], [dV.ptr, alkanes["dV"][1].ptr])
        V = dV.download().reshape((nb, nb), order="F")
        Vs = alkanes["dV"][1].download().reshape((small["mol"].nbasis,) * 2, order="F")
        _say("  form %d: (H2O)40 dE %.1e dV %.1e; C%d dE %.1e dV %.1e"
             % (form, abs(E[0] - E1), np.abs(V - V1).max(), small["n"], abs(E[1] - Es1), np.abs(Vs - Vs1).max()))
        assert abs(E[0] - E1) <= 1e-12 * abs(E1) and np.abs(V - V1).max() <= 1e-12
        assert abs(E[1] - Es1) <= 1e-12 * abs(Es1) and np.abs(Vs - Vs1).max() <= 1e-12


def test_pack_meta_gga():
    """r2SCAN (meta-GGA: four density launches per molecule, tau term in V) through a pack of
    three small molecules on two lanes, against the single-grid builds and the oracle."""
    xc = ca.ExchangeCorrelation()
    xc.Read("R2SCAN")
    mols = []
    for k, n in enumerate((1, 4, 2)):
        Z, X = wl.alkane_angstrom(n, 90 + k)
        xyz = X * ca.ANGSTROM_TO_BOHR
        mol = ca.Molecule(Z, xyz, "def2-svp")
        D, S = wl.workload_density(mol, 4 * n + 1, 60 + k)
        grid = ca.Grid(mol, "SG-1", 1, 0)
        grid.setPrecision(ca.CHX_PREC_MIXED)
        grid.setType(2)
        grid.getAO(0, 0)
        E1, V1 = ca.xc_build(grid, xc, D)
        nb = mol.nbasis
        dD, dV = hip.DeviceArray(nb * nb), hip.DeviceArray(nb * nb)
        dD.upload(np.asfortranarray(D).ravel(order="F"))
        mols.append(dict(n=n, Z=Z, xyz=xyz, mol=mol, D=D, grid=grid, E1=E1, V1=V1, dD=dD, dV=dV))
    pack = ca.Pack([m["grid"] for m in mols])
    for form in (1, 2):
        with pytest.raises(ca.ChxError) as e:  # stage launches are built for LDA / GGA grids
            pack.set_form(form)
        assert e.value.code == -5
    pack.set_lanes(2)
    for rep in range(2):
        E, N = pack.xc(xc, [m["dD"].ptr for m in mols], [m["dV"].ptr for m in mols])
        for k, m in enumerate(mols):
            V = m["dV"].download().reshape((m["mol"].nbasis,) * 2, order="F")
            _say("  r2SCAN pack rep %d C%d: dE %.1e dV %.1e" % (rep, m["n"], abs(E[k] - m["E1"]),
                                                               np.abs(V - m["V1"]).max()))
            assert abs(E[k] - m["E1"]) <= 1e-12 * abs(m["E1"])
            assert np.abs(V - m["V1"]).max() <= 1e-12
    m = mols[1]
    orc = ol.OracleSystem(m["Z"], m["xyz"], "SG-1", "def2-svp")
    orc.set_functional("R2SCAN")
    orc.get_ao()
    Eo, No, Vo = orc.xc(m["D"])
    V = m["dV"].download().reshape((m["mol"].nbasis,) * 2, order="F")
    _say("  r2SCAN pack C%d vs oracle: dE/atom %.2e max dV %.2e" % (m["n"], abs(E[1] - Eo) / len(m["Z"]),
                                                                   np.abs(V - Vo).max()))
    assert abs(E[1] - Eo) / len(m["Z"]) <= 1e-5
    assert np.abs(V - Vo).max() <= 5e-5


def test_pack_rejects_fp64_grid(alkanes):
    m = alkanes["mols"][0]
    g = ca.Grid(m["mol"], "SG-1", 1, 0)
    g.setType(1)
    g.getAO(0, 0)
    with pytest.raises(ca.ChxError) as e:
        ca.Pack([g])
    assert e.value.code == -5
