"""The many-molecule workload at full size: 64 linear alkanes C_nH_{2n+2}, n
seeded-uniform in 1..40 (34-970 AOs each), def2-SVP / SG-1 / PBE, mixed form,
as one pack. The pack against the same grids called one by one through the
single-grid C ABI (same kernels: E and N_e to 1e-12 relative, V to FP64
atomic-order rounding), again after the autotuner changed the Vxc work length
and lanes, and two mid-size members against the FP64 oracle (|dE_xc| <= 1e-5
Eh/atom, max |dV| <= 5e-5)."""
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
def alkanes64():
    xc = ca.ExchangeCorrelation()
    xc.Read("PBE")
    items = []
    for k, (n, Z, X) in enumerate(wl.alkane_set(64, 2, 40)):
        xyz = X * ca.ANGSTROM_TO_BOHR
        mol = ca.Molecule(Z, xyz, "def2-svp")
        D, S = wl.workload_density(mol, 4 * n + 1, 500 + k)
        grid = ca.Grid(mol, "SG-1", 1, 0)
        grid.setPrecision(ca.CHX_PREC_MIXED)
        grid.setType(1)
        grid.getAO(0, 0)
        nb = mol.nbasis
        dD, dV = hip.DeviceArray(nb * nb), hip.DeviceArray(nb * nb)
        dD.upload(np.asfortranarray(D).ravel(order="F"))
        E1, V1 = ca.xc_build(grid, xc, D)
        items.append(dict(n=n, Z=Z, xyz=xyz, mol=mol, D=D, grid=grid, dD=dD, dV=dV, E1=E1, V1=V1,
                          N1=grid.getNumElectrons()))
    _say("  64 alkanes: n = %s, %d AOs in all" % ([it["n"] for it in items], sum(it["mol"].nbasis for it in items)))
    pack = ca.Pack([it["grid"] for it in items])
    return dict(items=items, xc=xc, pack=pack)


def _run(a):
    its = a["items"]
    E, N = a["pack"].xc(a["xc"], [it["dD"].ptr for it in its], [it["dV"].ptr for it in its])
    V = [it["dV"].download().reshape((it["mol"].nbasis,) * 2, order="F") for it in its]
    return E, N, V


def test_alkanes64_pack_matches_single_grids(alkanes64):
    E, N, V = _run(alkanes64)
    worst = (0.0, 0.0)
    for k, it in enumerate(alkanes64["items"]):
        dE = abs(E[k] - it["E1"]) / abs(it["E1"])
        dV = np.abs(V[k] - it["V1"]).max() / max(1.0, np.abs(it["V1"]).max())
        worst = (max(worst[0], dE), max(worst[1], dV))
        assert dE <= 1e-12, (k, it["n"], dE)
        assert abs(N[k] - it["N1"]) <= 1e-12 * it["N1"]
        assert dV <= 1e-12, (k, it["n"], dV)  # same work items: only the order of FP64 atomics differs
    _say("  pack vs single grids: max rel dE %.1e, max dV %.1e" % worst)


def test_alkanes64_pack_autotuned(alkanes64):
    """After the autotuner picked the Vxc work length and the lane count: E and N_e are unchanged (they do
    not depend on either), V changes only by the FP32 partial sums of the new work items."""
    a = alkanes64
    its = a["items"]
    chunk, ms = a["pack"].autotune(a["xc"], [it["dD"].ptr for it in its], [it["dV"].ptr for it in its], reps=1)
    E, N, V = _run(a)
    _say("  autotuned: Vxc length %d, %d lanes, times %s" % (chunk, a["pack"].lanes(), ms))
    for k, it in enumerate(its):
        assert abs(E[k] - it["E1"]) <= 1e-12 * abs(it["E1"])
        assert np.abs(V[k] - it["V1"]).max() <= 5e-5
        assert np.abs(V[k] - V[k].T).max() <= 1e-12 * max(1.0, np.abs(V[k]).max())


def test_alkanes64_members_vs_oracle(alkanes64):
    its = alkanes64["items"]
    # two mid-size members: the molecules whose carbon counts are closest to 6 and to 9
    pick = sorted({min(range(len(its)), key=lambda k: (abs(its[k]["n"] - t), k)) for t in (6, 9)})
    for k in pick:
        it = its[k]
        orc = ol.OracleSystem(it["Z"], it["xyz"], "SG-1", "def2-svp")
        orc.set_functional("PBE")
        orc.get_ao()
        Eo, No, Vo = orc.xc(it["D"])
        na = len(it["Z"])
        dE, dV = abs(it["E1"] - Eo) / na, np.abs(it["V1"] - Vo).max()
        _say("  C%d (%d AOs) vs oracle: |dE|/atom %.2e max|dV| %.2e" % (it["n"], it["mol"].nbasis, dE, dV))
        assert dE <= 1e-5
        assert dV <= 5e-5
        assert abs(it["N1"] - No) <= 1e-6 * No
