"""Benchmark: one Kohn-Sham XC build per step on one MI355X.

Workload: (H2O)_16 cluster (seeded, 7.5 A cube, O-O >= 2.7 A) / def2-SVP
(384 AOs) / SG-2 grid / B3LYP (XC part), closed shell, D = 2 C C^T from a
seeded S-orthonormal 384 x 80 C. One step = one Vxc build in the reference's
order: getDensity(D) -> Evaluate("ev") -> getEnergy -> getFock, with D and V
in device memory (HBM) and E_xc returned to the host.

Both GPU forms are timed on the same inputs: the plain FP64 form and the
chosen mixed-precision form (split-BF16 MFMA contractions with FP32
accumulation; functional, E_xc, N_e and V sums in FP64). `value` is the
chosen form's time per build; `plain_fp64` holds the FP64 form's figures.

Prints one JSON line (rank 0): the metric, the per-kernel roofline of the
dominant kernel (HIP events on the library's stream), accuracy of both forms
against the FP64 CPU oracle, and the oracle's own time on this host (CPU
baseline).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_MFMA_PEAK_TFLOPS = 78.6   # MI355X dense FP64 matrix
# MI355X dense BF16 matrix peak (2.5 PFLOP/s) / 3: a split product takes three BF16 MFMAs.
BF16_SPLIT_PEAK_TFLOPS = 2500.0 / 3.0
HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="h2o16_def2svp_sg2_b3lyp")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run a plain streaming kernel (torch.neg on float32) that reads and writes this many "
                         "MiB, three times: the known-size reference for the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--ao-cache", type=int, default=-1, help="-1 auto, 0 recompute AO values, 1 cache them")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="meta-GGA workloads: skip the learned-functional leg (training + a build with it)")
    ap.add_argument("--surrogate-epochs", type=int, default=200)
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain FP64 leg (workloads whose FP64 AO tables do not fit in HBM); "
                         "speedup_vs_plain_fp64 is then null")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(the path would shard the grid batches across GPUs, broadcast D and all-reduce V)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "alkanes64":
        bench_pack(args)
        return

    import numpy as np

    import chinium_amd as ca
    from chinium_amd import hip
    from chinium_amd import workloads as wl

    w = wl.build_workload(args.workload)
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    n = mol.nbasis
    D, S = wl.workload_density(mol, w["nocc"], w["seed"])
    xc = ca.ExchangeCorrelation()
    xc.Read(w["functional"])
    dD = hip.DeviceArray(n * n)
    dD.upload(np.asfortranarray(D).ravel(order="F"))

    def measure(precision):
        """Builds the grid in the given form and times K XC builds (D and V in HBM)."""
        t0 = time.time()
        grid = ca.Grid(mol, w["grid"], 1, 0)
        grid.setPrecision(precision)
        grid.setAOCache(args.ao_cache)
        grid.setType(xc.family)
        grid.getAO(0, 0)
        t_setup = time.time() - t0
        stream = grid.stream()
        dV = hip.DeviceArray(n * n)
        tuned = grid.autotune(xc, dD.ptr, dV.ptr, reps=2) if grid.aoCached() else None

        def step():
            grid.getDensityDevice(dD.ptr, 1, n)
            xc.Evaluate("ev", grid)
            e = grid.getEnergy()
            grid.getFockDevice(dV.ptr, 1, n)
            return e

        for _ in range(args.warmup):
            step()
        hip.device_synchronize()
        e0, e1 = hip.Event(), hip.Event()
        wall0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            E = step()
        e1.record(stream)
        hip.device_synchronize()
        wall = time.perf_counter() - wall0
        ms_step = e0.elapsed_ms(e1) / args.steps
        # second timed pass with per-kernel HIP events on the library's stream
        grid.profile(True)
        grid.kernel_times(reset=True)
        for _ in range(args.steps):
            step()
        kt = grid.kernel_times(reset=True)
        grid.profile(False)
        V = dV.download().reshape((n, n), order="F")
        r = dict(grid=grid, E=E, V=V, N=grid.getNumElectrons(), ms=ms_step, wall_ms=1e3 * wall / args.steps,
                 setup_s=t_setup, kt=kt, stats=grid.stats(), cached=grid.aoCached(), tuned=tuned)
        del dV
        return r

    traffic = {}  # HBM bytes per launch measured with rocprofv3 --pmc for this workload
    tfile = os.path.join(ROOT, "profiles", "h2o16_def2svp_sg2_b3lyp_hbm_traffic.md")
    if args.workload == "h2o16_def2svp_sg2_b3lyp" and os.path.exists(tfile):
        for line in open(tfile):
            if line.startswith("traffic_bytes "):
                _, k, v = line.split()
                traffic[k] = int(v)

    # contractions per stage: GGA one density and one Vxc launch; meta-GGA four
    # density launches (A = phi, d_x phi, d_y phi, d_z phi) and one Vxc launch of four components
    nmodes = 4 if xc.family >= 2 else 1

    def roofline(r, peak, names):
        """Dominant contraction kernel: algorithmic FLOPs per launch / mean launch time."""
        flops_stage = r["stats"]["flops_per_build"] / 2.0  # each stage: half the build's contractions
        kt = r["kt"]
        dens_ms = kt["density"][0] / max(1, kt["density"][1])
        vxc_ms = kt["vxc"][0] / max(1, kt["vxc"][1])
        dom, dom_ms = ("density", dens_ms) if dens_ms >= vxc_ms else ("vxc", vxc_ms)
        launches = nmodes if dom == "density" else 1
        flops_kernel = flops_stage / launches
        dom_ms /= launches
        achieved = flops_kernel / (dom_ms * 1e-3) / 1e12
        return {"kernel": names[dom], "bound": "mfma", "achieved": round(achieved, 3), "peak": round(peak, 2),
                "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": traffic.get(names[dom]),
                "traffic_source": "profiles/h2o16_def2svp_sg2_b3lyp_hbm_traffic.md" if names[dom] in traffic else None,
                "avg_ms": round(dom_ms, 4), "flops_per_launch": flops_kernel}

    plain = None if args.no_plain else measure(ca.CHX_PREC_FP64)
    if plain is not None:
        del plain["grid"]  # the FP64 form's tables are not needed any more
    chosen = measure(ca.CHX_PREC_MIXED)
    st = chosen["stats"]
    flops_build = st["flops_per_build"]
    ms_step = chosen["ms"]
    nc = 4 if xc.family >= 2 else 1
    names64 = {"density": "k_density_c64<0>" if plain and plain["cached"] else "k_density_f64",
               "vxc": ("k_vxc_c64<%d>" % nc) if plain and plain["cached"] else "k_vxc_f64"}
    namesmx = {"density": "k_density_mx<0>", "vxc": "k_vxc_mx<%d>" % nc}

    out = {
        "metric": "Vxc build ms/SCF-iter & TFLOP/s, chosen vs plain FP64 GPU at |dE_xc|<=1e-5 Eh/atom",
        "value": round(ms_step, 4),
        "unit": "ms/SCF-iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3-split/f32-acc (f64 functional and sums)",
        "data": "synthetic (seeded geometry and density matrix)",
        "config": {"workload": wl.DESCRIPTIONS.get(args.workload, args.workload), "name": args.workload,
                   "points": st["points"], "nbasis": n, "batches": st["batches"],
                   "mean_alive_tiles_per_batch": round(st["mean_tiles"], 2),
                   "form": "mixed: split-BF16 MFMA contractions, FP32 accumulation, FP64 functional/E/N/V sums",
                   "ao_tables": "cached in HBM (bf16 hi/lo pairs)", "screening_threshold": 1e-14},
        "speedup_vs_plain_fp64": round(plain["ms"] / ms_step, 3) if plain else None,
        "autotune": {"vxc_chunk": chosen["tuned"][0], "ms": {str(k): round(v, 4) for k, v in chosen["tuned"][1].items()},
                     "plain_fp64_vxc_chunk": plain["tuned"][0] if plain and plain["tuned"] else None},
        "tflops": round(flops_build / (ms_step * 1e-3) / 1e12, 3),
        "gpoints_per_s": round(st["points"] / (ms_step * 1e-3) / 1e9, 4),
        "flops_per_build": flops_build,
        "kernels_ms": {k: round(v[0] / max(1, v[1]), 4) for k, v in chosen["kt"].items()},
        "wall_ms_per_step": round(chosen["wall_ms"], 4),
        "setup_s": round(chosen["setup_s"], 3),
        "E_xc": chosen["E"],
        "N_e": chosen["N"],
        "roofline": roofline(chosen, BF16_SPLIT_PEAK_TFLOPS, namesmx),
    }
    out["plain_fp64"] = None if plain is None else {
        "ms_per_step": round(plain["ms"], 4),
        "tflops": round(flops_build / (plain["ms"] * 1e-3) / 1e12, 3),
        "kernels_ms": {k: round(v[0] / max(1, v[1]), 4) for k, v in plain["kt"].items()},
        "roofline": roofline(plain, FP64_MFMA_PEAK_TFLOPS, names64),
        "ao_tables": "cached in HBM (FP64)" if plain["cached"] else "recomputed in kernels",
        "E_xc": plain["E"],
        "dE_vs_chosen": abs(plain["E"] - chosen["E"]),
        "max_dV_vs_chosen": float(np.abs(plain["V"] - chosen["V"]).max()),
    }
    E, V, N = chosen["E"], chosen["V"], chosen["N"]

    surr = None
    if xc.family >= 2 and not args.no_surrogate:
        # learned functional: trained on the r2SCAN samples of the H2O, (H2O)16 and alkane workload
        # grids (not this molecule), then attached for a mixed build of this workload
        from chinium_amd.surrogate import Surrogate, workload_samples
        t0 = time.perf_counter()
        rho_s, sig_s, tau_s, w_s = workload_samples(n_alkanes=8)
        t_samples = time.perf_counter() - t0
        net = Surrogate(11)
        t0 = time.perf_counter()
        loss = net.fit(xc, rho_s, sig_s, tau_s, w_s, epochs=args.surrogate_epochs)
        t_train = time.perf_counter() - t0
        xc.setSurrogate(net)
        chosen.pop("grid", None)  # one grid's AO tables in HBM at a time
        with_s = measure(ca.CHX_PREC_MIXED)
        with_s["fallback"] = with_s["grid"].fallbackCount()
        xc.setSurrogate(None)
        surr = dict(net=net, r=with_s)
        out["surrogate"] = {
            "network": "3-64-64-64-1 tanh on (ln rho, s/(1+s), alpha/(1+alpha)), BF16 weights, FP32 arithmetic",
            "training": {"samples": int(rho_s.size), "epochs": args.surrogate_epochs, "batch": 16384,
                         "sample_s": round(t_samples, 3), "train_s": round(t_train, 3),
                         "loss_first": float(loss[0]), "loss_last": float(loss[-1])},
            "ms_per_step": round(with_s["ms"], 4),
            "kernels_ms": {k: round(v[0] / max(1, v[1]), 4) for k, v in with_s["kt"].items()},
            "functional_ms_analytic": round(chosen["kt"]["functional"][0] / max(1, chosen["kt"]["functional"][1]), 4),
            "fallback_points": with_s["fallback"],
            "E_xc": with_s["E"],
            "speedup_vs_plain_fp64": round(plain["ms"] / with_s["ms"], 3) if plain else None,
        }

    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as ol
        nthreads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        orc = ol.OracleSystem(w["Z"], w["xyz_bohr"], w["grid"], w["basis"], nthreads)
        orc.set_functional(w["functional"])
        orc.get_ao()
        t1 = time.perf_counter()
        Eo, No, Vo = orc.xc(D)
        cpu_ms = 1e3 * (time.perf_counter() - t1)
        natoms = len(w["Z"])
        out["cpu_baseline"] = {"value": round(cpu_ms, 2), "unit": "ms/SCF-iter", "cores": nthreads, "kind": "port",
                               "sample": "full grid, one XC build (AO tables cached beforehand, as the reference "
                                         "does once per SCF)"}
        out["accuracy_vs_fp64_oracle"] = {"dE_xc": abs(E - Eo), "dE_xc_per_atom": abs(E - Eo) / natoms,
                                          "max_dV": float(np.abs(V - Vo).max()), "dN_e": abs(N - No)}
        if plain is not None:
            out["plain_fp64"]["accuracy_vs_fp64_oracle"] = {"dE_xc": abs(plain["E"] - Eo),
                                                            "max_dV": float(np.abs(plain["V"] - Vo).max())}
        if surr is not None:
            rs = surr["r"]
            out["surrogate"]["accuracy_vs_fp64_oracle"] = {"dE_xc": abs(rs["E"] - Eo),
                                                           "dE_xc_per_atom": abs(rs["E"] - Eo) / natoms,
                                                           "max_dV": float(np.abs(rs["V"] - Vo).max()),
                                                           "dN_e": abs(rs["N"] - No)}
    if args.copy_probe_mib:
        import torch
        a = torch.ones(args.copy_probe_mib * (1 << 18), dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.neg(a, out=b)
        torch.cuda.synchronize()
        out["copy_probe"] = {"kernel": "torch.neg float32", "read_bytes": a.numel() * 4, "write_bytes": b.numel() * 4}
    print(json.dumps(out), flush=True)


def bench_pack(args):
    """64 linear alkanes C_nH_2n+2 (n seeded-uniform in 1..40) / def2-SVP / SG-1 / PBE, mixed form:
    all molecules' XC builds as one pack (every molecule's kernels on staggered lanes, replayed as
    one hipGraph, Vxc work-item length and lane count chosen by the autotuner) against the same grids
    built one after another through the single-grid calls."""
    import numpy as np

    import chinium_amd as ca
    from chinium_amd import hip
    from chinium_amd import workloads as wl

    t0 = time.time()
    xc = ca.ExchangeCorrelation()
    xc.Read("PBE")
    items = []
    for k, (n, Z, X) in enumerate(wl.alkane_set(64, 2, 40)):
        mol = ca.Molecule(Z, X * ca.ANGSTROM_TO_BOHR, "def2-svp")
        D, S = wl.workload_density(mol, 4 * n + 1, 500 + k)
        grid = ca.Grid(mol, "SG-1", 1, 0)
        grid.setPrecision(ca.CHX_PREC_MIXED)
        grid.setType(1)
        grid.getAO(0, 0)
        nb = mol.nbasis
        dD, dV = hip.DeviceArray(nb * nb), hip.DeviceArray(nb * nb)
        dD.upload(np.asfortranarray(D).ravel(order="F"))
        items.append(dict(n=n, natoms=len(Z), mol=mol, grid=grid, dD=dD, dV=dV, stats=grid.stats()))
        print("alkane %d: C%d, %d AOs, %d points" % (k, n, nb, items[-1]["stats"]["points"]),
              file=sys.stderr, flush=True)
    t_setup = time.time() - t0
    points = sum(it["stats"]["points"] for it in items)
    flops = sum(it["stats"]["flops_per_build"] for it in items)

    # one after another: each grid through the single-grid C ABI, D and V in HBM
    def seq_step():
        E = []
        for it in items:
            nb = it["mol"].nbasis
            it["grid"].getDensityDevice(it["dD"].ptr, 1, nb)
            xc.Evaluate("ev", it["grid"])
            E.append(it["grid"].getEnergy())
            it["grid"].getFockDevice(it["dV"].ptr, 1, nb)
        return E

    for _ in range(args.warmup):
        seq_step()
    hip.device_synchronize()
    w0 = time.perf_counter()
    for _ in range(args.steps):
        E_seq = seq_step()
    hip.device_synchronize()
    ms_seq = 1e3 * (time.perf_counter() - w0) / args.steps
    V_seq = [it["dV"].download() for it in items]

    pack = ca.Pack([it["grid"] for it in items])
    Dp, Vp = [it["dD"].ptr for it in items], [it["dV"].ptr for it in items]
    chunk, tune_ms = pack.autotune(xc, Dp, Vp, reps=2)
    for _ in range(args.warmup):
        pack.xc(xc, Dp, Vp)
    hip.device_synchronize()
    e0, e1 = hip.Event(), hip.Event()
    stream = pack.stream()
    w0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        E_pack, N_pack = pack.xc(xc, Dp, Vp)
    e1.record(stream)
    hip.device_synchronize()
    wall_pack = 1e3 * (time.perf_counter() - w0) / args.steps
    ms_pack = e0.elapsed_ms(e1) / args.steps
    dV = max(float(np.abs(it["dV"].download() - v).max()) for it, v in zip(items, V_seq))
    dE = max(abs(a - b) / it["natoms"] for a, b, it in zip(E_pack, E_seq, items))

    out = {
        "metric": "Vxc build ms/SCF-iter & TFLOP/s, chosen vs plain FP64 GPU at |dE_xc|<=1e-5 Eh/atom",
        "value": round(ms_pack, 4),
        "unit": "ms/SCF-iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_pack, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3-split/f32-acc (f64 functional and sums)",
        "data": "synthetic (seeded geometries and density matrices)",
        "config": {"workload": "64 alkanes C_nH_2n+2, n in 1..40 / def2-SVP / SG-1 / PBE, one pack",
                   "name": "alkanes64", "molecules": len(items), "points": points,
                   "nbasis_total": sum(it["mol"].nbasis for it in items),
                   "form": "mixed, per-molecule kernels on staggered lanes, hipGraph replay",
                   "lanes": pack.lanes(), "vxc_chunk": chunk},
        "sequential_ms_per_step": round(ms_seq, 4),
        "speedup_vs_sequential": round(ms_seq / ms_pack, 3),
        "autotune_ms": {str(k): round(v, 4) for k, v in tune_ms.items()},
        "wall_ms_per_step": round(wall_pack, 4),
        "tflops": round(flops / (ms_pack * 1e-3) / 1e12, 3),
        "gpoints_per_s": round(points / (ms_pack * 1e-3) / 1e9, 4),
        "max_dE_per_atom_pack_vs_sequential": dE,
        "max_dV_pack_vs_sequential": dV,
        "setup_s": round(t_setup, 2),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
