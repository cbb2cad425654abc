"""Benchmark: one Kohn-Sham XC build per step on one MI355X.

Default workload: (H2O)_128 cluster (seeded, 15 A cube, O-O >= 2.7 A) /
def2-SVP (3072 AOs) / SG-2 grid (2.9 M points) / r2SCAN meta-GGA, closed
shell, D = 2 C C^T from a seeded S-orthonormal 3072 x 640 C. One step = one
XC build in the reference's order: getDensity(D) -> Evaluate("ev") ->
getEnergy -> getFock. `value` is the median over the timed steps with D and V
resident in HBM; the same build with D uploaded from and V downloaded to host
memory (the reference's by-value calls) is reported beside it.

Both GPU forms are timed on the same inputs: the plain FP64 form (fewer steps:
one of its builds takes seconds) and the chosen mixed-precision form
(split-BF16 MFMA contractions with FP32 accumulation; functional, E_xc, N_e
and V sums in FP64). The CPU baseline is the FP64 CPU restatement of the
reference loop nests on the host cores: on a large workload it runs a seeded
sample of its 128-point sub-grids and is scaled to the whole grid
("extrapolated"), since one full build of it takes about an hour.

Prints one JSON line: the metric, the per-kernel roofline of the dominant
kernel (HIP events on the library's stream), the mixed form's deviation from
the plain FP64 form, and the CPU baseline.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_MFMA_PEAK_TFLOPS = 78.6   # MI355X dense FP64 matrix
# MI355X dense BF16 matrix peak (2.5 PFLOP/s) / 3: a split product takes three BF16 MFMAs.
BF16_SPLIT_PEAK_TFLOPS = 2500.0 / 3.0
HBM_PEAK_GBS = 8000.0
DEFAULT_WORKLOAD = "h2o128_def2svp_sg2_r2scan"
CPU_SAMPLE_SUBGRIDS = 64


def _say(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def _traffic_table(workload):
    """HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated) recorded for this workload."""
    out = {}
    path = os.path.join(ROOT, "profiles", workload + "_hbm_traffic.md")
    if os.path.exists(path):
        for line in open(path):
            if line.startswith("traffic_bytes "):
                k, v = line[len("traffic_bytes "):].rsplit(None, 1)  # kernel names may hold spaces
                out[k.strip()] = int(v)
    return out, os.path.relpath(path, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=3, help="timed builds of the plain FP64 form")
    ap.add_argument("--plain-warmup", type=int, default=1)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=CPU_SAMPLE_SUBGRIDS,
                    help="sub-grids of the CPU baseline's sample on large workloads (sub-grids x nbasis^2 above 2e9): "
                         "every sub-grid carries all nbasis AOs, so each costs the same and a small sample fixes the rate")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run a plain streaming copy (hipMemcpy device to device) that reads and writes this "
                         "many MiB, three times: the known-size reference for the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--ao-cache", type=int, default=-1, help="-1 auto, 0 recompute AO values, 1 cache them")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="meta-GGA workloads: skip the learned-functional leg (training + builds with it)")
    ap.add_argument("--surrogate-epochs", type=int, default=400)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 leg")
    ap.add_argument("--no-host-io", action="store_true", help="skip the host-memory D/V leg")
    ap.add_argument("--scf", action="store_true",
                    help="instead of the build timing: the full Kohn-Sham SCF of the workload (r2SCAN: fitted J from "
                         "the device Coulomb engine, ADIIS -> CDIIS), mixed form from the core guess, then the plain "
                         "FP64 form from the mixed form's converged Fock matrix; one JSON line with both energies")
    ap.add_argument("--scf-mixed-tol", type=float, default=1e-6,
                    help="--scf: the mixed form's CDIIS tolerance on max|G| (the reference's is 1e-6)")
    ap.add_argument("--scf-max-iter", type=int, default=300, help="--scf: iterations per DIIS phase")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(the path would shard the grid batches across GPUs, broadcast D and all-reduce V)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "alkanes64":
        bench_pack(args)
        return
    if args.scf:
        bench_scf(args)
        return

    import numpy as np

    import chinium_amd as ca
    from chinium_amd import hip
    from chinium_amd import workloads as wl

    w = wl.build_workload(args.workload)
    natoms = len(w["Z"])
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    n = mol.nbasis
    _say("%s: %d atoms, %d AOs; density matrix" % (args.workload, natoms, n))
    D, S = wl.workload_density(mol, w["nocc"], w["seed"])
    xc = ca.ExchangeCorrelation()
    xc.Read(w["functional"])
    dD = hip.DeviceArray(n * n)
    dD.upload(np.asfortranarray(D).ravel(order="F"))

    def build(precision):
        t0 = time.time()
        grid = ca.Grid(mol, w["grid"], 1, 0)
        grid.setPrecision(precision)
        grid.setAOCache(args.ao_cache)
        grid.setType(xc.family)
        grid.getAO(0, 0)
        hip.device_synchronize()
        return grid, time.time() - t0

    def time_steps(grid, steps, warmup, dV):
        """Per-step times (HIP events on the library's stream) of `steps` builds with D and V in HBM."""
        stream = grid.stream()

        def step():
            grid.getDensityDevice(dD.ptr, 1, n)
            xc.Evaluate("ev", grid)
            e = grid.getEnergy()
            grid.getFockDevice(dV.ptr, 1, n)
            return e

        for _ in range(warmup):
            step()
        hip.device_synchronize()
        ev = [hip.Event() for _ in range(steps + 1)]
        wall0 = time.perf_counter()
        ev[0].record(stream)
        for k in range(steps):
            E = step()
            ev[k + 1].record(stream)
        hip.device_synchronize()
        wall = time.perf_counter() - wall0
        ms = [ev[k].elapsed_ms(ev[k + 1]) for k in range(steps)]
        return ms, E, 1e3 * wall / steps

    def time_host_io(grid, steps, warmup):
        """Per-step wall times of the by-value calls: D from host memory, V back to host memory."""
        def step():
            grid.getDensity([D])
            xc.Evaluate("ev", grid)
            e = grid.getEnergy()
            v = grid.getFock()[0]
            return e, v
        for _ in range(warmup):
            step()
        ms = []
        for _ in range(steps):
            t0 = time.perf_counter()
            step()
            ms.append(1e3 * (time.perf_counter() - t0))
        return ms

    def kernel_pass(grid, dV, reps):
        grid.profile(True)
        grid.kernel_times(reset=True)
        for _ in range(reps):
            grid.getDensityDevice(dD.ptr, 1, n)
            xc.Evaluate("ev", grid)
            grid.getEnergy()
            grid.getFockDevice(dV.ptr, 1, n)
        kt = grid.kernel_times(reset=True)
        grid.profile(False)
        return kt

    def measure(precision, steps, warmup, tune):
        grid, t_setup = build(precision)
        _say("%s form: grid ready in %.1f s (AO tables %s)" % ("mixed" if precision else "FP64", t_setup,
                                                               "cached" if grid.aoCached() else "recomputed"))
        dV = hip.DeviceArray(n * n)
        tuned = grid.autotuneTiles(xc, dD.ptr, dV.ptr, reps=2) if tune and precision == ca.CHX_PREC_MIXED else None
        ms, E, wall = time_steps(grid, steps, warmup, dV)
        _say("%s form: median %.3f ms per build over %d steps" % ("mixed" if precision else "FP64",
                                                                   statistics.median(ms), steps))
        kt = kernel_pass(grid, dV, max(1, min(3, steps)))
        V = dV.download().reshape((n, n), order="F")
        r = dict(grid=grid, E=E, V=V, N=grid.getNumElectrons(), ms=statistics.median(ms), ms_all=ms,
                 mean_ms=sum(ms) / len(ms), wall_ms=wall, setup_s=t_setup, kt=kt, stats=grid.stats(),
                 cached=grid.aoCached(), tuned=tuned, dV=dV, core_flops=grid.coreFlops())
        return r

    traffic, traffic_src = _traffic_table(args.workload)
    nc = 4 if xc.family >= 2 else 1

    def roofline(r, peak, names):
        """Dominant stage's kernel: algorithmic FLOPs per launch / mean launch time (HIP events)."""
        flops_stage = r["stats"]["flops_per_build"] / 2.0  # each stage: half the build's contractions
        kt = r["kt"]
        dens_ms = kt["density"][0] / max(1, kt["density"][1])
        vxc_ms = kt["vxc"][0] / max(1, kt["vxc"][1])
        launches = names["density_launches"]
        dom, dom_ms = ("density", dens_ms / launches) if dens_ms >= vxc_ms else ("vxc", vxc_ms)
        flops_kernel = flops_stage / (launches if dom == "density" else 1)
        if dom == "density" and nc == 1:  # LDA / GGA: the core batches' groups run in their own launch
            flops_kernel -= r["core_flops"]
        achieved = flops_kernel / (dom_ms * 1e-3) / 1e12
        name = names[dom]
        return {"kernel": name, "bound": "mfma", "achieved": round(achieved, 3), "peak": round(peak, 2),
                "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": traffic.get(name),
                "traffic_source": traffic_src if name in traffic else None,
                "avg_ms": round(dom_ms, 4), "flops_per_launch": flops_kernel,
                "flops_source": "chx_grid_stats: k*2*sum_b g_b n_b^2 over the screening plan, half per stage"}

    plain = None
    if not args.no_plain:
        plain = measure(ca.CHX_PREC_FP64, args.plain_steps, args.plain_warmup, tune=False)
        del plain["grid"], plain["dV"]  # free the FP64 form's tables before the mixed form's
    chosen = measure(ca.CHX_PREC_MIXED, args.steps, args.warmup, tune=True)
    host_io = None if args.no_host_io else time_host_io(chosen["grid"], max(5, min(args.steps, 20)), 1)
    st = chosen["stats"]
    flops_build = st["flops_per_build"]
    ms_step = chosen["ms"]
    names64 = {"density": ("k_density_c64<0..3>" if nc == 4 else "k_density_c64<0>") if plain and plain["cached"]
               else "k_density_f64",
               "vxc": ("k_vxc_c64<%d>" % nc) if plain and plain["cached"] else "k_vxc_f64",
               "density_launches": 4 if nc == 4 else 1}
    vxc_mx = chosen["grid"].vxcKernel() + "<%d>" % nc
    # the density stage: one launch (meta-GGA: rho, grad rho and tau fused, every group; LDA / GGA: the
    # groups of the other batches); the core batches' three-term launch is timed on its own
    namesmx = {"density": "k_density_mx_mgga" if nc == 4 else "k_density_mx<0>",
               "vxc": vxc_mx, "density_launches": 1}

    out = {
        "metric": "Vxc build ms/SCF-iter & TFLOP/s, chosen vs plain FP64 GPU at |dE_xc|<=1e-5 Eh/atom",
        "value": round(ms_step, 4),
        "unit": "ms/SCF-iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "ms_per_step_stat": "median of the timed steps (HIP events on the library's stream), D and V in HBM",
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3-split/f32-acc (f64 functional and sums)",
        "data": "synthetic (seeded geometry and density matrix)",
        "config": {"workload": wl.DESCRIPTIONS.get(args.workload, args.workload), "name": args.workload,
                   "points": st["points"], "nbasis": n, "atoms": natoms, "batches": st["batches"],
                   "mean_alive_tiles_per_batch": round(st["mean_tiles"], 2),
                   "form": "mixed: split-BF16 MFMA contractions, FP32 accumulation, FP64 functional/E/N/V sums",
                   "ao_tables": "cached in HBM (bf16 hi/lo pairs)", "screening_threshold": 1e-14},
        "ms_per_step_mean": round(chosen["mean_ms"], 4),
        "ms_per_step_min_max": [round(min(chosen["ms_all"]), 4), round(max(chosen["ms_all"]), 4)],
        "ms_per_step_host_io": None if host_io is None else round(statistics.median(host_io), 4),
        "host_io_note": "median wall time of the by-value calls getDensity({D}) / getFock() with D and V in host "
                        "memory (PCIe included); not the value",
        "speedup_vs_plain_fp64": round(plain["ms"] / ms_step, 3) if plain else None,
        "autotune": {"vxc_shape": chosen["tuned"][0], "vxc_chunk": chosen["tuned"][1],
                     "shapes": {"0": "128 x 128 blocks over 128-AO super-tile pairs",
                                "1": "256 x 256 blocks over 8-tile groups of each chunk's alive-tile union"},
                     "ms": {"%d/%d" % k: round(v, 4) for k, v in chosen["tuned"][2].items()}}
        if chosen["tuned"] else None,
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
    if plain is not None:
        out["plain_fp64"] = {
            "ms_per_step": round(plain["ms"], 4),
            "steps": args.plain_steps, "warmup": args.plain_warmup,
            "tflops": round(flops_build / (plain["ms"] * 1e-3) / 1e12, 3),
            "kernels_ms": {k: round(v[0] / max(1, v[1]), 4) for k, v in plain["kt"].items()},
            "roofline": roofline(plain, FP64_MFMA_PEAK_TFLOPS, names64),
            "ao_tables": "cached in HBM (FP64)" if plain["cached"] else "recomputed in kernels",
            "E_xc": plain["E"], "N_e": plain["N"],
        }
        out["accuracy_vs_plain_fp64"] = {
            "dE_xc": abs(plain["E"] - chosen["E"]), "dE_xc_per_atom": abs(plain["E"] - chosen["E"]) / natoms,
            "max_dV": float(np.abs(plain["V"] - chosen["V"]).max()), "dN_e": abs(plain["N"] - chosen["N"]),
            "note": "the plain FP64 form matches the FP64 oracle to 1e-13 Eh (tests/test_gpu_parity.py, "
                    "test_gpu_functionals.py)"}
    E, V, N = chosen["E"], chosen["V"], chosen["N"]

    if xc.family >= 2 and not args.no_surrogate:
        out["surrogate"] = surrogate_leg(args, xc, chosen, time_steps, plain, natoms)

    if not args.no_cpu_baseline:
        out["cpu_baseline"], acc = cpu_baseline(args, w, D, E, V, N, natoms)
        if acc:
            out["accuracy_vs_fp64_oracle"] = acc
    if args.copy_probe_mib:
        nb = args.copy_probe_mib * (1 << 20)
        a, b = hip.DeviceArray(nb // 8), hip.DeviceArray(nb // 8)
        for _ in range(3):
            hip.copy_device(b.ptr, a.ptr, nb)
        hip.device_synchronize()
        out["copy_probe"] = {"kernel": "the runtime's device-to-device copy (hipMemcpy)", "read_bytes": nb,
                             "write_bytes": nb}
    print(json.dumps(out), flush=True)


def surrogate_leg(args, xc, chosen, time_steps, plain, natoms):
    """Learned functional: a correction on an analytic LDA base (Slater + VWN5), trained on the r2SCAN
    samples of the other workloads' grids (all 64 alkanes, the water monomer and (H2O)16: not this
    molecule), then attached to the chosen form's grid for timed builds of this workload. Reports its
    E_xc error against the analytic r2SCAN, that of the LDA base alone (the method without the
    network), and the functional stage of each."""
    import statistics as _st

    import chinium_amd as ca
    from chinium_amd.surrogate import LDA_BASE, Surrogate, workload_samples
    t0 = time.perf_counter()
    rho_s, sig_s, tau_s, w_s = workload_samples(n_alkanes=64)
    t_samples = time.perf_counter() - t0
    net = Surrogate(11, base=LDA_BASE)
    t0 = time.perf_counter()
    loss = net.fit(xc, rho_s, sig_s, tau_s, w_s, epochs=args.surrogate_epochs)
    t_train = time.perf_counter() - t0
    del rho_s, sig_s, tau_s, w_s
    grid = chosen["grid"]

    def functional_ms(x):
        grid.profile(True)
        grid.kernel_times(reset=True)
        for _ in range(5):
            x.Evaluate("ev", grid)
        ms, n = grid.kernel_times(reset=True)["functional"]
        grid.profile(False)
        return ms / max(1, n)

    lda = ca.ExchangeCorrelation().from_codes(list(LDA_BASE))
    lda.Evaluate("ev", grid)
    E_lda = grid.getEnergy()
    t_lda, t_r2 = functional_ms(lda), functional_ms(xc)
    xc.setSurrogate(net)
    try:
        ms, E, _ = time_steps(grid, max(3, min(args.steps, 10)), 1, chosen["dV"])
        fall = grid.fallbackCount()
        t_net = functional_ms(xc)
    finally:
        xc.setSurrogate(None)
    err, err_b = abs(E - chosen["E"]) / natoms, abs(E_lda - chosen["E"]) / natoms
    return {
        "network": "3-64-64-64-1 tanh on (ln rho, s/(1+s), alpha/(1+alpha)) predicting F - F_LDA; FP32 weights, "
                   "split-BF16 MFMA inference",
        "base": "Slater + VWN5 (libxc 1, 7)",
        "training": {"samples": net.ndata, "epochs": args.surrogate_epochs, "batch": 16384,
                     "sample_s": round(t_samples, 3), "train_s": round(t_train, 3),
                     "loss_first": float(loss[0]), "loss_last": float(loss[-1])},
        "ms_per_step": round(_st.median(ms), 4),
        "fallback_points": fall,
        "E_xc": E,
        "dE_xc_per_atom_vs_analytic": err,
        "dE_xc_per_atom_base_alone": err_b,
        "gain_over_base": round(err_b / err, 1) if err > 0 else None,
        "functional_ms": {"analytic_r2scan": round(t_r2, 4), "lda_base": round(t_lda, 4),
                          "base_plus_network": round(t_net, 4)},
    }


def cpu_baseline(args, w, D, E, V, N, natoms):
    """The FP64 CPU restatement of the reference loop nests on this host's cores. A workload with more than
    4x `cpu_sample` sub-grids runs a seeded sample of them and scales the time by nsubgrids / sample."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    from chinium_amd import workloads as wl
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    orc = ol.OracleSystem(w["Z"], w["xyz_bohr"], w["grid"], w["basis"], nthreads)
    orc.set_functional(w["functional"])
    nsub = orc.nsubgrids
    sample = nsub if nsub * float(orc.nbasis) ** 2 <= 2e9 else min(nsub, args.cpu_sample)
    if sample < nsub:  # the first `sample` sub-grids of a seeded shuffle
        rng = wl.SplitMix64(4242)
        keys = sorted((rng.next_u64(), i) for i in range(nsub))
        orc.set_subset(sorted(i for _, i in keys[:sample]))
    _say("CPU baseline: %d of %d sub-grids on %d threads" % (sample, nsub, nthreads))
    orc.get_ao()
    t1 = time.perf_counter()
    Eo, No, Vo = orc.xc(D)
    cpu_ms = 1e3 * (time.perf_counter() - t1)
    scale = nsub / sample
    base = {"value": round(cpu_ms * scale, 2), "unit": "ms/SCF-iter", "cores": nthreads, "kind": "port",
            "measured_ms": round(cpu_ms, 2), "subgrids_timed": sample, "subgrids_total": nsub}
    if sample < nsub:
        base["sample"] = ("extrapolated: one XC build (density, functional, E_xc, Vxc) over a seeded sample of "
                          "%d of the %d 128-point sub-grids, scaled by %d/%d; every sub-grid carries all %d AOs, "
                          "so the scaling is exact in expectation; AO tables cached beforehand, as the reference "
                          "does once per SCF" % (sample, nsub, nsub, sample, orc.nbasis))
        return base, None
    base["sample"] = "full grid, one XC build (AO tables cached beforehand, as the reference does once per SCF)"
    acc = {"dE_xc": abs(E - Eo), "dE_xc_per_atom": abs(E - Eo) / natoms, "max_dV": float(np.abs(V - Vo).max()),
           "dN_e": abs(N - No)}
    return base, acc


def bench_scf(args):
    """Kohn-Sham SCF to convergence in both forms on the same molecule, grid type and fitted J:
    |E_mixed - E_FP64| per atom is the north star's SCF-level accuracy figure."""
    import numpy as np

    import chinium_amd as ca
    from chinium_amd import workloads as wl
    from chinium_amd.scf import RestrictedDFSCF

    w = wl.build_workload(args.workload)
    natoms = len(w["Z"])
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    xc = ca.ExchangeCorrelation()
    xc.Read(w["functional"])
    t0 = time.perf_counter()
    jf = ca.CoulombFit(mol)
    info = jf.info()
    jf.int1e()
    jf.metric()
    t_ints = time.perf_counter() - t0
    _say("Coulomb engine: %s; integrals and metric in %.1f s" % (info, t_ints))
    out = {"metric": "Kohn-Sham SCF total energy, chosen (mixed) vs plain FP64 GPU XC form",
           "unit": "Eh", "n_gpus": 1, "data": "synthetic (seeded geometry)",
           "config": {"workload": wl.DESCRIPTIONS.get(args.workload, args.workload), "name": args.workload,
                      "nbasis": mol.nbasis, "atoms": natoms, "nocc": w["nocc"], "coulomb": "density-fitted J, def2-svp-j, "
                      "Coulomb metric", "aux_functions": info["naux"], "primitive_pairs": info["pairs"],
                      "primitive_pairs_unscreened": info["pairs_unscreened"]},
           "integrals_s": round(t_ints, 2)}
    F0 = None
    for prec, tol, name in ((ca.CHX_PREC_MIXED, args.scf_mixed_tol, "mixed"), (ca.CHX_PREC_FP64, 1e-6, "fp64")):
        t0 = time.perf_counter()
        grid = ca.Grid(mol, w["grid"], 1, 0)
        grid.setPrecision(prec)
        grid.setType(xc.family)
        grid.getAO(0, 0)
        t_grid = time.perf_counter() - t0
        scf = RestrictedDFSCF(mol, w["nocc"], xc, grid, coulomb=jf)
        t0 = time.perf_counter()
        r = scf.run(F0=F0, tol=tol, max_iter=args.scf_max_iter, log=_say)
        t_scf = time.perf_counter() - t0
        F0 = r["F"]
        hist = r["history"]
        out[name] = {"E": r["E"], "E_xc": r["E_xc"], "iterations": r["iterations"], "diis_error": r["diis_error"],
                     "tol": tol, "start": "core guess" if name == "mixed" else "the mixed form's converged Fock matrix",
                     "last_dE": hist[-1][1] - hist[-2][1] if len(hist) > 1 else None,
                     "E_spread_last3": (max(h[1] for h in hist[-3:]) - min(h[1] for h in hist[-3:])
                                        if len(hist) >= 3 else None),
                     "scf_s": round(t_scf, 2), "grid_s": round(t_grid, 2),
                     "s_per_iter": {k: round(float(np.median(v)), 4) for k, v in scf.timing.items()}}
        _say("%s: E %.10f in %d iterations (%.1f s)" % (name, r["E"], r["iterations"], t_scf))
        del scf, grid
    out["dE_per_atom"] = abs(out["mixed"]["E"] - out["fp64"]["E"]) / natoms
    out["value"] = out["dE_per_atom"]
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

    # one after another: each grid through the single-grid C ABI, D and V in HBM; every grid's build
    # is enqueued (on the grid's own stream) before any energy is read, so the grids overlap
    def seq_step():
        for it in items:
            nb = it["mol"].nbasis
            it["grid"].getDensityDevice(it["dD"].ptr, 1, nb)
            xc.Evaluate("ev", it["grid"])
            it["grid"].getFockDevice(it["dV"].ptr, 1, nb)
        return [it["grid"].getEnergy() for it in items]

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
    form, lanes = pack.form(), pack.lanes()
    stream = pack.stream()

    def time_pack():
        for _ in range(args.warmup):
            pack.xc(xc, Dp, Vp)
        hip.device_synchronize()
        e0, e1 = hip.Event(), hip.Event()
        w0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            E, N = pack.xc(xc, Dp, Vp)
        e1.record(stream)
        hip.device_synchronize()
        return e0.elapsed_ms(e1) / args.steps, 1e3 * (time.perf_counter() - w0) / args.steps, E, N

    # every form at the tuned length and lanes for the record, then the autotuner's choice (the value)
    modes = {}
    for f in (0, 1, 2):
        pack.set_lanes(lanes)
        pack.set_form(f)
        modes["%s%s" % (ca.Pack.FORMS[f], "" if f == 1 else ", %d lanes" % lanes)] = round(time_pack()[0], 4)
    pack.set_lanes(lanes)
    pack.set_form(form)
    ms_pack, wall_pack, E_pack, N_pack = time_pack()
    _say("pack: autotuned %s; %s" % (ca.Pack.FORMS[form], modes))
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
                   "form": "mixed, " + {0: "per-molecule kernels on staggered lanes",
                                        1: "one persistent launch per stage over every molecule's items",
                                        2: "one launch per stage per lane, lanes staggered"}[form] +
                           ", hipGraph replay (autotuned)",
                   "lanes": None if form == 1 else lanes, "vxc_chunk": chunk,
                   "vxc_wide_from_npad": pack.vxcWideFrom()},
        "pack_forms_ms": modes,
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
