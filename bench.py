"""Benchmark: one Kohn-Sham XC build per step on one MI355X.

Workload: (H2O)_16 cluster (seeded, 7.5 A cube, O-O >= 2.7 A) / def2-SVP
(384 AOs) / SG-2 grid / B3LYP (XC part), closed shell, D = 2 C C^T from a
seeded S-orthonormal 384 x 80 C. One step = one Vxc build in the reference's
order: getDensity(D) -> Evaluate("ev") -> getEnergy -> getFock, with D and V
in device memory (HBM) and E_xc returned to the host.

Prints one JSON line (rank 0): the metric, the per-kernel roofline of the
dominant kernel (HIP events on the library's stream), accuracy against the
FP64 CPU oracle, and the oracle's own time on this host (CPU baseline).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_MFMA_PEAK_TFLOPS = 78.6   # MI355X dense FP64 matrix
HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="h2o16_def2svp_sg2_b3lyp")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ao-cache", type=int, default=-1, help="-1 auto, 0 recompute AO values, 1 cache them")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(the path would shard batches across GPUs and all-reduce V; see DESIGN.md)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import chinium_amd as ca
    from chinium_amd import hip
    from chinium_amd import workloads as wl

    w = wl.build_workload(args.workload)
    mol = ca.Molecule(w["Z"], w["xyz_bohr"], w["basis"])
    n = mol.nbasis
    D, S = wl.workload_density(mol, w["nocc"], w["seed"])
    t0 = time.time()
    grid = ca.Grid(mol, w["grid"], 1, 0)
    xc = ca.ExchangeCorrelation()
    xc.Read(w["functional"])
    grid.setAOCache(args.ao_cache)
    grid.setType(xc.family)
    grid.getAO(0, 0)
    t_setup = time.time() - t0
    st = grid.stats()
    stream = grid.stream()

    dD = hip.DeviceArray(n * n)
    dD.upload(np.asfortranarray(D).ravel(order="F"))
    dV = hip.DeviceArray(n * n)

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
    ms_total = e0.elapsed_ms(e1)
    ms_step = ms_total / args.steps

    # Second timed pass with per-kernel events (for the roofline).
    grid.profile(True)
    grid.kernel_times(reset=True)
    for _ in range(args.steps):
        step()
    kt = grid.kernel_times(reset=True)
    grid.profile(False)

    V = dV.download().reshape((n, n), order="F")
    N = grid.getNumElectrons()

    flops_build = st["flops_per_build"]
    flops_kernel = flops_build / 2.0  # each of the two contractions: 2 * sum_b g_b n_b^2
    dens_ms = kt["density"][0] / max(1, kt["density"][1])
    vxc_ms = kt["vxc"][0] / max(1, kt["vxc"][1])
    dom, dom_ms = ("density", dens_ms) if dens_ms >= vxc_ms else ("vxc", vxc_ms)
    achieved = flops_kernel / (dom_ms * 1e-3) / 1e12
    roofline = {"kernel": "k_density_f64" if dom == "density" else "k_vxc_f64", "bound": "mfma",
                "achieved": round(achieved, 3), "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP64_MFMA_PEAK_TFLOPS, 4), "traffic": None,
                "avg_ms": round(dom_ms, 4), "flops_per_launch": flops_kernel}

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
        "dtype": "f64",
        "data": "synthetic (seeded geometry and density matrix)",
        "config": {"workload": "(H2O)16 / def2-SVP / SG-2 / B3LYP", "name": args.workload,
                   "points": st["points"], "nbasis": n, "batches": st["batches"],
                   "mean_alive_tiles_per_batch": round(st["mean_tiles"], 2), "form": "plain FP64 (chosen = plain)",
                   "ao_tables": "cached in HBM" if grid.aoCached() else "recomputed in kernels",
                   "screening_threshold": 1e-14},
        "tflops": round(flops_build / (ms_step * 1e-3) / 1e12, 3),
        "gpoints_per_s": round(st["points"] / (ms_step * 1e-3) / 1e9, 4),
        "flops_per_build": flops_build,
        "kernels_ms": {k: round(v[0] / max(1, v[1]), 4) for k, v in kt.items()},
        "wall_ms_per_step": round(1e3 * wall / args.steps, 4),
        "setup_s": round(t_setup, 3),
        "E_xc": E,
        "N_e": N,
        "roofline": roofline,
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
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
