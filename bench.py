#!/usr/bin/env python3
"""CCSD(T) benchmark on one MI355X.

One step = one complete CCSD(T) solve (MP2 guess, CCSD iterations to
R_CONVERGENCE, (T)) of the workload's Hamiltonian, which is resident on the
device before the timed region starts.  Prints ONE JSON line.

    python bench.py [--steps K] [--warmup W] [--workload h2o8|h2o4|h2o]
                    [--form fp64|fp32|bf16x3] [--no-cpu]

The CPU baseline is the FP64 oracle (a C restatement of the reference's
loops, OpenMP over the outer loops) timed on a bounded sample of the same
workload: one CCSD iteration restricted to its first occupied rows and the
(T) of a few triplets, extrapolated to the full solve with the iteration
count the GPU run needed.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "(T)+CCSD wall s, (H2O)8/DZ: mixed+learned vs FP64 GPU at |ΔE|≤1µEh; TFLOP/s"

WORKLOADS = {
    "h2o8": dict(n=8, no=40, nv=152, name="(H2O)8/cc-pVDZ CCSD(T), all-electron (192 MOs, o=40, v=152)"),
    "h2o4": dict(n=4, no=20, nv=76, name="(H2O)4/cc-pVDZ CCSD(T), all-electron (96 MOs, o=20, v=76)"),
    "h2o": dict(n=1, no=5, nv=19, name="H2O/cc-pVDZ CCSD(T), all-electron (24 MOs, o=5, v=19)"),
}
CLUSTER_SEED = 8
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0 / 3.0}  # dense TFLOP/s


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def ccsd_flops(o, v):
    return 2.0 * o**2 * v**4 + 14.0 * o**3 * v**3 + 4.0 * o**4 * v**2 + 20.0 * o**2 * v**3


def triples_flops(o, v):
    ntrip = o * (o + 1) * (o + 2) / 6.0 - o  # i = j = k triplets are skipped
    return 12.0 * ntrip * (v**4 + o * v**3)


def cpu_baseline(no, nv, fock, blocks, niter, budget_s=25.0):
    """Bounded sample of the same workload on the host cores (oracle, OpenMP)."""
    from oracle import Oracle
    from ugacc_amd.hamiltonian import full_from_blocks
    ints = full_from_blocks(no, nv, blocks)
    orc = Oracle(fock, ints, no, omp=True)
    del ints
    threads = orc.threads
    rows = 1
    t0 = time.perf_counter()
    orc.iteration_sample(rows)
    t_iter = (time.perf_counter() - t0) * no / rows
    ntrip = orc.ntriplets()
    # triplets sampled: spread over the triplet list, within the budget
    t0 = time.perf_counter()
    k = 0
    samples = []
    while True:
        t = (k * 7919) % ntrip
        orc.tcorr_tjl_range(t, t + 1)
        samples.append(t)
        k += 1
        if time.perf_counter() - t0 > budget_s / 2 or k >= ntrip:
            break
    t_trip = (time.perf_counter() - t0) / k
    orc.close()
    total = niter * t_iter + ntrip * t_trip
    return dict(value=total, unit="s", cores=threads, kind="port",
                sample="FP64 oracle (OpenMP, %d threads): one CCSD iteration restricted to %d of %d occupied "
                       "rows (x%d) and (T) over %d of %d triplets, extrapolated to %d iterations + all triplets"
                       % (threads, rows, no, no // rows, k, ntrip, niter))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="h2o8", choices=sorted(WORKLOADS))
    ap.add_argument("--form", default="fp64", choices=["fp64", "fp32", "bf16x3"])
    ap.add_argument("--conv", type=float, default=1e-10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=CLUSTER_SEED)
    ap.add_argument("--synthetic", action="store_true",
                    help="seeded random Hamiltonian of the workload's dimensions instead of the water cluster")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (the CCSD/(T) sharding is a "
              "design note in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from ugacc_amd import ugx
    from ugacc_amd.synthetic import random_blocks

    wl = WORKLOADS[args.workload]
    no, nv = wl["no"], wl["nv"]
    if args.synthetic:
        log("generating synthetic Hamiltonian o=%d v=%d" % (no, nv))
        fock, blocks = random_blocks(no, nv, seed=args.seed)
        data = ("synthetic: seeded random closed-shell Hamiltonian (Gram-factor ERIs, diagonal Fock) of the "
                "workload's orbital dimensions")
    else:
        from ugacc_amd.chem import molecules
        from ugacc_amd.chem.scf import Reference
        atoms = molecules.to_bohr(molecules.water()) if wl["n"] == 1 else molecules.water_cluster(wl["n"], args.seed)
        t0 = time.perf_counter()
        ref = Reference(atoms, "cc-pvdz", log=None)
        log("front end (integrals, RHF, MO blocks; untimed): %.1f s, E_SCF = %.10f" % (time.perf_counter() - t0,
                                                                                       ref.escf))
        assert (ref.no, ref.nv) == (no, nv)
        fock, blocks = ref.fock, ref.blocks
        data = ("seeded water cluster (ring, O-O U[2.70,2.90] A, seed %d), cc-pVDZ, integrals and RHF from the "
                "project's host front end (untimed)" % args.seed if wl["n"] > 1 else
                "H2O at 0.9572 A / 104.52 deg, cc-pVDZ, RHF from the project's host front end (untimed)")
    ctx = ugx.Context(0)
    ctx.load_hamiltonian(no, nv, fock, blocks)
    prec = ugx.PREC_NAMES[args.form]
    opts = ugx.default_options()
    opts.wfn_ccsd_t = 1
    opts.maxiter = 100
    opts.r_convergence = args.conv
    opts.precision = prec

    for w in range(args.warmup):
        log("warmup %d" % w)
        r = ctx.ccsd(opts)
        log("warmup done: %d iterations, %.2f s CCSD + %.2f s (T)" % (r["niter"], r["seconds_ccsd"],
                                                                        r["seconds_triples"]))
    ctx.set_profiling(True)
    ctx.synchronize()
    walls, res = [], None
    for s in range(args.steps):
        t0 = time.perf_counter()
        res = ctx.ccsd(opts)
        ctx.synchronize()
        walls.append(time.perf_counter() - t0)
        log("step %d: %.3f s (%d iterations)" % (s, walls[-1], res["niter"]))
    stats = ctx.kernel_stats()
    ctx.set_profiling(False)
    wall = float(np.mean(walls))
    fl_ccsd = res["niter"] * ccsd_flops(no, nv)
    fl_t = triples_flops(no, nv)
    tflops = (fl_ccsd + fl_t) / wall * 1e-12

    # dominant kernel: the (T) W3 build GEMM (one launch per permutation and
    # triplet batch, tile 128x160 used by nothing else)
    dom = "triples_w"
    st = stats.get(dom, {"launches": 1, "ms": float("nan"), "gflop": 0.0})
    ach = st["gflop"] / st["ms"] if st["ms"] > 0 else float("nan")  # GFLOP/ms = TFLOP/s
    peak = PEAK[args.form]
    roof = dict(bound="mfma", achieved=ach, peak=peak, unit="TFLOP/s", frac=ach / peak,
                traffic=None, kernel="k_gemm (T) W3 build, tile 128x160",
                launches=st["launches"], avg_ms=st["ms"] / max(1, st["launches"]),
                flop_per_launch=st["gflop"] * 1e9 / max(1, st["launches"]))
    cpu = None
    if not args.no_cpu:
        log("cpu baseline sample")
        cpu = cpu_baseline(no, nv, fock, blocks, res["niter"])
    line = {
        "metric": METRIC, "value": wall, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f64" if args.form == "fp64" else "f32", "data": data,
        "config": {"workload": wl["name"], "no": no, "nv": nv, "form": args.form, "guess": "mp2",
                   "r_convergence": args.conv},
        "tflops": tflops, "iterations": res["niter"], "seconds_ccsd": res["seconds_ccsd"],
        "seconds_triples": res["seconds_triples"], "e_ccsd": res["e_ccsd"], "e_t": res["e_t"],
        "kernels": stats, "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
