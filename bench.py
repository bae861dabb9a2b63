#!/usr/bin/env python3
"""CCSD(T) benchmark on one MI355X: the mixed-precision, learned-guess form
against the plain FP64 form on the same Hamiltonian, in the same run.

One step = one complete CCSD(T) solve (initial guess, CCSD iterations, (T))
of the workload's Hamiltonian, which is resident on the device before the
timed region starts.  Prints ONE JSON line.

    python bench.py [--steps K] [--warmup W] [--workload h2o8|h2o4|h2o]
                    [--form mixed|fp64|fp32|bf16x3] [--guess learned|mp2] [--no-cpu]

--form mixed (the default, the headline): the CCSD iterations run with
split-BF16x3 contractions (FP32 accumulation) until ||R|| < 1e-5 (or the
low-precision residual stops falling), then in FP64 until ||R|| < 1e-7 (the
reference's default R_CONVERGENCE); (T) uses BF16x3 for W3 with FP64
energy sums; the initial T2 comes from the learned guess.  The FP64 form is
solved to ||R|| < 1e-10 in the same run, and value is the mixed form's wall
time; |dE_CCSD|, |dE_(T)| and |dE_total| against the FP64 form and the
speedup are reported with it.  The other forms time one form alone.

The CPU baseline is the FP64 oracle (a C restatement of the reference's
loops, OpenMP over the outer loops) timed on a bounded sample of the same
workload: one CCSD iteration restricted to its first occupied row and the
(T) of a sample of triplets, extrapolated to the full solve with the FP64
iteration count.
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
# published dense MI355X peaks of the arithmetic each form executes,
# TFLOP/s: BF16x3 runs on the BF16 MFMAs (three BF16 products per
# algorithmic product, so algorithmic flops can reach at most a third of the
# BF16 peak; that ceiling is reported beside the published-peak fraction)
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}
MFMA_PER_PRODUCT = {"fp64": 1, "fp32": 1, "bf16x3": 3}
FP64_CONV = 1e-10
# HBM bytes per launch of a kernel label on a workload, from two rocprofv3
# --pmc runs (FETCH_SIZE, WRITE_SIZE) scaled by the counters' reading for a
# plain copy of known size (fetch x2.0, write x1.0); the runs and their
# commands are in profiles/ (filled in when measured for the current
# kernels)
TRAFFIC = {("h2o8", "triples_w:bf16x3"): 179.2e9}


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def ladder_flops(o, v):
    """The packed symmetric / antisymmetric ladder: two GEMMs over the pairs
    i>=j, a>=b, e>=f and i>j, a>b, e>f, ~o^2 v^4 / 2 (a quarter of the full
    2 o^2 v^4)."""
    q, p, qs, ps = o * (o + 1) / 2.0, v * (v + 1) / 2.0, o * (o - 1) / 2.0, v * (v - 1) / 2.0
    return 2.0 * (q * p * p + qs * ps * ps)


def zmbij_flops(o, v):
    """Zmbij = sum_ef <mb|ef> tau_ij^ef on the same pairs: ~o^3 v^3 (half the
    full 2 o^3 v^3)."""
    q, p, qs, ps = o * (o + 1) / 2.0, v * (v + 1) / 2.0, o * (o - 1) / 2.0, v * (v - 1) / 2.0
    return 2.0 * o * v * (q * p + qs * ps)


def woooo_tau_flops(o, v):
    """The tau term of Woooo, sum_ef <mn|ef> tau_ij^ef, on the same pairs:
    ~o^4 v^2 (half the full 2 o^4 v^2)."""
    q, p, qs, ps = o * (o + 1) / 2.0, v * (v + 1) / 2.0, o * (o - 1) / 2.0, v * (v - 1) / 2.0
    return 2.0 * o * o * (q * p + qs * ps)


def ccsd_flops(o, v):
    """Algorithmic flops of one CCSD iteration (2 per multiply-add), with the
    ladder, Zmbij and the tau term of Woooo counted as executed (packed
    pairs)."""
    return (ladder_flops(o, v) + zmbij_flops(o, v) + woooo_tau_flops(o, v) + 12.0 * o**3 * v**3
            + 2.0 * o**4 * v**2 + 20.0 * o**2 * v**3)


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
    t0 = time.perf_counter()
    k = 0
    while True:
        t = (k * 7919) % ntrip
        orc.tcorr_tjl_range(t, t + 1)
        k += 1
        if time.perf_counter() - t0 > budget_s / 2 or k >= ntrip:
            break
    t_trip = (time.perf_counter() - t0) / k
    # serial leg (the reference is single-threaded): the same loops on one
    # thread over a few triplets; the serial CCSD iteration is the OpenMP
    # one scaled by the measured serial / OpenMP ratio of the (T) loops
    orc.set_threads(1)
    ks = 2
    t0 = time.perf_counter()
    for q in range(ks):
        t = (q * 7919 + 13) % ntrip
        orc.tcorr_tjl_range(t, t + 1)
    t_trip1 = (time.perf_counter() - t0) / ks
    orc.set_threads(threads)
    orc.close()
    total = niter * t_iter + ntrip * t_trip
    ratio = t_trip1 / t_trip
    serial = dict(value=niter * t_iter * ratio + ntrip * t_trip1, unit="s", cores=1, kind="port",
                  sample="FP64 oracle on 1 thread: (T) over %d triplets (%.2f s each, x%.1f the %d-thread time); "
                         "CCSD iterations estimated as the %d-thread row sample times that ratio; extrapolated "
                         "to %d iterations + all %d triplets" % (ks, t_trip1, ratio, threads, threads, niter, ntrip))
    return dict(value=total, unit="s", cores=threads, kind="port",
                sample="FP64 oracle (OpenMP, %d threads): one CCSD iteration restricted to %d of %d occupied "
                       "rows (x%d) and (T) over %d of %d triplets, extrapolated to %d iterations + all triplets"
                       % (threads, rows, no, no // rows, k, ntrip, niter),
                serial=serial, full_h2o=cpu_full_h2o())


def cpu_full_h2o():
    """The single-water configuration (o=5, v=19) solved completely by the
    oracle: CCSD to 1e-10 and (T), on one thread and on all."""
    from oracle import Oracle
    from ugacc_amd.chem import molecules
    from ugacc_amd.chem.scf import Reference
    ref = Reference(molecules.to_bohr(molecules.water()), "cc-pvdz")
    out = {}
    for label in ("serial", "openmp"):
        orc = Oracle(ref.fock, ref.full_ints(), ref.no, omp=True)
        threads = orc.threads
        if label == "serial":
            orc.set_threads(1)
        t0 = time.perf_counter()
        r = orc.ccsd(maxiter=100, conv=FP64_CONV)
        et = orc.tcorr_tjl()
        out[label + "_s"] = time.perf_counter() - t0
        out[label + "_threads"] = 1 if label == "serial" else threads
        orc.set_threads(threads)
        orc.close()
    out.update(iterations=r["niter"], e_ccsd=r["eccsd"], e_t=et,
               sample="complete FP64 oracle CCSD(T) of H2O/cc-pVDZ (o=5, v=19), R_CONVERGENCE 1e-10")
    return out


def _post_gpu(ugx, ref, reps=2):
    """The steps after a converged CCSD(T) on the GPU (FP64, R_CONVERGENCE
    1e-10): Lambda (HBAR first), the densities with their energies, the (T)
    gradient terms, and one dipole perturbation (mu_z, omega = 0) solved
    right- and left-hand with its linresp; seconds of the last of `reps`
    passes (the first one tunes tiles), each step bracketed by
    synchronisations."""
    c = ugx.Context(0)
    c.set_autotune(True)
    c.load_hamiltonian(ref.no, ref.nv, ref.fock, ref.blocks)
    r = c.ccsd(maxiter=100, r_convergence=FP64_CONV, wfn_ccsd_t=1)
    mu = ref.perturbation("Mu")[2]

    def timed(fn):
        c.synchronize()
        t0 = time.perf_counter()
        x = fn()
        c.synchronize()
        return x, time.perf_counter() - t0

    out = {}
    for _ in range(reps):
        lam, t_lam = timed(lambda: c.lambda_solve(maxiter=100, r_convergence=FP64_CONV))
        den, t_den = timed(lambda: c.density(which=()))
        _, t_tg = timed(lambda: c.tgrad(which=()))
        g, t_pc = timed(lambda: c.pert(mu, 0.0))
        xr, t_x = timed(lambda: g.solve(0, maxiter=100, r_convergence=FP64_CONV))
        yr, t_y = timed(lambda: g.solve(1, maxiter=100, r_convergence=FP64_CONV))
        alpha, t_lr = timed(lambda: g.linresp(g))
        g.release()
        out = dict(lambda_s=t_lam, lambda_iterations=lam["niter"], density_s=t_den, tgrad_s=t_tg,
                   pert_create_s=t_pc, pert_right_s=t_x, pert_right_iterations=xr["niter"], pert_left_s=t_y,
                   pert_left_iterations=yr["niter"], linresp_s=t_lr, alpha_zz=alpha,
                   e_density=den["e_one"] + den["e_two"], e_ccsd=r["e_ccsd"],
                   converged=bool(lam["converged"] and xr["converged"] and yr["converged"]))
    c.close()
    return out


def _post_oracle(ref, left_iterations):
    """The same steps with the oracle on one thread (the reference's
    serial loops, restated), at the same convergence; the left-hand solve
    (its Y2 terms are o^4 v^4 loops, ~2 s per iteration at H2O) is a
    three-iteration sample scaled to `left_iterations`."""
    from oracle import Oracle
    o = Oracle(ref.fock, ref.full_ints(), ref.no)
    assert o.ccsd(maxiter=100, conv=FP64_CONV)["converged"]
    o.build_tstar()
    mu = ref.perturbation("Mu")[2]
    out = {}
    t0 = time.perf_counter()
    o.lambda_init()
    lam = o.lambda_solve(maxiter=100, conv=FP64_CONV)
    out["lambda_s"] = time.perf_counter() - t0
    out["lambda_iterations"] = lam["niter"]
    t0 = time.perf_counter()
    o.density()
    out["density_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.tgrad(enable=False)
    out["tgrad_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    P = o.pert(mu, 0.0)
    out["pert_create_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    xr = P.solve(0, maxiter=100, conv=FP64_CONV)
    out["pert_right_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    yr = P.solve(1, maxiter=3, conv=FP64_CONV)
    out["pert_left_s"] = (time.perf_counter() - t0) * left_iterations / max(1, yr["niter"])
    out["pert_left_sample"] = "%d iterations timed, scaled to %d" % (yr["niter"], left_iterations)
    t0 = time.perf_counter()
    out["alpha_zz"] = P.linresp(P)
    out["linresp_s"] = time.perf_counter() - t0
    out.update(pert_right_iterations=xr["niter"], cores=1, kind="port")
    out.pop("alpha_zz")  # Y is not converged in the sample
    P.close()
    o.close()
    return out


def post_ccsd_measure(cpu=True):
    """Lambda, densities, (T) gradient terms and the linear response after a
    converged CCSD(T): the GPU at (H2O)4/cc-pVDZ (o=20, v=76) and at
    H2O/cc-pVDZ (o=5, v=19), where the serial oracle runs the same steps
    beside it (its (H2O)4 steps take minutes to hours: not sampled here)."""
    from ugacc_amd import ugx
    from ugacc_amd.chem import molecules
    from ugacc_amd.chem.scf import Reference
    out = {}
    h2o = Reference(molecules.to_bohr(molecules.water()), "cc-pvdz", log=None)
    for label, ref in (("h2o4", Reference(molecules.water_cluster(4, seed=1), "cc-pvdz", log=None)), ("h2o", h2o)):
        t0 = time.perf_counter()
        out[label] = _post_gpu(ugx, ref)
        log("post-CCSD steps %s on the GPU: %s (%.1f s with the CCSD and the tuning pass)"
            % (label, ", ".join("%s %.3f" % (k, v) for k, v in out[label].items() if k.endswith("_s")),
               time.perf_counter() - t0))
    if cpu:
        out["h2o_oracle"] = _post_oracle(h2o, out["h2o"]["pert_left_iterations"])
        log("post-CCSD steps h2o, oracle (1 thread): %s"
            % ", ".join("%s %.3f" % (k, v) for k, v in out["h2o_oracle"].items() if k.endswith("_s")))
    out["note"] = ("seconds per step after a converged FP64 CCSD(T) (R_CONVERGENCE 1e-10): Lambda with HBAR, "
                   "the densities and their energies, the (T) gradient terms, one dipole perturbation (mu_z, "
                   "omega 0) right- and left-hand and its linresp; GPU = the library (FP64), oracle = the "
                   "restated reference loops on one thread")
    return out


def library_shapes(wl, reps=10):
    """The vendor libraries through torch on the workload's GEMM shapes (one
    plain product each, random operands): BF16 torch.bmm (hipBLASLt) on the
    (T) W3 launch shapes with their 160-wide tile padding, batch 64 of
    v^2 x 160 x 2(v + o) and batch 4096 of 160 x 160 x 4(v + o); FP64 matmul
    (rocBLAS) on the ring (ov x ov x ov) and the packed ladder
    (o(o+1)/2 x v(v+1)/2 x v(v+1)/2); BF16 matmul (hipBLASLt) on the ring
    shape, whose BF16x3 form the CCSD's low-precision iterations run."""
    import torch
    o, v = wl["no"], wl["nv"]

    def run(shape_a, shape_b, dtype, batched):
        a = torch.randn(*shape_a, device="cuda", dtype=dtype)
        b = torch.randn(*shape_b, device="cuda", dtype=dtype)
        f = (lambda: torch.bmm(a, b)) if batched else (lambda: a @ b)
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        m, k = shape_a[-2], shape_a[-1]
        n = shape_b[-1]
        bat = shape_a[0] if batched else 1
        return dict(shape=[bat, m, n, k], ms=dt * 1e3, tflops=2.0 * bat * m * n * k / dt * 1e-12)

    v2, ov, pv, po = v * v, o * v, v * (v + 1) // 2, o * (o + 1) // 2
    return {
        "bf16_bmm_w3_za": run((64, v2, 2 * (v + o)), (64, 2 * (v + o), 160), torch.bfloat16, True),
        "bf16_bmm_w3_merged": run((4096, 160, 4 * (v + o)), (4096, 4 * (v + o), 160), torch.bfloat16, True),
        "fp64_ring": run((ov, ov), (ov, ov), torch.float64, False),
        "fp64_ladder": run((po, pv), (pv, pv), torch.float64, False),
        "bf16_ring": run((ov, ov), (ov, ov), torch.bfloat16, False),
        "note": "one BF16 product per call (the library kernel's BF16x3 W3 runs three); torch.bmm is hipBLASLt, "
                "the FP64 matmul rocBLAS",
    }


def cold_solve(ugx, no, nv, fock, blocks, opts, net, label):
    """The first ugx_ccsd after ugx_load_hamiltonian on a fresh context:
    every per-Hamiltonian conversion (integral copies in each precision,
    BF16 panels, (T) operand tables) and allocation happens inside it.  The
    tuned tiles come from the process's table (ugx_load_tile_table makes
    that the case in a new process)."""
    c = ugx.Context(0)
    c.set_autotune(True)
    if net is not None:
        c.set_guess_network(net)
    t0 = time.perf_counter()
    c.load_hamiltonian(no, nv, fock, blocks)
    c.synchronize()
    t1 = time.perf_counter()
    c.prepare(opts)
    c.synchronize()
    t2 = time.perf_counter()
    r = c.ccsd(opts)
    c.synchronize()
    t3 = time.perf_counter()
    c.close()
    log("%s cold: load %.2f s, prepare %.2f s, solve %.2f s" % (label, t1 - t0, t2 - t1, t3 - t2))
    return dict(load_s=t1 - t0, prepare_s=t2 - t1, first_solve_s=t3 - t1, converged=bool(r["converged"]),
                device_allocs_in_iterations=r["device_allocs"])


def run_form(ctx, opts, warmup, steps, label):
    for w in range(warmup):
        r = ctx.ccsd(opts)
        log("%s warmup %d: %d iterations (%d low), %.2f s CCSD + %.2f s (T)" % (
            label, w, r["niter"], r["niter_low"], r["seconds_ccsd"], r["seconds_triples"]))
    ctx.set_profiling(True)
    ctx.synchronize()
    walls, res = [], None
    for s in range(steps):
        t0 = time.perf_counter()
        res = ctx.ccsd(opts)
        ctx.synchronize()
        walls.append(time.perf_counter() - t0)
        if not res["converged"]:
            raise RuntimeError("%s form did not converge" % label)
        log("%s step %d: %.3f s (%d iterations, %d low; CCSD %.3f s, (T) %.3f s)" % (
            label, s, walls[-1], res["niter"], res["niter_low"], res["seconds_ccsd"], res["seconds_triples"]))
    stats = ctx.kernel_stats()
    ctx.set_profiling(False)
    return float(np.mean(walls)), res, stats


def roofline(stats, workload):
    """The kernel label with the most device time, against the published
    peak of the arithmetic it executes: achieved = algorithmic flops per
    launch / average launch time (HIP events on the library's stream)."""
    dom = max((k for k in stats if stats[k]["gflop"] > 0), key=lambda k: stats[k]["ms"])
    st = stats[dom]
    form = dom.split(":")[1] if ":" in dom else "fp64"
    ach = st["gflop"] / st["ms"]  # GFLOP/ms = TFLOP/s
    names = {"triples_w": "k_gemm_b3d (BF16x3) / k_gemm (FP64, FP32) (T) W3 grouped GEMM",
             "ladder": "k_gemm particle-particle ladder", "ring": "k_gemm ring contraction"}
    m = MFMA_PER_PRODUCT[form]
    return dict(bound="mfma", achieved=ach, peak=PEAK[form], unit="TFLOP/s", frac=ach / PEAK[form],
                traffic=TRAFFIC.get((workload, dom)),
                kernel="%s (%s)" % (names.get(dom.split(":")[0], dom), dom), launches=st["launches"],
                avg_ms=st["ms"] / max(1, st["launches"]), flop_per_launch=st["gflop"] * 1e9 / max(1, st["launches"]),
                mfma_products_per_flop=m, executed_tflops=ach * m, frac_of_form_ceiling=ach * m / PEAK[form])


def batch_problems():
    """The 64 uneven geometries: 32 distorted H2O monomers (seeds 100-131)
    and 32 (H2O)2 dimers (seeds 200-231), cc-pVDZ, host front end."""
    from ugacc_amd.chem import molecules
    from ugacc_amd.chem.scf import Reference
    t0 = time.perf_counter()
    geoms = [molecules.distorted_monomer(100 + k) for k in range(32)]
    geoms += [molecules.random_dimer(200 + k) for k in range(32)]
    problems = []
    for atoms in geoms:
        ref = Reference(atoms, "cc-pvdz", log=None)
        problems.append((ref.no, ref.nv, ref.fock, ref.blocks))
    log("front end for %d geometries (untimed): %.1f s" % (len(problems), time.perf_counter() - t0))
    return problems


def batch_cpu_baseline(problems):
    """Full FP64 oracle solves (OpenMP) of the first monomer and the first
    dimer, CCSD to 1e-10 and (T), extrapolated to the 64 problems."""
    from oracle import Oracle
    from ugacc_amd.hamiltonian import full_from_blocks
    times = {}
    threads = 1
    for q in (0, 32):
        no, nv, fock, blocks = problems[q]
        orc = Oracle(fock, full_from_blocks(no, nv, blocks), no, omp=True)
        threads = orc.threads
        t0 = time.perf_counter()
        orc.ccsd(maxiter=100, conv=FP64_CONV)
        orc.tcorr_tjl()
        times[q] = time.perf_counter() - t0
        orc.close()
    total = 32 * times[0] + 32 * times[32]
    return dict(value=len(problems) / total, unit="solves/s", cores=threads, kind="port",
                seconds_monomer=times[0], seconds_dimer=times[32],
                sample="FP64 oracle (OpenMP, %d threads): complete CCSD (1e-10) + (T) of one monomer (%.2f s) and one "
                       "dimer (%.2f s), extrapolated to 32 of each" % (threads, times[0], times[32]))


def batch_measure(args, problems, warmup, steps, cpu=True):
    """ugx_batch_solve of the 64 geometries, FP64 and mixed forms: solves/s,
    per-solve seconds (p50, p99), each result checked against the problem's
    own single-geometry FP64 solve; kernel statistics of the mixed form."""
    from ugacc_amd import ugx
    ctx = ugx.Context(0)
    ctx.set_autotune(not args.no_autotune)
    # single-geometry FP64 references, solved one after another on one
    # context (the second pass, after tuning, is timed: what the batch
    # replaces)
    for rep in range(2):
        single = []
        t0 = time.perf_counter()
        for no, nv, fock, blocks in problems:
            ctx.load_hamiltonian(no, nv, fock, blocks)
            r = ctx.ccsd(r_convergence=FP64_CONV, wfn_ccsd_t=1, maxiter=100)
            single.append(r["e_ccsd_final"] + r["e_t"])
        t_seq = time.perf_counter() - t0
    log("sequential single FP64 solves: %.3f s" % t_seq)
    graph = 0 if args.no_graph else 1
    pmode = args.batch_polish_mode
    forms = {"fp64": dict(precision=ugx.PREC_FP64, r_convergence=FP64_CONV),
             "mixed": dict(precision=ugx.PREC_NAMES[args.low], r_convergence=args.mixed_conv,
                           polish_threshold=args.polish,
                           polish_mode=ugx.POLISH_REFINE if pmode == "refine" else ugx.POLISH_PLAIN)}
    out, stats = {}, None
    for name, kw in forms.items():
        for w in range(warmup):
            ctx.batch_solve(problems, nworkers=args.workers, wfn_ccsd_t=1, maxiter=100, graph=graph, **kw)
        ctx.set_profiling(name == "mixed")
        walls, per, allocs = [], [], 0
        for st in range(steps):
            res, wall = ctx.batch_solve(problems, nworkers=args.workers, wfn_ccsd_t=1, maxiter=100, graph=graph, **kw)
            if not all(r["converged"] for r in res):
                raise RuntimeError("batch form %s did not converge" % name)
            walls.append(wall)
            per += [r["seconds_ccsd"] + r["seconds_triples"] for r in res]
            allocs = max(allocs, max(r["device_allocs"] for r in res))
            dE = max(abs(r["e_ccsd_final"] + r["e_t"] - e) for r, e in zip(res, single))
            log("batch %s step %d: %.3f s for %d solves, max |dE| vs single FP64 solves %.1e" % (
                name, st, wall, len(res), dE))
        if name == "mixed":
            stats = ctx.kernel_stats()
        ctx.set_profiling(False)
        wall = float(np.mean(walls))
        out[name] = dict(wall_s=wall, solves_per_s=len(problems) / wall, p50_s=float(np.percentile(per, 50)),
                         p99_s=float(np.percentile(per, 99)), max_dE_vs_single_fp64=dE,
                         iterations_mean=float(np.mean([r["niter"] for r in res])),
                         device_allocs_in_iterations=allocs)
    ctx.close()
    res = dict(metric="batched CCSD(T) solves/s, 32 distorted H2O + 32 (H2O)2, cc-pVDZ",
               value=out["mixed"]["solves_per_s"], unit="solves/s", steps=steps, warmup=warmup,
               workload="64 uneven geometries: 32 x (o=5, v=19) + 32 x (o=10, v=38)",
               workers=args.workers, graph=bool(graph), autotune=not args.no_autotune,
               mixed_r_convergence=args.mixed_conv, mixed_polish=pmode, fp64_r_convergence=FP64_CONV,
               forms=out, sequential_single_fp64_s=t_seq, kernels=stats,
               roofline=roofline(stats, "batch64") if stats else None,
               cpu_baseline=batch_cpu_baseline(problems) if cpu else None)
    return res


def run_batch(args):
    """--workload batch64: the batch line alone."""
    problems = batch_problems()
    b = batch_measure(args, problems, args.warmup, args.steps, cpu=not args.no_cpu)
    line = {"metric": b["metric"], "value": b["value"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": b["forms"]["mixed"]["wall_s"] * 1e3, "higher_is_better": True,
            "scaling": "strong", "vs_baseline": None, "dtype": "%s/f64" % args.low,
            "data": "seeded distorted monomers (O-H U[0.90,1.10] A, angle U[95,115] deg, seeds 100-131) and dimers "
                    "(O-O U[2.7,3.2] A, random rotations, seeds 200-231), cc-pVDZ, host front end (untimed)",
            "config": {k: b[k] for k in ("workload", "workers", "graph", "autotune", "mixed_r_convergence",
                                         "mixed_polish", "fp64_r_convergence")},
            "forms": b["forms"], "sequential_single_fp64_s": b["sequential_single_fp64_s"],
            "kernels": b["kernels"], "roofline": b["roofline"], "cpu_baseline": b["cpu_baseline"]}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="h2o8", choices=sorted(WORKLOADS) + ["batch64"],
                    help="batch64: 32 distorted H2O monomers + 32 (H2O)2 dimers solved as one batch")
    ap.add_argument("--workers", type=int, default=8, help="batch64: host worker threads (streams in flight)")
    ap.add_argument("--no-graph", action="store_true", help="batch64: eager iterations instead of hipGraph replays")
    ap.add_argument("--form", default="mixed", choices=["mixed", "fp64", "fp32", "bf16x3"])
    ap.add_argument("--low", default="bf16x3", choices=["bf16x3", "fp32"],
                    help="arithmetic of the low-precision phase of the mixed form")
    ap.add_argument("--guess", default="learned", choices=["learned", "mp2"])
    ap.add_argument("--mixed-conv", type=float, default=1e-7)
    ap.add_argument("--polish", type=float, default=1e-5)
    ap.add_argument("--polish-mode", default=None, choices=["refine", "plain"],
                    help="mixed form after the switch: refinement polish (FP64 residual at the switch point, then "
                         "low-precision difference residuals) or plain FP64 residuals; default refine for the "
                         "single-geometry workloads, plain for batch64 (launch-bound small systems, where the "
                         "refinement's extra launches cost more than its FP64 savings)")
    ap.add_argument("--no-plain-compare", action="store_true",
                    help="mixed form: skip the extra timed solve with the plain FP64 polish")
    ap.add_argument("--ref-steps", type=int, default=None,
                    help="timed solves of each comparison form (FP64 at 1e-10, FP64 at the mixed threshold, the "
                         "plain polish); default min(steps, 3)")
    ap.add_argument("--ref-warmup", type=int, default=1)
    ap.add_argument("--no-batch", action="store_true", help="skip the batch64 object of the default line")
    ap.add_argument("--no-cold", action="store_true", help="skip the cold (first-solve) measurements")
    ap.add_argument("--no-post", action="store_true",
                    help="skip the post-CCSD object (Lambda, densities, (T) gradient, response timings)")
    ap.add_argument("--batch-polish-mode", default="plain", choices=["refine", "plain"])
    ap.add_argument("--tiles", default=None, help="tuned tile table to load first (ugx_load_tile_table)")
    ap.add_argument("--save-tiles", default=None, help="write the tuned tile table here after the run")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="size-heuristic GEMM tiles instead of the autotuned ones (tuning runs in the warmup)")
    ap.add_argument("--seed", type=int, default=CLUSTER_SEED)
    ap.add_argument("--synthetic", action="store_true",
                    help="seeded random Hamiltonian of the workload's dimensions instead of the water cluster")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (torch) five times: the reference kernel whose "
                         "FETCH_SIZE / WRITE_SIZE counters calibrate the HBM traffic of the library's kernels")
    ap.add_argument("--library-shapes", action="store_true",
                    help="only time the vendor libraries on this workload's GEMM shapes: hipBLASLt BF16 batched "
                         "products of the two (T) W3 launch shapes, rocBLAS FP64 on the ring and ladder shapes")
    args = ap.parse_args()
    if args.library_shapes:
        print(json.dumps(library_shapes(WORKLOADS[args.workload])), flush=True)
        return
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": x.numel() * 4, "launches": 5}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (the CCSD/(T) sharding is a "
              "design note); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from ugacc_amd import ugx
    from ugacc_amd.synthetic import random_blocks

    if args.ref_steps is None:
        args.ref_steps = max(1, min(args.steps, 3))
    if args.workload == "batch64":
        args.batch_polish_mode = args.polish_mode or args.batch_polish_mode
        return run_batch(args)
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
    if args.tiles:
        ctx.load_tile_table(args.tiles)
    ctx.load_hamiltonian(no, nv, fock, blocks)
    ctx.set_autotune(not args.no_autotune)

    def options(prec, conv, guess, polish_mode=ugx.POLISH_PLAIN):
        o = ugx.default_options()
        o.wfn_ccsd_t = 1
        o.maxiter = 100
        o.r_convergence = conv
        o.precision = prec
        o.polish_threshold = args.polish
        o.polish_mode = polish_mode
        o.guess = guess
        return o

    net = None
    guess = ugx.GUESS_MP2
    if args.form == "mixed" and args.guess == "learned":
        from ugacc_amd.learned import WEIGHTS, load_network
        net = load_network(WEIGHTS)
        ctx.set_guess_network(net)
        guess = ugx.GUESS_MLP

    fp64_opts = options(ugx.PREC_FP64, FP64_CONV, ugx.GUESS_MP2)
    out = {}
    rs, rw = args.ref_steps, args.ref_warmup
    if args.form == "mixed":
        # the comparison forms: FP64 at the parity threshold 1e-10 and at the
        # mixed form's threshold (the equal-threshold ratio)
        wall64, res64, st64 = run_form(ctx, fp64_opts, rw, rs, "fp64 (1e-10)")
        fp64_eq = options(ugx.PREC_FP64, args.mixed_conv, ugx.GUESS_MP2)
        wall64e, res64e, _ = run_form(ctx, fp64_eq, rw, rs, "fp64 (%.0e)" % args.mixed_conv)
        if args.polish_mode is None:
            args.polish_mode = "refine"
        pmode = ugx.POLISH_REFINE if args.polish_mode == "refine" else ugx.POLISH_PLAIN
        if pmode == ugx.POLISH_REFINE and not args.no_plain_compare:
            # the same mixed form with the plain FP64 polish, for comparison
            wallp, resp, _ = run_form(ctx, options(ugx.PREC_NAMES[args.low], args.mixed_conv, guess), rw,
                                      rs, "mixed, plain FP64 polish")
            d_p = abs(resp["e_ccsd_final"] + resp["e_t"] - res64["e_ccsd_final"] - res64["e_t"])
            out.update(plain_polish_wall_s=wallp, plain_polish_iterations=resp["niter"],
                       plain_polish_seconds_ccsd=resp["seconds_ccsd"], plain_polish_speedup_vs_fp64=wall64 / wallp,
                       plain_polish_dE_total=d_p)
        mixed_opts = options(ugx.PREC_NAMES[args.low], args.mixed_conv, guess, pmode)
        wall, res, stats = run_form(ctx, mixed_opts, args.warmup, args.steps, "mixed")
        d_ccsd = abs(res["e_ccsd_final"] - res64["e_ccsd_final"])
        d_t = abs(res["e_t"] - res64["e_t"])
        d_tot = abs(res["e_ccsd_final"] + res["e_t"] - res64["e_ccsd_final"] - res64["e_t"])
        d_eq = abs(res64e["e_ccsd_final"] + res64e["e_t"] - res64["e_ccsd_final"] - res64["e_t"])
        out.update(fp64_same_threshold_wall_s=wall64e, fp64_same_threshold_iterations=res64e["niter"],
                   fp64_same_threshold_r_convergence=args.mixed_conv,
                   fp64_same_threshold_dE_total_vs_1e10=d_eq,
                   speedup_vs_fp64_same_threshold=wall64e / wall)
        out.update(fp64_wall_s=wall64, fp64_iterations=res64["niter"], fp64_seconds_ccsd=res64["seconds_ccsd"],
                   fp64_seconds_triples=res64["seconds_triples"], speedup_vs_fp64=wall64 / wall,
                   dE_ccsd=d_ccsd, dE_t=d_t, dE_total=d_tot, dE_ok=bool(d_tot <= 1e-6),
                   e_ccsd_fp64=res64["e_ccsd_final"], e_t_fp64=res64["e_t"],
                   fp64_tflops=(res64["niter"] * ccsd_flops(no, nv) + triples_flops(no, nv)) / wall64 * 1e-12,
                   fp64_kernels=st64)
        if args.save_tiles:
            ctx.save_tile_table(args.save_tiles)
        if not args.no_cold:
            cm = cold_solve(ugx, no, nv, fock, blocks, mixed_opts, net, "mixed")
            cf = cold_solve(ugx, no, nv, fock, blocks, fp64_opts, None, "fp64 (1e-10)")
            out["cold"] = dict(mixed=cm, fp64=cf, speedup_vs_fp64=cf["first_solve_s"] / cm["first_solve_s"],
                               note="first ugx_ccsd after ugx_load_hamiltonian on a fresh context (tuned tiles "
                                    "from the process's table): prepare (conversions, allocations) + solve")
        niter64 = res64["niter"]
        dtype = "%s/f64" % args.low
        config = {"workload": wl["name"], "no": no, "nv": nv, "form": "mixed", "low_precision": args.low,
                  "autotune": not args.no_autotune,
                  "polish_threshold": args.polish, "polish": args.polish_mode, "r_convergence": args.mixed_conv,
                  "guess": args.guess,
                  "fp64_r_convergence": FP64_CONV, "comparison_form_steps": rs, "comparison_form_warmup": rw}
    else:
        conv = FP64_CONV
        wall, res, stats = run_form(ctx, options(ugx.PREC_NAMES[args.form], conv, ugx.GUESS_MP2), args.warmup,
                                    args.steps, args.form)
        niter64 = res["niter"]
        dtype = "f64" if args.form == "fp64" else args.form
        config = {"workload": wl["name"], "no": no, "nv": nv, "form": args.form, "guess": "mp2",
                  "autotune": not args.no_autotune,
                  "r_convergence": conv}
    tflops = (res["niter"] * ccsd_flops(no, nv) + triples_flops(no, nv)) / wall * 1e-12
    cpu = None
    if not args.no_cpu:
        log("cpu baseline sample")
        cpu = cpu_baseline(no, nv, fock, blocks, niter64)
    line = {
        "metric": METRIC, "value": wall, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": dtype, "data": data, "config": config,
        "tflops": tflops, "iterations": res["niter"], "iterations_low": res["niter_low"],
        "seconds_ccsd": res["seconds_ccsd"], "seconds_triples": res["seconds_triples"],
        "e_ccsd": res["e_ccsd_final"], "e_t": res["e_t"],
    }
    line.update(out)
    line.update({"kernels": stats, "roofline": roofline(stats, args.workload), "cpu_baseline": cpu})
    if args.form == "mixed" and args.workload == "h2o8" and not args.no_post:
        line["post_ccsd"] = post_ccsd_measure(cpu=not args.no_cpu)
    if args.form == "mixed" and args.workload == "h2o8" and not args.no_batch:
        ctx.close()
        args.batch_polish_mode = args.batch_polish_mode or "plain"
        line["batch64"] = batch_measure(args, batch_problems(), rw, rs, cpu=not args.no_cpu)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
