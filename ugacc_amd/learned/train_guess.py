"""Trains the learned T2 guess on converged amplitudes that the library
itself produces on seeded water clusters, and evaluates it on held-out ones.

    python -m ugacc_amd.learned.train_guess train [--geoms 256] [--samples 8192] [--out PATH]
    python -m ugacc_amd.learned.train_guess evaluate [--clusters 16] [--weights PATH] [--out PATH]

Training set: (H2O)n, n = 2, 3, 4 in turn, seeds 1000, 1001, ... (the
cluster generator of ugacc_amd.chem.molecules), cc-pVDZ, all electrons.
Each geometry is solved to ||R|| < 1e-8 in FP64 on the GPU (MP2 guess); the
targets are y = (t2 - t2(MP2)) / y_scale on a sample of amplitudes, half
uniform and half drawn with probability proportional to |t2 - t2(MP2)|.
Features come from ugx_features, the network is trained by ugx_mlp_train
(FP32 MFMA forward/backward and Adam on the device).

Evaluation: held-out (H2O)8 clusters (seeds 5000, 5001, ...), each solved
from the MP2 guess and from the learned guess in FP64 (||R|| < 1e-10) and
in the mixed form (BF16x3, FP64 polish, ||R|| < 1e-7): iterations, CCSD
wall seconds, |dE| against the FP64 MP2-guess energy, and the 2-norm
distance of each guess from the converged T2.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from .. import ugx
from ..chem import molecules
from ..chem.scf import Reference
from . import WEIGHTS


def log(msg):
    print("[train_guess] " + msg, flush=True)


def training_geometries(n_geom, seed0=1000):
    return [(2 + k % 3, seed0 + k) for k in range(n_geom)]


def heldout_geometries(n, seed0=5000):
    return [(8, seed0 + k) for k in range(n)]


def denominators(eps, no, nv):
    eo, ev = eps[:no], eps[no:no + nv]
    return eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]


def front_end(n, seed):
    return Reference(molecules.water_cluster(n, seed), "cc-pvdz", log=None)


def collect(ctx, ref, samples, rng):
    """(features, t2 - t2(MP2)) of a sample of the converged amplitudes."""
    no, nv = ref.no, ref.nv
    ctx.load_hamiltonian(no, nv, ref.fock, ref.blocks)
    r = ctx.ccsd(r_convergence=1e-8, wfn_ccsd_t=0, maxiter=100)
    if not r["converged"]:
        raise RuntimeError("training geometry did not converge")
    _, t2 = ctx.amplitudes()
    delta = (t2 - ref.blocks["oovv"] / denominators(np.diag(ref.fock), no, nv)).ravel()
    N = delta.size
    k = min(samples, N)
    uni = rng.choice(N, k // 2, replace=False)
    w = np.abs(delta) / np.abs(delta).sum()
    imp = rng.choice(N, k - k // 2, replace=True, p=w)
    idx = np.unique(np.concatenate([uni, imp]))
    return ctx.features(idx=idx), delta[idx], r["niter"]


def train(args):
    ctx = ugx.Context(0)
    rng = np.random.default_rng(args.seed)
    Xs, ys, meta = [], [], []
    t0 = time.perf_counter()
    geoms = training_geometries(args.geoms)
    for g, (n, seed) in enumerate(geoms):
        ref = front_end(n, seed)
        X, d, it = collect(ctx, ref, args.samples, rng)
        Xs.append(X)
        ys.append(d)
        meta.append(dict(n=n, seed=seed, no=ref.no, nv=ref.nv, iterations=it, samples=len(d)))
        if g % 16 == 15 or g == len(geoms) - 1:
            log("collected %d/%d geometries, %d samples, %.0f s" % (g + 1, len(geoms), sum(len(y) for y in ys),
                                                                   time.perf_counter() - t0))
    ctx.close()
    # last 10% of the geometries: validation
    nval = max(1, len(geoms) // 10)
    Xtr, ytr = np.concatenate(Xs[:-nval]), np.concatenate(ys[:-nval])
    Xva, yva = np.concatenate(Xs[-nval:]), np.concatenate(ys[-nval:])
    shift = Xtr.mean(axis=0)
    scale = Xtr.std(axis=0)
    scale[scale <= 0] = 1.0
    y_scale = float(np.sqrt(np.mean(ytr.astype(np.float64) ** 2)))
    net = ugx.Network(0, seed=args.seed)
    net.set_scaling(shift, scale, y_scale)
    ytr_s, yva_s = (ytr / y_scale).astype(np.float32), (yva / y_scale).astype(np.float32)

    def val_loss():
        out = []
        for a in range(0, len(yva_s), 65536):
            b = min(len(yva_s), a + 65536)
            out.append(net.loss_grad(Xva[a:b], yva_s[a:b])[0] * (b - a))
        return float(np.sum(out) / len(yva_s))

    hist = []
    v0 = val_loss()
    log("training on %d samples (%d validation); initial validation loss %.4f" % (len(ytr_s), len(yva_s), v0))
    t1 = time.perf_counter()
    for lr, ep in ((2e-3, args.epochs // 2), (5e-4, args.epochs // 3), (1e-4, args.epochs - args.epochs // 2
                                                                          - args.epochs // 3)):
        h = net.train(Xtr, ytr_s, epochs=ep, batch=args.batch, lr=lr, seed=args.seed + len(hist))
        hist += [float(x) for x in h]
        log("lr %.0e: %d epochs, train loss %.4f, validation loss %.4f" % (lr, ep, h[-1], val_loss()))
    t_train = time.perf_counter() - t1
    out = args.out
    net.save(out)
    info = dict(geometries=meta, n_train=int(len(ytr_s)), n_val=int(len(yva_s)), y_scale=y_scale,
                shift=shift.tolist(), scale=scale.tolist(), epochs=args.epochs, batch=args.batch,
                train_loss=hist, val_loss_initial=v0, val_loss_final=val_loss(), seconds_train=t_train,
                seconds_total=time.perf_counter() - t0)
    with open(os.path.splitext(out)[0] + ".json", "w") as f:
        json.dump(info, f, indent=1)
    log("saved %s (train %.1f s)" % (out, t_train))


def evaluate(args):
    """Held-out (H2O)8 clusters: FP64 and mixed-precision solves from the MP2
    and the learned guess, and the distance of each guess from the converged
    T2."""
    from . import load_network
    ctx = ugx.Context(0)
    net = load_network(args.weights)
    ctx.set_guess_network(net)
    rows = []
    forms = (("fp64", ugx.PREC_FP64, args.conv), ("mixed", ugx.PREC_BF16X3, args.mixed_conv))
    for n, seed in heldout_geometries(args.clusters):
        t0 = time.perf_counter()
        ref = front_end(n, seed)
        tf = time.perf_counter() - t0
        no, nv = ref.no, ref.nv
        ctx.load_hamiltonian(no, nv, ref.fock, ref.blocks)
        row = dict(seed=seed, no=no, nv=nv, front_end_s=tf)
        for fname, prec, conv in forms:
            for gname, guess in (("mp2", ugx.GUESS_MP2), ("learned", ugx.GUESS_MLP)):
                r = ctx.ccsd(r_convergence=conv, wfn_ccsd_t=0, maxiter=100, guess=guess, precision=prec,
                             polish_mode=ugx.POLISH_REFINE if prec != ugx.PREC_FP64 else ugx.POLISH_PLAIN)
                if not r["converged"]:
                    raise RuntimeError("held-out cluster %d did not converge" % seed)
                key = "%s_%s" % (fname, gname)
                row["iter_" + key] = r["niter"]
                row["s_" + key] = r["seconds_ccsd"]
                row["e_" + key] = r["e_ccsd_final"]
                if fname == "fp64" and gname == "mp2":
                    t1c, t2c = ctx.amplitudes()
        ctx.ccsd(maxiter=0, wfn_ccsd_t=0, guess=ugx.GUESS_MLP)
        t1g, t2g = ctx.amplitudes()
        mp2 = ref.blocks["oovv"] / denominators(np.diag(ref.fock), no, nv)
        if args.ceiling:
            # what a perfect guess could save: the converged T2 with T1 = 0,
            # and the converged T1 with the MP2 T2 (FP64, same convergence)
            for key, a1, a2 in (("exact_t2", np.zeros_like(t1c), t2c), ("exact_t1", t1c, mp2)):
                ctx.set_amplitudes(a1, a2)
                r = ctx.ccsd(r_convergence=args.conv, wfn_ccsd_t=0, maxiter=100, guess=ugx.GUESS_USER)
                row["iter_fp64_" + key] = r["niter"]
        row["t2_err_mp2"] = float(np.linalg.norm(t2c - mp2))
        row["t2_err_learned"] = float(np.linalg.norm(t2c - t2g))
        row["t1_err_mp2"] = float(np.linalg.norm(t1c))  # canonical RHF: the MP2 guess has t1 = 0
        row["t1_err_learned"] = float(np.linalg.norm(t1c - t1g))
        row["dE_fp64"] = abs(row["e_fp64_mp2"] - row["e_fp64_learned"])
        row["dE_mixed"] = max(abs(row["e_mixed_mp2"] - row["e_fp64_mp2"]),
                              abs(row["e_mixed_learned"] - row["e_fp64_mp2"]))
        rows.append(row)
        log("seed %d: |dt1| %.3e -> %.3e, |dt2| %.3e -> %.3e; fp64 iterations %d -> %d (%.2f -> %.2f s); mixed "
            "%d -> %d (%.2f -> %.2f s); |dE| %.1e / %.1e" % (
                seed, row["t1_err_mp2"], row["t1_err_learned"], row["t2_err_mp2"], row["t2_err_learned"],
                row["iter_fp64_mp2"], row["iter_fp64_learned"],
                row["s_fp64_mp2"], row["s_fp64_learned"], row["iter_mixed_mp2"], row["iter_mixed_learned"],
                row["s_mixed_mp2"], row["s_mixed_learned"], row["dE_fp64"], row["dE_mixed"]))
    summary = dict(clusters=len(rows), conv=args.conv, mixed_conv=args.mixed_conv)
    for k in ("iter_fp64_mp2", "iter_fp64_learned", "iter_mixed_mp2", "iter_mixed_learned", "s_fp64_mp2",
              "s_fp64_learned", "s_mixed_mp2", "s_mixed_learned", "t2_err_mp2", "t2_err_learned", "t1_err_mp2",
              "t1_err_learned",
              "iter_fp64_exact_t2", "iter_fp64_exact_t1"):
        if k in rows[0]:
            summary["mean_" + k] = float(np.mean([r[k] for r in rows]))
    summary["max_dE_fp64"] = float(max(r["dE_fp64"] for r in rows))
    summary["max_dE_mixed"] = float(max(r["dE_mixed"] for r in rows))
    summary["rows"] = rows
    if args.out:
        with open(args.out, "w") as f:
            json.dump(summary, f, indent=1)
    log("means: " + ", ".join("%s %.4g" % (k[5:], v) for k, v in summary.items() if k.startswith("mean_")))
    return summary


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--geoms", type=int, default=256)
    t.add_argument("--samples", type=int, default=8192)
    t.add_argument("--epochs", type=int, default=30)
    t.add_argument("--batch", type=int, default=4096)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--out", default=WEIGHTS)
    e = sub.add_parser("evaluate")
    e.add_argument("--clusters", type=int, default=16)
    e.add_argument("--conv", type=float, default=1e-10)
    e.add_argument("--mixed-conv", type=float, default=1e-7)
    e.add_argument("--ceiling", action="store_true",
                   help="also solve from the exact converged T2 (T1 = 0) and from the exact T1 (MP2 T2)")
    e.add_argument("--weights", default=WEIGHTS)
    e.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    if args.cmd == "train":
        train(args)
    else:
        evaluate(args)


if __name__ == "__main__":
    sys.exit(main())
