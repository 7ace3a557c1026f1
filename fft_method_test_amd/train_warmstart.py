"""Train the warm-start U-Net on solver-generated pairs and evaluate it on held-out media.

    python -m fft_method_test_amd.train_warmstart --n 128 --train-seeds 10000:12000 --epochs 6 \\
        --out net.npz --bf16 --budget 1000
    python -m fft_method_test_amd.train_warmstart --init net.npz --epochs 0 --eval-n 128 \\
        --eval-seeds 20000:20200 --eval-tol 1e-8 --eval-precision mixed --report eval128.json

Pairs: for each training seed a periodic medium (overlapping radius-8 spheres or (12, 4, 4) prolate
ellipsoids, half each, porosity U[0.3, 0.8]; fft_method_test_amd.warmstart.medium_for_seed) and
the force densities of its three unit load directions from the GPU solver at --tol, each scaled to
unit RMS on the support.  The network (one forward pass -> nine channels) is trained by its own
forward / backward / Adam kernels on the squared error against them.  Evaluation: held-out seeds,
the full tensor solve cold and warm-started (fftperm_solve_tensor / fftperm_solve_tensor_warm),
Krylov iterations and wall time to --eval-tol, and the K difference between the two.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _seeds(spec):
    a, b = spec.split(":")
    return range(int(a), int(b))


def evaluate(net, n, seeds, tol, precision, workers=1, media=None):
    """Tensor solves cold vs warm on held-out media: per medium iterations and seconds."""
    import torch

    from . import permeability as P, warmstart as W

    g = P.Grid3D(n, n, n)
    green = P.GreenOperator(g, 1.0, 2, precision)
    rows = []
    src = media if media is not None else W.media(n, list(seeds), workers)
    for k, (sd, mask) in enumerate(src):
        sup = P.ForceSupport(P.BinaryMedium(g, mask.reshape(-1)), P.SUPPORT_INTERFACE, green)
        if sup.num_active_voxels() == 0:
            continue
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        K0, inf0 = P.solve_permeability_tensor(green, tol, 20000)
        t1 = time.perf_counter()
        K1, inf1, used = W.solve_tensor_warm(green, net, tol, 20000)
        t2 = time.perf_counter()
        rows.append(dict(seed=int(sd), cold_it=[int(x.iterations) for x in inf0],
                         warm_it=[int(x.iterations) for x in inf1], used=used, cold_s=t1 - t0, warm_s=t2 - t1,
                         k_rel_diff=float(np.linalg.norm(K1 - K0) / np.linalg.norm(K0)),
                         porosity=float(mask.mean())))
        r = rows[-1]
        log(f"[warmstart] eval {k + 1} seed {sd} n={n}: cold {r['cold_it']} ({r['cold_s']:.3f} s) warm {r['warm_it']} "
            f"({r['warm_s']:.3f} s) used {used} dK {r['k_rel_diff']:.1e}")
    green.close()
    cold = sum(sum(r["cold_it"]) for r in rows)
    warm = sum(sum(r["warm_it"]) for r in rows)
    ratios = [sum(r["warm_it"]) / max(1, sum(r["cold_it"])) for r in rows]
    return dict(n=n, count=len(rows), tol=tol, precision="mixed" if precision == 1 else "fp64",
                cold_iterations=cold, warm_iterations=warm, iteration_ratio=warm / max(1, cold),
                per_medium_ratio_mean=float(np.mean(ratios)) if ratios else None,
                per_medium_ratio_min=float(np.min(ratios)) if ratios else None,
                per_medium_ratio_max=float(np.max(ratios)) if ratios else None,
                cold_s=sum(r["cold_s"] for r in rows), warm_s=sum(r["warm_s"] for r in rows),
                time_ratio=sum(r["warm_s"] for r in rows) / max(1e-12, sum(r["cold_s"] for r in rows)),
                max_k_rel_diff=max((r["k_rel_diff"] for r in rows), default=None), rows=rows)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--train-seeds", default="10000:12000")
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-min", type=float, default=2e-5)
    ap.add_argument("--tol", type=float, default=1e-5, help="solver tolerance of the training targets")
    ap.add_argument("--target-precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--workers", type=int, default=1, help="processes generating the media")
    ap.add_argument("--init", default=None, help="start from saved parameters")
    ap.add_argument("--seed", type=int, default=0, help="initialisation and shuffling seed")
    ap.add_argument("--out", default=None)
    ap.add_argument("--bf16", action="store_true", help="save the parameters as BF16 bit patterns")
    ap.add_argument("--budget", type=float, default=0.0, help="seconds: stop training (and save) past this")
    ap.add_argument("--eval-n", type=int, default=0)
    ap.add_argument("--eval-seeds", default="20000:20200")
    ap.add_argument("--eval-tol", type=float, default=1e-8)
    ap.add_argument("--eval-precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--eval-bench", action="store_true",
                    help="also evaluate on the benchmark medium (256^3 Boolean spheres r 8, porosity 0.6, seed 2001)")
    ap.add_argument("--report", default=None, help="write the training / evaluation summary (JSON)")
    args = ap.parse_args()

    from . import _lib, warmstart as W

    t_start = time.perf_counter()
    _lib.load()
    net = W.WarmStartNet(args.seed)
    if args.init:
        net.load(args.init)
    report = dict(n=args.n, train_seeds=args.train_seeds, epochs=args.epochs, lr=args.lr, lr_min=args.lr_min,
                  tol=args.tol, target_precision=args.target_precision, init=args.init)
    if args.epochs > 0:
        prec = _lib.PRECISION_MIXED if args.target_precision == "mixed" else _lib.PRECISION_FP64
        t0 = time.perf_counter()
        data = W.generate(args.n, _seeds(args.train_seeds), tol=args.tol, workers=args.workers, precision=prec,
                          log=log)
        t1 = time.perf_counter()
        deadline = (t_start + args.budget) if args.budget > 0 else None
        hist = W.train(net, data, args.epochs, lr=args.lr, lr_min=args.lr_min, seed=args.seed, log=log,
                       deadline=deadline)
        t2 = time.perf_counter()
        report.update(pairs=len(data), pair_seconds=t1 - t0, train_seconds=t2 - t1, loss_history=hist)
        log(f"[warmstart] {len(data)} pairs in {t1 - t0:.1f} s, training {t2 - t1:.1f} s")
        del data
        if args.out:
            os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
            net.save(args.out, bf16=args.bf16)
            log(f"[warmstart] saved {args.out}")
    if args.eval_n > 0:
        prec = _lib.PRECISION_MIXED if args.eval_precision == "mixed" else _lib.PRECISION_FP64
        report["eval"] = evaluate(net, args.eval_n, _seeds(args.eval_seeds), args.eval_tol, prec, args.workers)
        e = report["eval"]
        log(f"[warmstart] eval n={e['n']} ({e['count']} media, {e['precision']}, tol {e['tol']:g}): iterations "
            f"warm/cold = {e['warm_iterations']}/{e['cold_iterations']} ({e['iteration_ratio']:.3f}), seconds "
            f"{e['warm_s']:.2f}/{e['cold_s']:.2f} ({e['time_ratio']:.3f})")
    if args.eval_bench:
        from . import geometry

        prec = _lib.PRECISION_MIXED if args.eval_precision == "mixed" else _lib.PRECISION_FP64
        m = geometry.boolean_spheres(256, 8.0, 0.6, 2001).reshape(-1)
        report["eval_bench"] = evaluate(net, 256, [], args.eval_tol, prec, media=[(2001, m)])
        e = report["eval_bench"]
        log(f"[warmstart] bench medium: iterations warm/cold = {e['warm_iterations']}/{e['cold_iterations']} "
            f"({e['iteration_ratio']:.3f}), seconds {e['warm_s']:.2f}/{e['cold_s']:.2f}")
    if args.report:
        os.makedirs(os.path.dirname(os.path.abspath(args.report)), exist_ok=True)
        with open(args.report, "w") as f:
            json.dump(report, f)
    print(json.dumps({k: (v if not isinstance(v, dict) else {a: b for a, b in v.items() if a != "rows"})
                      for k, v in report.items()}), flush=True)


if __name__ == "__main__":
    main()
