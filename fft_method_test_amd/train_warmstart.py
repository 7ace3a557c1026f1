"""Train the warm-start U-Net on solver-generated pairs and evaluate it on held-out media.

    python -m fft_method_test_amd.train_warmstart --n 64 --pairs 192 --epochs 10 --lr 5e-4 --out mse.npz
    python -m fft_method_test_amd.train_warmstart --objective residual --init mse.npz --n 64 \\
        --pairs 96 --epochs 6 --lr 2e-4 --bf16 --out fft_method_test_amd/data/warmstart_unet.npz

(the two commands that produced the committed network)

Pairs: random periodic media (overlapping spheres or prolate ellipsoids of aspect 3, porosity
U[0.3, 0.8], training seeds from --seed) and their direction-x interface force density from the
FP64 GPU solver at --tol.  Evaluation: held-out seeds (--eval-seed) at --eval-n, MINRES iterations
and wall time of the full tensor solve to --eval-tol, cold vs warm-started.
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


def evaluate(net, n, count, seed, tol, precision, media=None):
    """Tensor solves cold vs warm on `count` held-out media (or the given masks): per medium
    iterations and seconds."""
    import torch

    from . import permeability as P, warmstart as W

    rng = np.random.default_rng(seed)
    g = P.Grid3D(n, n, n)
    green = P.GreenOperator(g, 1.0, 2, precision)
    rows = []
    if media is not None:
        count = len(media)
    for i in range(count):
        mask = (W.random_medium(n, rng) if media is None else media[i]).reshape(-1)
        sup = P.ForceSupport(P.BinaryMedium(g, mask), P.SUPPORT_INTERFACE, green)
        if sup.num_active_voxels() == 0:
            continue
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        K0, inf0 = P.solve_permeability_tensor(green, tol, 20000)
        t1 = time.perf_counter()
        K1, inf1, used = W.solve_tensor_warm(green, net, tol, 20000)
        t2 = time.perf_counter()
        rows.append(dict(cold_it=[int(x.iterations) for x in inf0], warm_it=[int(x.iterations) for x in inf1],
                         used=used, cold_s=t1 - t0, warm_s=t2 - t1,
                         k_rel_diff=float(np.linalg.norm(K1 - K0) / np.linalg.norm(K0))))
        log(f"[warmstart] eval {i + 1}/{count} n={n}: cold {rows[-1]['cold_it']} ({t1 - t0:.2f} s) warm "
            f"{rows[-1]['warm_it']} ({t2 - t1:.2f} s) used {used}")
    green.close()
    cold = sum(sum(r["cold_it"]) for r in rows)
    warm = sum(sum(r["warm_it"]) for r in rows)
    return dict(n=n, count=len(rows), tol=tol, cold_iterations=cold, warm_iterations=warm,
                iteration_ratio=warm / max(1, cold), cold_s=sum(r["cold_s"] for r in rows),
                warm_s=sum(r["warm_s"] for r in rows), rows=rows)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--pairs", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--seed", type=int, default=10000)
    ap.add_argument("--init", default=None, help="start from saved parameters")
    ap.add_argument("--objective", choices=["mse", "residual"], default="mse",
                    help="mse: squared error against solver pairs; residual: relative residual of the "
                         "minimal-residual-scaled guess (needs only media, no solves)")
    ap.add_argument("--out", default=None)
    ap.add_argument("--bf16", action="store_true", help="save the parameters as BF16 bit patterns")
    ap.add_argument("--eval-n", type=int, default=0)
    ap.add_argument("--eval-count", type=int, default=4)
    ap.add_argument("--eval-seed", type=int, default=20000)
    ap.add_argument("--eval-tol", type=float, default=1e-6)
    ap.add_argument("--eval-precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--eval-bench", action="store_true",
                    help="also evaluate on the benchmark medium (256^3 Boolean spheres r 8, porosity 0.6, seed 2001)")
    ap.add_argument("--report", default=None, help="write the training / evaluation summary (JSON)")
    args = ap.parse_args()

    from . import _lib, warmstart as W

    _lib.load()
    net = W.WarmStartNet(args.seed)
    if args.init:
        net.load(args.init)
    report = dict(n=args.n, pairs=args.pairs, epochs=args.epochs, lr=args.lr, tol=args.tol, seed=args.seed)
    if args.objective == "residual" and args.pairs > 0 and args.epochs > 0:
        rng = np.random.default_rng(args.seed)
        t0 = time.perf_counter()
        masks = [W.random_medium(args.n, rng).reshape(-1) for _ in range(args.pairs)]
        t1 = time.perf_counter()
        hist = W.train_residual(net, masks, args.n, args.epochs, lr=args.lr, seed=args.seed, log=log)
        t2 = time.perf_counter()
        report.update(objective="residual", media_seconds=t1 - t0, train_seconds=t2 - t1, residual_history=hist)
        if args.out:
            os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
            net.save(args.out, bf16=args.bf16)
            log(f"[warmstart] saved {args.out}")
    elif args.pairs > 0 and args.epochs > 0:
        t0 = time.perf_counter()
        pairs = W.generate_pairs(args.n, args.pairs, args.seed, tol=args.tol, log=log)
        t1 = time.perf_counter()
        hist = W.train(net, pairs, args.n, args.epochs, lr=args.lr, seed=args.seed, log=log)
        t2 = time.perf_counter()
        report.update(pair_seconds=t1 - t0, train_seconds=t2 - t1, loss_history=hist)
        log(f"[warmstart] pairs {t1 - t0:.1f} s, training {t2 - t1:.1f} s")
        if args.out:
            os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
            net.save(args.out, bf16=args.bf16)
            log(f"[warmstart] saved {args.out}")
    if args.eval_n > 0:
        prec = _lib.PRECISION_MIXED if args.eval_precision == "mixed" else _lib.PRECISION_FP64
        report["eval"] = evaluate(net, args.eval_n, args.eval_count, args.eval_seed, args.eval_tol, prec)
        e = report["eval"]
        log(f"[warmstart] eval n={e['n']}: iterations warm/cold = {e['warm_iterations']}/{e['cold_iterations']} "
            f"({e['iteration_ratio']:.3f}), seconds {e['warm_s']:.2f}/{e['cold_s']:.2f}")
    if args.eval_bench:
        from . import geometry

        prec = _lib.PRECISION_MIXED if args.eval_precision == "mixed" else _lib.PRECISION_FP64
        m = geometry.boolean_spheres(256, 8.0, 0.6, 2001)
        report["eval_bench"] = evaluate(net, 256, 1, 0, args.eval_tol, prec, media=[m])
        e = report["eval_bench"]
        log(f"[warmstart] bench medium: iterations warm/cold = {e['warm_iterations']}/{e['cold_iterations']} "
            f"({e['iteration_ratio']:.3f}), seconds {e['warm_s']:.2f}/{e['cold_s']:.2f}")
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f)
    print(json.dumps({k: (v if not isinstance(v, dict) else {a: b for a, b in v.items() if a != "rows"})
                      for k, v in report.items()}))


if __name__ == "__main__":
    main()
