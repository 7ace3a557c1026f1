"""Mint the FP64 oracle k_xx of a subset of the 128-sample uneven-size batch (run in the build
container, not on a GPU).

    OMP_NUM_THREADS=2 python tests/golden/make_c4_fixtures.py

The batch is fft_method_test_amd.batch.c4_batch() (seeds 4000-4127, n in {64, 80, ..., 160},
porosity U[0.3, 0.8], radius-8 spheres or (12, 4, 4) prolate ellipsoids).  The subset is the first
sample of each (size, kind) pair for n in 80, 96, 112, 144, 160: both media kinds at every edge
that the static 32/64/128/256/512 plans do not cover.  Each record: the batch index, n, kind,
porosity, the SHA-256 of the mask bytes, NB, and the x column of K (K[r][0] = +mu <v_r>, mu = 1) of
the oracle's solve_one_direction (src/app/PermeabilityRunner.cpp:173-242) at MINRES tol 1e-9.
Output: tests/golden/c4_fixtures.json.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from fft_method_test_amd import batch  # noqa: E402

OUT = os.path.join(HERE, "c4_fixtures.json")
SIZES = (80, 96, 112, 144, 160)
TOL = 1e-9


def main():
    samples = batch.c4_batch()
    chosen = {}
    for i, (n, mask, kind, phi) in enumerate(samples):
        if n in SIZES and (n, kind) not in chosen:
            chosen[(n, kind)] = i
    recs = []
    if os.path.exists(OUT):
        with open(OUT) as f:
            recs = json.load(f)["samples"]
    done = {r["index"] for r in recs}
    for key in sorted(chosen, key=lambda k: k[0]):
        i = chosen[key]
        if i in done:
            continue
        n, mask, kind, phi = samples[i]
        t = time.time()
        o = oracle.Oracle((n, n, n))
        o.set_medium(mask)
        avg, info = o.solve_direction([1.0, 0.0, 0.0], TOL, 50000)
        rec = {"index": i, "seed": 4000 + i, "n": n, "kind": kind, "porosity_target": phi,
               "porosity": float(mask.mean()), "mask_sha256": hashlib.sha256(mask.tobytes()).hexdigest(),
               "nb": int(o.nb), "tol": TOL, "Kcol_x": [float(a) for a in avg],
               "iterations": int(info.iterations), "converged": int(info.converged), "seconds": time.time() - t}
        o.close()
        recs.append(rec)
        print(f"[mint] sample {i} n={n} {kind}: k_xx {avg[0]:.9e} iterations {info.iterations} "
              f"{time.time() - t:.0f} s", flush=True)
        with open(OUT, "w") as f:
            json.dump({"tol": TOL, "samples": recs}, f, indent=1)


if __name__ == "__main__":
    main()
