"""Mint FP64 permeability fixtures with the CPU oracle (run in the build container, not on a GPU).

    OMP_NUM_THREADS=6 python tests/golden/make_k_fixtures.py            # both cases
    OMP_NUM_THREADS=6 python tests/golden/make_k_fixtures.py c128       # one case

Cases (Boolean model of overlapping radius-8 lattice balls, porosity 0.6, seed 2001; the
generator of fft_method_test_amd.geometry):
  c128: 128^3 medium, full 3x3 K (three load directions) at MINRES tol 1e-12 and 1e-10.
  c256: the 256^3 medium of the bench workload, the x column of K (k_xx, k_yx, k_zx) at tol 1e-12.

The solves follow the reference path (PermeabilityRunner::solve_one_direction,
src/app/PermeabilityRunner.cpp:173-242; K = +mu <v>, :772-777) in the oracle's serial reduction
order.  Each record carries the SHA-256 of the mask bytes so that a test can check that the
generator reproduced the same medium.  Output: tests/golden/k_fixtures.json (merged per case).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from fft_method_test_amd import geometry  # noqa: E402

OUT = os.path.join(HERE, "k_fixtures.json")


def _mask(n):
    m = geometry.boolean_spheres(n, 8.0, 0.6, 2001).reshape(-1)
    return m, hashlib.sha256(m.tobytes()).hexdigest()


def _save(key, rec):
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    data[key] = rec
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    print(f"[mint] wrote {key} to {OUT}", flush=True)


def mint_c128():
    n = 128
    mask, sha = _mask(n)
    o = oracle.Oracle((n, n, n))
    o.set_medium(mask)
    rec = {"n": n, "radius": 8.0, "porosity_target": 0.6, "seed": 2001, "mask_sha256": sha, "nb": int(o.nb),
           "mu": 1.0, "p_radius": 2, "support": "interface", "reduction_order": "serial",
           "omp_threads_fft": int(os.environ.get("OMP_NUM_THREADS", "0") or 0)}
    for tol in (1e-12, 1e-10):
        t = time.time()
        K, infos = o.solve_tensor(tol, 50000)
        rec[f"K_tol{tol:.0e}"] = K.reshape(-1).tolist()
        rec[f"iterations_tol{tol:.0e}"] = [int(i.iterations) for i in infos]
        rec[f"converged_tol{tol:.0e}"] = [int(i.converged) for i in infos]
        rec[f"seconds_tol{tol:.0e}"] = time.time() - t
        print(f"[mint] c128 tol {tol:g}: K diag {np.diag(K)} iterations {rec[f'iterations_tol{tol:.0e}']} "
              f"{time.time() - t:.0f} s", flush=True)
    o.close()
    _save("c128", rec)


def mint_c256():
    n = 256
    mask, sha = _mask(n)
    o = oracle.Oracle((n, n, n))
    o.set_medium(mask)
    rec = {"n": n, "radius": 8.0, "porosity_target": 0.6, "seed": 2001, "mask_sha256": sha, "nb": int(o.nb),
           "mu": 1.0, "p_radius": 2, "support": "interface", "reduction_order": "serial", "direction": "x",
           "omp_threads_fft": int(os.environ.get("OMP_NUM_THREADS", "0") or 0)}
    tol = 1e-12
    t = time.time()
    avg, info = o.solve_direction([1.0, 0.0, 0.0], tol, 50000)
    rec["Kcol_x_tol1e-12"] = (1.0 * avg).tolist()  # K[r][0] = +mu <v_r>, mu = 1
    rec["iterations_tol1e-12"] = int(info.iterations)
    rec["converged_tol1e-12"] = int(info.converged)
    rec["seconds_tol1e-12"] = time.time() - t
    print(f"[mint] c256 column x: {avg} iterations {info.iterations} {time.time() - t:.0f} s", flush=True)
    o.close()
    _save("c256", rec)


if __name__ == "__main__":
    which = sys.argv[1:] or ["c128", "c256"]
    for w in which:
        {"c128": mint_c128, "c256": mint_c256}[w]()
