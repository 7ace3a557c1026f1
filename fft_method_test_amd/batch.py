"""Batched permeability solves of many microstructures of uneven sizes on one GPU.

Media of the same grid size form a size class.  A size class is solved by fftperm_solve_batch:
every Krylov iteration advances all of its media, the y and z FFT passes of an iteration are one
launch for the class, the vector and scalar kernels one launch each, and the iterations replay
captured hipGraphs (MIXED precision, CR on the FP32 operator with reliable FP64 updates; each
medium keeps its own support map, state and stopping test, so its result is the one its own
solve_one_direction gives).  Size classes are dealt largest-first to `workers` host threads, each
owning one context per size on its own HIP stream, so classes run concurrently on the GPU.
`mode="per_sample"` solves the media one at a time per worker instead (fftperm_solve_direction; the
only mode for FP64).
"""
from __future__ import annotations

import ctypes as C
import math
import queue
import threading
import time

import numpy as np

from . import _lib, geometry
from ._lib import check


def c4_sample(i, seed0=4000, sizes=(64, 80, 96, 112, 128, 144, 160)):
    """Sample i of the batch alone: (n, fluid mask uint8 n^3, kind, porosity)."""
    return _c4_sample((i, seed0, tuple(sizes)))


def _c4_sample(args):
    i, seed0, sizes = args
    rng = np.random.default_rng(seed0 + i)
    n = int(rng.choice(sizes))
    phi = float(rng.uniform(0.3, 0.8))
    kind = "spheres" if rng.uniform() < 0.5 else "ellipsoids"
    s = int(rng.integers(0, 2 ** 31 - 1))
    if kind == "spheres":
        mask = geometry.boolean_spheres(n, 8.0, phi, s)
    else:
        mask = geometry.boolean_ellipsoids(n, (12.0, 4.0, 4.0), phi, s)
    return n, mask.reshape(-1).astype(np.uint8), kind, phi


def c4_batch(count=128, seed0=4000, sizes=(64, 80, 96, 112, 128, 144, 160), workers=1):
    """The 128-sample batch: cubic n drawn uniformly from `sizes`, porosity U[0.3, 0.8], half
    Boolean spheres of radius 8, half prolate ellipsoids (semi-axes 12, 4, 4, random orientation);
    sample i uses seed seed0 + i.  Returns a list of (n, fluid mask uint8 n^3, kind, porosity).
    workers > 1 generates the media in spawned processes."""
    jobs = [(i, seed0, tuple(sizes)) for i in range(count)]
    if workers <= 1:
        return [_c4_sample(j) for j in jobs]
    import multiprocessing as mp

    with mp.get_context("spawn").Pool(workers) as pool:
        return pool.map(_c4_sample, jobs, chunksize=1)


class _Worker:
    def __init__(self, precision, mu, p_radius):
        self.precision, self.mu, self.p_radius = precision, mu, p_radius
        self.ctx = {}

    def context(self, n):
        h = self.ctx.get(n)
        if h is None:
            g = _lib.Grid(n, n, n, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0)
            h = C.c_void_p()
            check(_lib.load().fftperm_create(C.byref(g), self.mu, self.p_radius, self.precision, C.byref(h)))
            self.ctx[n] = h
        return h

    def close(self):
        for h in self.ctx.values():
            _lib.load().fftperm_destroy(h)
        self.ctx.clear()


BATCH_MAX = 64  # FFTPERM_BATCH_MAX


def solve_batch(samples, tol=1e-6, maxit=20000, precision=_lib.PRECISION_MIXED, workers=4, grad=(1.0, 0.0, 0.0),
                mu=1.0, p_radius=2, prepare=True, mode="batched", max_batch=BATCH_MAX):
    """Solve one load direction for every (n, mask, ...) sample.  Returns (results, seconds) where
    results[i] = dict(n, k (the column of K: +mu <v>), iterations, converged, seconds).  With
    prepare=True every context (and its autotuning) is created before the clock starts."""
    if mode not in ("batched", "per_sample"):
        raise ValueError("mode must be 'batched' or 'per_sample'")
    if mode == "batched" and precision != _lib.PRECISION_MIXED:
        raise ValueError("batched solves run in MIXED precision")
    L = _lib.load()
    ws = [_Worker(precision, mu, p_radius) for _ in range(workers)]
    g = np.asarray(grad, dtype=np.float64)
    results = [None] * len(samples)
    errors = []
    q = queue.Queue()
    if mode == "batched":
        groups = {}
        for i, s_ in enumerate(samples):
            groups.setdefault(s_[0], []).append(i)
        items = []
        for n, idx in groups.items():
            for s0 in range(0, len(idx), max_batch):
                items.append((n, idx[s0:s0 + max_batch]))
        items.sort(key=lambda it: -(it[0] ** 3) * len(it[1]))
    else:
        items = [(samples[i][0], [i]) for i in sorted(range(len(samples)), key=lambda i: -samples[i][0] ** 3)]
    if prepare:
        # contexts (FFT plans, Green tensor, z / y autotuning) and the x-pass autotuning, which
        # runs on the first medium of a grid shape, happen here, one at a time, off the clock
        first = {}
        for s_ in samples:
            first.setdefault(s_[0], s_[1])
        nb = C.c_uint64()
        for w in ws:
            for n, mask in sorted(first.items()):
                check(L.fftperm_set_medium(w.context(n), np.ascontiguousarray(mask, dtype=np.uint8),
                                           _lib.SUPPORT_INTERFACE, C.byref(nb)))
    for it in items:
        q.put(it)

    def run_batch(w, n, idx):
        h = w.context(n)
        masks = [np.ascontiguousarray(samples[i][1], dtype=np.uint8) for i in idx]
        ptrs = (C.c_void_p * len(idx))(*[m.ctypes.data for m in masks])
        k = np.zeros(3 * len(idx))
        infos = (_lib.SolveInfo * len(idx))()
        t0 = time.perf_counter()
        check(L.fftperm_solve_batch(h, len(idx), ptrs, _lib.SUPPORT_INTERFACE, g, tol, maxit, k,
                                    C.cast(infos, C.c_void_p)))
        dt = time.perf_counter() - t0
        for j, i in enumerate(idx):
            results[i] = dict(n=n, k=k[3 * j:3 * j + 3].tolist(), iterations=int(infos[j].iterations),
                              converged=bool(infos[j].converged), seconds=dt, batch=len(idx))

    def run_one(w, n, i):
        mask = samples[i][1]
        h = w.context(n)
        t0 = time.perf_counter()
        nb = C.c_uint64()
        check(L.fftperm_set_medium(h, np.ascontiguousarray(mask, dtype=np.uint8), _lib.SUPPORT_INTERFACE,
                                   C.byref(nb)))
        avg = np.zeros(3)
        info = _lib.SolveInfo()
        if nb.value == 0:
            results[i] = dict(n=n, k=[0.0, 0.0, 0.0], iterations=0, converged=False, seconds=0.0,
                              note="empty support")
            return
        check(L.fftperm_solve_direction(h, g, tol, maxit, None, avg, None, C.byref(info)))
        results[i] = dict(n=n, k=(mu * avg).tolist(), iterations=int(info.iterations),
                          converged=bool(info.converged), seconds=time.perf_counter() - t0)

    def run(w):
        while True:
            try:
                n, idx = q.get_nowait()
            except queue.Empty:
                return
            try:
                if mode == "batched":
                    run_batch(w, n, idx)
                else:
                    run_one(w, n, idx[0])
            except Exception as e:  # reported after the batch
                errors.append((idx, repr(e)))

    t0 = time.perf_counter()
    threads = [threading.Thread(target=run, args=(w,)) for w in ws]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    elapsed = time.perf_counter() - t0
    for w in ws:
        w.close()
    if errors:
        raise RuntimeError(f"{len(errors)} batch items failed, first: {errors[0]}")
    return results, elapsed


def work_estimate(n, iterations):
    """n^3 log n per apply x applies: the LPT weight a multi-GPU split would use."""
    return n ** 3 * math.log2(n) * (iterations + 2)


def main():
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser(description="Batched solve of the 128-sample uneven-size set (one load direction)")
    ap.add_argument("--count", type=int, default=128)
    ap.add_argument("--seed0", type=int, default=4000)
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--mode", choices=["batched", "per_sample"], default="batched")
    ap.add_argument("--gen-workers", type=int, default=1, help="processes generating the media")
    args = ap.parse_args()
    t0 = time.perf_counter()
    samples = c4_batch(args.count, args.seed0, workers=args.gen_workers)
    print(f"[batch] generated {len(samples)} media in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    prec = _lib.PRECISION_MIXED if args.precision == "mixed" else _lib.PRECISION_FP64
    res, sec = solve_batch(samples, tol=args.tol, precision=prec, workers=args.workers, mode=args.mode)
    its = [r["iterations"] for r in res]
    out = dict(samples=len(res), seconds=sec, samples_per_s=len(res) / sec, workers=args.workers, tol=args.tol,
               precision=args.precision, mode=args.mode, iterations_total=int(sum(its)), converged=int(sum(r["converged"] for r in res)),
               voxels_total=int(sum(r["n"] ** 3 for r in res)),
               k_xx=[r["k"][0] for r in res])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
