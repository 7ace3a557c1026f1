"""Measurements of the secondary workloads (one JSON line each):

    python -m fft_method_test_amd.measure_configs c1     # 32^3 sphere cell, full K at tol 1e-10, FP64 and mixed
    python -m fft_method_test_amd.measure_configs c2p    # 512^3 Boolean spheres r 8, porosity 0.35 (seed 2002)
    python -m fft_method_test_amd.measure_configs c3     # 256^3 Brinkman, spheres r 6, porosity 0.5 (seed 3001)
    python -m fft_method_test_amd.measure_configs apply  # FP32 apply of the bench medium, separate vs fused x+y
    python -m fft_method_test_amd.measure_configs probe  # HBM copy probes at 256^3 and 512^3

c2p solves direction x to tol 1e-6 (at most --maxit iterations) in FP64 and mixed precision and
reports seconds per iteration and the five-pass apply's HBM rate.  c3 runs the Brinkman iteration
for the three penalties 1e2, 1e4, 1e6 (direction x, absolute RMS-divergence tol 1e-8, at most
--maxit iterations) in FP64 and mixed precision and reports iterations, seconds per iteration and
the stencil sweeps' HBM rate against 127 N * 8 bytes per iteration (9 momentum, 4 divergence,
35 x 3 Jacobi, 9 projection arrays of N doubles).  Inputs are generated on the host, so the clock
starts with the medium already on the device for c2p and with the coefficient fields uploaded by
the Brinkman call for c3 (the host copy is inside its time).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def apply_bytes(N, Nh, L, nlines, s):
    """Algorithmic HBM bytes of one five-pass apply with s-byte reals: three half spectra written by
    the x forward, read and written by the y forward, the z pass and the y inverse, read by the x
    inverse; the Green tensor (6 reals per frequency) read by the z pass; the compact vector (FP64)
    and the line bitmask (1 bit per voxel + one int32 start per x line) read / written by the x
    passes."""
    spectra = (1 + 2 + 2 + 2 + 1) * 3 * Nh * 2 * s
    return spectra + 6 * Nh * s + 2 * (8 * L + N // 8 + 4 * nlines)


def _sync():
    import torch

    torch.cuda.synchronize()


def c1(args):
    from . import _lib, geometry, permeability as P

    n = 32
    mask = geometry.sphere_cell(n).reshape(-1)
    out = dict(config="c1", n=n, porosity=float(mask.mean()))
    for name, prec in (("fp64", _lib.PRECISION_FP64), ("mixed", _lib.PRECISION_MIXED)):
        g = P.Grid3D(n, n, n)
        green = P.GreenOperator(g, 1.0, 2, prec)
        sup = P.ForceSupport(P.BinaryMedium(g, mask), P.SUPPORT_INTERFACE, green)
        P.solve_permeability_tensor(green, 1e-10, 10000)  # warm-up (autotuning, graphs)
        _sync()
        t0 = time.perf_counter()
        K, infos = P.solve_permeability_tensor(green, 1e-10, 10000)
        _sync()
        dt = time.perf_counter() - t0
        out[name] = dict(seconds=dt, k_xx=float(K[0, 0]), iterations=[int(i.iterations) for i in infos],
                         nb=sup.num_active_voxels())
        green.close()
    print(json.dumps(out), flush=True)


def c2p(args):
    import torch

    from . import _lib, geometry, permeability as P

    L = _lib.load()
    n = 512
    t0 = time.perf_counter()
    mask = geometry.boolean_spheres(n, 8.0, 0.35, 2002).reshape(-1)
    gen_s = time.perf_counter() - t0
    mdev = torch.from_numpy(mask).to("cuda:0")
    out = dict(config="c2p", n=n, radius=8.0, porosity_target=0.35, seed=2002, porosity=float(mask.mean()),
               generate_s=gen_s, tol=args.tol, maxit=args.maxit)
    grad = np.array([1.0, 0.0, 0.0])
    for name, prec in (("fp64", _lib.PRECISION_FP64), ("mixed", _lib.PRECISION_MIXED)):
        _sync()
        t0 = time.perf_counter()
        green = P.GreenOperator(P.Grid3D(n, n, n), 1.0, 2, prec)
        nb = C.c_uint64()
        _lib.check(L.fftperm_set_medium_dev(green._h, C.c_void_p(mdev.data_ptr()), _lib.SUPPORT_INTERFACE,
                                            C.byref(nb)))
        _sync()
        setup = time.perf_counter() - t0
        avg = np.zeros(3)
        info = _lib.SolveInfo()
        _lib.check(L.fftperm_profile(green._h, 1))
        _sync()
        t0 = time.perf_counter()
        _lib.check(L.fftperm_solve_direction(green._h, grad, args.tol, args.maxit, None, avg, None, C.byref(info)))
        _sync()
        dt = time.perf_counter() - t0
        pr = _lib.ProfileStats()
        _lib.check(L.fftperm_profile_read(green._h, C.byref(pr)))
        pi = 1 if prec == _lib.PRECISION_MIXED else 0
        s = 4 if pi == 1 else 8
        a_ms = pr.apply_ms[pi] / max(1, pr.apply_launches[pi])
        a_bytes = apply_bytes(pr.N, pr.Nh, 3 * pr.nb, n * n, s)
        out[name] = dict(setup_s=setup, seconds=dt, iterations=int(info.iterations), converged=bool(info.converged),
                         final_rel_residual=float(info.final_rel_residual), s_per_iteration=dt / max(1, info.iterations),
                         k_xx=float(avg[0]), nb=int(nb.value), apply_ms=a_ms,
                         apply_GBs=a_bytes / (a_ms * 1e-3) / 1e9 if a_ms > 0 else None,
                         note="profiling on (eager launches, no graph replay)")
        green.close()
        del green
    print(json.dumps(out), flush=True)


def c3(args):
    from . import _lib, brinkman as B, geometry

    n = 256
    mask = geometry.boolean_spheres(n, 6.0, 0.5, 3001).reshape(n, n, n)
    ids = np.where(mask > 0, 1, 0).astype(np.uint32).reshape(-1)
    geo = B.Geometry(n, n, n, ids)
    db = B.single_fluid_database(1, 1.0)
    out = dict(config="c3", n=n, radius=6.0, porosity=float(mask.mean()), seed=3001, tol=1e-8, maxit=args.maxit,
               runs=[])
    for penalty in (1e2, 1e4, 1e6):
        for name, prec in (("fp64", _lib.PRECISION_FP64), ("mixed", _lib.PRECISION_MIXED)):
            cfg = B.FluidConfig(fluid_material_id=1, solid_penalty=penalty, tolerance=1e-8,
                                max_iterations=args.maxit, precision=prec)
            B.solve_brinkman(cfg, geo, db, [1.0, 0.0, 0.0]) if args.warm else None
            _sync()
            t0 = time.perf_counter()
            r = B.solve_brinkman(cfg, geo, db, [1.0, 0.0, 0.0])
            _sync()
            dt = time.perf_counter() - t0
            it = r.diagnostics.iterations
            bytes_it = 127 * n ** 3 * 8
            out["runs"].append(dict(penalty=penalty, precision=name, seconds=dt, iterations=it,
                                    residual=r.diagnostics.residual, converged=r.diagnostics.converged,
                                    s_per_iteration=dt / max(1, it), U_x=float(r.superficial_velocity[0]),
                                    model_GBs=bytes_it * it / dt / 1e9))
            print(json.dumps(out["runs"][-1]), file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


def apply(args):
    """The FP32 apply of the 256^3 bench medium (radius-8 spheres, porosity 0.6, seed 2001) timed by
    fftperm_time_apply with the x and y passes pinned to each form (FFTPERM_PASS_PLANE codes: 1
    separate passes, 2 / 3 fused over z planes with 8 / 16 waves per workgroup), three rounds."""
    from . import _lib, geometry, permeability as P

    L = _lib.load()
    n = args.n
    mask = geometry.boolean_spheres(n, 8.0, 0.6, 2001).reshape(-1)
    g = P.Grid3D(n, n, n)
    green = P.GreenOperator(g, 1.0, 2, _lib.PRECISION_MIXED)
    P.ForceSupport(P.BinaryMedium(g, mask), P.SUPPORT_INTERFACE, green)
    codes = (1, 2, 3) if args.plane == 0 else (args.plane,)
    ms = C.c_double()
    out = dict(config="apply", n=n, reps=args.reps, us_per_apply={str(c): [] for c in codes})
    for _ in range(3):
        for code in codes:
            _lib.check(L.fftperm_pin_variant(green._h, _lib.PASS_PLANE, _lib.PRECISION_MIXED, code))
            _lib.check(L.fftperm_time_apply(green._h, _lib.PRECISION_MIXED, args.reps, C.byref(ms)))
            out["us_per_apply"][str(code)].append(ms.value * 1e3)
    _lib.check(L.fftperm_pin_variant(green._h, _lib.PASS_PLANE, _lib.PRECISION_MIXED, 0))
    green.close()
    print(json.dumps(out), flush=True)


def probe(args):
    """fftperm_probe_copy at 256^3 and 512^3 (contiguous, y-pass and z-pass patterns)."""
    from . import _lib, permeability as P

    L = _lib.load()
    out = dict(config="probe", runs=[])
    for n, reps in ((256, 20), (512, 5)):
        green = P.GreenOperator(P.Grid3D(n, n, n), 1.0, 2, _lib.PRECISION_MIXED)
        rate = C.c_double()
        r = dict(n=n, reps=reps)
        for pattern, key in ((0, "stream_copy_GBs"), (1, "ypass_pattern_copy_GBs"), (2, "zpass_pattern_copy_GBs")):
            _lib.check(L.fftperm_probe_copy(green._h, pattern, reps, C.byref(rate)))
            r[key] = rate.value
        green.close()
        out["runs"].append(r)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("config", choices=["c1", "c2p", "c3", "apply", "probe"])
    ap.add_argument("--n", type=int, default=256, help="apply: grid edge")
    ap.add_argument("--plane", type=int, default=0, help="apply: only this FFTPERM_PASS_PLANE code (0: 1, 2 and 3)")
    ap.add_argument("--reps", type=int, default=20, help="apply: applies per timing")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--maxit", type=int, default=3000)
    ap.add_argument("--warm", action="store_true", help="c3: an untimed solve first")
    args = ap.parse_args()
    from . import _lib

    _lib.load()
    {"c1": c1, "c2p": c2p, "c3": c3, "apply": apply, "probe": probe}[args.config](args)


if __name__ == "__main__":
    main()
