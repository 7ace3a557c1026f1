#!/usr/bin/env python3
"""Benchmark: 256^3 permeability-tensor solves per second on one MI355X.

Workload (the config the metric is quoted on): a periodic 256^3 Boolean model of
overlapping radius-8 solid spheres at porosity 0.6 (seed 2001), full 3x3
permeability tensor from the three unit load directions.  One "step" = support
detection from the device-resident mask + the three direction solves (RHS,
MINRES, velocity recovery, domain average).  The Green tensor is precomputed
once per grid and viscosity (reported separately as setup).

The metric is quoted at <= 1e-6 relative K error, so both the plain FP64 path
(MINRES, the reference's solver) and the mixed-precision path (conjugate residuals
on the FP32 operator and FP32 Krylov vectors with reliable FP64 residual updates,
ending on the FP64 true residual; the three directions iterate in lockstep) are timed at the same relative-residual tolerance (--tol, default 1e-6)
and each one's K error is measured against an untimed FP64 solve of the same
medium at --ref-tol (1e-10): ||K - K_ref||_F / ||K_ref||_F.  The headline
value is the mixed-precision path with the learned warm start (the metric's
configuration; the network is fft_method_test_amd/data/warmstart_unet.npz) when
its K error meets 1e-6, else mixed precision alone, else the FP64 path.

Prints ONE JSON line.  --gpus N with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "256³ k-tensor solves/s, mixed+learned vs plain FP64 GPU, at ≤1e-6 rel k error"
K_ERR_BAR = 1e-6
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def bytes_zgreen(Nh, s_c=16, s_r=8, gfrac=1.0):
    # fused z pass: read + write 3 complex components, read the 6 real Green components of the
    # fraction gfrac of the kz planes (FP32 all-in-LDS z kernels read planes kz <= n/2 only and
    # take the others by mirror symmetry)
    return 2 * 3 * Nh * s_c + 6 * Nh * s_r * gfrac


def green_plane_fraction(n, fp32, z_tile):
    """Fraction of the Green tensor's kz planes a z launch reads: (n // 2 + 1) / n for the FP32
    all-in-LDS z kernels (tile codes ending in 1 or 2), 1 otherwise."""
    return (n // 2 + 1) / n if fp32 and z_tile % 10 in (1, 2) else 1.0


def bytes_apply(N, Nh, L, s_c=16, s_r=8, nlines=None, s_v=8, plane=False, gfrac=1.0):
    # x-fwd: line bitmask of the support (1 bit per voxel, one int32 start per x line) + compact
    # vector in (s_v bytes per value: FP64, or FP32 for the mixed solver's vectors), 3 half spectra
    # out; y fwd/inv and z: in + out; z adds the Green tensor; x-inv: spectra + bitmask in, compact
    # vector out.  plane: the x and y passes fused over z planes, so the half spectra cross HBM
    # only on their way into and out of the z pass.
    nl = nlines if nlines is not None else round(N ** (2.0 / 3.0))
    bits = N // 8 + 4 * nl
    xf = bits + s_v * L + 3 * Nh * s_c
    yp = 0 if plane else 2 * 3 * Nh * s_c
    zg = bytes_zgreen(Nh, s_c, s_r, gfrac)
    xi = 3 * Nh * s_c + bits + s_v * L
    return xf + 2 * yp + zg + xi


def run_precision(args, P, _lib, mask_dev, precision, tol=None, steps=None, warmup=None, label="", net=None):
    import ctypes as C

    import torch

    n = args.n
    tol = args.tol if tol is None else tol
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    grid = P.Grid3D(n, n, n)
    t0 = time.perf_counter()
    green = P.GreenOperator(grid, 1.0, 2, precision)
    setup_s = time.perf_counter() - t0
    nb = C.c_uint64()
    K = np.zeros(9)
    infos = (_lib.SolveInfo * 3)()
    used = (C.c_int * 3)()

    def step():
        _lib.check(_lib.load().fftperm_set_medium_dev(green._h, C.c_void_p(mask_dev.data_ptr()),
                                                       _lib.SUPPORT_INTERFACE, C.byref(nb)))
        if net is None:
            _lib.check(_lib.load().fftperm_solve_tensor(green._h, tol, args.maxit, K, C.cast(infos, C.c_void_p)))
        else:
            _lib.check(_lib.load().fftperm_solve_tensor_warm(green._h, net._h, tol, args.maxit, K,
                                                             C.cast(infos, C.c_void_p), C.cast(used, C.c_void_p)))

    for w in range(warmup):
        step()
        log(f"[bench] {label} warmup {w + 1}/{warmup} iterations={[int(i.iterations) for i in infos]}")
    # sampled live timing: one chunk of Krylov iterations in eight records HIP events around its
    # launches; the others replay graphs (timing every launch would cost ~3% of the solve)
    _lib.check(_lib.load().fftperm_profile(green._h, 2))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        step()
        log(f"[bench] {label} step {s + 1}/{steps} iterations={[int(i.iterations) for i in infos]}")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = _lib.ProfileStats()
    _lib.check(_lib.load().fftperm_profile_read(green._h, C.byref(prof)))
    res = dict(
        precision=precision, tol=tol, elapsed=elapsed, solves_per_s=steps / elapsed, ms_per_step=1e3 * elapsed / steps,
        K=K.reshape(3, 3).copy(), iterations=[int(i.iterations) for i in infos],
        sweeps=[int(i.refinement_sweeps) for i in infos], converged=[bool(i.converged) for i in infos],
        final_rel=[float(i.final_rel_residual) for i in infos], setup_s=setup_s, nb=int(nb.value),
        used_guess=[int(u) for u in used] if net is not None else None,
        prof=dict(z_ms=list(prof.zgreen_ms), z_n=list(prof.zgreen_launches), a_ms=list(prof.apply_ms),
                  a_n=list(prof.apply_launches), z_rhs=list(prof.zgreen_rhs), a_rhs=list(prof.apply_rhs),
                  N=int(prof.N), Nh=int(prof.Nh),
                  tiles=dict(zgreen=list(prof.zgreen_tile), xfwd=list(prof.xfwd_tile), xinv=list(prof.xinv_tile),
                             ypass=list(prof.ypass_tile))),
    )
    green.close()
    return res


# HBM traffic per launch of the fused z pass from two rocprofv3 --pmc runs (FETCH_SIZE x 2.0,
# WRITE_SIZE x 1.0: the factors a 2 GiB copy kernel reads for the same counters) on this workload,
# keyed by (precision, tile code as fftperm_profile_read reports it); the summary with the commands
# is profiles/passes_256_hbm_traffic.md.  A tile the autotuner picks that was not profiled reports null.
Z_TRAFFIC_MB = {
    ("fp64", 41): (856.6, 436.3),  # k_zgreen_s<double, 256, 4, 1, false>
    ("fp64", 81): (856.2, 427.8),  # k_zgreen_s<double, 256, 8, 1, false>
    ("fp64", 44): (856.4, 451.2),  # k_zgreen_r<double, 256, 4>
    ("fp32", 82): (329.5, 217.3),  # k_zgreen_s<float, 256, 8, 2, true>
    ("fp32", 81): (353.4, 243.2),  # k_zgreen_s<float, 256, 8, 1, true>
    ("fp32", 161): (322.0, 213.9),  # k_zgreen_s<float, 256, 16, 1, true>
    ("fp32", 41): (566.7, 266.7),  # k_zgreen_s<float, 256, 4, 1, true>
}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(args, mask, applies_per_solve, min_seconds=10.0):
    """Oracle (FP64 restatement of the reference path) on the host cores: repeated operator applies
    at the same size (a bounded sample of about min_seconds), scaled by the applies one tensor solve
    needs; plus one fully timed oracle tensor solve on the 32^3 sphere cell (tol 1e-10)."""
    import oracle
    from fft_method_test_amd import geometry

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n = args.n
    t0 = time.perf_counter()
    o = oracle.Oracle((n, n, n))  # the energy-consistent Green tensor and the FFT plans
    setup_s = time.perf_counter() - t0
    o.set_medium(mask.reshape(-1))
    x = np.random.default_rng(0).standard_normal(3 * o.nb)
    reps = 0
    t0 = time.perf_counter()
    while True:
        o.apply(x)
        reps += 1
        if time.perf_counter() - t0 >= min_seconds and reps >= 2:
            break
    t_apply = (time.perf_counter() - t0) / reps
    o.close()
    c1 = oracle.Oracle((32, 32, 32))
    c1.set_medium(geometry.sphere_cell(32).reshape(-1))
    t0 = time.perf_counter()
    K1, infos1 = c1.solve_tensor(1e-10, 10000)
    c1_s = time.perf_counter() - t0
    c1.close()
    return {
        "value": 1.0 / (t_apply * applies_per_solve),
        "unit": "k-tensor solves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{reps} FP64 VariationalOperator applies of the oracle at {n}^3 ({t_apply:.3f} s each: 6 C2C 3D "
                   f"FFTs + Green multiply, {threads} OpenMP threads), scaled by the {applies_per_solve} applies of "
                   f"one FP64 tensor solve at tol {args.tol:g} (RHS + MINRES iterations + recovery, 3 directions); "
                   f"Green-tensor setup excluded"),
        "fft_provider": ("the oracle's own mixed-radix Stockham C2C FFT (oracle/oracle.c; FFTW3 is not installed, "
                         "the reference's FFTW backend cannot be built)"),
        "cpu_model": _cpu_model(),
        "apply_s_at_n": t_apply,
        "setup_s_at_n": setup_s,
        "c1_sphere32_tensor_solve": {"seconds": c1_s, "tol": 1e-10,
                                     "iterations": [int(i.iterations) for i in infos1],
                                     "k_xx": float(K1[0, 0]), "threads": threads},
    }


def hbm_probes(args, P, _lib, reps=20):
    """Copy probes over a MIXED context's three padded FP32 half spectra at this grid
    (fftperm_probe_copy): the contiguous stream (what HBM delivers here, beside the 8 TB/s nominal
    peak) and the y / z passes' column access patterns, in GB/s of bytes read plus written."""
    import ctypes as C

    green = P.GreenOperator(P.Grid3D(args.n, args.n, args.n), 1.0, 2, _lib.PRECISION_MIXED)
    out = {}
    try:
        rate = C.c_double()
        for pattern, key in ((0, "stream_copy_GBs"), (1, "ypass_pattern_copy_GBs"), (2, "zpass_pattern_copy_GBs")):
            _lib.check(_lib.load().fftperm_probe_copy(green._h, pattern, reps, C.byref(rate)))
            out[key] = rate.value
    finally:
        green.close()
    out["what"] = (f"{reps} copies each of the three padded FP32 half spectra of a {args.n}^3 context (spectrum slot "
                   f"0 -> 1): contiguous 16-byte accesses; 256-byte row segments down the y columns; 64-byte row "
                   f"segments down the z columns (the y and z passes' access patterns)")
    return out


def oracle_fixture_column(args, mask):
    """The oracle's FP64 x column of K for this medium at tol 1e-12, minted in the build container
    (tests/golden/make_k_fixtures.py), when the medium is the one it was minted on."""
    import hashlib

    path = os.path.join(ROOT, "tests", "golden", "k_fixtures.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get("c256")
    if not rec or hashlib.sha256(mask.reshape(-1).tobytes()).hexdigest() != rec["mask_sha256"]:
        return None
    return np.asarray(rec["Kcol_x_tol1e-12"], dtype=np.float64)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--radius", type=float, default=8.0)
    ap.add_argument("--porosity", type=float, default=0.6)
    ap.add_argument("--seed", type=int, default=2001)
    ap.add_argument("--tol", type=float, default=1e-6, help="relative-residual tolerance of the timed solves")
    ap.add_argument("--ref-tol", type=float, default=1e-10,
                    help="tolerance of the untimed FP64 reference solve the K errors are measured against "
                         "(<= 0: no reference solve, K errors not reported)")
    ap.add_argument("--maxit", type=int, default=20000)
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both")
    ap.add_argument("--warmstart", default=os.path.join(ROOT, "fft_method_test_amd", "data", "warmstart_unet.npz"),
                    help="trained warm-start network for the mixed+learned run ('' to skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(value is the single-GPU throughput)", file=sys.stderr)
        sys.exit(2)

    import torch

    from fft_method_test_amd import _lib, geometry, permeability as P

    _lib.load()
    log(f"[bench] device: {_lib.device_info()}")
    t0 = time.perf_counter()
    mask = geometry.boolean_spheres(args.n, args.radius, args.porosity, args.seed)
    log(f"[bench] geometry {args.n}^3 porosity={mask.mean():.4f} ({time.perf_counter() - t0:.1f} s)")
    mask_dev = torch.from_numpy(mask.reshape(-1).copy()).to("cuda:0")
    torch.cuda.synchronize()

    ref = None
    if args.ref_tol > 0:
        ref = run_precision(args, P, _lib, mask_dev, _lib.PRECISION_FP64, tol=args.ref_tol, steps=1, warmup=0,
                            label=f"reference fp64 tol={args.ref_tol:g}")
    runs = {}
    if args.precision in ("fp64", "both"):
        runs["fp64"] = run_precision(args, P, _lib, mask_dev, _lib.PRECISION_FP64, label=f"fp64 tol={args.tol:g}")
    if args.precision in ("mixed", "both"):
        runs["mixed"] = run_precision(args, P, _lib, mask_dev, _lib.PRECISION_MIXED, label=f"mixed tol={args.tol:g}")
        if args.warmstart and os.path.exists(args.warmstart):
            from fft_method_test_amd import warmstart as W

            net = W.WarmStartNet(0)
            try:
                net.load(args.warmstart)
            except ValueError as e:
                log(f"[bench] warm-start network not usable ({e}); no mixed+learned run")
            else:
                runs["mixed_learned"] = run_precision(args, P, _lib, mask_dev, _lib.PRECISION_MIXED,
                                                      label=f"mixed+learned tol={args.tol:g}", net=net)
            net.close()

    def k_err(r):
        if ref is None:
            return None
        return float(np.linalg.norm(r["K"] - ref["K"]) / np.linalg.norm(ref["K"]))

    col_o = oracle_fixture_column(args, mask)

    def col_err(r):
        if col_o is None:
            return None
        return float(np.linalg.norm(r["K"][:, 0] - col_o) / np.linalg.norm(col_o))

    probes = hbm_probes(args, P, _lib)
    log(f"[bench] HBM copy probes: {probes}")
    errs = {k: k_err(r) for k, r in runs.items()}
    meets = {k: (e is not None and e <= K_ERR_BAR) for k, e in errs.items()}
    # headline: mixed precision with the learned warm start (the metric's configuration) when it
    # meets the K-error bar, else mixed precision alone, else the FP64 path
    if "mixed_learned" in runs and meets.get("mixed_learned"):
        head = runs["mixed_learned"]
        dtype = "f32 FFT/Green + f64 Krylov, bf16 U-Net guess"
    elif "mixed" in runs and meets.get("mixed"):
        head = runs["mixed"]
        dtype = "f32 FFT/Green + f64 Krylov"
    elif "fp64" in runs:
        head = runs["fp64"]
        dtype = "f64"
    else:
        head = runs["mixed"]
        dtype = "f32 FFT/Green + f64 Krylov"

    # roofline of the dominant kernel (fused z pass) from the live HIP-event timings
    pi = 1 if head["precision"] == _lib.PRECISION_MIXED else 0
    pr = head["prof"]
    Nh, N = pr["Nh"], pr["N"]
    sc, sr = (8, 4) if pi == 1 else (16, 8)
    # a launch that serves several right-hand sides (the three directions) moves their spectra and
    # reads the Green tensor once
    z_n = max(1, pr["z_n"][pi])
    sides_per_launch = pr["z_rhs"][pi] / z_n
    z_tile = pr["tiles"]["zgreen"][pi]
    gfrac = green_plane_fraction(args.n, pi == 1, z_tile)
    z_bytes = sides_per_launch * 2 * 3 * Nh * sc + 6 * Nh * sr * gfrac
    z_ms = pr["z_ms"][pi] / z_n
    achieved = z_bytes / (z_ms * 1e-3) / 1e9
    tr = Z_TRAFFIC_MB.get(("fp32" if pi == 1 else "fp64", z_tile)) if sides_per_launch == 1 else None
    traffic = None if tr is None else (tr[0] + tr[1]) * 1e6
    L = 3 * head["nb"]
    plane = pi == 1 and pr["tiles"]["xfwd"][pi] == 1000
    a_bytes = bytes_apply(N, Nh, L, sc, sr, s_v=4 if pi == 1 else 8, plane=plane, gfrac=gfrac)
    a_ms = pr["a_ms"][pi] / max(1, pr["a_rhs"][pi])  # per right-hand side

    out = {
        "metric": METRIC,
        "value": head["solves_per_s"],
        "unit": "k-tensor solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {
            "workload": (f"{args.n}^3 periodic Boolean model of overlapping radius-{args.radius:g} spheres, porosity "
                         f"{args.porosity:g} (seed {args.seed}); full 3x3 K from 3 load directions; Bignonnet variational "
                         f"solve to relative residual {args.tol:g} (FP64: MINRES; mixed: CR on the FP32 operator and FP32 "
                         f"vectors with reliable FP64 updates, the three directions in lockstep), p_radius 2, mu 1; step = support detection + 3 direction solves; K error "
                         f"against an untimed FP64 solve at tol {args.ref_tol:g}"),
            "n": args.n,
            "tol": args.tol,
            "porosity_measured": float(mask.mean()),
            "support_voxels": head["nb"],
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_zgreen_s (fused z FFT, symmetric Green multiply, z inverse FFT)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": (None if tr is None else
                               f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of z tile {z_tile}: {tr[0]} MB read + "
                               f"{tr[1]} MB written per launch (profiles/passes_256_hbm_traffic.md)"),
            "algorithmic_bytes_per_launch": z_bytes,
            "green_planes_fraction": gfrac,
            # the rate the same launch time would give with every Green plane read (the measure of
            # the pass before the FP32 kernels read the mirrored planes once): context, not `achieved`
            "full_tensor_equivalent_GBs": (sides_per_launch * 2 * 3 * Nh * sc + 6 * Nh * sr) / (z_ms * 1e-3) / 1e9,
            "sides_per_launch": sides_per_launch,
            "avg_launch_ms": z_ms,
            "launches": pr["z_n"][pi],
            "precision": "fp32" if pi == 1 else "fp64",
        },
        "hbm_probe": probes,
        "apply": {
            "avg_ms": a_ms,
            "algorithmic_bytes": a_bytes,
            "achieved_GBs": a_bytes / (a_ms * 1e-3) / 1e9,
            "frac": a_bytes / (a_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "k_error_bar": K_ERR_BAR,
        "runs": {
            k: {
                "solves_per_s": r["solves_per_s"], "ms_per_step": r["ms_per_step"], "tol": r["tol"],
                "k_rel_error": errs[k], "meets_k_error_bar": meets[k],
                "k_col_x_rel_error_vs_oracle": col_err(r), "iterations": r["iterations"],
                "fp64_residual_checks": r["sweeps"], "converged": r["converged"], "final_rel_residual": r["final_rel"],
                "setup_s": r["setup_s"], "variants": r["prof"]["tiles"], "K": r["K"].tolist(),
                "used_guess": r["used_guess"],
            } for k, r in runs.items()
        },
        "reference": None if ref is None else {
            "tol": ref["tol"], "solves_per_s": ref["solves_per_s"], "iterations": ref["iterations"],
            "K": ref["K"].tolist(), "k_col_x_rel_error_vs_oracle": col_err(ref),
        },
        "oracle_fixture": (None if col_o is None else
                           {"what": "CPU oracle FP64 x column of K at tol 1e-12 (tests/golden/k_fixtures.json)",
                            "Kcol_x": col_o.tolist()}),
        "speedup_mixed_vs_fp64": (runs["mixed"]["solves_per_s"] / runs["fp64"]["solves_per_s"]
                                  if "fp64" in runs and "mixed" in runs else None),
        "speedup_head_vs_fp64": (head["solves_per_s"] / runs["fp64"]["solves_per_s"] if "fp64" in runs else None),
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        ref_run = runs.get("fp64", head)
        applies = int(sum(it + 2 for it in ref_run["iterations"]))
        try:
            out["cpu_baseline"] = cpu_baseline(args, mask, applies)
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] cpu baseline failed: {e}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
