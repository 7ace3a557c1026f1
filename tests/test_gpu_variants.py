"""Every tile variant of the specialised passes, pinned one at a time, against the FP64 oracle.

The autotuner picks one variant per (pass, precision, length, count) at run time, so the parity
tests elsewhere only see the variant that won on the test box.  Here each variant is forced
through fftperm_pin_variant and the Green apply / variational operator are compared with the
oracle (FP64: relative Frobenius error < 1e-13 / 1e-12).  FP32 variants (the transforms of the
mixed-precision solves) are checked through the mixed K tensor of the 32^3 sphere cell: within
1e-6 of the oracle's K, the contract of the mixed path.
"""
import contextlib

import numpy as np
import pytest

from conftest import random_mask, rel

pytestmark = pytest.mark.gpu

# threads per line of the compile-time plan per length (fft_static.hpp StaticPlan / PlanInfo::threads):
# the largest radix's butterfly count for the power-of-two plans (32 = 8 x 4, 64 = 8 x 8, 128 = 16 x 8,
# 256 = 16 x 16, 512 = 8 x 8 x 8), 16 for the mixed plans 16 x r (one butterfly per thread)
_T = {32: 4, 64: 8, 128: 8, 256: 16, 512: 64, 80: 16, 96: 16, 112: 16, 144: 16, 160: 16}
POW2 = [32, 64, 128, 256, 512]
MIXED = [80, 96, 112, 144, 160]


def pass_codes(n, precision=0):
    """Codes (10 * lines per workgroup + threads per butterfly thread) of the x / column pass
    variants at length n, as fft_static.hpp enumerates them."""
    t = _T[n]
    base = max(2, 256 // t)
    cx_bytes = 16 if precision == 0 else 8
    out = []
    for v in range(5):
        lines = base if v < 2 else (base // 2 if v < 4 else base // 4)
        os_ = 2 if v in (0, 2) else 4
        lds = (n + lines * (n + 1)) * cx_bytes
        if lines >= 2 and lines % 2 == 0 and 64 <= os_ * lines * t <= 1024 and lds <= 100 * 1024:
            out.append(10 * lines + os_)
    return out


def xpass_codes(n, precision=0):
    """pass_codes(n) plus the merged-round x-pass variants valid at this precision (one transform of
    3 lines / 2 complex lines per tile; code 10 * lines + threads per butterfly thread + 5) and the
    wave-per-line-pair variant (code 88) at the lengths it serves."""
    t = _T[n]
    base = max(2, 256 // t)
    out = pass_codes(n, precision)
    cx_bytes = 16 if precision == 0 else 8
    for j in range(3):
        lines, os_ = base >> j, 1 << j
        lt = 3 * lines // 2
        threads = os_ * lt * t
        lds = (n + lt * (n + 1)) * cx_bytes
        if lines >= 2 and lines % 2 == 0 and 64 <= threads <= 1024 and threads % 64 == 0 and lds <= 100 * 1024:
            out.append(10 * lines + os_ + 5)
    if n in (128, 256, 512):
        out.append(88)  # one wave per line pair, 8 lines per workgroup
    return out


def col_codes(n, precision=0):
    """pass_codes(n) plus the wide column-pass variants (twice the base line count, OS 1 and 2)
    that fit the LDS budget at this precision."""
    t = _T[n]
    base = max(2, 256 // t)
    out = pass_codes(n, precision)
    cx_bytes = 16 if precision == 0 else 8
    lines = 2 * base
    for os_ in (1, 2):
        threads = os_ * lines * t
        lds = (n + lines * (n + 1)) * cx_bytes
        if 64 <= threads <= 1024 and lds <= 100 * 1024:
            out.append(10 * lines + os_)
    return out


@contextlib.contextmanager
def pinned(green, kind, precision, code):
    from fft_method_test_amd import _lib

    _lib.check(_lib.load().fftperm_pin_variant(green._h, kind, precision, code))
    try:
        yield
    finally:
        _lib.check(_lib.load().fftperm_pin_variant(green._h, kind, precision, 0))


def _pair(oracle_mod, shape, seed, precision=0):
    from fft_method_test_amd import permeability as P

    mask = random_mask(shape, 0.6, seed=seed)
    g = P.Grid3D(*shape)
    green = P.GreenOperator(g, 1.0, 2, precision)
    sup = P.ForceSupport(P.BinaryMedium(g, mask), P.SUPPORT_INTERFACE, green)
    o = oracle_mod.Oracle(shape)
    o.set_medium(mask)
    return green, sup, o


def _check_apply(green, sup, o, shape, seed):
    from fft_method_test_amd import permeability as P

    N = int(np.prod(shape))
    f = np.random.default_rng(seed).standard_normal(3 * N)
    assert rel(green.apply(f), o.green_apply(f)) < 1e-13
    x = np.random.default_rng(seed + 1).standard_normal(sup.size())
    assert rel(P.VariationalOperator(sup, green).apply(x), o.apply(x)) < 1e-12


# one shape per axis and length: the varied axis carries the specialised length
X_SHAPES = {n: (n, 6, 10) for n in POW2 + MIXED}
Y_SHAPES = {n: (6, n, 10) for n in POW2 + MIXED}
Z_SHAPES = {n: (6, 10, n) for n in POW2 + MIXED}


@pytest.mark.parametrize("n", POW2 + MIXED)
def test_x_pass_variants_fp64(fp, oracle_mod, n):
    from fft_method_test_amd import _lib

    shape = X_SHAPES[n]
    green, sup, o = _pair(oracle_mod, shape, seed=n)
    for code in xpass_codes(n, 0):
        for kind in (_lib.PASS_XFWD, _lib.PASS_XINV):
            with pinned(green, kind, _lib.PRECISION_FP64, code):
                _check_apply(green, sup, o, shape, seed=code)


@pytest.mark.parametrize("precision", [0, 1])
def test_x_wave_variant_odd_line_count(fp, oracle_mod, precision):
    """The wave-per-line-pair x passes with an odd number of x lines (the last wave holds one line)
    and a workgroup with idle waves: FP64 operator within 1e-12, FP32 operator within
    FP32_APPLY_BOUND of the oracle."""
    from fft_method_test_amd import _lib, permeability as P

    shape = (256, 5, 3)
    green, sup, o = _pair(oracle_mod, shape, seed=77, precision=precision)
    op = P.VariationalOperator(sup, green)
    x = np.random.default_rng(78).standard_normal(sup.size())
    ref = o.apply(x)
    for kind in (_lib.PASS_XFWD, _lib.PASS_XINV):
        with pinned(green, kind, precision, 88):
            if precision == 0:
                assert rel(op.apply(x), ref) < 1e-12
            else:
                assert rel(op.apply_fp32(x), ref) < FP32_APPLY_BOUND
    green.close()


def _run_mix_mask(n, ny, nz, seed):
    """Fluid mask whose x-line pairs (the unit of the wave passes) have support runs of three
    kinds: dense pairs (two solid voxels in three, all on the interface: longer than a wave's LDS
    line), pairs with one short solid block (runs that fit) and all-fluid pairs (empty runs)."""
    rng = np.random.default_rng(seed)
    m = np.ones((nz, ny, n), np.uint8)
    x = np.arange(n)
    for z in range(nz):
        for y in range(ny):
            kind = ((y + ny * z) // 2) % 3
            if kind == 0:
                m[z, y] = x % 3 == 0
            elif kind == 1:
                s = rng.integers(0, n - n // 8)
                m[z, y, s:s + n // 8] = 0
    return m.reshape(-1)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n", [128, 256, 512])
@pytest.mark.parametrize("ny,nz", [(4, 6), (5, 3)])
def test_x_wave_staged_and_direct_runs(fp, oracle_mod, precision, n, ny, nz):
    """The wave x passes stage a line pair's compact run through LDS when it fits one transform
    line (forward: source run; inverse: output run) and read / write it voxel by voxel otherwise.
    A medium with long, short and empty runs (and an odd line count) exercises both paths in
    one apply: FP64 operator within 1e-12, FP32 operator within FP32_APPLY_BOUND of the oracle."""
    from fft_method_test_amd import _lib, permeability as P

    shape = (n, ny, nz)
    mask = _run_mix_mask(n, ny, nz, seed=n + ny)
    o = oracle_mod.Oracle(shape)
    o.set_medium(mask)
    act, _ = o.support()
    per_line = np.bincount(act.astype(np.int64) // n, minlength=ny * nz)
    pairs = per_line[0::2] + np.append(per_line[1::2], [0] * (len(per_line) % 2))
    line_floats = (n + n // 16) * 2  # one padded transform line of FP32 complex values, in floats
    assert (pairs == 0).any() and ((pairs > 0) & (3 * pairs <= line_floats - 3)).any()
    assert (3 * pairs > line_floats).any()
    g = P.Grid3D(*shape)
    green = P.GreenOperator(g, 1.0, 2, precision)
    sup = P.ForceSupport(P.BinaryMedium(g, mask), P.SUPPORT_INTERFACE, green)
    op = P.VariationalOperator(sup, green)
    x = np.random.default_rng(n).standard_normal(sup.size())
    ref = o.apply(x)
    with pinned(green, _lib.PASS_XFWD, precision, 88), pinned(green, _lib.PASS_XINV, precision, 88):
        if precision == 0:
            assert rel(op.apply(x), ref) < 1e-12
        else:
            assert rel(op.apply_fp32(x), ref) < FP32_APPLY_BOUND
    green.close()


@pytest.mark.parametrize("n", POW2 + MIXED)
def test_y_pass_variants_fp64(fp, oracle_mod, n):
    from fft_method_test_amd import _lib

    shape = Y_SHAPES[n]
    green, sup, o = _pair(oracle_mod, shape, seed=n + 1)
    for code in col_codes(n, 0):
        with pinned(green, _lib.PASS_YPASS, _lib.PRECISION_FP64, code):
            _check_apply(green, sup, o, shape, seed=code)


@pytest.mark.parametrize("n", POW2 + MIXED)
def test_zgreen_variants_fp64(fp, oracle_mod, n):
    from fft_method_test_amd import _lib

    shape = Z_SHAPES[n]
    green, sup, o = _pair(oracle_mod, shape, seed=n + 2)
    for code in _lib.ZGREEN_CODES:
        with pinned(green, _lib.PASS_ZGREEN, _lib.PRECISION_FP64, code):
            _check_apply(green, sup, o, shape, seed=code)


def test_zgreen_variants_minres_alpha(fp, oracle_mod):
    """The fused z pass also returns the Lanczos alpha (Parseval): MINRES with each variant pinned
    follows the oracle's recurrence (first 12 residual estimates within 1e-9 relative)."""
    from fft_method_test_amd import _lib, permeability as P

    shape = (16, 12, 32)
    green, sup, o = _pair(oracle_mod, shape, seed=9)
    b = o.rhs([1.0, 0.0, 0.0])
    _, _, ref_hist = o.minres(b, tol=1e-10, maxit=12)
    for code in _lib.ZGREEN_CODES:
        with pinned(green, _lib.PASS_ZGREEN, _lib.PRECISION_FP64, code):
            res = P.MinresSolver(1e-10, 12).solve(P.VariationalOperator(sup, green), b)
            h = np.asarray(res.residual_history)
            assert rel(h[:12], ref_hist[:12]) < 1e-9, code


def test_fp32_variants_mixed_k_sphere32(fp, oracle_mod):
    from fft_method_test_amd import _lib, geometry as G, permeability as P

    n = 32
    mask = G.sphere_cell(n, 0.3).reshape(-1)
    o = oracle_mod.Oracle((n, n, n))
    o.set_medium(mask)
    k_ref, _ = o.solve_tensor(tol=1e-10, maxit=10000)
    g = P.Grid3D(n, n, n)
    green = P.GreenOperator(g, 1.0, 2, P.PRECISION_MIXED)
    sup = P.ForceSupport(P.BinaryMedium(g, mask), P.SUPPORT_INTERFACE, green)
    cases = [(_lib.PASS_ZGREEN, c) for c in _lib.ZGREEN_CODES]
    cases += [(k, c) for k in (_lib.PASS_XFWD, _lib.PASS_XINV) for c in xpass_codes(n, 1)]
    cases += [(_lib.PASS_YPASS, c) for c in col_codes(n, 1)]
    for kind, code in cases:
        with pinned(green, kind, _lib.PRECISION_MIXED, code):
            K, _ = P.solve_permeability_tensor(green, 1e-9, 10000)
            assert rel(K, k_ref) < 1e-6, (kind, code)


# FP32 operator error bound: the five FP32 sweeps round the spectrum at every pass and compute the
# butterflies in FP32; measured relative errors are 1e-7 .. 4e-7 on random vectors.
FP32_APPLY_BOUND = 1e-6


@pytest.mark.parametrize("n", POW2 + MIXED)
def test_fp32_operator_variants(fp, oracle_mod, n):
    """Every FP32 tile variant (x forward / inverse, column pass, z pass) pinned in turn: the FP32
    variational operator of the mixed sweeps within FP32_APPLY_BOUND of the FP64 oracle."""
    from fft_method_test_amd import _lib, permeability as P

    cases = []
    for shape, kinds, codes in ((X_SHAPES[n], (_lib.PASS_XFWD, _lib.PASS_XINV), xpass_codes(n, 1)),
                                (Y_SHAPES[n], (_lib.PASS_YPASS,), col_codes(n, 1)),
                                (Z_SHAPES[n], (_lib.PASS_ZGREEN,), _lib.ZGREEN_CODES)):
        green, sup, o = _pair(oracle_mod, shape, seed=n + 5, precision=_lib.PRECISION_MIXED)
        op = P.VariationalOperator(sup, green)
        x = np.random.default_rng(n).standard_normal(sup.size())
        ref = o.apply(x)
        for kind in kinds:
            for code in codes:
                with pinned(green, kind, _lib.PRECISION_MIXED, code):
                    err = rel(op.apply_fp32(x), ref)
                    cases.append((kind, code, err))
                    assert err < FP32_APPLY_BOUND, (kind, code, err)
        green.close()
    print("fp32 apply errors", n, max(c[2] for c in cases))


def test_fp32_operator_streamed_green_tensor(fp, oracle_mod):
    """A 256 x 256 x 224 grid: its FP32 half spectra (187 MB) fit the 256 MiB last-level cache but
    not beside the half of the FP32 Green tensor that the z pass reads (94 MB; the other planes by
    mirror symmetry), so the single-side z kernel reads the tensor with non-temporal loads.  Both
    of its tile shapes pinned: the FP32 operator within FP32_APPLY_BOUND of the oracle, and the
    same values from both tiles."""
    from fft_method_test_amd import _lib, permeability as P

    shape = (256, 256, 224)
    green, sup, o = _pair(oracle_mod, shape, seed=19, precision=_lib.PRECISION_MIXED)
    try:
        op = P.VariationalOperator(sup, green)
        x = np.random.default_rng(23).standard_normal(sup.size())
        ref = o.apply(x)
        ys = []
        for code in (82, 41):  # k_zgreen_s: 8 columns x 2 threads, 4 columns x 1 thread
            with pinned(green, _lib.PASS_ZGREEN, _lib.PRECISION_MIXED, code):
                y = op.apply_fp32(x)
            err = rel(y, ref)
            print(f"[stream-green] {shape} z code {code}: FP32 operator rel err {err:.2e}", flush=True)
            assert err < FP32_APPLY_BOUND
            ys.append(y)
        assert rel(ys[0], ys[1]) < 1e-6
    finally:
        green.close()
        o.close()


def test_fp32_operator_rejected_in_fp64_context(fp, oracle_mod):
    from fft_method_test_amd import permeability as P

    green, sup, _ = _pair(oracle_mod, (8, 8, 8), seed=3)
    with pytest.raises(ValueError, match="MIXED contexts only"):
        P.VariationalOperator(sup, green).apply_fp32(np.zeros(sup.size()))
