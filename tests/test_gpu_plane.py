"""FP32 x and y passes fused over z planes (grids with nx = ny = 256).

The fused forward pass gathers a z plane's rows from the FP32 compact vector, runs the x and y
transforms of the plane on chip and writes the y-transformed half spectrum; the fused inverse does
the reverse and scatters to the FP32 compact vector with the support sums.  Checked here through
the FP32-vector apply that the default mixed solver iterates (fftperm_operator_apply_fp32v):

* with the fused passes pinned, against the separate x and y passes and against the FP64 oracle
  operator (src/bignonnet/VariationalOperator.cpp:19-41), on sparse (Boolean spheres) and dense
  (random, full-solid support) supports, including planes without any support voxel;
* the permeability column of a mixed solve with the fused passes against the oracle's FP64 K;
* the profile reports the fused form while it is pinned.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import random_mask, rel

pytestmark = pytest.mark.gpu

# FP32 operator error bound, as for every FP32 tile variant (measured 1e-7 .. 4e-7)
FP32_APPLY_BOUND = 1e-6
# fftperm_pin_variant codes of the fused passes: 8 and 16 waves per workgroup
PLANE_CODES = (2, 3)


def _spheres_slab(nz, seed, porosity=0.6):
    from fft_method_test_amd import geometry as G

    return G.boolean_spheres(256, 8.0, porosity, seed).reshape(256, 256, 256)[:nz].reshape(-1).copy()


def _setup(oracle_mod, shape, mask):
    from fft_method_test_amd import _lib, permeability as P

    g = P.Grid3D(*shape)
    green = P.GreenOperator(g, 1.0, 2, _lib.PRECISION_MIXED)
    sup = P.ForceSupport(P.BinaryMedium(g, mask), P.SUPPORT_INTERFACE, green)
    o = oracle_mod.Oracle(shape)
    o.set_medium(mask)
    return green, sup, o


def _pin(green, code):
    from fft_method_test_amd import _lib

    _lib.check(_lib.load().fftperm_pin_variant(green._h, _lib.PASS_PLANE, _lib.PRECISION_MIXED, code))


def _cases():
    # (name, shape, mask builder)
    return [
        ("spheres_z16", (256, 256, 16), lambda: _spheres_slab(16, 41)),
        ("random_z8", (256, 256, 8), lambda: random_mask((256, 256, 8), 0.6, 5)),
        ("halfempty_z12", (256, 256, 12), None),
    ]


@pytest.mark.parametrize("case", [c[0] for c in _cases()])
def test_plane_apply_against_separate_passes_and_oracle(fp, oracle_mod, case):
    from fft_method_test_amd import _lib, permeability as P

    name, shape, build = next(c for c in _cases() if c[0] == case)
    if build is None:
        # alternate planes fully fluid (no support voxel in them) and spheres
        m = _spheres_slab(shape[2], 43).reshape(shape[2], 256, 256)
        m[::2] = 1
        mask = m.reshape(-1).copy()
    else:
        mask = build()
    green, sup, o = _setup(oracle_mod, shape, mask)
    try:
        op = P.VariationalOperator(sup, green)
        x = np.random.default_rng(7).standard_normal(sup.size())
        ref = o.apply(x)
        prof = _lib.ProfileStats()
        _pin(green, 1)
        y_sep = op.apply_fp32v(x)
        _lib.check(_lib.load().fftperm_profile_read(green._h, C.byref(prof)))
        assert list(prof.xfwd_tile)[1] != 1000
        e_sep = rel(y_sep, ref)
        assert e_sep < FP32_APPLY_BOUND
        ys = {}
        for code in PLANE_CODES:  # 8 and 16 waves per workgroup
            _pin(green, code)
            y_plane = op.apply_fp32v(x)
            _lib.check(_lib.load().fftperm_profile_read(green._h, C.byref(prof)))
            assert list(prof.xfwd_tile)[1] == 1000 and list(prof.ypass_tile)[1] == 1000
            assert np.array_equal(y_plane, op.apply_fp32v(x)), "fused passes are not deterministic"
            e_plane, e_ps = rel(y_plane, ref), rel(y_plane, y_sep)
            print(f"[plane] {name} code {code}: nb {sup.size() // 3} fused {e_plane:.2e} separate {e_sep:.2e} "
                  f"fused-vs-separate {e_ps:.2e}", flush=True)
            assert e_plane < FP32_APPLY_BOUND
            ys[code] = y_plane
        # the two workgroup shapes run the same arithmetic on the same data
        assert np.array_equal(ys[PLANE_CODES[0]], ys[PLANE_CODES[1]])
    finally:
        _pin(green, 0)
        green.close()
        o.close()


def test_plane_full_solid_support(fp, oracle_mod):
    """Full-solid support (every solid voxel carries force): whole x lines on the support."""
    from fft_method_test_amd import permeability as P

    shape = (256, 256, 6)
    mask = random_mask(shape, 0.5, 9)
    g = P.Grid3D(*shape)
    from fft_method_test_amd import _lib

    green = P.GreenOperator(g, 1.0, 2, _lib.PRECISION_MIXED)
    sup = P.ForceSupport(P.BinaryMedium(g, mask), P.SUPPORT_FULLSOLID, green)
    o = oracle_mod.Oracle(shape)
    o.set_medium(mask, support=oracle_mod.SUPPORT_FULLSOLID)
    try:
        op = P.VariationalOperator(sup, green)
        x = np.random.default_rng(3).standard_normal(sup.size())
        ref = o.apply(x)
        for code in PLANE_CODES:
            _pin(green, code)
            err = rel(op.apply_fp32v(x), ref)
            print(f"[plane] full-solid support nb {sup.size() // 3} code {code}: fused {err:.2e}", flush=True)
            assert err < FP32_APPLY_BOUND
    finally:
        _pin(green, 0)
        green.close()
        o.close()


def test_plane_mixed_k_column_against_oracle(fp, oracle_mod):
    """x column of K from a mixed solve (CR on FP32 vectors, reliable FP64 updates) with the fused
    passes, against the oracle's FP64 MINRES column at tol 1e-10."""
    from fft_method_test_amd import permeability as P

    shape = (256, 256, 8)
    mask = _spheres_slab(8, 47)
    green, sup, o = _setup(oracle_mod, shape, mask)
    try:
        avg_o, info_o = o.solve_direction([1.0, 0.0, 0.0], 1e-10, 20000)
        res = {}
        for code in (2, 3, 1):
            _pin(green, code)
            avg, info, _ = P.solve_one_direction(green, sup, [1.0, 0.0, 0.0], 1e-9, 20000)
            res[code] = (avg, int(info.iterations), bool(info.converged))
        for code, (avg, its, conv) in res.items():
            err = rel(avg, avg_o)
            print(f"[plane] K column, plane code {code}: rel err {err:.2e} iterations {its} "
                  f"(oracle {info_o.iterations})", flush=True)
            assert conv and err < 1e-6
        for code in PLANE_CODES:
            assert abs(res[code][1] - res[1][1]) <= 0.05 * res[1][1] + 5
    finally:
        _pin(green, 0)
        green.close()
        o.close()
