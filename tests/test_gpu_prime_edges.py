"""Grid edges with prime factors above 13 (the reference's FFTW backend plans any length,
src/fft/FFTWBackend.cpp:176-192): the runtime FFT plan runs such a prime as a stage of direct DFT
sums.  Checked against numpy.fft (the FFT hooks) and the FP64 oracle (whose mixed-radix FFT takes
any prime; pinned against numpy in tests/test_oracle_math.py): Green apply, RHS, operator, a
permeability column at tol 1e-12 (the contract: within 1e-10), and the FP32 operator of a MIXED
context within the FP32 bound."""
import numpy as np
import pytest

from conftest import random_mask, rel

pytestmark = pytest.mark.gpu

PRIME_SHAPES = [(17, 12, 10), (19, 17, 23), (34, 8, 6), (101, 6, 4), (8, 6, 37)]


def _pair(oracle_mod, shape, mask, precision=0):
    from fft_method_test_amd import permeability as P

    g = P.Grid3D(*shape)
    green = P.GreenOperator(g, 1.0, 2, precision)
    sup = P.ForceSupport(P.BinaryMedium(g, mask), P.SUPPORT_INTERFACE, green)
    o = oracle_mod.Oracle(shape)
    o.set_medium(mask)
    return green, sup, o


@pytest.mark.parametrize("shape", PRIME_SHAPES)
def test_fft_hooks_prime_edges(fp, shape):
    from fft_method_test_amd import permeability as P

    nx, ny, nz = shape
    N = nx * ny * nz
    green = P.GreenOperator(P.Grid3D(*shape))
    try:
        f = np.random.default_rng(3).standard_normal(3 * N)
        spec = green.fft_forward(f).reshape(3, nz, ny, nx)
        ref = np.stack([np.fft.fftn(f[c * N:(c + 1) * N].reshape(nz, ny, nx)) for c in range(3)])
        assert rel(spec, ref) < 1e-13
        assert rel(green.fft_inverse(spec.reshape(-1)), f) < 1e-13
    finally:
        green.close()


@pytest.mark.parametrize("shape", PRIME_SHAPES)
def test_operator_and_k_prime_edges(fp, oracle_mod, shape):
    from fft_method_test_amd import permeability as P

    mask = random_mask(shape, 0.65, seed=sum(shape))
    green, sup, o = _pair(oracle_mod, shape, mask)
    try:
        act_o, _ = o.support()
        assert np.array_equal(sup.active_voxels().astype(np.uint64), act_o)
        N = int(np.prod(shape))
        f = np.random.default_rng(1).standard_normal(3 * N)
        assert rel(green.apply(f), o.green_apply(f)) < 1e-12
        grad = [1.0, -0.3, 0.2]
        x = np.random.default_rng(5).standard_normal(sup.size())
        assert rel(P.VariationalOperator(sup, green).apply(x), o.apply(x)) < 1e-12
        avg, info, _ = P.solve_one_direction(green, sup, grad, 1e-12, 20000)
        avg_o, info_o = o.solve_direction(grad, 1e-12, 20000)
        err = rel(avg, avg_o)
        print(f"[prime] {shape}: nb {sup.size() // 3} K column rel err {err:.2e} iterations "
              f"{info.iterations}/{info_o.iterations}", flush=True)
        assert info.converged and err < 1e-10
    finally:
        green.close()
        o.close()


@pytest.mark.parametrize("shape", [(17, 12, 10), (101, 6, 4)])
def test_fp32_operator_prime_edges(fp, oracle_mod, shape):
    from fft_method_test_amd import _lib, permeability as P

    mask = random_mask(shape, 0.6, seed=7)
    green, sup, o = _pair(oracle_mod, shape, mask, _lib.PRECISION_MIXED)
    try:
        x = np.random.default_rng(2).standard_normal(sup.size())
        op = P.VariationalOperator(sup, green)
        ref = o.apply(x)
        assert rel(op.apply_fp32(x), ref) < 1e-6
        assert rel(op.apply_fp32v(x), ref) < 1e-6
    finally:
        green.close()
        o.close()
