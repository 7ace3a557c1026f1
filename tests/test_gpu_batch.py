"""GPU tests of the batched solves: every medium of a batch gets exactly the result of a solo solve
of that medium (fftperm_solve_direction), whether the batch runs size class by size class through
fftperm_solve_batch (one launch per y / z pass for a whole class, lockstep CR, graph replay) or
sample by sample on concurrent workers."""
import ctypes as C

import numpy as np
import pytest

from conftest import random_mask

pytestmark = pytest.mark.gpu


def _samples():
    from fft_method_test_amd import geometry as G

    out = []
    for i, n in enumerate((32, 48, 64, 32, 40, 64, 32, 48)):
        m = G.boolean_spheres(n, 5.0, 0.5 + 0.04 * i, 100 + i).reshape(-1) if i % 2 else random_mask((n,) * 3, 0.6, i)
        out.append((n, m.astype(np.uint8), "t", 0.0))
    return out


def _solo(mask, n, precision, tol):
    from fft_method_test_amd import _lib

    L = _lib.load()
    g = _lib.Grid(n, n, n, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0)
    h = C.c_void_p()
    _lib.check(L.fftperm_create(C.byref(g), 1.0, 2, precision, C.byref(h)))
    nb = C.c_uint64()
    _lib.check(L.fftperm_set_medium(h, mask, _lib.SUPPORT_INTERFACE, C.byref(nb)))
    avg = np.zeros(3)
    info = _lib.SolveInfo()
    _lib.check(L.fftperm_solve_direction(h, np.array([1.0, 0.0, 0.0]), tol, 20000, None, avg, None, C.byref(info)))
    L.fftperm_destroy(h)
    return avg, info


def test_per_sample_batch_matches_solo_solves_fp64(fp):
    from fft_method_test_amd import _lib, batch

    samples = _samples()
    res, _ = batch.solve_batch(samples, tol=1e-8, precision=_lib.PRECISION_FP64, workers=3, mode="per_sample")
    for (n, mask, _, _), r in zip(samples, res):
        avg, info = _solo(mask, n, _lib.PRECISION_FP64, 1e-8)
        assert r["iterations"] == info.iterations
        assert r["k"] == (1.0 * avg).tolist()


@pytest.mark.parametrize("max_batch", [64, 2])
def test_batched_size_classes_match_solo_solves_mixed(fp, max_batch):
    """Size classes of three, two, two and one media (or chunks of at most two): iterations and K
    column bit-identical to each medium's own MIXED solve."""
    from fft_method_test_amd import _lib, batch

    samples = _samples()
    res, _ = batch.solve_batch(samples, tol=1e-7, precision=_lib.PRECISION_MIXED, workers=2, mode="batched",
                               max_batch=max_batch)
    for (n, mask, _, _), r in zip(samples, res):
        avg, info = _solo(mask, n, _lib.PRECISION_MIXED, 1e-7)
        print(f"[batch] n={n} batch={r['batch']} iterations {r['iterations']} solo {info.iterations}", flush=True)
        assert r["converged"] and info.converged
        assert r["iterations"] == info.iterations
        assert r["k"] == (1.0 * avg).tolist()


def test_solve_batch_arguments(fp):
    """fftperm_solve_batch refuses an FP64 context, an empty or oversized batch and a bad support
    mode; a medium without support gets k = 0 and converged = 0 while the others solve."""
    from fft_method_test_amd import _lib

    L = _lib.load()
    n = 16
    g = _lib.Grid(n, n, n, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0)
    masks = [random_mask((n,) * 3, 0.6, 3).astype(np.uint8), np.ones(n ** 3, dtype=np.uint8)]
    ptrs = (C.c_void_p * 2)(*[m.ctypes.data for m in masks])
    k = np.zeros(6)
    infos = (_lib.SolveInfo * 2)()
    grad = np.array([1.0, 0.0, 0.0])
    h64 = C.c_void_p()
    _lib.check(L.fftperm_create(C.byref(g), 1.0, 2, _lib.PRECISION_FP64, C.byref(h64)))
    assert L.fftperm_solve_batch(h64, 2, ptrs, 1, grad, 1e-6, 100, k, None) == _lib.EINVAL
    L.fftperm_destroy(h64)
    h = C.c_void_p()
    _lib.check(L.fftperm_create(C.byref(g), 1.0, 2, _lib.PRECISION_MIXED, C.byref(h)))
    assert L.fftperm_solve_batch(h, 0, ptrs, 1, grad, 1e-6, 100, k, None) == _lib.EINVAL
    assert L.fftperm_solve_batch(h, 65, ptrs, 1, grad, 1e-6, 100, k, None) == _lib.EINVAL
    assert L.fftperm_solve_batch(h, 2, ptrs, 7, grad, 1e-6, 100, k, None) == _lib.EINVAL
    _lib.check(L.fftperm_solve_batch(h, 2, ptrs, 1, grad, 1e-6, 2000, k, C.cast(infos, C.c_void_p)))
    L.fftperm_destroy(h)
    avg, info = _solo(masks[0], n, _lib.PRECISION_MIXED, 1e-6)
    assert k[:3].tolist() == (1.0 * avg).tolist() and infos[0].iterations == info.iterations
    assert k[3:].tolist() == [0.0, 0.0, 0.0] and infos[1].converged == 0
