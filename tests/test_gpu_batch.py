"""GPU test of the batched solver: concurrent workers and graph-replayed MINRES give every
sample exactly the result of a solo solve of that sample."""
import ctypes as C

import numpy as np
import pytest

from conftest import random_mask

pytestmark = pytest.mark.gpu


def test_batch_matches_solo_solves(fp):
    from fft_method_test_amd import _lib, batch, geometry as G

    samples = []
    for i, n in enumerate((32, 48, 64, 32, 40, 64)):
        m = G.boolean_spheres(n, 5.0, 0.5 + 0.05 * i, 100 + i).reshape(-1) if i % 2 else random_mask((n,) * 3, 0.6, i)
        samples.append((n, m.astype(np.uint8), "t", 0.0))
    res, _ = batch.solve_batch(samples, tol=1e-8, precision=_lib.PRECISION_FP64, workers=3)
    L = _lib.load()
    for (n, mask, _, _), r in zip(samples, res):
        g = _lib.Grid(n, n, n, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0)
        h = C.c_void_p()
        _lib.check(L.fftperm_create(C.byref(g), 1.0, 2, _lib.PRECISION_FP64, C.byref(h)))
        nb = C.c_uint64()
        _lib.check(L.fftperm_set_medium(h, mask, _lib.SUPPORT_INTERFACE, C.byref(nb)))
        avg = np.zeros(3)
        info = _lib.SolveInfo()
        _lib.check(L.fftperm_solve_direction(h, np.array([1.0, 0.0, 0.0]), 1e-8, 20000, None, avg, None,
                                             C.byref(info)))
        L.fftperm_destroy(h)
        assert r["iterations"] == info.iterations
        assert r["k"] == (1.0 * avg).tolist()

