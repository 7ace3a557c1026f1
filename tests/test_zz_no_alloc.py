"""fftperm_support_noalloc (the allocation-free support map) against the context path it mirrors,
fftperm_set_medium_dev + fftperm_get_support, through the C ABI.

Bit-exact on index work: the ascending active list, voxel_to_local (-1 off the support) and the
counts must equal the context path's on every call.  Two calls in a row, the second on a smaller
grid (the static scratch keeps the first call's contents, so the second must not depend on them),
plus both support modes, a capacity too small for the support, a grid over the static limit and a
device-side counts output.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from fft_method_test_amd import _lib as L

    lib = L.load()
    f = lib.fftperm_support_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_uint64] * 3 + [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    return L, lib


def _context_support(shape, mask_dev, mode):
    """The existing entry points: fftperm_set_medium_dev + fftperm_get_support on a context."""
    from fft_method_test_amd import permeability as P

    L, lib = _lib()
    green = P.GreenOperator(P.Grid3D(*shape))
    nb = C.c_uint64()
    L.check(lib.fftperm_set_medium_dev(green._h, C.c_void_p(mask_dev.data_ptr()), mode, C.byref(nb)))
    N = int(np.prod(shape))
    act = np.zeros(max(nb.value, 1), dtype=np.uint64)
    v2l = np.zeros(N, dtype=np.int64)
    L.check(lib.fftperm_get_support(green._h, act.ctypes.data_as(C.c_void_p), v2l.ctypes.data_as(C.c_void_p)))
    green.close()
    return int(nb.value), act[: nb.value].astype(np.int64), v2l


def _noalloc(shape, mask_dev, mode, capacity=None, want_counts_dev=False):
    import torch

    L, lib = _lib()
    nx, ny, nz = shape
    N = nx * ny * nz
    cap = N if capacity is None else capacity
    act = torch.full((max(cap, 1),), -7, dtype=torch.int64, device="cuda")
    v2l = torch.full((N,), -7, dtype=torch.int64, device="cuda")
    cnt = torch.zeros(2, dtype=torch.int64, device="cuda") if want_counts_dev else None
    torch.cuda.synchronize()
    nb, nf = C.c_uint64(), C.c_uint64()
    rc = lib.fftperm_support_noalloc(nx, ny, nz, C.c_void_p(mask_dev.data_ptr()), mode, C.c_void_p(v2l.data_ptr()),
                                     C.c_void_p(act.data_ptr()), cap,
                                     C.c_void_p(cnt.data_ptr()) if cnt is not None else None, None,
                                     C.byref(nb), C.byref(nf))
    torch.cuda.synchronize()
    return rc, int(nb.value), int(nf.value), act.cpu().numpy(), v2l.cpu().numpy(), \
        (cnt.cpu().numpy() if cnt is not None else None)


@pytest.mark.parametrize("mode", [0, 1])
def test_noalloc_support_matches_context_path(fp, mode):
    import torch

    from conftest import random_mask

    for call, shape in enumerate([(64, 48, 40), (24, 20, 16)]):
        mask = random_mask(shape, 0.62, seed=31 + call + 10 * mode)
        mask_dev = torch.from_numpy(mask).to("cuda")
        nb_ref, act_ref, v2l_ref = _context_support(shape, mask_dev, mode)
        rc, nb, nf, act, v2l, cnt = _noalloc(shape, mask_dev, mode, want_counts_dev=True)
        print(f"[noalloc] call {call} shape {shape} mode {mode}: rc {rc} nb {nb} (context {nb_ref})", flush=True)
        assert rc == 0
        assert nb == nb_ref and nf == int(mask.sum())
        assert np.array_equal(act[:nb], act_ref)
        assert np.all(act[nb:] == -7)  # nothing written past the support
        assert np.array_equal(v2l, v2l_ref)
        assert cnt.tolist() == [nb, nf]


def test_noalloc_support_limits(fp):
    import torch

    from conftest import random_mask

    L, lib = _lib()
    shape = (16, 16, 16)
    mask = random_mask(shape, 0.6, seed=5)
    mask_dev = torch.from_numpy(mask).to("cuda")
    nb_ref, act_ref, _ = _context_support(shape, mask_dev, 1)
    # capacity one short: the call reports FFTPERM_ERANGE, writes only capacity entries, counts stay exact
    rc, nb, nf, act, v2l, _ = _noalloc(shape, mask_dev, 1, capacity=nb_ref - 1)
    assert rc == 2 and nb == nb_ref
    assert np.array_equal(act[: nb_ref - 1], act_ref[: nb_ref - 1])
    # a grid over the static scratch's limit is refused before any launch
    nbv = C.c_uint64()
    rc = lib.fftperm_support_noalloc(1024, 1024, 1025, C.c_void_p(mask_dev.data_ptr()), 1, None,
                                     C.c_void_p(mask_dev.data_ptr()), 0, None, None, C.byref(nbv), None)
    assert rc == 2 and b"FFTPERM_NOALLOC_MAX_VOXELS" in lib.fftperm_last_error()
    rc = lib.fftperm_support_noalloc(16, 16, 16, C.c_void_p(mask_dev.data_ptr()), 7, None, None, 0, None, None,
                                     None, None)
    assert rc == 1
    # and a normal call afterwards still matches
    rc, nb, nf, act, v2l, _ = _noalloc(shape, mask_dev, 1)
    assert rc == 0 and nb == nb_ref and np.array_equal(act[:nb], act_ref)
