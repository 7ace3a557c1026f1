"""HBM copy probes (fftperm_probe_copy): the contiguous stream and the y / z passes' column patterns
over a MIXED context's FP32 spectrum slots report a finite, positive rate; error paths."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def test_copy_probes_and_errors(fp):
    from fft_method_test_amd import _lib, permeability as P

    L = _lib.load()
    green = P.GreenOperator(P.Grid3D(64, 48, 40), 1.0, 2, _lib.PRECISION_MIXED)
    rate = C.c_double()
    try:
        for pattern in (0, 1, 2):
            _lib.check(L.fftperm_probe_copy(green._h, pattern, 3, C.byref(rate)))
            print(f"[probe] pattern {pattern}: {rate.value:.0f} GB/s", flush=True)
            assert 0.0 < rate.value < 1e5
        assert L.fftperm_probe_copy(green._h, 3, 1, C.byref(rate)) == _lib.EINVAL
        assert L.fftperm_probe_copy(green._h, 0, 0, C.byref(rate)) == _lib.EINVAL
    finally:
        green.close()
    g64 = P.GreenOperator(P.Grid3D(16, 16, 16), 1.0, 2, _lib.PRECISION_FP64)
    try:
        assert L.fftperm_probe_copy(g64._h, 0, 1, C.byref(rate)) == _lib.EINVAL
        assert "MIXED" in L.fftperm_last_error().decode()
    finally:
        g64.close()
