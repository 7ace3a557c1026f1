"""CPU checks of the C ABI library (loads, exports every declared symbol; no compute call)
and of the host-side modules (geometry generators, voxel/velocity file formats)."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    hdr_dir = os.path.join(ROOT, "include")
    for fn in os.listdir(hdr_dir):
        if fn.endswith(".h"):
            text = open(os.path.join(hdr_dir, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names.update(re.findall(r"\b(fftperm_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import ctypes as C

    from fft_method_test_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfftperm.so not built (run __graft_entry__.build())")
    lib = C.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.EXPORTED_SYMBOLS)
    lib.fftperm_last_error.restype = C.c_char_p
    assert lib.fftperm_last_error() == b""  # touches no device


def test_geometry_generators_deterministic():
    from fft_method_test_amd import geometry as G

    a = G.boolean_spheres(48, 4.0, 0.6, 2001)
    b = G.boolean_spheres(48, 4.0, 0.6, 2001)
    c = G.boolean_spheres(48, 4.0, 0.6, 2002)
    assert a.dtype == np.uint8 and a.shape == (48, 48, 48)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert abs(a.mean() - 0.6) < 0.05
    s = G.sphere_cell(32)
    assert abs(s.mean() - 0.8887) < 1e-3  # r = 0.3 L, voxel-centre test
    e = G.boolean_ellipsoids(40, (6.0, 2.0, 2.0), 0.7, 3)
    assert abs(e.mean() - 0.7) < 0.08


def test_fgeo_round_trip_and_errors(tmp_path):
    from fft_method_test_amd import voxelio

    ids = np.random.default_rng(0).integers(0, 5, size=(3, 4, 5)).astype(np.uint32)
    for vt in (1, 2):
        p = tmp_path / f"a{vt}.fgeo"
        voxelio.write_fgeo(p, ids, value_type=vt)
        assert os.path.getsize(p) == 32 + ids.size * (4 if vt == 1 else 2)
        assert np.array_equal(voxelio.read_fgeo(p), ids)
        assert np.array_equal(voxelio.load(str(p), fluid_id=3), (ids == 3).astype(np.uint8))
    bad = tmp_path / "bad.fgeo"
    bad.write_bytes(b"FGEX" + bytes(28))
    with pytest.raises(RuntimeError, match="invalid magic"):
        voxelio.read_fgeo(bad)
    trunc = tmp_path / "t.fgeo"
    voxelio.write_fgeo(trunc, ids)
    trunc.write_bytes(trunc.read_bytes()[:-2])
    with pytest.raises(RuntimeError, match="payload size mismatch"):
        voxelio.read_fgeo(trunc)


def test_geom_text_round_trip(tmp_path):
    from fft_method_test_amd import voxelio

    ids = np.random.default_rng(1).integers(0, 3, size=(2, 3, 4)).astype(np.uint32)
    p = tmp_path / "g.geom"
    voxelio.write_geom_text(p, ids)
    txt = "# comment\n; other\n" + p.read_text()
    p.write_text(txt)
    assert np.array_equal(voxelio.read_geom_text(p), ids)


def test_fvel_layout(tmp_path):
    """FVEL v1 header (src/app/PermeabilityRunner.cpp:109-159): 4+1+1+2+3*8+8+8+3*8 = 72 bytes."""
    from fft_method_test_amd import voxelio

    v = np.random.default_rng(2).standard_normal((3, 24))
    p = tmp_path / "u.fvel"
    voxelio.write_fvel(p, (2, 3, 4), 0.7, 1.5, (1.0, 0.0, 0.0), v)
    raw = p.read_bytes()
    assert raw[:4] == b"FVEL" and raw[4] == 1 and len(raw) == 72 + 24 * 3 * 8
    assert np.frombuffer(raw[72:96], dtype="<f8").tolist() == [v[0, 0], v[1, 0], v[2, 0]]
    d = voxelio.read_fvel(p)
    assert d["shape"] == (2, 3, 4) and d["porosity"] == 0.7 and np.array_equal(d["velocity"], v)


def test_bench_refuses_multi_gpu():
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_c4_batch_generator_is_seeded():
    """The uneven-size batch: sizes from {64, ..., 160}, porosity in [0.3, 0.8], seeded."""
    from fft_method_test_amd import batch

    a = batch.c4_batch(3, 4000)
    b = batch.c4_batch(3, 4000)
    assert [s[0] for s in a] == [s[0] for s in b]
    assert all(np.array_equal(x[1], y[1]) for x, y in zip(a, b))
    assert all(s[0] in (64, 80, 96, 112, 128, 144, 160) for s in a)
    assert all(0.3 <= s[3] <= 0.8 for s in a)


REF_SRC = "/root/reference/src"


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference sources not present")
@pytest.mark.parametrize("precision", ["float", "double"])
def test_cpp_binding_compiles_against_reference_headers(precision):
    """fft_method_test_amd/host/*.hpp implement the reference's IFFTBackend, ILinearOperator,
    MinresSolver::solve, solve_one_direction and solveBrinkman plug points: compile them against
    the reference's own headers in both of its precisions (src/core/Types.hpp:14-21)."""
    import subprocess
    import tempfile

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = ('#include "fftperm_reference.hpp"\n#include "fftperm_brinkman.hpp"\n'
           "static_assert(!std::is_abstract_v<permeability::hip::HipFFTBackend>);\n"
           "static_assert(!std::is_abstract_v<permeability::hip::HipVariationalOperator>);\n"
           "int main() { return 0; }\n")
    with tempfile.NamedTemporaryFile("w", suffix=".cpp", delete=False) as f:
        f.write(src)
    flags = ["-DFFT_METHOD_PRECISION_DOUBLE"] if precision == "double" else []
    cmd = ["g++", "-std=c++20", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", *flags, "-I" + REF_SRC,
           "-I/root/reference/include", "-I" + os.path.join(root, "include"),
           "-I" + os.path.join(root, "fft_method_test_amd", "host"), f.name]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    os.unlink(f.name)
    assert r.returncode == 0, r.stdout
