"""The reference's own classes driving the GPU through the C++ binding (fft_method_test_amd/host).

oracle/_ref/dropin_check_f64 and _f32 are built by oracle/Makefile from the reference sources
(Scalar = double and the reference's default Scalar = float) plus oracle/ref_dropin_check.cpp,
linked against libfftperm.so.  Each run computes the permeability tensor three ways -- the
reference's GreenOperator/MinresSolver path with HipFFTBackend as its FFT backend, the
reference's MinresSolver over HipVariationalOperator, and the all-device binding -- plus the
reference's solveBrinkman against the binding's GPU branch; every result is checked here against
the oracle (or, for Brinkman, against the reference's own host loop in the same process).
"""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = {64: os.path.join(ROOT, "oracle", "_ref", "dropin_check_f64"),
       32: os.path.join(ROOT, "oracle", "_ref", "dropin_check_f32")}


def _run(bits, mask, n, tol, maxit, tmp_path, penalty=1e4, b_maxit=60):
    exe = BIN[bits]
    if not os.path.exists(exe):
        pytest.skip("drop-in check not built (needs the reference sources at build time)")
    path = str(tmp_path / f"mask{bits}.u8")
    np.ascontiguousarray(mask, dtype=np.uint8).tofile(path)
    print(f"[dropin] {os.path.basename(exe)} n={n} tol={tol}", flush=True)
    r = subprocess.run([exe, path, str(n), str(n), str(n), repr(tol), str(maxit), "1.0", "2", repr(penalty),
                        str(b_maxit)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=600)
    assert r.returncode == 0, r.stderr + r.stdout
    return json.loads(r.stdout)


def _media():
    from fft_method_test_amd import geometry as G

    return {"sphere32": (32, G.sphere_cell(32).reshape(-1)),
            "boolean24": (24, G.boolean_spheres(24, 4.0, 0.6, 11).reshape(-1))}


@pytest.mark.parametrize("name", ["sphere32", "boolean24"])
def test_dropin_fp64(oracle_mod, tmp_path, name):
    n, mask = _media()[name]
    out = _run(64, mask, n, 1e-12, 5000, tmp_path)
    assert out["scalar_bytes"] == 8 and out["support_equal"] is True
    assert out["operator_sizes_equal"] is True and out["operator_apply_rel_diff"] <= 1e-13
    o = oracle_mod.Oracle((n, n, n))
    o.set_medium(mask)
    Ko, io = o.solve_tensor(1e-12, 5000)
    for path in ("reference_over_hip_fft", "reference_minres_over_hip_operator", "hip_tensor"):
        K = np.array(out[path]["K"]).reshape(3, 3)
        err = rel(K, Ko)
        print(f"[dropin] {name} fp64 {path}: rel K err {err:.2e} iterations {out[path]['iterations']} "
              f"oracle {[int(i.iterations) for i in io]}", flush=True)
        assert err <= 1e-10
    assert out["hip_history_len"] == out["hip_tensor"]["iterations"][0] + 1
    b = out["brinkman"]
    # same Jacobi sweeps in FP64: equal iteration counts, U to rounding of the final mean
    assert b["hip_iterations"] == b["reference_iterations"] and b["porosity_equal"] is True
    assert rel(np.array(b["hip_U"]), np.array(b["reference_U"])) <= 1e-12
    assert all(out["errors"].values()), out["errors"]


def test_dropin_fp32_reference_build(oracle_mod, tmp_path):
    """The reference's default build (Scalar = float): the binding widens to double at the ABI and
    rounds back.  One apply of HipVariationalOperator matches the reference's float
    VariationalOperator (over HipFFTBackend) to float rounding, and the all-device path, which
    computes in FP64 and rounds only b, x and <u>, keeps K within 1e-5 of the FP64 oracle at the same
    tol.  The reference's own float MinresSolver does not reach tol 1e-6 on this cell within 5000
    iterations over either operator (float Lanczos loses orthogonality; 139 iterations in FP64), so
    the two paths that run it are only required to return finite tensors: their K errors (2.3e-3
    and 1.0e-1 measured) are the reference's float solver's, not the binding's."""
    n, mask = _media()["sphere32"]
    out = _run(32, mask, n, 1e-6, 5000, tmp_path)
    assert out["scalar_bytes"] == 4 and out["support_equal"] is True
    assert out["operator_sizes_equal"] is True and out["operator_apply_rel_diff"] <= 1e-5
    o = oracle_mod.Oracle((n, n, n))
    o.set_medium(mask)
    Ko, _ = o.solve_tensor(1e-6, 5000)
    for path in ("hip_tensor", "reference_over_hip_fft", "reference_minres_over_hip_operator"):
        K = np.array(out[path]["K"]).reshape(3, 3)
        err = rel(K, Ko)
        print(f"[dropin] sphere32 fp32 {path}: rel K err {err:.2e} iterations {out[path]['iterations']}", flush=True)
        assert np.all(np.isfinite(K))
        if path == "hip_tensor":
            assert err <= 1e-5
    assert all(out["errors"].values()), out["errors"]
