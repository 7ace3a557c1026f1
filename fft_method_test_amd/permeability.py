"""Host-side mirror of the reference's Bignonnet permeability classes, over libfftperm.so.

Class and method names, argument meaning and error behaviour follow the reference's
C++ API (namespace permeability); the compute runs on the GPU through the C ABI of
include/fftperm.h.  One fftperm context holds what the reference spreads over
GreenOperator (grid, Green tensor, FFT plans) and ForceSupport (medium + support map),
so ForceSupport takes the GreenOperator it attaches to as a third argument.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import PRECISION_FP64, PRECISION_MIXED, SUPPORT_FULLSOLID, SUPPORT_INTERFACE, check, f64, ptr

__all__ = [
    "Grid3D", "BinaryMedium", "GreenOperator", "ForceSupport", "BackgroundForce", "VariationalRHS",
    "VariationalOperator", "MinresSolver", "SolverResult", "VelocityRecovery", "PermeabilityTensor",
    "Diagnostics", "solve_one_direction", "solve_permeability_tensor",
    "SUPPORT_FULLSOLID", "SUPPORT_INTERFACE", "PRECISION_FP64", "PRECISION_MIXED",
]


class Grid3D:
    """src/core/Grid3D.hpp:14-93 (x-fastest flat index, periodic wrap, l = h*n default)."""

    def __init__(self, nx, ny, nz, hx=1.0, hy=1.0, hz=1.0, lx=0.0, ly=0.0, lz=0.0):
        if nx <= 0 or ny <= 0 or nz <= 0:
            raise ValueError("Grid dimensions must be positive")
        self._n = (int(nx), int(ny), int(nz))
        self._h = (float(hx), float(hy), float(hz))
        self._l = (lx if lx > 0 else hx * nx, ly if ly > 0 else hy * ny, lz if lz > 0 else hz * nz)

    def nx(self): return self._n[0]
    def ny(self): return self._n[1]
    def nz(self): return self._n[2]
    def hx(self): return self._h[0]
    def hy(self): return self._h[1]
    def hz(self): return self._h[2]
    def lx(self): return self._l[0]
    def ly(self): return self._l[1]
    def lz(self): return self._l[2]

    @property
    def shape(self):
        return self._n

    def total_size(self):
        nx, ny, nz = self._n
        return nx * ny * nz

    def flat_index(self, i, j, k):
        return (k * self._n[1] + j) * self._n[0] + i

    def unflatten(self, idx):
        i = idx % self._n[0]
        idx //= self._n[0]
        j = idx % self._n[1]
        return i, j, idx // self._n[1]

    @staticmethod
    def _wrap(v, n):
        m = v % n  # Python's % is already non-negative for n > 0
        return m

    def periodic_index(self, i, n):
        return self._wrap(i, n)

    def periodic_neighbor(self, i, j, k, di, dj, dk):
        nx, ny, nz = self._n
        return self.flat_index(self._wrap(i + di, nx), self._wrap(j + dj, ny), self._wrap(k + dk, nz))

    def _c(self):
        nx, ny, nz = self._n
        return _lib.Grid(nx, ny, nz, *self._h, *self._l)


class BinaryMedium:
    """src/geometry/BinaryMedium.cpp:7-43: uint8 mask, 1 = fluid, 0 = solid."""

    def __init__(self, grid: Grid3D, fluid_mask):
        m = np.ascontiguousarray(np.asarray(fluid_mask).reshape(-1) != 0, dtype=np.uint8)
        if m.size != grid.total_size():
            raise ValueError("BinaryMedium mask size must match grid size")
        self._grid = grid
        self._mask = m

    def grid(self):
        return self._grid

    def size(self):
        return self._mask.size

    def mask(self):
        return self._mask

    def is_fluid(self, idx):
        return bool(self._mask[idx])

    def is_solid(self, idx):
        return not self.is_fluid(idx)

    def porosity(self):
        return float(np.count_nonzero(self._mask)) / float(self._mask.size) if self._mask.size else 0.0


class GreenOperator:
    """GreenOperator (src/bignonnet/GreenOperator.cpp:27-65) + its FFT backend, on the GPU.

    precision: PRECISION_FP64 (the reference at FFT_METHOD_PRECISION=DOUBLE) or
    PRECISION_MIXED (FP32 transforms inside MINRES, FP64 everywhere else)."""

    def __init__(self, grid: Grid3D, viscosity=1.0, p_radius=2, precision=PRECISION_FP64):
        L = _lib.load()
        self._grid = grid
        self.viscosity = float(viscosity)
        self.p_radius = max(0, int(p_radius))
        self.precision = precision
        h = C.c_void_p()
        g = grid._c()
        # a grid edge the FFT plans cannot take is reported before any device work
        check(L.fftperm_check_edges(*grid.shape))
        check(L.fftperm_create(C.byref(g), float(viscosity), int(p_radius), int(precision), C.byref(h)))
        self._h = h
        self._tensor = None

    def set_refinement(self, fp32_below=2.0, inner_floor=3e-6, max_sweeps=64):
        """MIXED mode precision schedule: sweeps to max(inner_floor, 0.5 tol / r); the first one applies
        the FP64 operator while its recurrence residual is above fp32_below (>= 1: FP32 throughout)."""
        check(_lib.load().fftperm_set_refinement(self._h, float(fp32_below), float(inner_floor), int(max_sweeps)))

    def set_mixed_solver(self, solver=_lib.MIXED_CR_FP32_VECTORS, reliable_delta=0.1):
        """MIXED mode solver: CR with reliable updates on FP32 (default) or FP64 vectors, CG with
        reliable updates, or refined MINRES sweeps."""
        check(_lib.load().fftperm_set_mixed_solver(self._h, int(solver), float(reliable_delta)))

    def set_tensor_batch(self, mode=1):
        """Iterate the three load directions of a MIXED tensor solve in lockstep, each direction's
        FFT passes back to back (1, the default) or one launch per pass for the three (2), or one
        direction after the other (0); the results are identical."""
        check(_lib.load().fftperm_set_tensor_batch(self._h, int(mode)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().fftperm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid(self):
        return self._grid

    def spectral_size(self):
        return self._grid.total_size()

    def _tensor6(self):
        if self._tensor is None:
            t = np.zeros(6 * self._grid.total_size())
            check(_lib.load().fftperm_green_tensor(self._h, t))
            self._tensor = t.reshape(6, -1)
        return self._tensor

    def tensor_components(self, idx):
        N = self._grid.total_size()
        if idx < 0 or idx >= N:
            raise IndexError("GreenOperator::tensor_components index out of range")
        return tuple(self._tensor6()[:, idx])

    def tensor(self):
        """All six reference-layout arrays (xx, xy, xz, yy, yz, zz) of N."""
        return self._tensor6().copy()

    def has_nan_in_tensor(self):
        return not bool(np.all(np.isfinite(self._tensor6())))

    def apply(self, force):
        N = self._grid.total_size()
        f = f64(force, 3 * N)
        u = np.zeros(3 * N)
        check(_lib.load().fftperm_green_apply(self._h, f, u))
        return u

    # IFFTBackend hooks (src/fft/IFFTBackend.hpp:12-50)
    def fft_forward(self, real3):
        N = self._grid.total_size()
        out = np.zeros(6 * N)
        check(_lib.load().fftperm_fft_forward(self._h, f64(real3, 3 * N), out))
        return out.view(np.complex128)

    def fft_inverse(self, spec3):
        N = self._grid.total_size()
        s = np.ascontiguousarray(spec3, dtype=np.complex128).reshape(-1)
        if s.size != 3 * N:
            raise ValueError("spectrum size mismatch")
        out = np.zeros(3 * N)
        check(_lib.load().fftperm_fft_inverse(self._h, s.view(np.float64), out))
        return out

    def spectral_green(self, t6, force_spec):
        N = self._grid.total_size()
        fs = np.ascontiguousarray(force_spec, dtype=np.complex128).reshape(-1)
        out = np.zeros(6 * N)
        check(_lib.load().fftperm_spectral_green(self._h, f64(t6, 6 * N), fs.view(np.float64), out))
        return out.view(np.complex128)


class ForceSupport:
    """ForceSupport (src/bignonnet/ForceSupport.cpp:9-47), built on the GPU."""

    kInvalid = -1

    def __init__(self, medium: BinaryMedium, mode=SUPPORT_INTERFACE, green: GreenOperator = None):
        if green is None:
            raise ValueError("ForceSupport needs the GreenOperator whose context holds the medium")
        if green.grid().total_size() != medium.size():
            raise ValueError("ForceSupport: medium and Green operator grids differ")
        self._green = green
        self._medium = medium
        nb = C.c_uint64()
        check(_lib.load().fftperm_set_medium(green._h, medium.mask(), int(mode), C.byref(nb)))
        self._nb = int(nb.value)
        self._active = None
        self._v2l = None

    def _fetch(self):
        if self._active is None:
            act = np.zeros(max(1, self._nb), dtype=np.uint64)
            v2l = np.zeros(self._medium.size(), dtype=np.int64)
            check(_lib.load().fftperm_get_support(self._green._h, ptr(act), ptr(v2l)))
            self._active = act[: self._nb]
            self._v2l = v2l

    def grid(self):
        return self._medium.grid()

    def size(self):
        return 3 * self._nb

    def num_active_voxels(self):
        return self._nb

    def active_voxels(self):
        self._fetch()
        return self._active

    def voxel_to_local(self):
        self._fetch()
        return self._v2l

    def contains_voxel(self, v):
        self._fetch()
        return 0 <= v < self._v2l.size and self._v2l[v] != self.kInvalid

    def local_index(self, v):
        if not self.contains_voxel(v):
            raise IndexError("Voxel is not in active support")
        return int(self._v2l[v])

    def global_index(self, local):
        if local < 0 or local >= self._nb:
            raise IndexError("Support local index out of range")
        return int(self.active_voxels()[local])


class BackgroundForce:
    """src/bignonnet/BackgroundForce.cpp:14-60 (f0 = -g on fluid, -NF/NB g on the support)."""

    def __init__(self, medium: BinaryMedium, support: ForceSupport):
        if medium.size() != support.grid().total_size():
            raise ValueError("BackgroundForce requires medium and support on same grid")
        self._support = support

    def build(self, macroscopic_gradient):
        N = self._support.grid().total_size()
        out = np.zeros(3 * N)
        check(_lib.load().fftperm_background_force(self._support._green._h, f64(macroscopic_gradient, 3), out))
        return out


class VariationalRHS:
    """src/bignonnet/VariationalRHS.cpp:15-31."""

    def __init__(self, support: ForceSupport, green: GreenOperator, background_force: BackgroundForce):
        self._support, self._green = support, green

    def build(self, macroscopic_gradient):
        b = np.zeros(self._support.size())
        check(_lib.load().fftperm_rhs(self._green._h, f64(macroscopic_gradient, 3), b))
        return b


class VariationalOperator:
    """ILinearOperator of src/bignonnet/VariationalOperator.cpp:19-41 (FP64 apply)."""

    def __init__(self, support: ForceSupport, green: GreenOperator):
        self._support, self._green = support, green

    def size(self):
        return self._support.size()

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        if x.size != self.size():
            raise ValueError("VariationalOperator::apply input size mismatch")
        y = np.zeros(self.size())
        check(_lib.load().fftperm_operator_apply(self._green._h, x, y))
        return y

    def apply_fp32(self, x):
        """The FP32 apply the MIXED-mode sweeps use (MIXED contexts only; no reference counterpart)."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        if x.size != self.size():
            raise ValueError("VariationalOperator::apply input size mismatch")
        y = np.zeros(self.size())
        check(_lib.load().fftperm_operator_apply_fp32(self._green._h, x, y))
        return y

    def apply_fp32v(self, x):
        """The FP32 apply on FP32 compact vectors that the default MIXED solver iterates (MIXED
        contexts only; no reference counterpart)."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        if x.size != self.size():
            raise ValueError("VariationalOperator::apply input size mismatch")
        y = np.zeros(self.size())
        check(_lib.load().fftperm_operator_apply_fp32v(self._green._h, x, y))
        return y


@dataclass
class SolverResult:
    """src/solver/SolverResult.hpp:10-16."""
    solution: np.ndarray
    converged: bool = False
    iterations: int = 0
    final_rel_residual: float = 1.0
    residual_history: np.ndarray = field(default_factory=lambda: np.zeros(0))
    refinement_sweeps: int = 0


class MinresSolver:
    """MinresSolver (src/solver/MinresSolver.cpp:39-218), device-resident."""

    def __init__(self, tolerance, max_iterations):
        if not tolerance > 0:
            raise ValueError("MinresSolver tolerance must be positive")
        self.tolerance = float(tolerance)
        self.max_iterations = int(max_iterations)

    def solve(self, op: VariationalOperator, rhs, initial_guess=None) -> SolverResult:
        n = op.size()
        b = np.ascontiguousarray(rhs, dtype=np.float64).reshape(-1)
        if b.size != n:
            raise ValueError("MinresSolver::solve rhs size mismatch")
        x0 = None
        if initial_guess is not None and len(initial_guess) > 0:
            x0 = np.ascontiguousarray(initial_guess, dtype=np.float64).reshape(-1)
            if x0.size != n:
                raise ValueError("MinresSolver::solve initial guess size mismatch")
        x = np.zeros(n)
        hist = np.zeros(self.max_iterations + 70)
        info = _lib.SolveInfo()
        check(_lib.load().fftperm_minres(op._green._h, b, ptr(x0), self.tolerance, self.max_iterations, x,
                                         ptr(hist), C.byref(info)))
        return SolverResult(solution=x, converged=bool(info.converged), iterations=int(info.iterations),
                            final_rel_residual=float(info.final_rel_residual),
                            residual_history=hist[: info.history_len].copy(),
                            refinement_sweeps=int(info.refinement_sweeps))


class VelocityRecovery:
    """src/bignonnet/VelocityRecovery.cpp:12-55."""

    def __init__(self, support: ForceSupport, green: GreenOperator, background_force: BackgroundForce):
        self._support, self._green = support, green

    def recover(self, macroscopic_gradient, compact_solution):
        x = np.ascontiguousarray(compact_solution, dtype=np.float64).reshape(-1)
        if x.size != self._support.size():
            raise ValueError("VelocityRecovery::recover compact solution size mismatch")
        N = self._support.grid().total_size()
        u = np.zeros(3 * N)
        avg = np.zeros(3)
        check(_lib.load().fftperm_recover(self._green._h, f64(macroscopic_gradient, 3), x, ptr(u), avg))
        return u

    @staticmethod
    def average_over_domain(velocity):
        v = np.asarray(velocity).reshape(3, -1)
        return v.sum(axis=1) * (1.0 / v.shape[1])

    def average_over_support(self, velocity):
        v = np.asarray(velocity).reshape(3, -1)
        act = self._support.active_voxels().astype(np.int64)
        if act.size == 0:
            return np.zeros(3)
        return v[:, act].sum(axis=1) * (1.0 / act.size)


class PermeabilityTensor:
    """src/post/PermeabilityTensor.cpp:9-32 (K column j = -<v^(j)>)."""

    def __init__(self):
        self._k = np.zeros(9)

    def set_column(self, column, avg_velocity):
        if column < 0 or column >= 3:
            raise IndexError("PermeabilityTensor::set_column index out of range")
        for row in range(3):
            self._k[row * 3 + column] = -avg_velocity[row]

    def data(self):
        return self._k.copy()

    def to_string(self, indent=""):
        rows = []
        for r in range(3):
            rows.append("  ".join(f"{self._k[3 * r + c]:.6e}" for c in range(3)))
        return ("\n" + indent).join(rows)


class Diagnostics:
    """src/post/Diagnostics.cpp:26-64 (FP64 on the GPU)."""

    @staticmethod
    def residual_norm(op: VariationalOperator, x, b):
        out = C.c_double()
        check(_lib.load().fftperm_residual_norm(op._green._h, f64(x, op.size()), f64(b, op.size()), C.byref(out)))
        return out.value

    @staticmethod
    def relative_symmetry_gap(op: VariationalOperator, x, y):
        out = C.c_double()
        check(_lib.load().fftperm_symmetry_gap(op._green._h, f64(x, op.size()), f64(y, op.size()), C.byref(out)))
        return out.value


def solve_one_direction(green: GreenOperator, support: ForceSupport, grad, tolerance, max_iterations,
                        initial_guess=None, want_velocity=False):
    """solve_one_direction (src/app/PermeabilityRunner.cpp:173-242): returns (avg_velocity, info, u|None)."""
    if support.num_active_voxels() == 0:
        raise RuntimeError("Support set is empty; cannot run variational solve")
    N = green.grid().total_size()
    avg = np.zeros(3)
    u = np.zeros(3 * N) if want_velocity else None
    info = _lib.SolveInfo()
    x0 = None if initial_guess is None else f64(initial_guess, support.size())
    check(_lib.load().fftperm_solve_direction(green._h, f64(grad, 3), float(tolerance), int(max_iterations),
                                              ptr(x0), avg, ptr(u), C.byref(info)))
    return avg, info, u


def solve_permeability_tensor(green: GreenOperator, tolerance, max_iterations):
    """PermeabilityRunner::run for gradient=all: K (3x3) = +mu <v_r>^(c), plus per-direction infos."""
    K = np.zeros(9)
    infos = (_lib.SolveInfo * 3)()
    check(_lib.load().fftperm_solve_tensor(green._h, float(tolerance), int(max_iterations), K,
                                           C.cast(infos, C.c_void_p)))
    return K.reshape(3, 3), list(infos)
