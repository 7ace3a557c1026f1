"""Brinkman penalisation API mirroring common::fluid (include/brinkman/FluidBrinkman.hpp:15-55).

Host side: material-phase resolution and coefficient fields (src/brinkman/FluidBrinkman.cpp:67-191),
K assembly (:509-553).  The iteration itself (:282-507) runs on the GPU through
fftperm_brinkman_solve.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from . import _lib
from ._lib import PRECISION_FP64, check, f64, ptr


@dataclass
class FluidConfig:
    """common::FluidConfig (include/app/config/Config.hpp:36-51), the fields the solver reads."""
    fluid_material_id: int = 0
    solid_penalty: float = 1.0e8
    forcing_magnitude: float = 1.0
    voxel_size: float = 1.0
    max_iterations: int = 600
    require_single_solid_material_id: bool = True
    tolerance: float = 1.0e-8
    relaxation: float = 0.9
    precision: int = PRECISION_FP64


@dataclass
class Geometry:
    """common::Geometry (include/core/Geometry.hpp:9-18): uint32 material ids, x fastest."""
    nx: int
    ny: int
    nz: int
    material_ids: np.ndarray

    def voxel_count(self):
        return self.nx * self.ny * self.nz


@dataclass
class MaterialProperties:
    values: Dict[str, float] = field(default_factory=dict)
    tags: Dict[str, str] = field(default_factory=dict)


class MaterialDatabase:
    """common::MaterialDatabase (include/materials/MaterialDatabase.hpp), the queries the solver uses."""

    def __init__(self):
        self._m: Dict[int, MaterialProperties] = {}

    def insert(self, material_id: int, props: MaterialProperties):
        self._m[int(material_id)] = props

    def has(self, material_id):
        return int(material_id) in self._m

    def at(self, material_id):
        if int(material_id) not in self._m:
            raise KeyError(f"Unknown material id: {material_id}")
        return self._m[int(material_id)]

    def has_tag(self, material_id, key):
        return int(material_id) in self._m and key in self._m[int(material_id)].tags

    def tag(self, material_id, key):
        return self.at(material_id).tags[key]


@dataclass
class SolveDiagnostics:
    iterations: int = 0
    residual: float = 0.0
    converged: bool = False


@dataclass
class SolveResult:
    velocity: Optional[np.ndarray]  # (3, N) SoA or None
    superficial_velocity: np.ndarray
    porosity: float
    diagnostics: SolveDiagnostics
    viscosity: float


@dataclass
class PermeabilityResult:
    tensor: np.ndarray
    load_diagnostics: List[SolveDiagnostics]
    porosity: float
    viscosity: float


def resolve_phase_mask(config: FluidConfig, geometry: Geometry, db: MaterialDatabase):
    """resolvePhaseMask (:67-131): fluid/solid id sets from phase tags, else from fluid_material_id."""
    ids = sorted(set(int(v) for v in np.unique(geometry.material_ids)))
    uses_tags = any(db.has_tag(i, "phase") for i in ids)
    fluid, solid = set(), set()
    if uses_tags:
        for i in ids:
            if not db.has_tag(i, "phase"):
                raise RuntimeError(f"Material id {i} is missing required tag phase=fluid|solid")
            ph = db.tag(i, "phase")
            if ph == "fluid":
                fluid.add(i)
            elif ph == "solid":
                solid.add(i)
            else:
                raise RuntimeError(f"Unsupported phase tag for material id {i}: {ph}")
    else:
        fluid.add(int(config.fluid_material_id))
        solid.update(i for i in ids if i != config.fluid_material_id)
    if not fluid:
        raise RuntimeError("No fluid material ids found. Set phase=fluid or fluid_material_id.")
    if not solid:
        raise RuntimeError("No solid material ids found for Brinkman solve.")
    if config.require_single_solid_material_id and len(solid) != 1:
        raise RuntimeError(f"Expected exactly one solid material id, found {len(solid)}")
    return fluid, solid


def build_coefficient_fields(config: FluidConfig, geometry: Geometry, db: MaterialDatabase, fluid_ids):
    """buildCoefficientFields (:152-191): alpha = penalty (0 in fluid), viscosity = 1 (fluid's in fluid)."""
    ids = np.asarray(geometry.material_ids, dtype=np.uint32).reshape(-1)
    alpha = np.full(ids.size, float(config.solid_penalty))
    visc = np.ones(ids.size)
    count, vsum = 0, 0.0
    for fid in sorted(fluid_ids):
        props = db.at(fid).values
        if "viscosity" not in props:
            raise RuntimeError(f"Missing viscosity for fluid material id: {fid}")
        sel = ids == fid
        alpha[sel] = 0.0
        visc[sel] = props["viscosity"]
        c = int(np.count_nonzero(sel))
        count += c
        vsum += props["viscosity"] * c
    porosity = count / ids.size if ids.size else 0.0
    mean_visc = vsum / count if count else 1.0
    return alpha, visc, porosity, mean_visc


def solve_brinkman(config: FluidConfig, geometry: Geometry, db: MaterialDatabase, force, want_velocity=False):
    """solveBrinkman (:282-507) on the GPU."""
    fluid, _ = resolve_phase_mask(config, geometry, db)
    alpha, visc, porosity, mean_visc = build_coefficient_fields(config, geometry, db, fluid)
    N = geometry.voxel_count()
    U = np.zeros(3)
    u = np.zeros(3 * N) if want_velocity else None
    info = _lib.BrinkmanInfo()
    check(_lib.load().fftperm_brinkman_solve(geometry.nx, geometry.ny, geometry.nz, alpha, visc,
                                             float(config.relaxation), float(config.tolerance),
                                             int(config.max_iterations), f64(force, 3), int(config.precision), U,
                                             ptr(u), C.byref(info)))
    diag = SolveDiagnostics(int(info.iterations), float(info.residual), bool(info.converged))
    return SolveResult(None if u is None else u.reshape(3, N), U, porosity, diag, mean_visc)


def compute_permeability_tensor(config: FluidConfig, geometry: Geometry, db: MaterialDatabase,
                                on_load_case_solved: Optional[Callable] = None) -> PermeabilityResult:
    """computePermeabilityTensor (:509-553): 3 load cases, K_ij = mu U_i / F_j h^2."""
    K = np.zeros(9)
    scale = config.voxel_size * config.voxel_size
    diags = []
    porosity, visc = 0.0, 1.0
    for j in range(3):
        force = np.zeros(3)
        force[j] = config.forcing_magnitude
        r = solve_brinkman(config, geometry, db, force, want_velocity=on_load_case_solved is not None)
        porosity, visc = r.porosity, r.viscosity
        diags.append(r.diagnostics)
        for i in range(3):
            K[i * 3 + j] = r.viscosity * r.superficial_velocity[i] / force[j] * scale
        if on_load_case_solved is not None:
            on_load_case_solved(j, force, r)
    return PermeabilityResult(K.reshape(3, 3), diags, porosity, visc)


def single_fluid_database(fluid_id: int, viscosity: float) -> MaterialDatabase:
    """The database PermeabilityRunner builds without a material file (src/app/PermeabilityRunner.cpp:261-281)."""
    db = MaterialDatabase()
    db.insert(fluid_id, MaterialProperties(values={"viscosity": float(viscosity)}))
    return db
