"""Command-line runner: the counterpart of the reference's `fft_method` executable.

    python -m fft_method_test_amd.runner --input sample.fgeo --gradient all --tol 1e-10 --maxit 10000 \\
        --output out/result.txt [--write-velocity-fields] [--precision mixed]

Configuration follows SimulationConfig (src/app/SimulationConfig.hpp, SimulationConfig.cpp):
the same defaults, the same `--config` file sections ([global] keys, [analysis], [fluid],
[bignonnet] / [fluid.bignonnet], [brinkman] / [fluid.brinkman]), the same CLI flags applied after
the file, and the same validation messages (ValueError where the reference throws
std::invalid_argument, RuntimeError for std::runtime_error).  The run follows
PermeabilityRunner::run (src/app/PermeabilityRunner.cpp:472-823) and run_brinkman (:246-469):
same routes, same output files (`<stem>_permeability.txt`, `<stem>_velocity.txt`,
`<stem>_axis<k>.fvel` / `<stem>.fvel`) with the same text layout.

Differences, all on purpose:
* every solve runs on the GPU through libfftperm.so; `--compute-backend cpu` is refused (this
  library has no CPU backend), and the Brinkman route runs on the GPU as well;
* `--precision fp64|mixed` (config key `precision` in the Bignonnet sections) selects the
  context's precision mode; fp64 is the reference's DOUBLE build;
* `--threads`, `--fft-threads` and `--parallel-load-cases` are accepted and validated but have no
  effect (the load cases run one after another on the one GPU), and `--progress-interval` is
  accepted but the device-resident MINRES reports no per-iteration progress.
"""
from __future__ import annotations

import datetime as _dt
import math
import os
import sys
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

USAGE = ("Usage: python -m fft_method_test_amd.runner [--config <file>] [--input <geometry>] "
         "[--material-file <file>] [--output <file>] [--log <file>] [--analysis fluid] "
         "[--fluid-mode permeability|pressure_gradient] [--fluid-solver bignonnet|brinkman] "
         "[--compute-backend gpu] [--pressure-gradient <gx,gy,gz>] [--write-velocity-fields] "
         "[--support interface|fullsolid] [--mu <value>] [--fluid-id <value>] [--voxel-size <value>] "
         "[--tol <value>] [--maxit <value>] [--threads <value>] [--fft-threads <value>] "
         "[--p-radius <value>] [--gradient x|y|z|all] [--parallel-load-cases] "
         "[--solid-penalty <value>] [--relaxation <value>] [--forcing-magnitude <value>] "
         "[--progress-interval <value>] [--precision fp64|mixed]")


@dataclass
class SimulationConfig:
    """src/app/SimulationConfig.hpp:21-68, field for field (plus `precision`)."""
    geometry_file: str = "data/sample_geometry.geom"
    material_file: str = ""
    output_file: str = "output.txt"
    log_file: str = ""
    analysis: str = "fluid"
    fluid_mode: str = "permeability"  # permeability | pressure_gradient
    compute_backend: str = "gpu"
    pressure_gradient: Tuple[float, float, float] = (1.0, 0.0, 0.0)
    support_mode: str = "interface"  # interface | fullsolid
    gradient_mode: str = "all"  # x | y | z | all
    mu: float = 1.0
    mu_explicit: bool = False
    fluid_id: int = 1
    fluid_id_explicit: bool = False
    voxel_size: float = 1.0
    tolerance: float = 1e-10
    max_iterations: int = 400
    p_radius: int = 2
    progress_interval: int = 0
    threads: int = 0
    fft_threads: int = 0
    parallel_load_cases: bool = False
    write_velocity_fields: bool = False
    solver: str = "bignonnet"
    brinkman_solid_penalty: float = 1.0e8
    brinkman_relaxation: float = 0.9
    brinkman_tolerance: float = 1.0e-8
    brinkman_forcing_magnitude: float = 1.0
    brinkman_write_velocity_fields: bool = False
    brinkman_require_single_solid_id: bool = True
    precision: str = "fp64"  # fp64 | mixed


# ---------------------------------------------------------------- parsing helpers

def _parse_bool(v: str) -> bool:
    return v.strip().lower() in ("1", "true", "yes", "on")


def _parse_analysis(v: str) -> str:
    lo = v.strip().lower()
    if lo in ("fluid", "elastic"):
        return lo
    raise ValueError("Invalid analysis type: " + v)


def _parse_fluid_mode(v: str) -> str:
    lo = v.strip().lower()
    if lo in ("permeability", "pressure_gradient"):
        return lo
    raise ValueError("Invalid fluid analysis mode: " + v)


def _parse_backend(v: str) -> str:
    lo = v.strip().lower()
    if lo == "gpu":
        return lo
    if lo == "cpu":
        raise ValueError("compute_backend=cpu is not available: this library runs every solve on the GPU "
                         "(use gpu)")
    raise ValueError("Invalid compute backend: " + v + ". Use cpu|gpu")


def _parse_gradient(v: str, source: str) -> str:
    lo = v.strip().lower()
    if lo in ("x", "y", "z", "all"):
        return lo
    raise ValueError("Invalid " + source + " value: " + v + ". Use x|y|z|all")


def _parse_vector3(v: str, source: str) -> Tuple[float, float, float]:
    toks = v.replace(",", " ").split()
    try:
        vals = [float(t) for t in toks[:3]]
    except ValueError:
        vals = []
    if len(vals) < 3:
        raise ValueError("Invalid " + source + " value: " + v + ". Expected three numbers")
    if len(toks) > 3:
        raise ValueError("Invalid " + source + " value: " + v + ". Expected exactly three numbers")
    return (vals[0], vals[1], vals[2])


def _parse_support(v: str, flag: bool) -> str:
    lo = v.lower()
    if lo in ("interface", "fullsolid"):
        return lo
    if flag:
        raise ValueError("Invalid --support value. Use interface|fullsolid")
    raise ValueError("Invalid support value: " + v)


def _parse_precision(v: str) -> str:
    lo = v.strip().lower()
    if lo in ("fp64", "mixed"):
        return lo
    raise ValueError("Invalid precision: " + v + ". Use fp64|mixed")


def parse_config_file(path: str) -> Dict[str, Dict[str, str]]:
    """parseConfigFile (src/app/SimulationConfig.cpp:99-147): `key = value` lines, `[section]`
    headers (lower-cased), `#` / `;` comments; keys before any header land in section ''."""
    try:
        f = open(path)
    except OSError:
        raise RuntimeError("Cannot open config file: " + path)
    sections: Dict[str, Dict[str, str]] = {}
    section = ""
    with f:
        for line_no, line in enumerate(f, 1):
            cut = [i for i in (line.find("#"), line.find(";")) if i >= 0]
            if cut:
                line = line[: min(cut)]
            line = line.strip()
            if not line:
                continue
            if line[0] == "[":
                close = line.find("]")
                if close < 0:
                    raise RuntimeError(f"Config line {line_no}: missing closing ']'")
                section = line[1:close].strip().lower()
                continue
            eq = line.find("=")
            if eq < 0:
                raise RuntimeError(f"Config line {line_no}: expected key = value")
            key = line[:eq].strip().lower()
            val = line[eq + 1:].strip()
            if not key or not val:
                raise RuntimeError(f"Config line {line_no}: empty key or value")
            sections.setdefault(section, {})[key] = val
    return sections


def _apply_sections(cfg: SimulationConfig, m: Dict[str, Dict[str, str]]) -> None:
    """applyGlobal, applyBignonnet, applyBrinkman, applyAnalysis, applyFluidCommon,
    applyFluidBignonnet, applyFluidBrinkman in the reference's order (SimulationConfig.cpp:150-317)."""
    def get(sec, key):
        return m.get(sec, {}).get(key)

    g = lambda k: get("", k)  # noqa: E731
    if g("geometry_file") is not None: cfg.geometry_file = g("geometry_file")
    if g("material_file") is not None: cfg.material_file = g("material_file")
    if g("output_file") is not None: cfg.output_file = g("output_file")
    if g("log_file") is not None: cfg.log_file = g("log_file")
    if g("solver") is not None: cfg.solver = g("solver")
    if g("compute_backend") is not None: cfg.compute_backend = _parse_backend(g("compute_backend"))
    if g("threads") is not None: cfg.threads = int(g("threads"))
    if g("max_iterations") is not None: cfg.max_iterations = int(g("max_iterations"))
    if g("progress_interval") is not None: cfg.progress_interval = int(g("progress_interval"))
    if g("voxel_size") is not None: cfg.voxel_size = float(g("voxel_size"))

    def bignonnet(sec):
        if get(sec, "mu") is not None:
            cfg.mu, cfg.mu_explicit = float(get(sec, "mu")), True
        if get(sec, "fluid_id") is not None:
            cfg.fluid_id, cfg.fluid_id_explicit = int(get(sec, "fluid_id")), True
        if get(sec, "tolerance") is not None: cfg.tolerance = float(get(sec, "tolerance"))
        if get(sec, "p_radius") is not None: cfg.p_radius = int(get(sec, "p_radius"))
        if get(sec, "fft_threads") is not None: cfg.fft_threads = int(get(sec, "fft_threads"))
        if get(sec, "parallel_load_cases") is not None: cfg.parallel_load_cases = _parse_bool(get(sec, "parallel_load_cases"))
        if get(sec, "support") is not None: cfg.support_mode = _parse_support(get(sec, "support"), False)
        if get(sec, "gradient") is not None: cfg.gradient_mode = _parse_gradient(get(sec, "gradient"), sec + ".gradient")
        if get(sec, "precision") is not None: cfg.precision = _parse_precision(get(sec, "precision"))

    def brinkman(sec):
        if get(sec, "solid_penalty") is not None: cfg.brinkman_solid_penalty = float(get(sec, "solid_penalty"))
        if get(sec, "relaxation") is not None: cfg.brinkman_relaxation = float(get(sec, "relaxation"))
        if get(sec, "tolerance") is not None: cfg.brinkman_tolerance = float(get(sec, "tolerance"))
        if get(sec, "forcing_magnitude") is not None: cfg.brinkman_forcing_magnitude = float(get(sec, "forcing_magnitude"))
        if get(sec, "write_velocity_fields") is not None:
            cfg.brinkman_write_velocity_fields = _parse_bool(get(sec, "write_velocity_fields"))
            cfg.write_velocity_fields = cfg.brinkman_write_velocity_fields
        if get(sec, "require_single_solid_id") is not None:
            cfg.brinkman_require_single_solid_id = _parse_bool(get(sec, "require_single_solid_id"))

    bignonnet("bignonnet")
    brinkman("brinkman")
    if get("analysis", "type") is not None:
        cfg.analysis = _parse_analysis(get("analysis", "type"))
    f = lambda k: get("fluid", k)  # noqa: E731
    if f("mode") is not None: cfg.fluid_mode = _parse_fluid_mode(f("mode"))
    if f("solver") is not None: cfg.solver = f("solver")
    if f("compute_backend") is not None: cfg.compute_backend = _parse_backend(f("compute_backend"))
    if f("pressure_gradient") is not None: cfg.pressure_gradient = _parse_vector3(f("pressure_gradient"), "fluid.pressure_gradient")
    if f("mu") is not None: cfg.mu, cfg.mu_explicit = float(f("mu")), True
    if f("fluid_id") is not None: cfg.fluid_id, cfg.fluid_id_explicit = int(f("fluid_id")), True
    if f("voxel_size") is not None: cfg.voxel_size = float(f("voxel_size"))
    if f("tolerance") is not None:
        cfg.tolerance = float(f("tolerance"))
        cfg.brinkman_tolerance = cfg.tolerance
    if f("max_iterations") is not None: cfg.max_iterations = int(f("max_iterations"))
    if f("progress_interval") is not None: cfg.progress_interval = int(f("progress_interval"))
    if f("threads") is not None: cfg.threads = int(f("threads"))
    if f("fft_threads") is not None: cfg.fft_threads = int(f("fft_threads"))
    if f("parallel_load_cases") is not None: cfg.parallel_load_cases = _parse_bool(f("parallel_load_cases"))
    if f("write_velocity_fields") is not None:
        cfg.write_velocity_fields = _parse_bool(f("write_velocity_fields"))
        cfg.brinkman_write_velocity_fields = cfg.write_velocity_fields
    bignonnet("fluid.bignonnet")
    brinkman("fluid.brinkman")


def from_cli(argv: List[str]) -> SimulationConfig:
    """SimulationConfig::from_cli (src/app/SimulationConfig.cpp:303-440): the --config file first,
    then every CLI flag in order, then validation.  argv excludes the program name."""
    cfg = SimulationConfig()

    def value(i, flag):
        if i + 1 >= len(argv):
            raise ValueError("Missing value for " + flag)
        return argv[i + 1]

    i = 0
    while i < len(argv):
        if argv[i] == "--config":
            _apply_sections(cfg, parse_config_file(value(i, "--config")))
            i += 2
            continue
        i += 1

    setters = {
        "--input": lambda v: setattr(cfg, "geometry_file", v),
        "--material-file": lambda v: setattr(cfg, "material_file", v),
        "--output": lambda v: setattr(cfg, "output_file", v),
        "--log": lambda v: setattr(cfg, "log_file", v),
        "--analysis": lambda v: setattr(cfg, "analysis", _parse_analysis(v)),
        "--fluid-mode": lambda v: setattr(cfg, "fluid_mode", _parse_fluid_mode(v)),
        "--fluid-solver": lambda v: setattr(cfg, "solver", v),
        "--solver": lambda v: setattr(cfg, "solver", v),
        "--compute-backend": lambda v: setattr(cfg, "compute_backend", _parse_backend(v)),
        "--pressure-gradient": lambda v: setattr(cfg, "pressure_gradient", _parse_vector3(v, "--pressure-gradient")),
        "--support": lambda v: setattr(cfg, "support_mode", _parse_support(v, True)),
        "--voxel-size": lambda v: setattr(cfg, "voxel_size", float(v)),
        "--maxit": lambda v: setattr(cfg, "max_iterations", int(v)),
        "--threads": lambda v: setattr(cfg, "threads", int(v)),
        "--fft-threads": lambda v: setattr(cfg, "fft_threads", int(v)),
        "--p-radius": lambda v: setattr(cfg, "p_radius", int(v)),
        "--gradient": lambda v: setattr(cfg, "gradient_mode", _parse_gradient(v, "--gradient")),
        "--solid-penalty": lambda v: setattr(cfg, "brinkman_solid_penalty", float(v)),
        "--relaxation": lambda v: setattr(cfg, "brinkman_relaxation", float(v)),
        "--forcing-magnitude": lambda v: setattr(cfg, "brinkman_forcing_magnitude", float(v)),
        "--progress-interval": lambda v: setattr(cfg, "progress_interval", int(v)),
        "--precision": lambda v: setattr(cfg, "precision", _parse_precision(v)),
    }
    i = 0
    while i < len(argv):
        a = argv[i]
        if a == "--config":
            i += 2
            continue
        if a == "--write-velocity-fields":
            cfg.write_velocity_fields = cfg.brinkman_write_velocity_fields = True
        elif a == "--parallel-load-cases":
            cfg.parallel_load_cases = True
        elif a == "--mu":
            cfg.mu, cfg.mu_explicit = float(value(i, a)), True
            i += 1
        elif a == "--fluid-id":
            cfg.fluid_id, cfg.fluid_id_explicit = int(value(i, a)), True
            i += 1
        elif a == "--tol":
            cfg.tolerance = float(value(i, a))
            cfg.brinkman_tolerance = cfg.tolerance  # applies to both solvers
            i += 1
        elif a in setters:
            setters[a](value(i, a))
            i += 1
        else:
            raise ValueError("Unknown CLI option: " + a)
        i += 1

    if cfg.fft_threads == 0:
        cfg.fft_threads = cfg.threads
    if cfg.fft_threads == 0:
        cfg.fft_threads = 1
    if cfg.analysis != "fluid":
        raise ValueError("analysis must be 'fluid' (other analyses are not implemented yet)")
    if cfg.solver not in ("bignonnet", "brinkman"):
        raise ValueError("solver must be 'bignonnet' or 'brinkman'")
    if cfg.fluid_mode == "pressure_gradient" and sum(g * g for g in cfg.pressure_gradient) <= 0.0:
        raise ValueError("pressure_gradient must be non-zero when fluid mode is pressure_gradient")
    if cfg.mu_explicit and cfg.mu <= 0.0:
        raise ValueError("mu must be positive")
    if cfg.voxel_size <= 0.0:
        raise ValueError("voxel_size must be > 0")
    if cfg.tolerance <= 0.0:
        raise ValueError("tolerance must be positive")
    if cfg.max_iterations == 0:
        raise ValueError("max_iterations must be >= 1")
    if cfg.fft_threads <= 0:
        raise ValueError("fft_threads must be >= 1")
    if cfg.threads < 0:
        raise ValueError("threads must be >= 0")
    if cfg.p_radius < 0:
        raise ValueError("p_radius must be >= 0")
    return cfg


# ---------------------------------------------------------------- output helpers

def fmt_default(x: float) -> str:
    """A double written by an ostream in its default format (6 significant digits, %g)."""
    return f"{float(x):g}"


def fmt_sci(x: float) -> str:
    """std::scientific with setprecision(6)."""
    return f"{float(x):.6e}"


def _stem(base_output: str):
    p = base_output if base_output else "output.txt"
    parent = os.path.dirname(p)
    name = os.path.basename(p)
    stem = os.path.splitext(name)[0]  # std::filesystem::path::stem: a leading dot is not an extension
    return parent, (stem if stem else "output")


def text_output_path(base_output: str, suffix: str) -> str:
    """text_output_path (src/app/PermeabilityRunner.cpp:99-104): <parent>/<stem>_<suffix>.txt."""
    parent, stem = _stem(base_output)
    return os.path.join(parent, f"{stem}_{suffix}.txt") if parent else f"{stem}_{suffix}.txt"


def velocity_field_output_path(base_output: str, suffix: str) -> str:
    """velocity_field_output_path (:86-92): <stem>.fvel, or <stem>_<suffix>.fvel."""
    parent, stem = _stem(base_output)
    name = f"{stem}.fvel" if not suffix else f"{stem}_{suffix}.fvel"
    return os.path.join(parent, name) if parent else name


def velocity_field_output_path_axis(base_output: str, axis: int) -> str:
    return velocity_field_output_path(base_output, f"axis{axis + 1}")


def tensor_string(K, indent: str = "") -> str:
    """PermeabilityTensor::to_string (src/post/PermeabilityTensor.cpp:20-28)."""
    k = np.asarray(K, dtype=np.float64).reshape(9)
    return ("\n" + indent).join("  ".join(fmt_sci(k[3 * r + c]) for c in range(3)) for r in range(3))


def read_material_database(path: str):
    """MaterialDatabase::read (src/materials/MaterialDatabase.cpp:44-118): one material per line,
    `<id> key=value ... [phase=fluid|solid]`, `#` / `;` comments."""
    from .brinkman import MaterialDatabase, MaterialProperties

    try:
        f = open(path)
    except OSError:
        raise RuntimeError("Cannot open material file: " + path)
    db = MaterialDatabase()
    count = 0
    with f:
        for line_no, line in enumerate(f, 1):
            cut = [i for i in (line.find("#"), line.find(";")) if i >= 0]
            if cut:
                line = line[: min(cut)]
            toks = line.split()
            if not toks:
                continue
            try:
                mid = int(toks[0])
                if mid < 0 or mid > 0xFFFFFFFF:
                    raise ValueError
            except ValueError:
                raise RuntimeError(f"Invalid material line {line_no}: expected leading integer ID")
            props = MaterialProperties()
            for tok in toks[1:]:
                eq = tok.find("=")
                if eq <= 0 or eq == len(tok) - 1:
                    raise RuntimeError(f"Invalid material property token at line {line_no}: {tok}")
                key, val = tok[:eq], tok[eq + 1:]
                if not all(ch.isalnum() or ch in "_." for ch in key):
                    raise RuntimeError(f"Invalid material property name at line {line_no}: {key}")
                lk = key.lower()
                if lk == "phase":
                    lv = val.lower()
                    if lv not in ("fluid", "solid"):
                        raise RuntimeError(f"Invalid phase value at line {line_no}: expected fluid|solid")
                    props.tags[lk] = lv
                    continue
                try:
                    props.values[lk] = float(val)
                except ValueError:
                    raise RuntimeError(f"Invalid numeric value for property '{key}' at line {line_no}")
            if not props.values and not props.tags:
                raise RuntimeError(f"Material line {line_no} has no properties")
            db.insert(mid, props)
            count += 1
    if count == 0:
        raise RuntimeError("Material file has no material definitions: " + path)
    return db


class Logger:
    """`[YYYY-MM-DD HH:MM:SS] [LEVEL] message` lines on stderr and, once opened, in a log file."""

    def __init__(self):
        self._f = None

    def open(self, path):
        d = os.path.dirname(path)
        if d:
            os.makedirs(d, exist_ok=True)
        if self._f:
            self._f.close()
        self._f = open(path, "a")

    def _emit(self, level, msg):
        line = f"[{_dt.datetime.now():%Y-%m-%d %H:%M:%S}] [{level}] {msg}"
        print(line, file=sys.stderr, flush=True)
        if self._f:
            self._f.write(line + "\n")
            self._f.flush()

    def info(self, msg):
        self._emit("INFO", msg)

    def warn(self, msg):
        self._emit("WARN", msg)

    def close(self):
        if self._f:
            self._f.close()
            self._f = None


_INDENT = " " * 29  # width of the "[YYYY-MM-DD HH:MM:SS] [INFO] " prefix


def _ensure_parent(path):
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)


# ---------------------------------------------------------------- runner

class PermeabilityRunner:
    """PermeabilityRunner (src/app/PermeabilityRunner.cpp:244-823) over the GPU library."""

    def __init__(self, config: SimulationConfig, log: Optional[Logger] = None):
        self.config = config
        self.log = log or Logger()
        self.result = None  # K (3x3) or the macroscopic velocity, for callers

    def run(self) -> int:
        cfg = self.config
        log = self.log
        log.info("Entering PermeabilityRunner")
        if cfg.analysis != "fluid":
            raise RuntimeError("Only fluid analysis is implemented in this version")
        if cfg.solver == "brinkman":
            log.info("Execution route: fluid -> brinkman -> " + cfg.fluid_mode)
            return self._run_brinkman()
        return self._run_bignonnet()

    # -- Bignonnet ------------------------------------------------------------------------------
    def _run_bignonnet(self) -> int:
        from . import _lib, permeability as P, voxelio

        cfg, log = self.config, self.log
        t_start = time.perf_counter()
        log.info("Execution route: fluid -> bignonnet -> " + cfg.fluid_mode)
        log.info("Execution backend route: gpu")
        log.info("Loading geometry: " + cfg.geometry_file)
        mask = voxelio.load(cfg.geometry_file, cfg.fluid_id)
        if cfg.material_file:
            db = read_material_database(cfg.material_file)
            fluid_mat, fluid_visc = 0, 0.0
            for mid in range(1, 65536):
                if not db.has(mid):
                    continue
                v = db.at(mid).values.get("viscosity")
                if v is not None and v > 0.0:
                    fluid_mat, fluid_visc = mid, v
                    break
            if fluid_mat == 0:
                raise RuntimeError(f"material_file '{cfg.material_file}' has no material with viscosity > 0 "
                                   "(fluid phase)")
            if not cfg.mu_explicit:
                cfg.mu = fluid_visc
                log.info(f"mu set from material {fluid_mat} (viscosity = {fmt_default(fluid_visc)})")
            if not cfg.fluid_id_explicit:
                cfg.fluid_id = fluid_mat
                mask = voxelio.load(cfg.geometry_file, cfg.fluid_id)
                log.info(f"fluid_id set from material file: {cfg.fluid_id}")
        if cfg.mu <= 0.0:
            raise RuntimeError("mu must be positive; provide --mu or set viscosity in material_file")
        nz, ny, nx = mask.shape
        h = cfg.voxel_size
        grid = P.Grid3D(nx, ny, nz, h, h, h)
        _lib.check(_lib.load().fftperm_check_edges(nx, ny, nz))
        medium = P.BinaryMedium(grid, mask)
        prec = _lib.PRECISION_MIXED if cfg.precision == "mixed" else _lib.PRECISION_FP64
        green = P.GreenOperator(grid, cfg.mu, cfg.p_radius, prec)
        try:
            mode = P.SUPPORT_INTERFACE if cfg.support_mode == "interface" else P.SUPPORT_FULLSOLID
            support = P.ForceSupport(medium, mode, green)
            if support.num_active_voxels() == 0:
                raise RuntimeError("Support set is empty; cannot run variational solve")
            log.info("Solver: bignonnet")
            log.info("Fluid mode: " + cfg.fluid_mode)
            log.info("Compute backend: gpu (" + _lib.device_info() + ")")
            log.info("Input geometry file: " + cfg.geometry_file)
            log.info("Input material file: " + (cfg.material_file or "(none)"))
            log.info("Output file: " + cfg.output_file)
            log.info(f"Fluid properties: fluid_id={cfg.fluid_id} mu={fmt_default(cfg.mu)}")
            log.info(f"Stop criteria (MINRES): tolerance={fmt_default(cfg.tolerance)} "
                     f"max_iterations={cfg.max_iterations}")
            log.info(f"Bignonnet parameters: support={cfg.support_mode} gradient={cfg.gradient_mode} "
                     f"p_radius={cfg.p_radius} precision={cfg.precision}")
            log.info(f"Grid: {nx} x {ny} x {nz}")
            log.info("Voxel size: " + fmt_default(h))
            log.info("Porosity: " + fmt_default(medium.porosity()))
            log.info(f"Support voxels (NB): {support.num_active_voxels()}")
            if cfg.fluid_mode == "pressure_gradient":
                self._bignonnet_pressure_gradient(P, green, support, medium, grid, t_start)
            else:
                self._bignonnet_permeability(P, green, support, medium, grid, t_start)
        finally:
            green.close()
        return 0

    def _solve(self, P, green, support, col, grad):
        cfg = self.config
        t0 = time.perf_counter()
        avg, info, u = P.solve_one_direction(green, support, grad, cfg.tolerance, cfg.max_iterations,
                                             want_velocity=cfg.write_velocity_fields)
        dt = time.perf_counter() - t0
        self.log.info(f"Completed load case axis={col} iterations={int(info.iterations)} "
                      f"relres={float(info.final_rel_residual):.2e} "
                      f"converged={'yes' if info.converged else 'no'} elapsed={dt:.2f} s")
        return avg, info, u

    def _bignonnet_pressure_gradient(self, P, green, support, medium, grid, t_start):
        from . import voxelio

        cfg, log = self.config, self.log
        grad = np.asarray(cfg.pressure_gradient, dtype=np.float64)
        log.info("Starting pressure-gradient solve grad=(" + ", ".join(f"{g:.4e}" for g in grad) + ")")
        avg, info, u = self._solve(P, green, support, 0, grad)
        if cfg.write_velocity_fields and u is not None:
            path = velocity_field_output_path(cfg.output_file, "")
            _ensure_parent(path)
            voxelio.write_fvel(path, grid.shape, medium.porosity(), cfg.mu, grad, u)
            log.info("Velocity field written to " + path)
        if not info.converged:
            log.warn("Solve did not converge for pressure-gradient response, final rel residual="
                     f"{fmt_default(info.final_rel_residual)}")
        log.info("Macroscopic velocity U=(" + ", ".join(fmt_sci(a) for a in avg) + ")")
        self.result = np.asarray(avg)
        if cfg.output_file:
            path = text_output_path(cfg.output_file, "velocity")
            _ensure_parent(path)
            with open(path, "w") as f:
                f.write("# FFT_Method fluid pressure-gradient response\n")
                f.write("# solver = bignonnet\n")
                f.write(f"# geometry_file = {cfg.geometry_file}\n")
                f.write(f"# grid = {grid.nx()} x {grid.ny()} x {grid.nz()}\n")
                f.write(f"# voxel_size = {fmt_default(cfg.voxel_size)}\n")
                f.write(f"# porosity = {fmt_default(medium.porosity())}\n")
                f.write(f"# mu = {fmt_default(cfg.mu)}\n")
                f.write("# pressure_gradient = " + " ".join(fmt_default(g) for g in grad) + "\n")
                f.write("# macroscopic_velocity = " + " ".join(fmt_default(a) for a in avg) + "\n")
                f.write(f"# iterations = {int(info.iterations)}\n")
                f.write(f"# final_rel_residual = {fmt_default(info.final_rel_residual)}\n")
                f.write(f"# converged = {'true' if info.converged else 'false'}\n")
            log.info("Results written to " + path)
        log.info(f"Bignonnet total elapsed={time.perf_counter() - t_start:.2f} s")

    def _bignonnet_permeability(self, P, green, support, medium, grid, t_start):
        from . import voxelio

        cfg, log = self.config, self.log
        axes = {"x": [0], "y": [1], "z": [2], "all": [0, 1, 2]}[cfg.gradient_mode]
        K = P.PermeabilityTensor()
        for col in axes:
            grad = np.zeros(3)
            grad[col] = 1.0
            log.info(f"Starting load case axis={col} grad=(" + ", ".join(f"{g:.2e}" for g in grad) + ")")
            avg, info, u = self._solve(P, green, support, col, grad)
            if not info.converged:
                log.warn(f"Solve did not converge for load case {col}, final rel residual="
                         f"{fmt_default(info.final_rel_residual)}")
            K.set_column(col, -cfg.mu * np.asarray(avg))
            if cfg.write_velocity_fields and u is not None:
                path = velocity_field_output_path_axis(cfg.output_file, col)
                _ensure_parent(path)
                voxelio.write_fvel(path, grid.shape, medium.porosity(), cfg.mu, grad, u)
                log.info("Velocity field written to " + path)
        self.result = K.data().reshape(3, 3)
        log.info("Permeability tensor K (Darcy sign, K > 0):\n" + _INDENT + K.to_string(_INDENT))
        if cfg.output_file:
            path = text_output_path(cfg.output_file, "permeability")
            _ensure_parent(path)
            with open(path, "w") as f:
                f.write("# FFT_Method permeability result\n")
                f.write(f"# geometry_file = {cfg.geometry_file}\n")
                f.write(f"# grid = {grid.nx()} x {grid.ny()} x {grid.nz()}\n")
                f.write(f"# voxel_size = {fmt_default(cfg.voxel_size)}\n")
                f.write(f"# porosity = {fmt_default(medium.porosity())}\n")
                f.write(f"# mu = {fmt_default(cfg.mu)}\n")
                f.write("#\n")
                f.write("# Permeability tensor K (Darcy sign, K > 0):\n")
                f.write(K.to_string() + "\n")
            log.info("Results written to " + path)
        log.info(f"Bignonnet total elapsed={time.perf_counter() - t_start:.2f} s")

    # -- Brinkman -------------------------------------------------------------------------------
    def _run_brinkman(self) -> int:
        from . import _lib, brinkman as B, voxelio

        cfg, log = self.config, self.log
        log.info("Loading geometry: " + cfg.geometry_file)
        path = cfg.geometry_file
        ids = voxelio.read_fgeo(path) if path.endswith(".fgeo") else voxelio.read_geom_text(path)
        nz, ny, nx = ids.shape
        geo = B.Geometry(nx, ny, nz, ids.reshape(-1))
        log.info(f"Grid: {nx} x {ny} x {nz}")
        if cfg.material_file:
            log.info("Loading material database: " + cfg.material_file)
            db = read_material_database(cfg.material_file)
        else:
            if cfg.mu <= 0.0:
                raise RuntimeError("Brinkman solver requires mu > 0; provide --mu or a material_file with viscosity")
            db = B.single_fluid_database(cfg.fluid_id, cfg.mu)
            log.info(f"Auto-built material DB: fluid_id={cfg.fluid_id} viscosity={cfg.mu:.6f}")
        fc = B.FluidConfig(fluid_material_id=cfg.fluid_id, solid_penalty=cfg.brinkman_solid_penalty,
                           forcing_magnitude=cfg.brinkman_forcing_magnitude, voxel_size=cfg.voxel_size,
                           max_iterations=int(cfg.max_iterations),
                           require_single_solid_material_id=cfg.brinkman_require_single_solid_id,
                           tolerance=cfg.brinkman_tolerance, relaxation=cfg.brinkman_relaxation,
                           precision=_lib.PRECISION_MIXED if cfg.precision == "mixed" else _lib.PRECISION_FP64)
        log.info("Solver: brinkman")
        log.info("Fluid mode: " + cfg.fluid_mode)
        log.info(f"Brinkman parameters: solid_penalty={fmt_default(cfg.brinkman_solid_penalty)} "
                 f"relaxation={fmt_default(cfg.brinkman_relaxation)} tolerance={fmt_default(cfg.brinkman_tolerance)} "
                 f"max_iterations={cfg.max_iterations}")
        if cfg.fluid_mode == "pressure_gradient":
            gp = np.asarray(cfg.pressure_gradient, dtype=np.float64)
            force = -gp
            log.info("Pressure-gradient response: grad_p=(" + ", ".join(f"{g:.4e}" for g in gp) + ")"
                     " equivalent_force=(-grad_p)=(" + ", ".join(f"{g:.4e}" for g in force) + ")")
            r = B.solve_brinkman(fc, geo, db, force, want_velocity=cfg.write_velocity_fields)
            d = r.diagnostics
            log.info("Macroscopic velocity U=(" + ", ".join(fmt_sci(a) for a in r.superficial_velocity) + ")"
                     f" iterations={d.iterations} residual={fmt_sci(d.residual)} converged={'yes' if d.converged else 'no'}")
            self.result = np.asarray(r.superficial_velocity)
            if cfg.write_velocity_fields:
                vpath = velocity_field_output_path(cfg.output_file, "")
                _ensure_parent(vpath)
                voxelio.write_fvel(vpath, (nx, ny, nz), r.porosity, r.viscosity, force, r.velocity)
                log.info("Velocity field written to " + vpath)
            if cfg.output_file:
                tpath = text_output_path(cfg.output_file, "velocity")
                _ensure_parent(tpath)
                with open(tpath, "w") as f:
                    f.write("# FFT_Method fluid pressure-gradient response\n")
                    f.write("# solver = brinkman\n")
                    f.write(f"# geometry_file = {cfg.geometry_file}\n")
                    f.write(f"# grid = {nx} x {ny} x {nz}\n")
                    f.write(f"# voxel_size = {fmt_default(cfg.voxel_size)}\n")
                    f.write(f"# porosity = {fmt_default(r.porosity)}\n")
                    f.write(f"# viscosity = {fmt_default(r.viscosity)}\n")
                    f.write("# pressure_gradient = " + " ".join(fmt_default(g) for g in gp) + "\n")
                    f.write("# macroscopic_velocity = " + " ".join(fmt_default(a) for a in r.superficial_velocity) + "\n")
                    f.write(f"# iterations = {d.iterations}\n")
                    f.write(f"# residual = {fmt_default(d.residual)}\n")
                    f.write(f"# converged = {'true' if d.converged else 'false'}\n")
                log.info("Results written to " + tpath)
            return 0

        def on_case(axis, force, solve):
            if not cfg.write_velocity_fields:
                return
            vpath = velocity_field_output_path_axis(cfg.output_file, axis)
            _ensure_parent(vpath)
            voxelio.write_fvel(vpath, (nx, ny, nz), solve.porosity, solve.viscosity, force, solve.velocity)
            log.info("Velocity field written to " + vpath)

        res = B.compute_permeability_tensor(fc, geo, db, on_case if cfg.write_velocity_fields else None)
        self.result = np.asarray(res.tensor)
        s = tensor_string(res.tensor, _INDENT)
        log.info("Permeability tensor K (Darcy sign, K > 0):\n" + _INDENT + s)
        if cfg.output_file:
            tpath = text_output_path(cfg.output_file, "permeability")
            _ensure_parent(tpath)
            with open(tpath, "w") as f:
                f.write("# FFT_Method permeability result\n")
                f.write("# solver = brinkman\n")
                f.write(f"# geometry_file = {cfg.geometry_file}\n")
                f.write(f"# grid = {nx} x {ny} x {nz}\n")
                f.write(f"# voxel_size = {fmt_default(cfg.voxel_size)}\n")
                f.write(f"# porosity = {fmt_default(res.porosity)}\n")
                f.write(f"# viscosity = {fmt_default(res.viscosity)}\n")
                f.write("#\n")
                f.write("# Permeability tensor K (Darcy sign, K > 0):\n")
                f.write(tensor_string(res.tensor, _INDENT) + "\n")
            log.info("Results written to " + tpath)
        return 0


def read_permeability_file(path: str):
    """Parse a `<stem>_permeability.txt` written by this runner or the reference: (header dict, K 3x3)."""
    header, rows = {}, []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if s.startswith("#"):
                body = s[1:].strip()
                if " = " in body:
                    k, v = body.split(" = ", 1)
                    header[k.strip()] = v.strip()
            elif s:
                rows.append([float(t) for t in s.split()])
    K = np.asarray(rows, dtype=np.float64)
    if K.shape != (3, 3):
        raise ValueError(f"{path}: expected a 3x3 tensor, got shape {K.shape}")
    return header, K


def main(argv=None) -> int:
    """main (src/main.cpp:43-89): log startup, parse, run; any exception -> 'Fatal error' + usage, 1."""
    argv = list(sys.argv[1:] if argv is None else argv)
    log = Logger()
    try:
        if "--log" in argv:
            i = argv.index("--log")
            if i + 1 < len(argv):
                log.open(argv[i + 1])
        log.info("=" * 60)
        log.info("FFT_Method startup")
        log.info("Parsing command line and configuration")
        cfg = from_cli(argv)
        if cfg.log_file and ("--log" not in argv):
            log.open(cfg.log_file)
            log.info("Log file opened from configuration: " + cfg.log_file)
        log.info("Analysis path: " + cfg.analysis)
        log.info(f"Fluid path: solver={cfg.solver} mode={cfg.fluid_mode}")
        log.info("Compute backend: " + cfg.compute_backend)
        log.info("Initializing runner")
        return PermeabilityRunner(cfg, log).run()
    except Exception as ex:  # mirrors the catch-all of main
        print(f"Fatal error: {ex}", file=sys.stderr)
        print(USAGE, file=sys.stderr)
        return 1
    finally:
        log.close()


if __name__ == "__main__":
    sys.exit(main())
