"""Voxel geometry and velocity-field files in the reference's formats.

.fgeo v1 (src/io/VoxelIO.cpp:89-159): little-endian header
    'FGEO' | u8 version=1 | u8 value_type (1 = uint32, 2 = uint16) | u16 reserved |
    u64 nx | u64 ny | u64 nz | payload of nx*ny*nz material ids, x fastest.
.geom text (src/io/VoxelIO.cpp:48-87): '#'/';' comments, first content line "nx ny nz",
    then the ids, whitespace separated.
FVEL v1 (src/app/PermeabilityRunner.cpp:109-159): 'FVEL' | u8 version=1 | u8 0 | u16 0 |
    u64 nx, ny, nz | f64 porosity | f64 mu | f64 driving[3] | f64 (ux,uy,uz) per voxel.
Arrays are (nz, ny, nx) == flat index (k*ny + j)*nx + i.
"""
from __future__ import annotations

import struct

import numpy as np

_FGEO_HEADER = struct.Struct("<4sBBHQQQ")
_FVEL_HEADER = struct.Struct("<4sBBHQQQdd3d")


def write_fgeo(path, ids, value_type: int = 2) -> None:
    ids = np.asarray(ids)
    if ids.ndim != 3:
        raise ValueError("ids must be (nz, ny, nx)")
    nz, ny, nx = ids.shape
    if value_type not in (1, 2):
        raise ValueError("value_type must be 1 (uint32) or 2 (uint16)")
    dt = "<u4" if value_type == 1 else "<u2"
    with open(path, "wb") as f:
        f.write(_FGEO_HEADER.pack(b"FGEO", 1, value_type, 0, nx, ny, nz))
        f.write(np.ascontiguousarray(ids, dtype=dt).tobytes())


def read_fgeo(path) -> np.ndarray:
    """VoxelIO::load_geometry for .fgeo: raw uint32 material ids, shape (nz, ny, nx)."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < _FGEO_HEADER.size:
        raise RuntimeError(f"VoxelIO::load_fgeo file too small: {path}")
    magic, version, vt, _, nx, ny, nz = _FGEO_HEADER.unpack_from(data)
    if magic != b"FGEO":
        raise RuntimeError(f"VoxelIO::load_fgeo invalid magic in: {path}")
    if version != 1:
        raise RuntimeError(f"VoxelIO::load_fgeo unsupported version {version} in: {path}")
    if vt not in (1, 2):
        raise RuntimeError(f"VoxelIO::load_fgeo unsupported value_type {vt} in: {path}")
    esz = 4 if vt == 1 else 2
    count = nx * ny * nz
    if len(data) != _FGEO_HEADER.size + count * esz:
        raise RuntimeError(f"VoxelIO::load_fgeo payload size mismatch in: {path}")
    arr = np.frombuffer(data, dtype="<u4" if vt == 1 else "<u2", offset=_FGEO_HEADER.size, count=count)
    return arr.astype(np.uint32).reshape(nz, ny, nx)


def read_fgeo_mask(path, fluid_id: int = 1) -> np.ndarray:
    """VoxelIO::load_fgeo: uint8 mask, 1 where the id equals fluid_id."""
    return (read_fgeo(path) == fluid_id).astype(np.uint8)


def read_geom_text(path) -> np.ndarray:
    """VoxelIO::load_geometry for text files: uint32 ids, shape (nz, ny, nx)."""
    dims = None
    vals = []
    with open(path) as f:
        for line in f:
            cut = min([i for i in (line.find("#"), line.find(";")) if i >= 0], default=-1)
            if cut >= 0:
                line = line[:cut]
            toks = line.split()
            if not toks:
                continue
            if dims is None:
                if len(toks) < 3:
                    raise RuntimeError("VoxelIO::load_text expected first non-comment line: nx ny nz")
                dims = tuple(int(t) for t in toks[:3])
                continue
            vals.extend(int(t) for t in toks)
            if len(vals) >= dims[0] * dims[1] * dims[2]:
                break
    if dims is None:
        raise RuntimeError("VoxelIO::load_text expected first non-comment line: nx ny nz")
    nx, ny, nz = dims
    if len(vals) < nx * ny * nz:
        raise RuntimeError("VoxelIO::load_text insufficient voxel values")
    return np.array(vals[: nx * ny * nz], dtype=np.uint32).reshape(nz, ny, nx)


def write_geom_text(path, ids) -> None:
    ids = np.asarray(ids)
    nz, ny, nx = ids.shape
    with open(path, "w") as f:
        f.write(f"{nx} {ny} {nz}\n")
        for k in range(nz):
            for j in range(ny):
                f.write(" ".join(str(int(v)) for v in ids[k, j]) + "\n")


def load(path, fluid_id: int = 1) -> np.ndarray:
    """VoxelIO::load: dispatch on the .fgeo extension; returns the uint8 fluid mask."""
    ids = read_fgeo(path) if str(path).endswith(".fgeo") else read_geom_text(path)
    return (ids == fluid_id).astype(np.uint8)


def write_fvel(path, shape, porosity, mu, driving, velocity) -> None:
    """write_velocity_field_binary: velocity is SoA (3, N) or (3N,)."""
    nx, ny, nz = shape
    v = np.asarray(velocity, dtype=np.float64).reshape(3, -1)
    with open(path, "wb") as f:
        f.write(_FVEL_HEADER.pack(b"FVEL", 1, 0, 0, nx, ny, nz, float(porosity), float(mu), *map(float, driving)))
        f.write(np.ascontiguousarray(v.T).tobytes())


def read_fvel(path):
    with open(path, "rb") as f:
        data = f.read()
    magic, ver, _, _, nx, ny, nz, por, mu, d0, d1, d2 = _FVEL_HEADER.unpack_from(data)
    if magic != b"FVEL" or ver != 1:
        raise RuntimeError("not an FVEL v1 file")
    v = np.frombuffer(data, dtype="<f8", offset=_FVEL_HEADER.size).reshape(-1, 3).T.copy()
    return dict(shape=(nx, ny, nz), porosity=por, mu=mu, driving=(d0, d1, d2), velocity=v)
