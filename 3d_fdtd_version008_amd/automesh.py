"""Automatic nonuniform mesh (MeshMode::AUTO_NONUNIFORM) over the C ABI of
include/fdtd_mesh.h.

`AutoMeshConfig` and `MeshOverride` mirror the reference's AutoMeshConfig
and MeshOverrideRegion (src/auto_mesh_generator.hpp:53-117, same defaults);
`generate` is AutoMeshGenerator::generate (:156-253), returning the grid
lines and cell widths of each axis as GridSpacing holds them;
`auto_mesh_spec` is make_context's AUTO_NONUNIFORM branch
(src/params.hpp:189-217 with register_structure_bounds_for_auto_mesh,
:540-553): it turns a SceneSpec into one on the generated grid.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import runtime

SQRT2_REFERENCE = 1.41421356237  # AutoMeshConfig::max_grading_ratio default


@dataclass
class AutoMeshConfig:
    lambda_min: float = 300e-9
    f_max: float = 0.0
    mesh_accuracy: int = 2
    ppw_override: float = 0.0
    dx_min: float = 1e-9
    dx_max: float = 100e-9
    max_grading_ratio: float = SQRT2_REFERENCE
    pml_graded: bool = False
    override_enabled: bool = False
    overrides: List["MeshOverride"] = field(default_factory=list)

    def ppw_target(self) -> float:
        """AutoMeshConfig::get_ppw_target (src/auto_mesh_generator.hpp:104-111)."""
        if self.ppw_override > 0:
            return self.ppw_override
        if self.mesh_accuracy <= 1:
            return 6.0
        if self.mesh_accuracy == 2:
            return 10.0
        return 14.0 + (self.mesh_accuracy - 3) * 4.0


@dataclass
class MeshOverride:
    x0: float = 0.0
    x1: float = 0.0
    y0: float = 0.0
    y1: float = 0.0
    z0: float = 0.0
    z1: float = 0.0
    dx: float = 10e-9
    dy: float = 10e-9
    dz: float = 10e-9
    enabled: bool = True
    priority: int = 0


@dataclass
class MeshBox:
    """AutoMeshStructureBounds: a box in physical coordinates and its index."""
    x_min: float
    x_max: float
    y_min: float
    y_max: float
    z_min: float
    z_max: float
    n: float


class _Cfg(C.Structure):
    _fields_ = [("lambda_min", C.c_double), ("f_max", C.c_double), ("mesh_accuracy", C.c_int),
                ("ppw_override", C.c_double), ("dx_min", C.c_double), ("dx_max", C.c_double),
                ("max_grading_ratio", C.c_double), ("pml_graded", C.c_int), ("override_enabled", C.c_int)]


class _Box(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("x_min", "x_max", "y_min", "y_max", "z_min", "z_max", "n")]


class _Override(C.Structure):
    _fields_ = [("enabled", C.c_int)] + [(n, C.c_double) for n in
                                         ("x0", "x1", "y0", "y1", "z0", "z1", "dx", "dy", "dz")] + [
        ("priority", C.c_int)]


_bound = False


def _lib():
    global _bound
    lib = runtime.load()
    if not _bound:
        lib.fdtd_auto_mesh_axis.restype = C.c_int
        lib.fdtd_auto_mesh_axis.argtypes = [C.POINTER(_Cfg), C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(_Box),
                                            C.c_int, C.POINTER(_Override), C.c_double, C.POINTER(C.c_double),
                                            C.c_long, C.POINTER(C.c_long)]
        lib.fdtd_structure_mesh_bounds.restype = C.c_int
        lib.fdtd_structure_mesh_bounds.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_double, C.POINTER(_Box)]
        _bound = True
    return lib


def _c_cfg(cfg: AutoMeshConfig) -> _Cfg:
    return _Cfg(cfg.lambda_min, cfg.f_max, int(cfg.mesh_accuracy), cfg.ppw_override, cfg.dx_min, cfg.dx_max,
                cfg.max_grading_ratio, int(bool(cfg.pml_graded)), int(bool(cfg.override_enabled)))


def axis_lines(cfg: AutoMeshConfig, axis: int, domain: float, npml: int, boxes: Sequence[MeshBox] = (),
               background_n: float = 1.0) -> np.ndarray:
    """Total-domain grid lines of one axis (fdtd_auto_mesh_axis)."""
    lib = _lib()
    c = _c_cfg(cfg)
    bx = (_Box * max(1, len(boxes)))(*[_Box(b.x_min, b.x_max, b.y_min, b.y_max, b.z_min, b.z_max, b.n)
                                       for b in boxes])
    ov = (_Override * max(1, len(cfg.overrides)))(*[
        _Override(int(bool(o.enabled)), o.x0, o.x1, o.y0, o.y1, o.z0, o.z1, o.dx, o.dy, o.dz, int(o.priority))
        for o in cfg.overrides])
    n = C.c_long()
    args = (C.byref(c), axis, float(domain), int(npml), len(boxes), bx, len(cfg.overrides), ov, float(background_n))
    runtime._check(lib.fdtd_auto_mesh_axis(*args, None, 0, C.byref(n)))
    out = np.empty(n.value, dtype=np.float64)
    runtime._check(lib.fdtd_auto_mesh_axis(*args, out.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n)))
    return out


@dataclass
class Grid:
    """GridSpacing as AutoMeshGenerator::generate returns it."""
    x_bounds: np.ndarray
    y_bounds: np.ndarray
    z_bounds: np.ndarray
    npml: int

    @property
    def dx(self) -> np.ndarray:
        return np.diff(self.x_bounds)

    @property
    def dy(self) -> np.ndarray:
        return np.diff(self.y_bounds)

    @property
    def dz(self) -> np.ndarray:
        return np.diff(self.z_bounds)


def generate(cfg: AutoMeshConfig, domain: Sequence[float], npml: int, boxes: Sequence[MeshBox] = (),
             background_n: float = 1.0) -> Grid:
    lines = [axis_lines(cfg, a, domain[a], npml, boxes, background_n) for a in range(3)]
    return Grid(lines[0], lines[1], lines[2], int(npml))


def structure_box(st) -> MeshBox:
    """register_structure_bounds_for_auto_mesh for one scene.Structure."""
    lib = _lib()
    p = (C.c_double * 6)(*[float(st.params[q]) if q < len(st.params) else 0.0 for q in range(6)])
    b = _Box()
    runtime._check(lib.fdtd_structure_mesh_bounds(int(st.type), p, float(st.n), C.byref(b)))
    return MeshBox(b.x_min, b.x_max, b.y_min, b.y_max, b.z_min, b.z_max, b.n)


def auto_mesh_spec(spec, lambda0: float, cfg: Optional[AutoMeshConfig] = None, domain: Optional[Sequence[float]] = None):
    """A copy of `spec` on the auto-generated grid, as make_context builds it:
    lambda_min = lambda0, dx_min / dx_max default to lambda0/100 and
    lambda0/4 when not positive, the structures' boxes, background index
    1.0 (the generate() call passes 1.0), the physical domain `domain`
    (default n*_core * d*_uniform, the reference's fallback)."""
    cfg = dataclasses.replace(cfg or AutoMeshConfig())
    cfg.lambda_min = lambda0
    if cfg.dx_min <= 0:
        cfg.dx_min = lambda0 / 100.0
    if cfg.dx_max <= 0:
        cfg.dx_max = lambda0 / 4.0
    if domain is None:
        domain = (spec.nx_core * spec.dx_uniform, spec.ny_core * spec.dy_uniform, spec.nz_core * spec.dz_uniform)
    npml = spec.npml
    g = generate(cfg, domain, npml, [structure_box(s) for s in spec.structures], 1.0)
    n = [len(b) - 1 - 2 * npml for b in (g.x_bounds, g.y_bounds, g.z_bounds)]
    return spec.copy(nx_core=n[0], ny_core=n[1], nz_core=n[2], dx=g.dx, dy=g.dy, dz=g.dz,
                     x_bounds=g.x_bounds, y_bounds=g.y_bounds, z_bounds=g.z_bounds)
