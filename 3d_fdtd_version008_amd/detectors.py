"""Detector output files in the reference's formats, written from the
device's records (so the reference's Py_Visualization scripts read them
unchanged).

TLS diagnostics (records from Scene.tls_records, one every save_every steps
from step 0):

- <root>/populations/inversion_time.csv, PopulationTimeSeriesDetector
  (src/two_level_detectors.hpp:17-109): header
  "t, total_Nu, total_Ng, inversion, inv_fraction", rows written with
  std::setprecision(18) (the %.18g conversion);
- <root>/polarization/polarization_time.csv, PolarizationDiagnostics
  (src/two_level_detectors.hpp:205-263): header
  "t, Pz_center, dPz_dt_center, total_P_energy", %.18g;
- <root>/tls_global_history.csv, TLSGlobalHistory
  (src/two_level_detectors.hpp:265-335): header
  "step,t_fs,Nu_total,Ng_total,inversion_total,Nu_percent,Ng_percent,inv_norm",
  the step, t in fs with 6 decimals, totals in scientific notation with 12
  digits, fractions with 8 decimals.

Each value is derived from the record exactly as the reference derives it
from its own sums (inversion = Nu - Ng, fraction = inversion / (Nu + Ng)
when the sum exceeds 1e-30, and so on).
"""
from __future__ import annotations

import os
from typing import Sequence

import numpy as np


def _g18(x: float) -> str:
    return "%.18g" % x


def write_tls_csv(root: str, records: np.ndarray, dt: float, save_every: int) -> Sequence[str]:
    """Writes the three TLS CSV files under `root`; returns their paths."""
    rec = np.asarray(records, dtype=np.float64).reshape(-1, 5)
    pop_dir = os.path.join(root, "populations")
    pol_dir = os.path.join(root, "polarization")
    os.makedirs(pop_dir, exist_ok=True)
    os.makedirs(pol_dir, exist_ok=True)
    p_inv = os.path.join(pop_dir, "inversion_time.csv")
    p_pol = os.path.join(pol_dir, "polarization_time.csv")
    p_hist = os.path.join(root, "tls_global_history.csv")
    with open(p_inv, "w") as fi, open(p_pol, "w") as fp, open(p_hist, "w") as fh:
        fi.write("t, total_Nu, total_Ng, inversion, inv_fraction\n")
        fp.write("t, Pz_center, dPz_dt_center, total_P_energy\n")
        fh.write("step,t_fs,Nu_total,Ng_total,inversion_total,Nu_percent,Ng_percent,inv_norm\n")
        for r, (nu, ng, u_int, pz, dpz) in enumerate(rec):
            n = r * save_every
            t = n * dt
            inversion = nu - ng
            denom = nu + ng
            frac = inversion / denom if denom > 1e-30 else 0.0
            fi.write(f"{_g18(t)},{_g18(nu)},{_g18(ng)},{_g18(inversion)},{_g18(frac)}\n")
            fp.write(f"{_g18(t)},{_g18(pz)},{_g18(dpz)},{_g18(u_int)}\n")
            ntot = nu + ng
            nu_pc, ng_pc, inv_norm = 0.0, 0.0, 0.5
            if ntot > 1e-30:
                nu_pc = nu / ntot
                ng_pc = ng / ntot
                inv_norm = (inversion / ntot + 1.0) / 2.0
            t_fs = n * dt * 1e15
            fh.write(f"{n},{t_fs:.6f},{nu:.12e},{ng:.12e},{inversion:.12e},{nu_pc:.8f},{ng_pc:.8f},{inv_norm:.8f}\n")
    return [p_inv, p_pol, p_hist]


# ---- grid helpers: GridSpacing (src/global_function.hpp:92-158) ----

def _off(bounds, npml: int) -> float:
    return float(bounds[npml]) if npml < len(bounds) else 0.0


def cell_center_physical(bounds, npml: int, i: int) -> float:
    """GridSpacing::cell_center_physical_{x,y,z} (src/global_function.hpp:123-134)."""
    if i + 1 >= len(bounds):
        return float(bounds[-1]) - _off(bounds, npml)
    return 0.5 * (float(bounds[i]) + float(bounds[i + 1])) - _off(bounds, npml)


def physical_to_index(bounds, npml: int, phys: float) -> int:
    """GridSpacing::physical_to_index_{x,y,z} (src/global_function.hpp:137-157)."""
    import bisect
    total = phys + _off(bounds, npml)
    if total <= bounds[0]:
        return 0
    if total >= bounds[-1]:
        return len(bounds) - 2
    return bisect.bisect_right(list(bounds), total) - 1


COMPONENT_NAMES = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "E_mag", "H_mag", "Sx", "Sy", "Sz", "S_mag"]


def component_values(comp: int, f: dict) -> np.ndarray:
    """Detectors::get_field_value (src/detectors/idetector.hpp:79-109) on
    FP64 arrays f["Ex"] ... f["Hz"], same operations in the same order."""
    Ex, Ey, Ez, Hx, Hy, Hz = (np.asarray(f[n], dtype=np.float64) for n in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"))
    if comp < 6:
        return (Ex, Ey, Ez, Hx, Hy, Hz)[comp]
    if comp == 6:
        return np.sqrt(Ex * Ex + Ey * Ey + Ez * Ez)
    if comp == 7:
        return np.sqrt(Hx * Hx + Hy * Hy + Hz * Hz)
    sx, sy, sz = Ey * Hz - Ez * Hy, Ez * Hx - Ex * Hz, Ex * Hy - Ey * Hx
    if comp == 8:
        return sx
    if comp == 9:
        return sy
    if comp == 10:
        return sz
    return np.sqrt(sx * sx + sy * sy + sz * sz)


def _g(x: float, p: int) -> str:
    return "%.*g" % (p, x)


# ---- PointFieldDetector (src/detectors/point_field_detector.hpp:41-179) ----

def write_point_probe(root: str, name: str, setup, cell, components: Sequence[int], traces: Sequence[np.ndarray],
                      save_every: int, n_steps: int, float64: bool = True) -> str:
    """<root>/<name>/<comp>_ts.bin (one value per record, native byte order,
    float64 or float32) and metadata.json, for a probe at total-grid cell
    `cell` (the placement of PointFieldDetector::initialize, which
    Setup.probe(c)[0] gives) whose component traces are `traces`."""
    d = os.path.join(root, name)
    os.makedirs(d, exist_ok=True)
    xb, yb, zb = (setup.array(n, shaped=False) for n in ("x_bounds", "y_bounds", "z_bounds"))
    npml = setup.npml
    nx, ny, nz = setup.shape
    i0, j0, k0 = (int(v) for v in cell)
    for c, tr in zip(components, traces):
        np.asarray(tr, dtype=np.float64 if float64 else np.float32).tofile(os.path.join(d, COMPONENT_NAMES[c] + "_ts.bin"))
    px, py, pz = (cell_center_physical(b, npml, q) for b, q in ((xb, i0), (yb, j0), (zb, k0)))
    ox, oy, oz = _off(xb, npml), _off(yb, npml), _off(zb, npml)
    names = [COMPONENT_NAMES[c] for c in components]
    lines = ["{", '  "detector_type": "PointFieldDetector",', f'  "NxT": {nx},', f'  "NyT": {ny},', f'  "NzT": {nz},',
             f'  "i0": {i0},', f'  "j0": {j0},', f'  "k0": {k0},',
             f'  "probe_x_physical_m": {_g(px, 15)},', f'  "probe_y_physical_m": {_g(py, 15)},',
             f'  "probe_z_physical_m": {_g(pz, 15)},',
             f'  "probe_x_physical_nm": {_g(px * 1e9, 15)},', f'  "probe_y_physical_nm": {_g(py * 1e9, 15)},',
             f'  "probe_z_physical_nm": {_g(pz * 1e9, 15)},',
             f'  "pml_offset_x_m": {_g(ox, 15)},', f'  "pml_offset_y_m": {_g(oy, 15)},', f'  "pml_offset_z_m": {_g(oz, 15)},',
             f'  "save_every": {save_every},', f'  "n_steps": {n_steps},', f'  "dt": {_g(setup.dt, 18)},',
             f'  "dtype": "{"float64" if float64 else "float32"}",',
             '  "components": [' + ", ".join(f'"{n}"' for n in names) + "],",
             '  "files": [' + ", ".join(f'"{n}_ts.bin"' for n in names) + "],",
             '  "note": "time series of field components at single point"', "}"]
    with open(os.path.join(d, "metadata.json"), "w") as f:
        f.write("\n".join(lines) + "\n")
    return d


def _slice_index(setup, plane: int, position: float, domain_min=(0.0, 0.0, 0.0)):
    """Slice index and its cell-centre coordinate for a plane (0 XY, 1 XZ,
    2 YZ) at a physical position, as FieldMovie2D / MeshDetector::initialize
    place it (src/detectors/field_movie_2d.hpp:70-92): physical_to_index of
    position - domain_min along the normal, clamped to [1, N-2]."""
    nx, ny, nz = setup.shape
    npml = setup.npml
    xb, yb, zb = (setup.array(n, shaped=False) for n in ("x_bounds", "y_bounds", "z_bounds"))
    b, n, dm = ((zb, nz, domain_min[2]), (yb, ny, domain_min[1]), (xb, nx, domain_min[0]))[plane]
    idx = max(1, min(physical_to_index(b, npml, position - dm), n - 2))
    return idx, cell_center_physical(b, npml, idx)


def _plane_of(v: np.ndarray, plane: int, index: int) -> np.ndarray:
    """The slice in the detectors' loop order (dim1 outer, dim2 inner)."""
    return np.ascontiguousarray((v[:, :, index], v[:, index, :], v[index, :, :])[plane])


# ---- FieldMovie2D (src/detectors/field_movie_2d.hpp:41-213) ----

class FieldMovie2DWriter:
    """Frames of one field component on an XY (0), XZ (1) or YZ (2) plane:
    <root>/<name>/frame_NNNN.raw (values in the reference's loop order,
    float64 or float32) and metadata.json.  The slice index comes from the
    physical position as in FieldMovie2D::initialize (clamped to [1, N-2])."""

    def __init__(self, root: str, name: str, setup, component: int, slice_plane: int, slice_position: float,
                 save_every: int, n_steps: int, float64: bool, n_phys, L_phys, domain_min=(0.0, 0.0, 0.0),
                 frame_pattern: str = "frame_%04d.raw"):
        self.dir = os.path.join(root, name)
        os.makedirs(self.dir, exist_ok=True)
        self.component, self.plane, self.float64 = component, slice_plane, float64
        self.save_every, self.frame_pattern, self.frame_id = save_every, frame_pattern, 0
        nx, ny, nz = setup.shape
        npml = setup.npml
        xb, yb, zb = (setup.array(n, shaped=False) for n in ("x_bounds", "y_bounds", "z_bounds"))
        self.index, coord = _slice_index(setup, slice_plane, slice_position, domain_min)
        pname = ("XY", "XZ", "YZ")[slice_plane]
        dims = ((nx, ny, "x", "y"), (nx, nz, "x", "z"), (ny, nz, "y", "z"))[slice_plane]
        lines = ["{", '  "detector_type": "FieldMovie2D",', f'  "field_component": "{COMPONENT_NAMES[component]}",',
                 f'  "slice_plane": "{pname}",', f'  "slice_index": {self.index},',
                 f'  "slice_physical_m": {_g(coord, 15)},', f'  "slice_physical_nm": {_g(coord * 1e9, 15)},',
                 f'  "NxT": {nx},', f'  "NyT": {ny},', f'  "NzT": {nz},',
                 f'  "Nx_phys": {n_phys[0]},', f'  "Ny_phys": {n_phys[1]},', f'  "Nz_phys": {n_phys[2]},',
                 f'  "npml": {npml},', f'  "save_every": {save_every},', f'  "n_steps": {n_steps},',
                 f'  "frame_pattern": "{frame_pattern}",',
                 f'  "pml_offset_x_m": {_g(_off(xb, npml), 15)},', f'  "pml_offset_y_m": {_g(_off(yb, npml), 15)},',
                 f'  "pml_offset_z_m": {_g(_off(zb, npml), 15)},',
                 f'  "Lx_phys_m": {_g(L_phys[0], 15)},', f'  "Ly_phys_m": {_g(L_phys[1], 15)},',
                 f'  "Lz_phys_m": {_g(L_phys[2], 15)},', f'  "dtype": "{"float64" if float64 else "float32"}",',
                 f'  "slice_dim1": {dims[0]},', f'  "slice_dim2": {dims[1]},', f'  "dim1_label": "{dims[2]}",',
                 f'  "dim2_label": "{dims[3]}",', '  "note": "row-major storage, dim1 varies fastest"', "}"]
        with open(os.path.join(self.dir, "metadata.json"), "w") as f:
            f.write("\n".join(lines) + "\n")

    def frame(self, fields: dict) -> str:
        """Writes the next frame from FP64 arrays fields["Ex"] ... ["Hz"]."""
        v = component_values(self.component, fields)
        path = os.path.join(self.dir, self.frame_pattern % self.frame_id)
        self.frame_id += 1
        _plane_of(v, self.plane, self.index).astype(np.float64 if self.float64 else np.float32).tofile(path)
        return path


# ---- mesh description files (src/params.hpp:297-503, written at set-up) ----

STRUCTURE_TYPE_NAMES = {0: "box", 1: "sphere", 2: "cylinder"}


def _g6(x: float) -> str:
    # an ostream with default flags and precision 6
    return "%g" % x


def write_mesh_files(root: str, setup, spec, lambda0: float, mode: str = "UNIFORM") -> Sequence[str]:
    """<root>/grid_spacing.txt, coordinate_mapping.txt and mesh_info.json
    in the reference's formats (src/params.hpp:297-503): spacing and bounds
    one per line at 6 significant digits, the coordinate diagnostics at 6,
    mesh_info.json at 15.  `mode` is "UNIFORM" or "AUTO_NONUNIFORM"; the
    source marker is the first dipole of `spec` (the domain centre when
    there is none), placed by GridSpacing::physical_to_index."""
    os.makedirs(root, exist_ok=True)
    d = [setup.array(n, shaped=False) for n in ("dx", "dy", "dz")]
    b = [setup.array(n, shaped=False) for n in ("x_bounds", "y_bounds", "z_bounds")]
    npml = setup.npml
    nt = setup.shape
    p_grid = os.path.join(root, "grid_spacing.txt")
    with open(p_grid, "w") as f:
        f.write(f"# Mode: {mode}\n")
        for label, arr in zip(("dx", "dy", "dz", "x_bounds", "y_bounds", "z_bounds"), d + b):
            f.write(f"# {label}\n")
            f.write("".join(_g6(v) + "\n" for v in arr.tolist()))

    tot_min = [float(a[0]) for a in b]
    tot_max = [float(a[-1]) for a in b]
    core_min = [float(a[npml]) for a in b]
    core_max = [float(a[n - npml]) for a, n in zip(b, nt)]
    offs = [_off(a, npml) for a in b]
    p_coord = os.path.join(root, "coordinate_mapping.txt")
    with open(p_coord, "w") as f:
        f.write("# Coordinate System Mapping Diagnostics\n# =====================================\n\n")
        f.write("# Total Domain (includes PML)\n")
        for q, ax in enumerate("xyz"):
            f.write(f"total_{ax}_min_nm = {_g6(tot_min[q] * 1e9)}\ntotal_{ax}_max_nm = {_g6(tot_max[q] * 1e9)}\n")
        f.write("\n# Core Region indices\n")
        for q, ax in enumerate("ijk"):
            f.write(f"{ax}_core_start = {npml}\n{ax}_core_end = {nt[q] - npml}\n")
        f.write("\n# Core Region in total coordinates (nm)\n")
        for q, ax in enumerate("xyz"):
            f.write(f"core_{ax}_min_total_nm = {_g6(core_min[q] * 1e9)}\ncore_{ax}_max_total_nm = {_g6(core_max[q] * 1e9)}\n")
        f.write("\n# PML Offsets (nm)\n")
        for q, ax in enumerate("xyz"):
            f.write(f"pml_offset_{ax}_nm = {_g6(offs[q] * 1e9)}\n")
        f.write("\n# Physical domain size (nm)\n")
        for q, ax in enumerate("xyz"):
            f.write(f"phys_size_{ax}_nm = {_g6((core_max[q] - core_min[q]) * 1e9)}\n")

    size = [core_max[q] - core_min[q] for q in range(3)]
    target = [s / 2.0 for s in size]
    if spec.dipoles:
        s0 = spec.dipoles[0]
        target = [s0.x - spec.domain_min[0], s0.y - spec.domain_min[1], s0.z - spec.domain_min[2]]
    src = [physical_to_index(b[q], npml, target[q]) for q in range(3)]
    actual = [cell_center_physical(b[q], npml, src[q]) for q in range(3)]
    g = lambda x: _g(x, 15)  # noqa: E731
    lines = ["{", f'  "mode": "{mode}",', f'  "NxT": {nt[0]},', f'  "NyT": {nt[1]},', f'  "NzT": {nt[2]},',
             f'  "npml": {npml},']
    for ax, arr in zip("xyz", d):
        lines += [f'  "d{ax}_min_nm": {g(float(arr.min()) * 1e9)},', f'  "d{ax}_max_nm": {g(float(arr.max()) * 1e9)},']
    lines += [f'  "domain_{ax}_nm": {g(tot_max[q] * 1e9)},' for q, ax in enumerate("xyz")]
    lines += [f'  "dt_fs": {g(setup.dt * 1e15)},', f'  "lambda0_nm": {g(lambda0 * 1e9)},']
    lines += [f'  "source_{ax}_target_nm": {g(target[q] * 1e9)},' for q, ax in enumerate("xyz")]
    lines += [f'  "source_{ax}": {src[q]},' for q, ax in enumerate("ijk")]
    lines += [f'  "source_{ax}_actual_nm": {g(actual[q] * 1e9)},' for q, ax in enumerate("xyz")]
    for q, ax in enumerate("xyz"):
        lines += [f'  "phys_domain_{ax}_min_nm": 0.0,', f'  "phys_domain_{ax}_max_nm": {g(size[q] * 1e9)},']
    lines += ['  "grid_type": "bounds",'] + [f'  "grid_{ax}_count": {len(b[q])},' for q, ax in enumerate("xyz")]
    text = "\n".join(lines) + f'\n  "num_structures": {len(spec.structures)}'
    if spec.structures:
        items = []
        for s in spec.structures:
            par = ", ".join(g(float(s.params[j] if j < len(s.params) else 0.0) * 1e9) for j in range(6))
            items.append(f'    {{"type": "{STRUCTURE_TYPE_NAMES[s.type]}", "n": {g(s.n)}, "params_nm": [{par}]}}')
        text += ",\n  \"structures\": [\n" + ",\n".join(items) + "\n  ]\n}\n"
    else:
        text += ",\n  \"structures\": []\n}\n"
    p_info = os.path.join(root, "mesh_info.json")
    with open(p_info, "w") as f:
        f.write(text)
    return [p_grid, p_coord, p_info]


# ---- MeshDetector (src/detectors/mesh_detector.hpp:19-263) ----

def write_mesh_detector(root: str, name: str, setup, slice_plane: int, slice_position: float, n_phys, L_phys,
                        domain_min=(0.0, 0.0, 0.0), float64: bool = True, mesh_lines: bool = True,
                        spacing_arrays: bool = True) -> str:
    """<root>/<name>/refractive_index.raw (the slice of the cell-centre index
    grid, Setup.array("n_grid"), as make_n_grid_from_scene builds it),
    metadata.json, and optionally x/y/z_bounds.txt and dx/dy/dz_array.txt,
    as the detector package writes them at set-up (src/params.hpp:695-739)."""
    d = os.path.join(root, name)
    os.makedirs(d, exist_ok=True)
    nx, ny, nz = setup.shape
    npml = setup.npml
    idx, coord = _slice_index(setup, slice_plane, slice_position, domain_min)
    n_grid = setup.array("n_grid")
    _plane_of(n_grid, slice_plane, idx).astype(np.float64 if float64 else np.float32).tofile(
        os.path.join(d, "refractive_index.raw"))
    b = [setup.array(q, shaped=False) for q in ("x_bounds", "y_bounds", "z_bounds")]
    g = lambda x: _g(x, 15)  # noqa: E731
    dims = ((nx, ny, "x", "y"), (nx, nz, "x", "z"), (ny, nz, "y", "z"))[slice_plane]
    lines = ["{", '  "detector_type": "MeshDetector",', f'  "slice_plane": "{("XY", "XZ", "YZ")[slice_plane]}",',
             f'  "slice_index": {idx},', f'  "slice_physical_m": {g(coord)},', f'  "slice_physical_nm": {g(coord * 1e9)},',
             f'  "NxT": {nx},', f'  "NyT": {ny},', f'  "NzT": {nz},', f'  "Nx_phys": {n_phys[0]},',
             f'  "Ny_phys": {n_phys[1]},', f'  "Nz_phys": {n_phys[2]},', f'  "npml": {npml},']
    lines += [f'  "pml_offset_{ax}_m": {g(_off(b[q], npml))},' for q, ax in enumerate("xyz")]
    lines += [f'  "L{ax}_phys_m": {g(L_phys[q])},' for q, ax in enumerate("xyz")]
    lines += [f'  "dtype": "{"float64" if float64 else "float32"}",', f'  "slice_dim1": {dims[0]},',
              f'  "slice_dim2": {dims[1]},', f'  "dim1_label": "{dims[2]}",', f'  "dim2_label": "{dims[3]}",',
              '  "note": "row-major storage, dim1 varies fastest"', "}"]
    with open(os.path.join(d, "metadata.json"), "w") as f:
        f.write("\n".join(lines) + "\n")
    if mesh_lines:
        for q, ax in enumerate("xyz"):
            with open(os.path.join(d, f"{ax}_bounds.txt"), "w") as f:
                f.write(f"# {ax.upper()} cell boundaries (meters)\n" + "".join(g(v) + "\n" for v in b[q].tolist()))
    if spacing_arrays:
        for ax in "xyz":
            arr = setup.array(f"d{ax}", shaped=False)
            with open(os.path.join(d, f"d{ax}_array.txt"), "w") as f:
                f.write(f"# d{ax} spacing array (meters)\n" + "".join(g(v) + "\n" for v in arr.tolist()))
    return d


# ---- 1-D mesh profiles (src/mesh_1d_profile_export.hpp:14-66) ----

def write_mesh_1d_profiles(out_dir: str, setup) -> Sequence[str]:
    """mesh_dx_profile.csv and mesh_dy_profile.csv: physical cell centres
    and widths in nm over the physical domain, 10 significant digits."""
    os.makedirs(out_dir, exist_ok=True)
    npml = setup.npml
    paths = []
    for ax in "xy":
        b = setup.array(f"{ax}_bounds", shaped=False)
        d = setup.array(f"d{ax}", shaped=False)
        start = float(b[npml])
        rows = [f"{_g((0.5 * (float(b[i]) + float(b[i + 1])) - start) * 1e9, 10)},{_g(float(d[i]) * 1e9, 10)}\n"
                for i in range(npml, len(d) - npml)]
        p = os.path.join(out_dir, f"mesh_d{ax}_profile.csv")
        with open(p, "w") as f:
            f.write(f"{ax}_center_nm,d{ax}_nm\n" + "".join(rows))
        paths.append(p)
    return paths


# ---- PopulationTimeSeriesDetector frames (src/two_level_detectors.hpp:17-109) ----

class PopulationFrames:
    """<root>/<name>/populations_Nu.bin and populations_Ng.bin: the full-grid
    FP64 densities in idx3 order at the first recorded step and at the end of
    the run (two frames each), and metadata_populations.json, as
    PopulationTimeSeriesDetector writes them.  Its CSV is write_tls_csv's
    populations/inversion_time.csv."""

    def __init__(self, root: str, setup, save_every: int, n_steps: int, name: str = "populations"):
        self.dir = os.path.join(root, name)
        os.makedirs(self.dir, exist_ok=True)
        self.shape, self.every, self.n_steps, self.dt = setup.shape, save_every, n_steps, setup.dt
        self.first_done = False
        self.paths = [os.path.join(self.dir, "populations_Nu.bin"), os.path.join(self.dir, "populations_Ng.bin")]
        for p in self.paths:  # the detector opens (and truncates) both at construction
            open(p, "wb").close()

    def _append(self, nu, ng):
        for p, a in zip(self.paths, (nu, ng)):
            with open(p, "ab") as f:
                np.ascontiguousarray(a, dtype=np.float64).tofile(f)

    def record(self, n: int, nu, ng) -> None:
        """After step n's population update; the first recorded step's frame."""
        if n % self.every != 0 or self.first_done:
            return
        self._append(nu, ng)
        self.first_done = True

    def finalize(self, nu, ng) -> None:
        self._append(nu, ng)
        nx, ny, nz = self.shape
        lines = ["{", f'  "NxT": {nx},', f'  "NyT": {ny},', f'  "NzT": {nz},', f'  "saveEvery": {self.every},',
                 f'  "nSteps": {self.n_steps},', f'  "dt": {_g18(self.dt)},', '  "dtype": "float64",',
                 '  "binary_frames": 2,', '  "binary_mode": "first_and_last",',
                 '  "binary_data_units": "m^-3 (density)",', '  "csv_total_units": "atoms (integrated)",',
                 '  "what": "population_densities_Nu_and_Ng_vs_time"', "}"]
        with open(os.path.join(self.dir, "metadata_populations.json"), "w") as f:
            f.write("\n".join(lines) + "\n")
