"""Scenario description (Python side of include/fdtd_scene.h).

The fields mirror the reference's compile-time configuration
(3D_FDTD_version008 src/user_config.hpp:11-211, src/tls_materials.hpp:23-25)
as gathered by Config::sim (src/params.hpp:823-868).  ``SceneSpec.to_c()``
produces the C struct consumed by the scene layer of the HIP library and by
the CPU checkers, keeping every referenced array alive on the Python object.

The named workloads at the bottom are the benchmark configurations of this
path (uniform CPML sphere case, graded-mesh gain slab, the ~100M-cell
precision form) plus the small cases the parity tests use.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

# Integer encodings, src/user_config.hpp:53-66, 121, 135 and src/params.hpp:598-663.
SHAPE_BOX, SHAPE_SPHERE, SHAPE_CYLINDER_Z = 0, 1, 2
WF_RICKER, WF_GAUSS_MOD_SINE, WF_RICKER_GAUSS2, WF_CW = 0, 1, 2, 3
POL = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5}
DIR = {"+X": 0, "-X": 1, "+Y": 2, "-Y": 3, "+Z": 4, "-Z": 5}
COMP = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5, "E_mag": 6, "H_mag": 7,
        "Sx": 8, "Sy": 9, "Sz": 10, "S_mag": 11}
BC_PEC, BC_CPML = 0, 1

C0 = 299792458.0


class _Structure(C.Structure):
    _fields_ = [("type", C.c_int), ("n", C.c_double), ("params", C.c_double * 6), ("tls_material", C.c_int)]


class _TLSMaterial(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("lambda0", "gamma", "tau", "N0", "inversion_fraction")]


class _Dipole(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("x", "y", "z", "amplitude", "frequency", "wavelength", "tau",
                                         "df_fwhm", "t0_factor")] + [("waveform", C.c_int), ("polarization", C.c_int)]


class _Plane(C.Structure):
    _fields_ = ([(n, C.c_double) for n in ("injection_position", "amplitude", "frequency", "wavelength", "tau",
                                           "df_fwhm", "t0_factor")]
                + [("waveform", C.c_int), ("direction", C.c_int), ("polarization", C.c_int)]
                + [(n, C.c_double) for n in ("x_min", "x_max", "y_min", "y_max", "z_min", "z_max")])


class _Probe(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("n_components", C.c_int),
                ("components", C.c_int * 12)]


class _DFT(C.Structure):
    _fields_ = [("component", C.c_int)] + [(n, C.c_int) for n in ("i0", "i1", "j0", "j1", "k0", "k1")] + [
        ("n_freqs", C.c_int), ("freqs", C.c_double * 8)]


class CSceneSpec(C.Structure):
    """Binary layout of fdtd_scene_spec (include/fdtd_scene.h)."""
    _fields_ = [
        ("nx_core", C.c_int), ("ny_core", C.c_int), ("nz_core", C.c_int), ("npml", C.c_int),
        ("dx_uniform", C.c_double), ("dy_uniform", C.c_double), ("dz_uniform", C.c_double),
        ("dx", C.POINTER(C.c_double)), ("dy", C.POINTER(C.c_double)), ("dz", C.POINTER(C.c_double)),
        ("domain_x_min", C.c_double), ("domain_y_min", C.c_double), ("domain_z_min", C.c_double),
        ("S", C.c_double),
        ("boundary_type", C.c_int),
        ("cpml_m", C.c_double), ("cpml_rerr", C.c_double), ("cpml_alpha0", C.c_double), ("cpml_kappa_max", C.c_double),
        ("cpml_alpha_linear", C.c_int),
        ("background_n", C.c_double),
        ("n_structures", C.c_int), ("structures", C.POINTER(_Structure)),
        ("tls_enabled", C.c_int), ("tls_clamp", C.c_int),
        ("n_tls_materials", C.c_int), ("tls_materials", C.POINTER(_TLSMaterial)),
        ("n_dipoles", C.c_int), ("dipoles", C.POINTER(_Dipole)),
        ("n_planes", C.c_int), ("planes", C.POINTER(_Plane)),
        ("n_probes", C.c_int), ("probes", C.POINTER(_Probe)),
        ("save_every", C.c_int),
        ("n_dfts", C.c_int), ("dfts", C.POINTER(_DFT)),
        ("subpixel_samples", C.c_int),
        ("stability_interval", C.c_int),
        ("device_bake", C.c_int),
        ("x_bounds", C.POINTER(C.c_double)), ("y_bounds", C.POINTER(C.c_double)),
        ("z_bounds", C.POINTER(C.c_double)),
    ]


@dataclass
class Structure:
    type: int
    n: float
    params: Sequence[float]
    tls_material: int = -1


@dataclass
class TLSMaterial:
    lambda0: float
    gamma: float
    tau: float
    N0: float
    inversion_fraction: float


@dataclass
class Dipole:
    x: float
    y: float
    z: float
    amplitude: float = 1e-7
    frequency: float = 0.0
    wavelength: float = 1.0e-6
    tau: float = 3e-15
    df_fwhm: float = 0.0
    t0_factor: float = 3.0
    waveform: int = WF_GAUSS_MOD_SINE
    polarization: int = 2


@dataclass
class PlaneWave:
    injection_position: float
    amplitude: float = 1.0
    frequency: float = 0.0
    wavelength: float = 1.0e-6
    tau: float = 3e-15
    df_fwhm: float = 0.0
    t0_factor: float = 3.0
    waveform: int = WF_GAUSS_MOD_SINE
    direction: int = 4
    polarization: int = 0
    x_min: float = 0.0
    x_max: float = 0.0
    y_min: float = 0.0
    y_max: float = 0.0
    z_min: float = 0.0
    z_max: float = 0.0


@dataclass
class Probe:
    x: float
    y: float
    z: float
    components: Sequence[int] = (2,)


@dataclass
class DFTMonitor:
    component: int
    box: Sequence[int]          # i0, i1, j0, j1, k0, k1 (half-open, total grid)
    freqs: Sequence[float]


@dataclass
class SceneSpec:
    nx_core: int
    ny_core: int
    nz_core: int
    npml: int = 10
    dx_uniform: float = 20e-9
    dy_uniform: float = 20e-9
    dz_uniform: float = 20e-9
    dx: Optional[np.ndarray] = None   # full total-grid spacing arrays (graded meshes)
    dy: Optional[np.ndarray] = None
    dz: Optional[np.ndarray] = None
    domain_min: Sequence[float] = (0.0, 0.0, 0.0)
    S: float = 0.99
    boundary_type: int = BC_CPML
    cpml_m: float = 3.0
    cpml_rerr: float = 1e-8
    cpml_alpha0: float = 0.05
    cpml_kappa_max: float = 5.0
    cpml_alpha_linear: int = 1
    background_n: float = 1.0
    structures: List[Structure] = field(default_factory=list)
    tls_enabled: int = 1
    tls_clamp: int = 0
    tls_materials: List[TLSMaterial] = field(default_factory=list)
    dipoles: List[Dipole] = field(default_factory=list)
    planes: List[PlaneWave] = field(default_factory=list)
    probes: List[Probe] = field(default_factory=list)
    save_every: int = 1
    dfts: List[DFTMonitor] = field(default_factory=list)
    subpixel_samples: int = 8
    stability_interval: int = 10
    device_bake: int = 0            # 1: bake materials on the device at scene creation
    x_bounds: Optional[np.ndarray] = None  # grid lines of an auto-meshed grid (with dx/dy/dz)
    y_bounds: Optional[np.ndarray] = None
    z_bounds: Optional[np.ndarray] = None
    n_steps: int = 100
    name: str = "scene"

    # ---- derived geometry (host-side, no library needed) ----
    def total_dims(self):
        if self.boundary_type == BC_PEC:
            return self.nx_core, self.ny_core, self.nz_core
        return tuple(c + 2 * self.npml for c in (self.nx_core, self.ny_core, self.nz_core))

    def n_cells(self) -> int:
        a, b, c = self.total_dims()
        return a * b * c

    def copy(self, **kw) -> "SceneSpec":
        return dataclasses.replace(self, **kw)

    def to_c(self) -> CSceneSpec:
        s = CSceneSpec()
        keep = []
        s.nx_core, s.ny_core, s.nz_core, s.npml = self.nx_core, self.ny_core, self.nz_core, self.npml
        s.dx_uniform, s.dy_uniform, s.dz_uniform = self.dx_uniform, self.dy_uniform, self.dz_uniform
        for nm in ("dx", "dy", "dz", "x_bounds", "y_bounds", "z_bounds"):
            arr = getattr(self, nm)
            if arr is not None:
                a = np.ascontiguousarray(arr, dtype=np.float64)
                keep.append(a)
                setattr(s, nm, a.ctypes.data_as(C.POINTER(C.c_double)))
        s.domain_x_min, s.domain_y_min, s.domain_z_min = self.domain_min
        s.S = self.S
        s.boundary_type = self.boundary_type
        s.cpml_m, s.cpml_rerr, s.cpml_alpha0, s.cpml_kappa_max = (self.cpml_m, self.cpml_rerr, self.cpml_alpha0,
                                                                  self.cpml_kappa_max)
        s.cpml_alpha_linear = self.cpml_alpha_linear
        s.background_n = self.background_n

        def arr(ctype, items, fill):
            if not items:
                return 0, None
            a = (ctype * len(items))()
            for i, it in enumerate(items):
                fill(a[i], it)
            keep.append(a)
            return len(items), C.cast(a, C.POINTER(ctype))

        def f_struct(d, it):
            d.type, d.n, d.tls_material = it.type, it.n, it.tls_material
            for q in range(6):
                d.params[q] = it.params[q] if q < len(it.params) else 0.0

        def f_copy(d, it):
            for nm, _ in d._fields_:
                setattr(d, nm, getattr(it, nm))

        def f_probe(d, it):
            d.x, d.y, d.z = it.x, it.y, it.z
            d.n_components = len(it.components)
            for q, c in enumerate(it.components):
                d.components[q] = c

        def f_dft(d, it):
            d.component = it.component
            d.i0, d.i1, d.j0, d.j1, d.k0, d.k1 = it.box
            d.n_freqs = len(it.freqs)
            for q, f in enumerate(it.freqs):
                d.freqs[q] = f

        s.n_structures, s.structures = arr(_Structure, self.structures, f_struct)
        s.tls_enabled, s.tls_clamp = self.tls_enabled, self.tls_clamp
        s.n_tls_materials, s.tls_materials = arr(_TLSMaterial, self.tls_materials, f_copy)
        s.n_dipoles, s.dipoles = arr(_Dipole, self.dipoles, f_copy)
        s.n_planes, s.planes = arr(_Plane, self.planes, f_copy)
        s.n_probes, s.probes = arr(_Probe, self.probes, f_probe)
        s.save_every = self.save_every
        s.n_dfts, s.dfts = arr(_DFT, self.dfts, f_dft)
        s.subpixel_samples = self.subpixel_samples
        s.stability_interval = self.stability_interval
        s.device_bake = self.device_bake
        s._keep = keep  # keep arrays alive with the struct
        return s


# ---------------------------------------------------------------------------
# Meshes

def graded_axis(n_total: int, fine: float, coarse: float, fine_len: float, ratio: float = 1.2,
                center_frac: float = 0.5) -> np.ndarray:
    """Spacing array of n_total cells: a run of `fine` cells covering `fine_len`
    centred at `center_frac` of the axis, graded geometrically by `ratio` per
    cell to `coarse`, then `coarse` cells to fill the axis.  The grading is the
    documented closed form used for the graded-mesh workloads (the reference
    builds such arrays with its auto-mesher, src/auto_mesh_generator.hpp:386-898;
    this path only consumes them)."""
    n_fine = int(round(fine_len / fine))
    ramp = []
    d = fine
    while True:
        d = min(d * ratio, coarse)
        ramp.append(d)
        if d >= coarse:
            break
    n_rest = n_total - n_fine - 2 * len(ramp)
    if n_rest < 2:
        raise ValueError("axis too short for the graded profile")
    n_lo = int(round(n_rest * center_frac))
    n_hi = n_rest - n_lo
    out = [coarse] * n_lo + ramp[::-1] + [fine] * n_fine + ramp + [coarse] * n_hi
    return np.asarray(out, dtype=np.float64)


def cell_center_phys(idx: int, d: float, npml: int) -> float:
    """Physical coordinate of the centre of total-grid cell idx on a uniform axis."""
    return (idx - npml + 0.5) * d


# ---------------------------------------------------------------------------
# Named workloads

def uniform_sphere_case(core: int = 64, npml: int = 10, d: float = 20e-9, steps: int = 1000,
                        probes_ijk=None) -> SceneSpec:
    """The plain CPU case: core^3 uniform Yee grid, vacuum with a centred
    dielectric sphere (eps_r 2.25, radius core/6.4 cells i.e. 10 cells at 64),
    npml-cell CPML (m=3, kappa_max=5, alpha_max=0.05), centred Ez point dipole
    driven by a Gaussian-modulated sinusoid (lambda0 1 um, 3 fs), S = 0.99."""
    half = core * d / 2
    r = (10.0 / 64.0) * core * d
    cidx = npml + core // 2
    if probes_ijk is None:
        s = core / 64.0
        probes_ijk = [(cidx, cidx, cidx + int(round(10 * s))), (cidx - int(round(15 * s)), cidx, cidx),
                      (cidx, cidx + int(round(15 * s)), cidx),
                      (cidx + int(round(13 * s)), cidx + int(round(13 * s)), cidx - int(round(12 * s)))]
    probes = [Probe(cell_center_phys(i, d, npml), cell_center_phys(j, d, npml), cell_center_phys(k, d, npml), (2,))
              for (i, j, k) in probes_ijk]
    src_pos = cell_center_phys(cidx, d, npml)
    return SceneSpec(
        nx_core=core, ny_core=core, nz_core=core, npml=npml, dx_uniform=d, dy_uniform=d, dz_uniform=d,
        S=0.99, cpml_m=3.0, cpml_rerr=1e-8, cpml_alpha0=0.05, cpml_kappa_max=5.0,
        structures=[Structure(SHAPE_SPHERE, 1.5, (half, half, half, r, 0, 0))],
        dipoles=[Dipole(src_pos, src_pos, src_pos, amplitude=1e-7, wavelength=1.0e-6, tau=3e-15, t0_factor=3.0,
                        waveform=WF_GAUSS_MOD_SINE, polarization=POL["Ez"])],
        probes=probes, save_every=1, n_steps=steps, name=f"uniform_sphere_{core}")


def gain_slab_case(n_total: int = 256, npml: int = 12, steps: int = 5000, lateral: float = 20e-9,
                   fine: float = 10e-9, coarse: float = 40e-9, fine_len: float = 400e-9,
                   n_cuboids: int = 0, seed: int = 1, shape=None) -> SceneSpec:
    """Graded-mesh gain case: z spacing 10 nm in a central 400 nm slab graded
    at ratio 1.2 per cell to 40 nm; x, y uniform (`lateral`).  A 200 nm two-level
    gain slab (N 1e24 m^-3, 1.0 um transition, 20 THz linewidth -> gamma =
    2 pi 20e12 s^-1, tau21 1 ns, 30% initial inversion, host n 1.5) sits on an
    eps_r 3.5 substrate; an Ex current sheet (plane-wave source, -Z) above the
    slab drives it.  Probes in the slab and above it; a DFT monitor on a z
    plane at 3 frequencies.  With n_cuboids > 0, seeded random dielectric
    cuboids (eps_r U[1,12], edges U{8..40} cells) are added (precision form)."""
    if shape is None:
        shape = (n_total, n_total, n_total)
    nxt, nyt, nzt = shape
    dz = graded_axis(nzt, fine, coarse, fine_len)
    dx = np.full(nxt, lateral)
    dy = np.full(nyt, lateral)
    zb = np.concatenate([[0.0], np.cumsum(dz)])
    pml_z = zb[npml]
    # fine region location in physical z
    kf0 = int(np.argmax(dz <= fine * 1.0000001))
    z_fine0 = zb[kf0] - pml_z
    z_fine1 = z_fine0 + fine_len
    z_mid = 0.5 * (z_fine0 + z_fine1)
    Lx = (nxt - 2 * npml) * lateral
    Ly = (nyt - 2 * npml) * lateral
    Lz = zb[nzt - npml] - pml_z
    big = 1.0  # extends past the domain
    slab_z0, slab_z1 = z_mid - 100e-9, z_mid + 100e-9
    structures = [
        Structure(SHAPE_BOX, math.sqrt(3.5), (-big, big, -big, big, -big, slab_z0)),
        Structure(SHAPE_BOX, 1.5, (-big, big, -big, big, slab_z0, slab_z1), tls_material=0),
    ]
    if n_cuboids:
        rng = np.random.default_rng(seed)
        zcore = nzt - 2 * npml
        for _ in range(n_cuboids):
            ex, ey, ez = (int(v) for v in rng.integers(8, 41, size=3))
            i0 = int(rng.integers(0, max(1, nxt - 2 * npml - ex)))
            j0 = int(rng.integers(0, max(1, nyt - 2 * npml - ey)))
            k0 = int(rng.integers(0, max(1, zcore - ez)))
            eps = float(rng.uniform(1.0, 12.0))
            zc = zb[npml + k0] - pml_z
            zc1 = zb[npml + min(k0 + ez, zcore)] - pml_z
            structures.append(Structure(SHAPE_BOX, math.sqrt(eps),
                                        (i0 * lateral, (i0 + ex) * lateral, j0 * lateral, (j0 + ey) * lateral,
                                         zc, zc1)))
    tls = [TLSMaterial(lambda0=1.0e-6, gamma=2 * math.pi * 20e12, tau=1e-9, N0=1e24, inversion_fraction=0.3)]
    src_z = min(z_fine1 + 300e-9, Lz - 2 * coarse)
    planes = [PlaneWave(injection_position=src_z, amplitude=1.0e3, wavelength=1.0e-6, tau=3e-15, t0_factor=3.0,
                        waveform=WF_GAUSS_MOD_SINE, direction=DIR["-Z"], polarization=POL["Ex"])]
    probes = [Probe(Lx / 2, Ly / 2, z_mid, (0, 2)), Probe(Lx / 2 + 3 * lateral, Ly / 2, z_mid + 50e-9, (0, 2, 4)),
              Probe(Lx / 2, Ly / 2, src_z - 100e-9, (0,)), Probe(Lx / 3, Ly / 4, slab_z0 - 60e-9, (0, 1))]
    kmon = int(np.searchsorted(zb, pml_z + src_z - 150e-9)) - 1
    kmon = min(max(kmon, npml), nzt - npml - 1)
    dfts = [DFTMonitor(0, (npml, nxt - npml, npml, nyt - npml, kmon, kmon + 1),
                       (C0 / 0.95e-6, C0 / 1.0e-6, C0 / 1.05e-6))]
    return SceneSpec(
        nx_core=nxt - 2 * npml, ny_core=nyt - 2 * npml, nz_core=nzt - 2 * npml, npml=npml,
        dx_uniform=lateral, dy_uniform=lateral, dz_uniform=fine, dx=dx, dy=dy, dz=dz,
        S=0.99, cpml_m=3.0, cpml_rerr=1e-8, cpml_alpha0=0.05, cpml_kappa_max=5.0,
        structures=structures, tls_materials=tls, planes=planes, probes=probes, dfts=dfts,
        save_every=1, n_steps=steps, name=f"gain_slab_{nxt}x{nyt}x{nzt}" + (f"_cub{n_cuboids}" if n_cuboids else ""))


def precision_case(steps: int = 4000, shape=(512, 512, 384), n_cuboids: int = 64, seed: int = 1) -> SceneSpec:
    """~100M-cell precision-form geometry: the gain-slab structure plus 64
    seeded random dielectric cuboids on a 512x512x384 graded grid."""
    sp = gain_slab_case(npml=12, steps=steps, shape=shape, n_cuboids=n_cuboids, seed=seed)
    sp.name = f"precision_{shape[0]}x{shape[1]}x{shape[2]}"
    return sp


def _random_domain(rng, lo: int, hi: int, gain: Optional[bool], nblocks: Optional[int], steps: int,
                   name: str) -> SceneSpec:
    """A seeded random domain: each total axis log-uniform in [lo, hi] cells,
    CPML 8..12 cells, a uniform mesh (spacing U[10, 40] nm) or a z axis graded
    10 -> 40 nm, `nblocks` dielectric blocks (eps_r U[1, 12]; None: 1..8), an
    Ez point dipole (lambda U[0.8, 1.6] um), a two-level gain slab with a DFT
    plane above it when `gain` (None: with probability 1/4), one Ez probe."""
    dims = [int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))))) for _ in range(3)]
    npml = int(rng.integers(8, 13))
    graded = bool(rng.random() < 0.5)
    if graded:
        lat = 20e-9
        dx, dy = np.full(dims[0], lat), np.full(dims[1], lat)
        dz = graded_axis(dims[2], 10e-9, 40e-9, max(4, dims[2] // 5) * 10e-9)
    else:
        d = float(rng.uniform(10e-9, 40e-9))
        dx, dy, dz = np.full(dims[0], d), np.full(dims[1], d), np.full(dims[2], d)
    bounds = [np.concatenate([[0.0], np.cumsum(a)]) for a in (dx, dy, dz)]
    core_lo = [b[npml] for b in bounds]
    core_len = [b[n - npml] - b[npml] for b, n in zip(bounds, dims)]
    structures = []
    nb = int(rng.integers(1, 9)) if nblocks is None else nblocks
    for _ in range(nb):
        lo_ = [float(rng.uniform(0, 0.7)) * L for L in core_len]
        hi_ = [lo_[a] + float(rng.uniform(0.1, 0.3)) * core_len[a] for a in range(3)]
        eps = float(rng.uniform(1.0, 12.0))
        structures.append(Structure(SHAPE_BOX, math.sqrt(eps), (lo_[0], hi_[0], lo_[1], hi_[1], lo_[2], hi_[2])))
    tls, dfts = [], []
    if gain is None:
        gain = bool(rng.random() < 0.25)
    zmid = 0.5 * core_len[2]
    if gain:
        big = 1.0
        th = 0.2 * core_len[2]
        structures.append(Structure(SHAPE_BOX, 1.5, (-big, big, -big, big, zmid - th / 2, zmid + th / 2),
                                    tls_material=0))
        tls = [TLSMaterial(lambda0=1.0e-6, gamma=2 * math.pi * 20e12, tau=1e-9, N0=1e24, inversion_fraction=0.3)]
        kmon = int(np.searchsorted(bounds[2], core_lo[2] + zmid + 0.3 * core_len[2])) - 1
        kmon = min(max(kmon, npml), dims[2] - npml - 1)
        dfts = [DFTMonitor(2, (npml, dims[0] - npml, npml, dims[1] - npml, kmon, kmon + 1),
                           (C0 / 0.95e-6, C0 / 1.0e-6, C0 / 1.05e-6))]
    src = [0.5 * core_len[0], 0.5 * core_len[1], zmid + 0.25 * core_len[2]]
    lam = float(rng.uniform(0.8e-6, 1.6e-6))
    dip = Dipole(src[0], src[1], src[2], amplitude=1e-7, wavelength=lam, tau=3e-15, t0_factor=3.0,
                 waveform=WF_GAUSS_MOD_SINE, polarization=POL["Ez"])
    probes = [Probe(0.5 * core_len[0], 0.4 * core_len[1], zmid, (2,))]
    return SceneSpec(
        nx_core=dims[0] - 2 * npml, ny_core=dims[1] - 2 * npml, nz_core=dims[2] - 2 * npml, npml=npml,
        dx_uniform=float(dx[0]), dy_uniform=float(dy[0]), dz_uniform=float(dz.min()), dx=dx, dy=dy, dz=dz,
        S=0.99, structures=structures, tls_materials=tls, dipoles=[dip], probes=probes, dfts=dfts,
        save_every=1, n_steps=steps, name=f"{name}_{dims[0]}x{dims[1]}x{dims[2]}")


def batch_case(index: int, seed: int = 1, steps: int = 2000, lo: int = 32, hi: int = 160) -> SceneSpec:
    """One domain of the batching workload, seeded by (seed, index): axes
    log-uniform in [lo, hi] = [32, 160] cells, 1..8 blocks, gain in a quarter
    of the domains (see _random_domain)."""
    rng = np.random.default_rng([seed, index])
    return _random_domain(rng, lo, hi, None, None, steps, f"batch_{seed}_{index}")


def serve_case(index: int, kind: int, seed: int = 1, lo: int = 48, hi: int = 224, steps: int = 500) -> SceneSpec:
    """One request of the serving workload, seeded by (seed, index): kind 0 a
    dielectric scene (one block), 1 dielectric + gain slab (one block), 2 a
    multi-block scene (4..8 blocks); axes log-uniform in [lo, hi] cells."""
    rng = np.random.default_rng([seed, 1000003 + index])
    gain = kind == 1
    nblocks = 1 if kind in (0, 1) else int(rng.integers(4, 9))
    return _random_domain(rng, lo, hi, gain, nblocks, steps, f"serve_{seed}_{index}_k{kind}")
