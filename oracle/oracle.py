"""ctypes front-end for the CPU checkers.

TEST INFRASTRUCTURE ONLY: the FP64 oracle (liboracle.so, a C restatement of
the reference's serial step) and, where it was built, the compiled reference
harness (_ref/libref.so).  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")
# the reference harness with the device library bound in as main() would bind it
REF_DROPIN_SO = os.path.join(HERE, "_ref", "libref_dropin.so")

_libs = {}


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and _ref/libref.so where the reference tree exists)."""
    out = subprocess.run(["make", "-C", HERE, "-s"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout, out.stderr)


def available(impl: str = "oracle") -> bool:
    return os.path.exists(ORACLE_SO if impl == "oracle" else REF_SO)


def _lib(impl: str):
    if impl in _libs:
        return _libs[impl]
    path = ORACLE_SO if impl == "oracle" else REF_SO
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run make -C oracle)")
    lib = C.CDLL(path)
    p = "oracle_" if impl == "oracle" else "ref_"
    vp = C.c_void_p
    sig = {
        "create": (vp, [vp]),
        "destroy": (None, [vp]),
        "step": (C.c_int, [vp, C.c_int]),
        "info": (None, [vp, C.POINTER(C.c_long)]),
        "dt": (C.c_double, [vp]),
        "last_max_E": (C.c_double, [vp]),
        "get_array": (C.c_long, [vp, C.c_char_p, vp]),
        "get_region_map": (C.c_long, [vp, vp]),
        "probe_records": (C.c_long, [vp, C.c_int]),
        "probe_index": (C.c_int, [vp, C.c_int, C.POINTER(C.c_long)]),
        "get_probe": (C.c_long, [vp, C.c_int, C.c_int, vp]),
        "set_array": (C.c_int, [vp, C.c_char_p, vp]),
    }
    if impl == "ref":
        d, i, l = C.c_double, C.c_int, C.c_long
        sig.update({
            "add_point_detector": (i, [vp, C.c_char_p, C.c_char_p, d, d, d, i, vp, i, l, i, d, d, d]),
            "add_movie": (i, [vp, C.c_char_p, C.c_char_p, i, i, d, i, l, i, l, l, l, d, d, d, d, d, d]),
            "add_mesh_detector": (i, [vp, C.c_char_p, C.c_char_p, i, d, i, i, i, l, l, l, d, d, d, d, d, d]),
            "dump_mesh_profiles": (i, [vp, C.c_char_p]),
            "n_grid": (l, [vp, vp]),
            "add_tls_detectors": (i, [vp, C.c_char_p, l, l, l, l, l]),
            "time_steps": (C.c_double, [vp, C.c_int]),
            "create_prebaked": (vp, [vp, vp]),
            "set_threads": (None, [C.c_int]),
        })
    if impl == "oracle":
        sig.update({
            "get_region": (C.c_int, [vp, C.c_int, vp]),
            "get_source": (C.c_int, [vp, C.c_int, C.c_long, vp]),
            "get_dft": (C.c_long, [vp, C.c_int, vp]),
            "set_storage": (C.c_int, [vp, C.c_int, vp, C.c_ulonglong]),
            "round_store": (C.c_double, [C.c_double, C.c_int, C.c_int]),
            "round_blocks": (None, [vp, C.c_long, C.c_int, C.c_int]),
            "time_steps": (C.c_double, [vp, C.c_int]),
            "last_conservation": (C.c_double, [vp]),
        })
    fns = {}
    for name, (res, args) in sig.items():
        f = getattr(lib, p + name)
        f.restype = res
        f.argtypes = args
        fns[name] = f
    f = getattr(lib, p + "threads")
    f.restype = C.c_int
    fns["threads"] = f
    _libs[impl] = fns
    return fns


class CpuSim:
    """One CPU simulation (impl 'oracle' = the C restatement, 'ref' = the
    compiled reference harness)."""

    def __init__(self, spec, impl: str = "oracle", materials=None):
        """materials (reference harness only): the 12 MaterialGrids arrays
        (aEx ... bHz, FP64, idx3 order) used in place of the reference's
        serial bake, for timing large grids."""
        self.impl = impl
        self.f = _lib(impl)
        self._cspec = spec.to_c()
        if materials is not None:
            if impl != "ref":
                raise ValueError("materials are taken by the reference harness only")
            self._mats = [np.ascontiguousarray(m, dtype=np.float64).ravel() for m in materials]
            ptrs = (C.c_void_p * 12)(*[m.ctypes.data for m in self._mats])
            self.h = self.f["create_prebaked"](C.addressof(self._cspec), ptrs)
            self._mats = None
        else:
            self.h = self.f["create"](C.addressof(self._cspec))
        if not self.h:
            raise RuntimeError("create failed")
        info = self.info()
        self.shape = (info[0], info[1], info[2])
        self.npml = info[3]

    def close(self):
        if self.h:
            self.f["destroy"](self.h)
            self.h = None

    def add_point_detector(self, out_root, name, x, y, z, components, save_every, n_steps, float64=True,
                           domain_min=(0.0, 0.0, 0.0)):
        """The reference's PointFieldDetector writing its own files (ref only)."""
        comps = (C.c_int * len(components))(*components)
        self.f["add_point_detector"](self.h, out_root.encode(), name.encode(), x, y, z, len(components),
                                     C.cast(comps, C.c_void_p), save_every, n_steps, int(float64), *domain_min)

    def add_movie(self, out_root, name, component, plane, position, save_every, n_steps, float64, n_phys, L_phys,
                  domain_min=(0.0, 0.0, 0.0)):
        """The reference's FieldMovie2D writing its own files (ref only)."""
        self.f["add_movie"](self.h, out_root.encode(), name.encode(), component, plane, position, save_every, n_steps,
                            int(float64), *n_phys, *L_phys, *domain_min)

    def add_mesh_detector(self, out_root, name, plane, position, float64, mesh_lines, spacing, n_phys, L_phys,
                          domain_min=(0.0, 0.0, 0.0)):
        """The reference's MeshDetector writing its files now (ref only)."""
        self.f["add_mesh_detector"](self.h, out_root.encode(), name.encode(), plane, position, int(float64),
                                    int(mesh_lines), int(spacing), *n_phys, *L_phys, *domain_min)

    def dump_mesh_profiles(self, out_dir):
        """The reference's mesh_export::dump_mesh_1d_profiles (ref only)."""
        if self.f["dump_mesh_profiles"](self.h, out_dir.encode()) != 0:
            raise RuntimeError("dump_mesh_1d_profiles failed")

    def add_tls_detectors(self, out_root, ic, jc, kc, save_every, n_steps):
        """main's population, polarization and TLS-history detectors (ref only;
        files are complete after close())."""
        if self.f["add_tls_detectors"](self.h, out_root.encode(), ic, jc, kc, save_every, n_steps) != 0:
            raise RuntimeError("no TLS region")

    def n_grid(self) -> np.ndarray:
        """The reference's make_n_grid_from_scene (ref only)."""
        n = self.f["n_grid"](self.h, None)
        out = np.empty(n, dtype=np.float64)
        self.f["n_grid"](self.h, out.ctypes.data_as(C.c_void_p))
        return out.reshape(self.shape)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        out = (C.c_long * 14)()
        self.f["info"](self.h, out)
        return list(out)

    @property
    def dt(self) -> float:
        return self.f["dt"](self.h)

    @property
    def steps_done(self) -> int:
        return self.info()[4]

    def step(self, n: int) -> int:
        return self.f["step"](self.h, n)

    def time_steps(self, n: int) -> float:
        return self.f["time_steps"](self.h, n)

    def array(self, name: str, shaped: bool = True) -> np.ndarray:
        n = self.f["get_array"](self.h, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        out = np.empty(n, dtype=np.float64)
        self.f["get_array"](self.h, name.encode(), out.ctypes.data_as(C.c_void_p))
        if shaped and n == self.shape[0] * self.shape[1] * self.shape[2]:
            return out.reshape(self.shape)
        return out

    def set_array(self, name: str, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        if self.f["set_array"](self.h, name.encode(), a.ctypes.data_as(C.c_void_p)) != 0:
            raise KeyError(name)

    def region_map(self) -> np.ndarray:
        n = self.f["get_region_map"](self.h, None)
        out = np.empty(max(n, 0), dtype=np.int32)
        if n > 0:
            self.f["get_region_map"](self.h, out.ctypes.data_as(C.c_void_p))
            return out.reshape(self.shape)
        return out

    def probe_index(self, p: int):
        ijk = (C.c_long * 3)()
        nc = self.f["probe_index"](self.h, p, ijk)
        return tuple(ijk), nc

    def probe(self, p: int, c: int = 0) -> np.ndarray:
        n = self.f["probe_records"](self.h, p)
        out = np.empty(max(n, 0), dtype=np.float64)
        if n > 0:
            self.f["get_probe"](self.h, p, c, out.ctypes.data_as(C.c_void_p))
        return out

    def region(self, r: int) -> np.ndarray:
        out = np.empty(17, dtype=np.float64)
        if self.f["get_region"](self.h, r, out.ctypes.data_as(C.c_void_p)) != 0:
            raise IndexError(r)
        return out

    def source(self, q: int, n: int = 0) -> np.ndarray:
        out = np.empty(19, dtype=np.float64)
        if self.f["get_source"](self.h, q, n, out.ctypes.data_as(C.c_void_p)) != 0:
            raise IndexError(q)
        return out

    def dft(self, m: int) -> np.ndarray:
        n = self.f["get_dft"](self.h, m, None)
        out = np.empty(n, dtype=np.float64)
        self.f["get_dft"](self.h, m, out.ctypes.data_as(C.c_void_p))
        return out

    STORAGE = {"fp64": 0, "fp16": 1, "fp16_stochastic": 2, "fp16_residual8": 3, "fp16_psi32": 4,
               "bfp16_32": 5, "bfp16_64": 6, "bfp16_8": 7}

    def set_storage(self, fmt: str, scales18=None, seed: int = 1) -> None:
        """Emulated storage rounding of fields and psi (oracle only): see
        oracle_set_storage in fdtd_oracle.c."""
        sc = None if scales18 is None else np.ascontiguousarray(scales18, dtype=np.float64)
        ptr = None if sc is None else sc.ctypes.data_as(C.c_void_p)
        if self.f["set_storage"](self.h, self.STORAGE[fmt], ptr, seed) != 0:
            raise ValueError(fmt)

    def stability(self):
        info = self.info()
        return {"stable": bool(info[8]), "fail_step": info[9], "warnings": info[10],
                "max_E": self.f["last_max_E"](self.h)}


def threads(impl: str = "oracle") -> int:
    return _lib(impl)["threads"]()


def set_threads(n: int, impl: str = "ref") -> None:
    """OpenMP threads of the reference harness's parallel loops."""
    _lib(impl)["set_threads"](n)


REF_MESH_SO = os.path.join(HERE, "_ref", "libref_mesh.so")
REF_MESH_TESTS = (os.path.join(HERE, "_ref", "test_auto_mesh"), os.path.join(HERE, "_ref", "verify_mesh_override"))


def ref_auto_mesh(cfg, domain, npml, boxes=(), background_n=1.0, cap=20000):
    """The reference's own AutoMeshGenerator::generate (compiled from the
    reference tree into _ref/libref_mesh.so): x, y and z grid lines for an
    AutoMeshConfig-like `cfg` (attributes as in automesh.AutoMeshConfig)."""
    if "mesh" not in _libs:
        lib = C.CDLL(REF_MESH_SO)
        f = lib.ref_auto_mesh
        d = C.c_double
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, d, d, d, C.c_long, C.c_int, C.c_void_p, C.c_int, C.c_void_p, d, C.c_void_p,
                      C.c_long, C.c_void_p]
        _libs["mesh"] = f
    f = _libs["mesh"]
    c = np.array([cfg.lambda_min, cfg.f_max, cfg.mesh_accuracy, cfg.ppw_override, cfg.dx_min, cfg.dx_max,
                  cfg.max_grading_ratio, float(bool(cfg.pml_graded)), float(bool(cfg.override_enabled))])
    st = np.array([[b.x_min, b.x_max, b.y_min, b.y_max, b.z_min, b.z_max, b.n] for b in boxes],
                  dtype=np.float64).reshape(-1, 7)
    ov = np.array([[float(bool(o.enabled)), o.x0, o.x1, o.y0, o.y1, o.z0, o.z1, o.dx, o.dy, o.dz, o.priority]
                   for o in cfg.overrides], dtype=np.float64).reshape(-1, 11)
    out = np.zeros(3 * cap)
    counts = np.zeros(3, dtype=np.int64)
    rc = f(c.ctypes.data, float(domain[0]), float(domain[1]), float(domain[2]), int(npml), len(st),
           st.ctypes.data, len(ov), ov.ctypes.data, float(background_n), out.ctypes.data, cap, counts.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"ref_auto_mesh returned {rc}")
    return [out[a * cap:a * cap + counts[a]].copy() for a in range(3)]


def ref_dropin_compare(spec, nsteps: int, precision: int = 0, device_bake: bool = False):
    """Runs the reference's own setup for `spec`, hands its arrays to the
    device library through the C ABI (fdtd_create, fdtd_set_sources,
    fdtd_add_probe, fdtd_step) and steps both for nsteps; returns (number
    of probe channels, relative L2 of Ex Ey Ez Hx Hy Hz Nu Ng Pz Pz_prev, the
    largest over the probe traces and the largest over the 12 material
    arrays).  precision: 0 FP64, 1 FP32, 2 FP16; device_bake: the binding
    passes the reference's structure list and the device bakes."""
    if not os.path.exists(REF_DROPIN_SO):
        raise FileNotFoundError(f"{REF_DROPIN_SO} is not built (run make -C oracle after the device library)")
    lib = C.CDLL(REF_DROPIN_SO)
    f = lib.ref_dropin_compare
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
    cs = spec.to_c()
    rel = (C.c_double * 12)()
    rc = f(C.addressof(cs), nsteps, precision, int(device_bake), rel)
    if rc < 0:
        raise RuntimeError(f"drop-in run failed ({rc})")
    return rc, list(rel)
