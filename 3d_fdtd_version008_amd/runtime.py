"""ctypes bindings of lib3dfdtd_hip.so (include/fdtd_hip.h, include/fdtd_scene.h).

This is the Python side of the drop-in boundary: ``Setup`` runs the
reference-equivalent host setup for a SceneSpec (no device needed) and
``Scene`` owns one device simulation and steps it on the GPU.  There is no
CPU fallback: if the library is missing or no device is present, creating a
Scene raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib3dfdtd_hip.so")

FP64, FP32, FP16 = 0, 1, 2
# status codes of the C ABI (include/fdtd_hip.h)
FDTD_OK, FDTD_EINVAL, FDTD_EDEVICE, FDTD_ENOMEM, FDTD_EUNSTABLE = 0, -1, -2, -3, -4
PRECISION = {"fp64": FP64, "fp32": FP32, "fp16": FP16}
FIELD_ID = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5, "Nu": 6, "Ng": 7, "Pz": 8, "Pz_prev": 9}
PSI_NAMES = ["psi_Ex_y", "psi_Ex_z", "psi_Ey_z", "psi_Ey_x", "psi_Ez_x", "psi_Ez_y",
             "psi_Hx_y", "psi_Hx_z", "psi_Hy_z", "psi_Hy_x", "psi_Hz_x", "psi_Hz_y"]
MATERIAL_NAMES = ["aEx", "bEx", "aEy", "bEy", "aEz", "bEz", "aHx", "bHx", "aHy", "bHy", "aHz", "bHz"]
for _q, _n in enumerate(PSI_NAMES):
    FIELD_ID[_n] = 10 + _q

# The C ABI's exported entry points (include/fdtd_hip.h, include/fdtd_scene.h).
EXPORTS = [
    "fdtd_create", "fdtd_destroy", "fdtd_last_error", "fdtd_device_bytes", "fdtd_download_material",
    "fdtd_add_tls_monitor", "fdtd_read_tls_monitor",
    "fdtd_set_sources", "fdtd_add_probe",
    "fdtd_add_dft", "fdtd_set_monitor", "fdtd_step", "fdtd_synchronize", "fdtd_download", "fdtd_upload",
    "fdtd_read_probe", "fdtd_read_dft", "fdtd_read_check", "fdtd_info", "fdtd_step_timed", "fdtd_step_timed_every",
    "fdtd_step_noalloc",
    "fdtd_scene_step_timed_every",
    "fdtd_setup_build", "fdtd_setup_free", "fdtd_setup_info", "fdtd_setup_dt", "fdtd_setup_array",
    "fdtd_setup_source_values", "fdtd_setup_probe", "fdtd_scene_create", "fdtd_scene_destroy", "fdtd_scene_step",
    "fdtd_scene_steps_done", "fdtd_scene_handle", "fdtd_scene_step_timed", "fdtd_scene_step_async",
    "fdtd_scene_step_wait", "fdtd_scene_step_query", "fdtd_step_async", "fdtd_step_wait", "fdtd_step_query",
    "fdtd_set_fp16_scales",
    "fdtd_get_fp16_scales", "fdtd_set_kernel_mode", "fdtd_set_device_pool", "fdtd_device_pool_bytes",
    "fdtd_stage_results",
    "fdtd_batch_create", "fdtd_batch_destroy", "fdtd_batch_add", "fdtd_batch_remove", "fdtd_batch_count",
    "fdtd_batch_set_tuning", "fdtd_batch_get_tuning", "fdtd_batch_set_graph", "fdtd_batch_step",
    "fdtd_scene_batch_create", "fdtd_scene_batch_destroy", "fdtd_scene_batch_add", "fdtd_scene_batch_remove",
    "fdtd_scene_batch_count", "fdtd_scene_batch_step", "fdtd_scene_batch_autotune", "fdtd_scene_batch_handle",
    "fdtd_scene_set_stop", "fdtd_auto_mesh_axis", "fdtd_structure_mesh_bounds", "fdtd_is_stable",
    "fdtd_setup_fp16_scales",
]


class FdtdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"fdtd error {code}: {msg}")
        self.code = code


class CheckRecord(C.Structure):
    _fields_ = [("step", C.c_long), ("max_E", C.c_double), ("max_H", C.c_double), ("max_P", C.c_double),
                ("nan_E", C.c_int), ("nan_H", C.c_int), ("nan_P", C.c_int), ("nan_N", C.c_int),
                ("pop_total", C.c_double), ("pop_initial", C.c_double)]


_lib = None


def library_path() -> str:
    return LIB_PATH


def load() -> C.CDLL:
    """Load lib3dfdtd_hip.so; raises if it is not built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    vp, i, l, d = C.c_void_p, C.c_int, C.c_long, C.c_double
    sigs = {
        "fdtd_last_error": (C.c_char_p, []),
        "fdtd_set_step_graphs": (i, [vp, i]),
        "fdtd_destroy": (i, [vp]),
        "fdtd_device_bytes": (i, [vp, C.POINTER(C.c_size_t)]),
        "fdtd_download_material": (i, [vp, i, C.c_void_p]),
        "fdtd_add_tls_monitor": (i, [vp, vp, vp, vp, i, i, i, i, l]),
        "fdtd_read_tls_monitor": (i, [vp, vp, l, C.POINTER(C.c_long)]),
        "fdtd_set_sources": (i, [vp, i, vp]),
        "fdtd_add_probe": (i, [vp, i, i, i, i, i, l]),
        "fdtd_add_dft": (i, [vp, i, vp, i, vp]),
        "fdtd_set_monitor": (i, [vp, i]),
        "fdtd_step": (i, [vp, l, i, vp]),
        "fdtd_synchronize": (i, [vp]),
        "fdtd_download": (i, [vp, i, vp]),
        "fdtd_upload": (i, [vp, i, vp]),
        "fdtd_read_probe": (i, [vp, i, vp, l, C.POINTER(l)]),
        "fdtd_read_dft": (i, [vp, i, vp]),
        "fdtd_read_check": (i, [vp, vp, i, C.POINTER(i)]),
        "fdtd_info": (i, [vp, C.POINTER(l)]),
        "fdtd_setup_build": (i, [vp, C.POINTER(vp)]),
        "fdtd_setup_free": (None, [vp]),
        "fdtd_setup_info": (i, [vp, C.POINTER(l)]),
        "fdtd_setup_dt": (d, [vp]),
        "fdtd_setup_array": (l, [vp, C.c_char_p, vp]),
        "fdtd_setup_source_values": (i, [vp, l, i, vp]),
        "fdtd_setup_probe": (i, [vp, i, C.POINTER(l)]),
        "fdtd_scene_create": (i, [vp, i, l, C.POINTER(vp)]),
        "fdtd_scene_destroy": (i, [vp]),
        "fdtd_scene_step": (i, [vp, i]),
        "fdtd_scene_step_timed": (i, [vp, i, vp]),
        "fdtd_scene_step_timed_every": (i, [vp, i, i, vp]),
        "fdtd_step_timed_every": (i, [vp, l, i, vp, i, vp]),
        "fdtd_step_noalloc": (i, [vp, l, i, vp]),
        "fdtd_scene_step_async": (i, [vp, i]),
        "fdtd_scene_step_wait": (i, [vp]),
        "fdtd_step_async": (i, [vp, l, i, vp]),
        "fdtd_step_wait": (i, [vp]),
        "fdtd_step_query": (i, [vp]),
        "fdtd_scene_step_query": (i, [vp]),
        "fdtd_step_timed": (i, [vp, l, i, vp, vp]),
        "fdtd_set_fp16_scales": (i, [vp, vp]),
        "fdtd_set_kernel_mode": (i, [vp, i, i]),
        "fdtd_set_device_pool": (i, [l]),
        "fdtd_stage_results": (i, [vp]),
        "fdtd_device_pool_bytes": (i, [C.POINTER(l), C.POINTER(l)]),
        "fdtd_get_fp16_scales": (i, [vp, vp]),
        "fdtd_scene_steps_done": (l, [vp]),
        "fdtd_scene_handle": (vp, [vp]),
        "fdtd_batch_set_tuning": (i, [vp, vp]),
        "fdtd_batch_get_tuning": (i, [vp, vp]),
        "fdtd_batch_set_graph": (i, [vp, i]),
        "fdtd_scene_batch_create": (i, [i, C.POINTER(vp)]),
        "fdtd_scene_batch_destroy": (i, [vp]),
        "fdtd_scene_batch_add": (i, [vp, vp]),
        "fdtd_scene_batch_remove": (i, [vp, vp]),
        "fdtd_scene_batch_count": (i, [vp]),
        "fdtd_scene_batch_step": (i, [vp, i, C.POINTER(d)]),
        "fdtd_scene_batch_autotune": (i, [vp, i, vp]),
        "fdtd_scene_batch_handle": (vp, [vp]),
        "fdtd_scene_set_stop": (i, [vp, l]),
        "fdtd_is_stable": (i, [vp]),
        "fdtd_setup_fp16_scales": (i, [vp, vp]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> int:
    if rc < 0:
        raise FdtdError(rc, load().fdtd_last_error().decode(errors="replace"))
    return rc


def set_device_pool(max_cached_bytes: int):
    """Process-wide device memory cache (fdtd_set_device_pool): destroyed
    handles' blocks are kept for reuse, up to that many bytes; 0 turns it off
    and frees what it holds."""
    _check(load().fdtd_set_device_pool(int(max_cached_bytes)))


def device_pool_bytes() -> int:
    """Bytes the device memory cache holds now."""
    v = C.c_long(0)
    _check(load().fdtd_device_pool_bytes(C.byref(v), None))
    return v.value


def device_pool_cap() -> int:
    """The device memory cache's cap (0: off)."""
    v = C.c_long(0)
    _check(load().fdtd_device_pool_bytes(None, C.byref(v)))
    return v.value


class Setup:
    """Host setup of a scene (grid, dt, CPML profiles, bake, TLS, sources,
    probes) computed by the library's C++ restatement of the reference's
    make_context / bake_and_bind / package path."""

    def __init__(self, spec):
        self.lib = load()
        self.spec = spec
        self._c = spec.to_c()
        h = C.c_void_p()
        rc = self.lib.fdtd_setup_build(C.addressof(self._c), C.byref(h))
        if rc != 0:
            raise FdtdError(rc, "setup build failed (invalid spec)")
        self.h = h
        info = (C.c_long * 10)()
        self.lib.fdtd_setup_info(self.h, info)
        self.shape = (info[0], info[1], info[2])
        self.npml = info[3]
        self.n_sources = info[4]
        self.n_channels = info[5]
        self.n_dfts = info[6]
        self.tls_on = bool(info[7])
        self.n_regions = info[8]
        self.bc_type = info[9]
        self.dt = self.lib.fdtd_setup_dt(self.h)

    def array(self, name: str, shaped: bool = True) -> np.ndarray:
        n = self.lib.fdtd_setup_array(self.h, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        out = np.empty(n, dtype=np.float64)
        if n:
            self.lib.fdtd_setup_array(self.h, name.encode(), out.ctypes.data_as(C.c_void_p))
        N = self.shape[0] * self.shape[1] * self.shape[2]
        if shaped and n == N:
            return out.reshape(self.shape)
        return out

    def source_values(self, n0: int, nsteps: int) -> np.ndarray:
        out = np.empty((nsteps, max(self.n_sources, 1)), dtype=np.float64)
        if self.n_sources:
            _check(self.lib.fdtd_setup_source_values(self.h, n0, nsteps, out.ctypes.data_as(C.c_void_p)))
        return out[:, :self.n_sources]

    def probe(self, c: int):
        v = (C.c_long * 4)()
        _check(self.lib.fdtd_setup_probe(self.h, c, v))
        return tuple(v[:3]), v[3]

    def fp16_scales(self):
        """The 18 initial FP16 storage scales a scene of this setup starts
        from, or None when the library keeps its defaults (no source)."""
        out = np.empty(18, dtype=np.float64)
        return out if _check(self.lib.fdtd_setup_fp16_scales(self.h, out.ctypes.data_as(C.c_void_p))) == 1 else None

    def close(self):
        if getattr(self, "h", None):
            self.lib.fdtd_setup_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """One device simulation of a Setup at a given precision."""

    def __init__(self, setup: Setup, precision="fp64", probe_capacity: Optional[int] = None):
        self.lib = load()
        self.setup = setup
        self.precision = PRECISION[precision] if isinstance(precision, str) else int(precision)
        cap = probe_capacity if probe_capacity is not None else max(16, setup.spec.n_steps + 1)
        h = C.c_void_p()
        _check(self.lib.fdtd_scene_create(setup.h, self.precision, cap, C.byref(h)))
        self.sc = h
        self.h = self.lib.fdtd_scene_handle(self.sc)
        self.shape = setup.shape

    def close(self):
        if getattr(self, "sc", None):
            self.lib.fdtd_scene_destroy(self.sc)
            self.sc = None
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stop(self, n_stop: int):
        """Batch steps never take this scene past step n_stop (-1: no limit)."""
        _check(self.lib.fdtd_scene_set_stop(self.sc, n_stop))

    def steps_done(self) -> int:
        return self.lib.fdtd_scene_steps_done(self.sc)

    def step(self, nsteps: int, check: bool = True) -> int:
        rc = self.lib.fdtd_scene_step(self.sc, nsteps)
        if check:
            _check(rc)
        return rc

    def step_timed(self, nsteps: int, every: int = 1):
        """Steps and returns (H-kernel ms, E-kernel ms, call ms, steps timed)
        from HIP events around the kernels of every `every`-th step (the
        kernel times are sums over the steps timed)."""
        ms = (C.c_double * 4)()
        _check(self.lib.fdtd_scene_step_timed_every(self.sc, nsteps, every, ms))
        return tuple(ms)

    def step_async(self, nsteps: int):
        """Enqueues nsteps steps and returns at once (see step_wait)."""
        _check(self.lib.fdtd_scene_step_async(self.sc, nsteps))

    def step_wait(self):
        """Waits for the enqueued steps and applies the stability check."""
        _check(self.lib.fdtd_scene_step_wait(self.sc))

    def step_ready(self) -> bool:
        """True when the enqueued steps have all run (never blocks)."""
        r = self.lib.fdtd_scene_step_query(self.sc)
        _check(min(r, 0))
        return r == 1

    def set_step_graphs(self, on: bool = True):
        """Record each stepping call's launches in a HIP graph replayed for
        later calls of the same shape (fdtd_set_step_graphs); bitwise the
        direct launches' results."""
        _check(self.lib.fdtd_set_step_graphs(self.h, int(on)))

    def set_kernel_mode(self, mode: int, param: int = 0):
        """0: separate H/E kernels, `param` = wavefront k-width (16, 32, 64;
        0 keeps it); 1: fused step kernel, `param` = x-planes per work-group;
        2: dependency-ordered step, `param` = its E tiles' plane lag (0: 2)."""
        _check(self.lib.fdtd_set_kernel_mode(self.h, mode, param))

    def synchronize(self):
        _check(self.lib.fdtd_synchronize(self.h))

    def field(self, name: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        _check(self.lib.fdtd_download(self.h, FIELD_ID[name], out.ctypes.data_as(C.c_void_p)))
        return out

    def upload(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        _check(self.lib.fdtd_upload(self.h, FIELD_ID[name], a.ctypes.data_as(C.c_void_p)))

    def probe(self, channel: int) -> np.ndarray:
        cnt = C.c_long()
        _check(self.lib.fdtd_read_probe(self.h, channel, None, 0, C.byref(cnt)))
        out = np.empty(cnt.value, dtype=np.float64)
        if cnt.value:
            _check(self.lib.fdtd_read_probe(self.h, channel, out.ctypes.data_as(C.c_void_p), cnt.value, C.byref(cnt)))
        return out

    def add_tls_monitor(self, i: int, j: int, k: int, save_every: int = 1, capacity: int = 100000):
        """Records the TLS diagnostics (total Nu, total Ng, interaction energy,
        Pz and dPz/dt at cell (i, j, k)) every save_every steps on the device."""
        self._dxyz = [np.ascontiguousarray(self.setup.array(n, shaped=False)) for n in ("dx", "dy", "dz")]
        p = [a.ctypes.data_as(C.c_void_p) for a in self._dxyz]
        _check(self.lib.fdtd_add_tls_monitor(self.h, p[0], p[1], p[2], i, j, k, save_every, capacity))

    def tls_records(self) -> np.ndarray:
        """(records, 5): total Nu, total Ng, U_int, Pz and dPz/dt at the cell."""
        cnt = C.c_long()
        _check(self.lib.fdtd_read_tls_monitor(self.h, None, 0, C.byref(cnt)))
        out = np.empty((cnt.value, 5), dtype=np.float64)
        if cnt.value:
            _check(self.lib.fdtd_read_tls_monitor(self.h, out.ctypes.data_as(C.c_void_p), cnt.value, C.byref(cnt)))
        return out

    def stage_results(self):
        """Queue a host copy of the probe records and DFT accumulators after
        the steps enqueued so far; probe() and dft() read it until the next
        step (fdtd_stage_results)."""
        _check(self.lib.fdtd_stage_results(self.h))

    def dft(self, m: int) -> np.ndarray:
        d = self.setup.spec.dfts[m]
        i0, i1, j0, j1, k0, k1 = d.box
        n = (i1 - i0) * (j1 - j0) * (k1 - k0) * len(d.freqs) * 2
        out = np.empty(n, dtype=np.float64)
        _check(self.lib.fdtd_read_dft(self.h, m, out.ctypes.data_as(C.c_void_p)))
        return out

    def stable(self) -> bool:
        """False once the stability policy has stopped this simulation."""
        return _check(self.lib.fdtd_is_stable(self.h)) == 1

    def checks(self):
        n = C.c_int()
        buf = (CheckRecord * 4096)()
        _check(self.lib.fdtd_read_check(self.h, buf, 4096, C.byref(n)))
        return [buf[q] for q in range(min(n.value, 4096))]

    def info(self):
        v = (C.c_long * 8)()
        _check(self.lib.fdtd_info(self.h, v))
        return {"shape": tuple(v[:3]), "precision": v[3], "n_materials": v[4], "mat_id_bytes": v[5],
                "gain_cells": v[6], "npml": v[7]}

    def device_bytes(self) -> int:
        b = C.c_size_t()
        _check(self.lib.fdtd_device_bytes(self.h, C.byref(b)))
        return b.value

    def material(self, name: str) -> np.ndarray:
        """MaterialGrids array `name` (aEx ... bHz) as the device holds it."""
        q = MATERIAL_NAMES.index(name)
        out = np.empty(self.shape, dtype=np.float64)
        _check(self.lib.fdtd_download_material(self.h, q, out.ctypes.data_as(C.c_void_p)))
        return out


class Batch:
    """Scenes of one precision stepped together: one persistent H launch and
    one persistent E launch per step for all members (fdtd_scene_batch_*)."""

    def __init__(self, precision="fp32"):
        self.lib = load()
        self.precision = PRECISION[precision] if isinstance(precision, str) else int(precision)
        h = C.c_void_p()
        _check(self.lib.fdtd_scene_batch_create(self.precision, C.byref(h)))
        self.sb = h
        self.b = self.lib.fdtd_scene_batch_handle(self.sb)
        self.members = []

    def add(self, scene: "Scene") -> int:
        slot = _check(self.lib.fdtd_scene_batch_add(self.sb, scene.sc))
        self.members.append(scene)
        return slot

    def remove(self, scene: "Scene"):
        _check(self.lib.fdtd_scene_batch_remove(self.sb, scene.sc))
        self.members.remove(scene)

    def __len__(self):
        return len(self.members)

    def step(self, nsteps: int) -> float:
        """Advances every member nsteps steps; returns the device milliseconds."""
        ms = C.c_double()
        _check(self.lib.fdtd_scene_batch_step(self.sb, nsteps, C.byref(ms)))
        return ms.value

    def set_graph(self, on: bool = True):
        _check(self.lib.fdtd_batch_set_graph(self.b, int(on)))

    def set_tuning(self, rows: int, wg_per_cu: int, chunk: int, planes: int = 1):
        t = (C.c_int * 4)(rows, wg_per_cu, chunk, planes)
        _check(self.lib.fdtd_batch_set_tuning(self.b, t))

    def tuning(self):
        """(rows, wg_per_cu, chunk, planes) of the current tiling."""
        t = (C.c_int * 4)()
        _check(self.lib.fdtd_batch_get_tuning(self.b, t))
        return tuple(t)

    def autotune(self, trial_steps: int = 5):
        """Times every candidate tiling on real steps and keeps the fastest;
        returns ((rows, wg_per_cu, chunk, planes), steps taken)."""
        t = (C.c_int * 4)()
        taken = _check(self.lib.fdtd_scene_batch_autotune(self.sb, trial_steps, t))
        return tuple(t), taken

    def close(self):
        if getattr(self, "sb", None):
            self.lib.fdtd_scene_batch_destroy(self.sb)
            self.sb = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
