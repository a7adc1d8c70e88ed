"""fdtd_step_noalloc (the step that allocates no device memory) against
fdtd_step, through the C ABI: two handles built from one setup advance the
same steps with the same source values, one through each entry point, in two
calls in a row (the second shorter), and must agree bitwise in every field,
gain array, probe trace and monitor record.  The no-alloc handle's device
footprint does not change; a call whose scratch need exceeds the maximum is
refused without stepping."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
MAX_BYTES = 1048576  # FDTD_NOALLOC_SCRATCH_BYTES


def _say(msg):
    print(msg, flush=True)


def _step(lib, fn, sc, n0, n, vals):
    v = np.ascontiguousarray(vals, dtype=np.float64)
    return fn(sc.h, n0, n, v.ctypes.data_as(C.c_void_p) if v.size else None)


def _records(sc):
    return [(r.step, r.max_E, r.max_H, r.max_P, r.nan_E, r.nan_H, r.pop_total) for r in sc.checks()]


def _extras(sc, lib):
    out = []
    for which in range(6, 10):
        a = np.empty(sc.shape, dtype=np.float64)
        assert lib.fdtd_download(sc.h, which, a.ctypes.data_as(C.c_void_p)) == 0
        out.append(a)
    return out


CASES = [
    ("gain_slab_40_fp64", lambda pkg: pkg.scene.gain_slab_case(n_total=40, npml=6, steps=60, fine_len=100e-9,
                                                                coarse=20e-9, fine=5e-9), "fp64"),
    ("gain_slab_40_fp16", lambda pkg: pkg.scene.gain_slab_case(n_total=40, npml=6, steps=60, fine_len=100e-9,
                                                                coarse=20e-9, fine=5e-9), "fp16"),
    ("uniform_sphere_fp32", lambda pkg: pkg.scene.uniform_sphere_case(core=24, npml=6, steps=60), "fp32"),
]


@pytest.mark.parametrize("name,make,precision", CASES, ids=[c[0] for c in CASES])
def test_step_noalloc_matches_step(pkg, name, make, precision):
    lib = pkg.load()
    sp = make(pkg)
    setup = pkg.Setup(sp)
    calls = [(0, 37), (37, 13)]  # second call smaller
    _say(f"[noalloc] {name} {precision} shape={setup.shape} sources={setup.n_sources} calls={calls}")
    a = pkg.Scene(setup, precision)
    b = pkg.Scene(setup, precision)
    for n0, n in calls:
        vals = setup.source_values(n0, n)
        assert _step(lib, lib.fdtd_step, a, n0, n, vals) == 0
        before = b.device_bytes()
        rc = _step(lib, lib.fdtd_step_noalloc, b, n0, n, vals)
        assert rc == 0, pkg.runtime.load().fdtd_last_error()
        assert b.device_bytes() == before
        assert _records(a) == _records(b)
        _say(f"[noalloc] {name} steps {n0}..{n0 + n}: {len(_records(b))} monitor records equal")
    for f in FIELDS:
        assert np.array_equal(a.field(f), b.field(f)), f
    for q, (x, y) in enumerate(zip(_extras(a, lib), _extras(b, lib))):
        assert np.array_equal(x, y), q
    nch = sum(len(p.components) for p in sp.probes)
    for c in range(nch):
        assert np.array_equal(a.probe(c), b.probe(c)), c
    assert a.stable() and b.stable()
    _say(f"[noalloc] {name}: fields, gain arrays and {nch} probe channels bitwise equal")


def test_step_noalloc_refuses_oversized_call(pkg):
    lib = pkg.load()
    sp = pkg.scene.uniform_sphere_case(core=16, npml=4, steps=10)
    setup = pkg.Setup(sp)
    assert setup.n_sources >= 1
    n = MAX_BYTES // (8 * setup.n_sources) + 1  # source values alone exceed the maximum
    _say(f"[noalloc] oversized call: {n} steps x {setup.n_sources} sources")
    sc = pkg.Scene(setup, "fp64")
    vals = np.zeros((n, setup.n_sources))
    rc = _step(lib, lib.fdtd_step_noalloc, sc, 0, n, vals)
    assert rc == -1, rc
    assert b"scratch" in lib.fdtd_last_error()
    for f in FIELDS:
        assert not sc.field(f).any(), f
    # the handle still steps afterwards
    vals = setup.source_values(0, 5)
    assert _step(lib, lib.fdtd_step_noalloc, sc, 0, 5, vals) == 0
