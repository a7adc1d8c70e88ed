"""The stability monitor's verdicts on the device against the reference's
StabilityMonitor::check_stability (src/stability_monitor.hpp:42-144), run by
the compiled reference harness (oracle/_ref/libref.so) on the same inputs:

- a NaN planted in Ez (plain data through fdtd_upload) fails the run at the
  first monitor step after it, in every storage format, as the reference's
  NaN/Inf scan does;
- a Courant number above 1 grows the fields until the E-threshold warnings
  pass 10: the FP64 device run stops at the reference's step with the
  reference's warning count and maximum; FP32 and FP16 stop no later;
- at every monitor step of the golden cases the device's FP64 max |E| is
  the reference's (bitwise), and the verdict is "stable" as in the fixture.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _first_failure(sc):
    """Step of the first record whose NaN/Inf flags are set, else None."""
    for r in sc.checks():
        if r.nan_E or r.nan_H or r.nan_P or r.nan_N:
            return r.step
    return None


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp16"])
def test_nan_in_field_fails_at_reference_step(pkg, oracle_mod, precision):
    import oracle

    if not oracle.available("ref"):
        pytest.skip("oracle/_ref/libref.so not built")
    sp = pkg.scene.uniform_sphere_case(core=16, npml=6, steps=60)
    _say(f"[case] nan upload {precision}")
    setup = pkg.Setup(sp)
    sc = pkg.Scene(setup, precision)
    ref = oracle.CpuSim(sp, "ref")
    sc.step(15)
    assert ref.step(15) == 0
    ez = ref.array("Ez").copy()
    ez[14, 13, 12] = np.nan
    ref.set_array("Ez", ez)
    sc.upload("Ez", ez)
    rc = sc.step(20, check=False)
    assert rc == pkg.runtime.FDTD_EUNSTABLE
    assert not sc.stable()
    assert ref.step(20) == 1
    st = ref.stability()
    assert not st["stable"]
    assert _first_failure(sc) == st["fail_step"] == 20
    # a later call still reports the failed verdict of the handle
    assert not sc.stable()
    sc.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp16"])
def test_inf_in_field_fails_at_reference_step(pkg, oracle_mod, precision):
    import oracle

    if not oracle.available("ref"):
        pytest.skip("oracle/_ref/libref.so not built")
    sp = pkg.scene.gain_slab_case(n_total=40, npml=6, steps=60, fine_len=100e-9, coarse=20e-9, fine=5e-9)
    _say(f"[case] inf upload {precision}")
    setup = pkg.Setup(sp)
    sc = pkg.Scene(setup, precision)
    ref = oracle.CpuSim(sp, "ref")
    sc.step(21)
    assert ref.step(21) == 0
    hy = ref.array("Hy").copy()
    hy[5, 20, 7] = -np.inf
    ref.set_array("Hy", hy)
    sc.upload("Hy", hy)
    assert sc.step(10, check=False) == pkg.runtime.FDTD_EUNSTABLE
    assert ref.step(10) == 1
    assert _first_failure(sc) == ref.stability()["fail_step"] == 30
    sc.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp16"])
def test_courant_above_one_stops_like_reference(pkg, oracle_mod, precision):
    import oracle

    if not oracle.available("ref"):
        pytest.skip("oracle/_ref/libref.so not built")
    sp = pkg.scene.uniform_sphere_case(core=20, npml=6, steps=300).copy(S=1.2)
    _say(f"[case] S=1.2 {precision}")
    ref = oracle.CpuSim(sp, "ref")
    assert ref.step(300) == 1
    st = ref.stability()
    setup = pkg.Setup(sp)
    sc = pkg.Scene(setup, precision)
    rc = sc.step(300, check=False)
    assert rc == pkg.runtime.FDTD_EUNSTABLE
    recs = sc.checks()
    # the verdict: the first monitor step at which the reference's policy fails
    warn, fail = 0, None
    for r in recs:
        if r.nan_E or r.nan_H:
            fail = r.step
            break
        warn += r.max_E > 1e10
        if warn > 10:
            fail = r.step
            break
    _say(f"  reference fails at {st['fail_step']} ({st['warnings']} warnings); device at {fail}")
    assert fail is not None
    if precision == "fp64":
        assert fail == st["fail_step"] == 140 and st["warnings"] == 11
        last = [r for r in recs if r.step == fail][0]
        assert last.max_E == st["max_E"]
    else:
        assert fail <= st["fail_step"]
    sc.close()


CASES = ["uniform_sphere", "gain_slab", "gain_clamp_dipole", "plane_waves", "pec_cavity", "graded_cuboids",
         "auto_mesh_cylinder"]


@pytest.mark.parametrize("name", CASES)
def test_monitor_max_E_matches_golden(pkg, name):
    """The last monitor step's max |E| of the FP64 device run equals the one
    the compiled reference recorded in the fixture, and both are stable."""
    sys.path.insert(0, GOLDEN)
    import cases

    sp = cases.CASES[name]()
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    steps = int(g["steps"])
    _say(f"[golden monitor] {name} steps={steps}")
    setup = pkg.Setup(sp)
    sc = pkg.Scene(setup, "fp64")
    assert sc.step(steps, check=False) == 0
    assert bool(g["stable"]) and sc.stable()
    recs = sc.checks()
    assert [r.step for r in recs] == list(range(0, steps, 10))
    assert recs[-1].max_E == float(g["max_E"])
    assert not any(r.nan_E or r.nan_H or r.nan_P or r.nan_N for r in recs)
    sc.close()
