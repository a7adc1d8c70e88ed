"""The precision form at its configuration: the 512x512x384 graded grid with
the gain slab, substrate and 64 seeded dielectric cuboids (about 100M cells),
stepped for the configuration's 4000 steps in FP64, FP32 and FP16 storage.

- FP32 and FP16 against the FP64 device path (the oracle at this size: the
  FP64 path is bitwise the compiled reference on the golden cases and within
  1e-12 of the CPU oracle at full size elsewhere), per probe vector and per
  DFT spectrum, at the bars written below;
- the kernel schedules agree bit for bit at this size (FP32 and FP16);
- the same structure on a reduced grid (96x96x80, 64 cuboids) against the
  CPU oracle: FP64 within 1e-12, FP32 within 2e-4.
"""
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
STEPS = 4000
# relative L2 against the FP64 device path after 4000 steps
FP32_PROBE, FP32_SPECTRUM = 1e-5, 1e-5
FP16_PROBE, FP16_SPECTRUM = 2e-2, 1e-3


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _probe_vectors(sc, spec):
    out, c = [], 0
    for pr in spec.probes:
        n = len(pr.components)
        out.append(np.stack([sc.probe(c + q) for q in range(n)]))
        c += n
    return out


@pytest.fixture(scope="module")
def c3_runs(pkg):
    """Probe vectors and DFT spectra of the three storage formats."""
    spec = pkg.scene.precision_case(steps=STEPS)
    setup = pkg.Setup(spec)
    res = {}
    for prec in ("fp64", "fp32", "fp16"):
        sc = pkg.Scene(setup, prec, probe_capacity=STEPS + 16)
        done = 0
        while done < STEPS:
            k = min(1000, STEPS - done)
            sc.step(k)
            done += k
            _say(f"[c3] {prec} {done}/{STEPS}")
        res[prec] = (_probe_vectors(sc, spec), [sc.dft(m) for m in range(len(spec.dfts))])
        sc.close()
    return spec, res


@pytest.mark.parametrize("prec,bar_probe,bar_spec", [("fp32", FP32_PROBE, FP32_SPECTRUM),
                                                     ("fp16", FP16_PROBE, FP16_SPECTRUM)])
def test_c3_error_against_fp64(c3_runs, prec, bar_probe, bar_spec):
    spec, res = c3_runs
    ref_p, ref_d = res["fp64"]
    got_p, got_d = res[prec]
    errs = {f"probe{p}": rel_l2(a, b) for p, (a, b) in enumerate(zip(got_p, ref_p))}
    errs.update({f"dft{m}": rel_l2(a, b) for m, (a, b) in enumerate(zip(got_d, ref_d))})
    _say(f"  {prec} vs fp64 after {STEPS} steps: " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()))
    for p, a in enumerate(ref_p):
        assert np.abs(a).max() > 0, f"probe {p} saw no field"
    bad = {k: v for k, v in errs.items() if not v <= (bar_probe if k.startswith("probe") else bar_spec)}
    assert not bad, bad


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_c3_schedules_agree_bitwise(pkg, precision):
    """At full size, every schedule (two-kernel step at several wavefront
    shapes and tile orders, the fused step kernel and the dependency-ordered
    step) gives the same bits as the default."""
    spec = pkg.scene.precision_case(steps=30)
    setup = pkg.Setup(spec)
    modes = [(0, 32), (0, 64), (0, 0x100000 + 32), (1, 0), (2, 0)]
    scenes = []
    for mode in modes:
        _say(f"[c3 schedules] {precision} mode={mode}")
        sc = pkg.Scene(setup, precision)
        sc.set_kernel_mode(*mode)
        sc.step(30)
        scenes.append(sc)
    for name in FIELDS + ("Nu", "Pz"):
        base = scenes[0].field(name)
        if name == "Ex":
            assert np.abs(base).max() > 0
        for mode, sc in zip(modes[1:], scenes[1:]):
            assert np.array_equal(base, sc.field(name)), f"mode {mode}: {name} differs"
        del base
    for sc in scenes:
        sc.close()


def test_c3_structure_reduced_against_oracle(pkg, oracle_mod):
    """The precision geometry (gain slab, substrate, 64 cuboids) on a reduced
    96x96x80 grid: FP64 within 1e-12 and FP32 within 2e-4 of the CPU oracle."""
    spec = pkg.scene.precision_case(steps=300, shape=(96, 96, 80))
    _say(f"[c3 reduced] {spec.name} cells={spec.n_cells()}")
    ref = oracle_mod.CpuSim(spec, "oracle")
    ref.step(spec.n_steps)
    setup = pkg.Setup(spec)
    for prec, tol in (("fp64", 1e-12), ("fp32", 2e-4)):
        sc = pkg.Scene(setup, prec)
        sc.step(spec.n_steps)
        errs = {}
        if prec == "fp64":
            for f in FIELDS + ("Nu", "Ng", "Pz", "Pz_prev"):
                errs[f] = rel_l2(sc.field(f), ref.array(f))
        else:
            for grp, names in (("E", FIELDS[:3]), ("H", FIELDS[3:])):
                errs[grp] = rel_l2(np.stack([sc.field(n) for n in names]), np.stack([ref.array(n) for n in names]))
        c = 0
        for p, pr in enumerate(spec.probes):
            n = len(pr.components)
            a = np.stack([sc.probe(c + q) for q in range(n)])
            b = np.stack([ref.probe(p, q) for q in range(n)])
            errs[f"probe{p}"] = rel_l2(a, b)
            c += n
        errs["dft0"] = rel_l2(sc.dft(0), ref.dft(0))
        _say(f"  {prec}: " + " ".join(f"{k}={v:.1e}" for k, v in errs.items()))
        bad = {k: v for k, v in errs.items() if not v <= tol}
        assert not bad, bad
        sc.close()
