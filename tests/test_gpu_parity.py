"""GPU parity: the HIP path through the C ABI against the FP64 CPU oracle on
the same seeded inputs.

Tolerances: FP64 device path <= 1e-12 relative L2 on every probe trace and
final field (the reference's own FMA-contraction noise is ~1e-14 after 1000
steps); FP32 storage/arithmetic <= 2e-4 relative L2 on fields and probes.
"""
import math
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


_ORACLE_CACHE = {}


def _run_pair(pkg, oracle_mod, sp, precision="fp64", steps=None, mode=None):
    steps = sp.n_steps if steps is None else steps
    _say(f"[case] {sp.name} {precision} mode={mode} steps={steps} cells={sp.n_cells()}")
    setup = pkg.Setup(sp)
    sc = pkg.Scene(setup, precision)
    if mode is not None:
        sc.set_kernel_mode(*mode)
    sc.step(steps)
    key = (sp.name, steps)
    ref = _ORACLE_CACHE.get(key)
    if ref is None:
        ref = oracle_mod.CpuSim(sp, "oracle")
        ref.step(steps)
        _ORACLE_CACHE.clear()
        _ORACLE_CACHE[key] = ref
    return setup, sc, ref


def _channels(sp):
    """Probe channel c of the device path -> (probe, component index) of the oracle."""
    out = []
    for p, pr in enumerate(sp.probes):
        for q in range(len(pr.components)):
            out.append((p, q))
    return out


def _compare(setup, sc, ref, tol, extra=(), vector=False):
    """Relative L2 per field (vector=True: of the stacked E and H vector fields,
    so a component that is pure round-off noise by symmetry is judged against
    the field it belongs to) and per probe channel."""
    worst = {}
    if vector:
        for grp, names in (("E", FIELDS[:3]), ("H", FIELDS[3:])):
            a = np.stack([sc.field(n) for n in names])
            b = np.stack([ref.array(n) for n in names])
            worst[grp] = rel_l2(a, b)
        names = tuple(extra)
    else:
        names = FIELDS + tuple(extra)
    for name in names:
        a, b = sc.field(name), ref.array(name)
        worst[name] = rel_l2(a, b)
    for c, (p, q) in enumerate(_channels(setup.spec)):
        worst[f"probe{c}"] = rel_l2(sc.probe(c), ref.probe(p, q))
    _say("  " + " ".join(f"{k}={v:.2e}" for k, v in worst.items()))
    bad = {k: v for k, v in worst.items() if not v <= tol}
    assert not bad, f"exceeds {tol}: {bad}"


# (0, w): H and E kernels with wavefronts of w consecutive k; (1, c): fused
# step kernel with c x-planes per work-group
MODES = [(0, 16), (0, 32), (0, 64), (1, 16), (1, 7), (1, 128)]
MODES2 = [(0, 16), (1, 16)]


@pytest.mark.parametrize("mode", MODES, ids=lambda m: f"mode{m[0]}_c{m[1]}")
def test_uniform_sphere_fp64_full(pkg, oracle_mod, mode):
    """The plain CPU case at full size: 64^3 core + 10-cell CPML, 1000 steps,
    for the two-kernel step and the fused step at several tile shapes."""
    sp = pkg.scene.uniform_sphere_case(core=64, npml=10, steps=1000)
    setup, sc, ref = _run_pair(pkg, oracle_mod, sp, "fp64", mode=mode)
    _compare(setup, sc, ref, 1e-12)
    # probe record count = steps (save_every 1, n = 0 included)
    assert len(sc.probe(0)) == sp.n_steps


@pytest.mark.parametrize("mode", MODES2, ids=lambda m: f"mode{m[0]}")
def test_psi_state_fp64(pkg, oracle_mod, mode):
    sp = pkg.scene.uniform_sphere_case(core=16, npml=6, steps=150)
    setup, sc, ref = _run_pair(pkg, oracle_mod, sp, "fp64", mode=mode)
    for name in pkg.runtime.PSI_NAMES:
        assert rel_l2(sc.field(name), ref.array(name)) <= 1e-12, name


@pytest.mark.parametrize("mode", MODES, ids=lambda m: f"mode{m[0]}_c{m[1]}")
def test_gain_slab_fp64(pkg, oracle_mod, mode):
    sp = pkg.scene.gain_slab_case(n_total=40, npml=6, steps=400, fine_len=100e-9, coarse=20e-9, fine=5e-9)
    setup, sc, ref = _run_pair(pkg, oracle_mod, sp, "fp64", mode=mode)
    _compare(setup, sc, ref, 1e-12, extra=("Nu", "Ng", "Pz", "Pz_prev"))


def test_gain_cuboids_nonuniform_fp64(pkg, oracle_mod):
    sp = pkg.scene.gain_slab_case(n_total=48, npml=6, steps=300, fine_len=100e-9, coarse=20e-9, fine=5e-9,
                                  n_cuboids=6, seed=3)
    setup, sc, ref = _run_pair(pkg, oracle_mod, sp, "fp64")
    _compare(setup, sc, ref, 1e-12, extra=("Nu", "Ng", "Pz"))


@pytest.mark.parametrize("mode", MODES2, ids=lambda m: f"mode{m[0]}")
def test_fp32_uniform(pkg, oracle_mod, mode):
    sp = pkg.scene.uniform_sphere_case(core=32, npml=8, steps=400)
    setup, sc, ref = _run_pair(pkg, oracle_mod, sp, "fp32", mode=mode)
    _compare(setup, sc, ref, 2e-4, vector=True)


@pytest.mark.parametrize("mode", MODES2, ids=lambda m: f"mode{m[0]}")
def test_fp32_gain(pkg, oracle_mod, mode):
    sp = pkg.scene.gain_slab_case(n_total=40, npml=6, steps=400, fine_len=100e-9, coarse=20e-9, fine=5e-9)
    setup, sc, ref = _run_pair(pkg, oracle_mod, sp, "fp32", mode=mode)
    _compare(setup, sc, ref, 2e-4, extra=("Pz",), vector=True)
    assert rel_l2(sc.field("Nu"), ref.array("Nu")) <= 1e-9


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_step_kernels_agree_bitwise(pkg, precision):
    """The two-kernel step (at every wavefront k-width, with 32-bit and 64-bit
    element offsets: 0x20000 + width selects the latter; with the E kernel
    walking its tiles backwards, the default, or forwards: 0x100000) and the
    fused step kernel run the same operations per value, so their results
    are identical in every storage format; so does the dependency-ordered
    step (mode 2, at E-tile lags 2, 1 and 3; 44 planes: queues of 6 and 5)."""
    sp = pkg.scene.gain_slab_case(n_total=44, npml=6, steps=120, fine_len=100e-9, coarse=20e-9, fine=5e-9,
                                  n_cuboids=4, seed=5)
    setup = pkg.Setup(sp)
    runs = {}
    for mode in [(0, 16), (0, 32), (0, 64), (0, 0x20000 + 32), (0, 0x20000 + 16 + 256 * 64), (1, 16), (1, 5),
                 (0, 0x100000 + 32), (0, 0x100000 + 0x20000 + 64), (1, 4096 * 1 + 65536 * 2 + 3),
                 (1, 4096 * 3 + 65536 * 1 + 9), (1, 4096 * 4 + 7), (2, 0), (2, 1), (2, 3)]:
        _say(f"[case] {sp.name} {precision} mode={mode}")
        sc = pkg.Scene(setup, precision)
        sc.set_kernel_mode(*mode)
        sc.step(sp.n_steps)
        runs[mode] = {n: sc.field(n) for n in FIELDS + ("Nu", "Pz") + tuple(pkg.runtime.PSI_NAMES)}
        sc.close()
    base = runs[(0, 16)]
    for mode, got in runs.items():
        diff = [n for n in base if not np.array_equal(base[n], got[n])]
        assert not diff, f"mode {mode} differs in {diff}"


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp16"])
def test_step_kernels_agree_bitwise_ragged(pkg, precision):
    """A grid whose rows span several 64-cell k-segments, the last one
    partial (NzT = 150), with a partial last group of rows (NyT = 30) and a
    partial last chunk of planes: the fused kernel carries state across
    segments and planes, and agrees with the two-kernel step (every
    wavefront k-width, whose wavefronts straddle the row end) bit for bit."""
    sp = pkg.scene.gain_slab_case(shape=(20, 30, 150), npml=6, steps=60, fine_len=300e-9, coarse=20e-9, fine=5e-9,
                                  n_cuboids=3, seed=9)
    setup = pkg.Setup(sp)
    runs = {}
    # (1, planes + 4096 * cells_per_lane + 65536 * rows): the fused kernel
    # with 1..4 cells per lane, i.e. 3, 2 or 1 k-chunk wavefronts per row
    # (dead lanes past NzT in the last chunk), and 1..4 rows per block (7 is
    # taken as 4)
    for mode in [(0, 32), (0, 16), (0, 64), (0, 0x100000 + 32), (1, 3), (1, 20), (1, 4096 * 1 + 65536 * 2 + 4),
                 (1, 4096 * 2 + 65536 * 7 + 6), (1, 4096 * 3 + 65536 * 1 + 20), (1, 4096 * 4 + 65536 * 3 + 1),
                 (2, 0), (2, 1), (2, 5)]:
        _say(f"[case] ragged {precision} mode={mode}")
        sc = pkg.Scene(setup, precision)
        sc.set_kernel_mode(*mode)
        sc.step(sp.n_steps)
        runs[mode] = {n: sc.field(n) for n in FIELDS + ("Nu", "Pz") + tuple(pkg.runtime.PSI_NAMES)}
        sc.close()
    base = runs[(0, 32)]
    assert np.abs(base["Ex"]).max() > 0
    for mode, got in runs.items():
        diff = [n for n in base if not np.array_equal(base[n], got[n])]
        assert not diff, f"mode {mode} differs in {diff}"


@pytest.mark.parametrize("shape", [(6, 24, 40), (9, 20, 70), (17, 33, 65)], ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp16"])
def test_flow_step_small_and_uneven_queues(pkg, precision, shape):
    """The dependency-ordered step (mode 2) splits the planes into eight
    queues: fewer planes than queues (6: every queue one plane, all E tiles in
    the edge launch), a remainder of one (9) and one queue short (17), each
    with a partial last band of rows and a partial last k-tile; bitwise the
    two-kernel step, over calls that end between monitor steps."""
    sp = pkg.scene.gain_slab_case(shape=shape, npml=2, steps=45, fine_len=60e-9, coarse=20e-9, fine=5e-9,
                                  n_cuboids=2, seed=11)
    setup = pkg.Setup(sp)
    runs = {}
    for mode in [(0, 32), (2, 0), (2, 2)]:
        _say(f"[case] flow {shape} {precision} mode={mode}")
        sc = pkg.Scene(setup, precision)
        sc.set_kernel_mode(*mode)
        for n in (20, 13, 12):
            sc.step(n)
        runs[mode] = {n: sc.field(n) for n in FIELDS + ("Nu", "Pz") + tuple(pkg.runtime.PSI_NAMES)}
        runs[mode]["checks"] = [(r.step, r.max_E, r.max_H) for r in sc.checks()]
        sc.close()
    base = runs[(0, 32)]
    assert np.abs(base["Ex"]).max() > 0
    for mode, got in runs.items():
        diff = [n for n in base if not (np.array_equal(base[n], got[n]) if n != "checks" else base[n] == got[n])]
        assert not diff, f"mode {mode} differs in {diff}"


GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


@pytest.mark.parametrize("mode", [(0, 16), (0, 64), (1, 8), (2, 0)], ids=lambda m: f"mode{m[0]}_c{m[1]}")
@pytest.mark.parametrize("name", ["uniform_sphere", "gain_slab", "gain_clamp_dipole", "plane_waves", "pec_cavity",
                                  "graded_cuboids", "auto_mesh_cylinder"])
def test_fp64_matches_reference_golden(pkg, name, mode):
    """The device FP64 path against the compiled reference's own outputs."""
    sys.path.insert(0, GOLDEN)
    import cases

    sp = cases.CASES[name]()
    g = np.load(__import__("os").path.join(GOLDEN, f"{name}.npz"))
    _say(f"[golden] {name} steps={int(g['steps'])}")
    setup = pkg.Setup(sp)
    sc = pkg.Scene(setup, "fp64")
    sc.set_kernel_mode(*mode)
    sc.step(int(g["steps"]))
    worst = {}
    for f in cases.FIELDS:
        worst[f] = rel_l2(sc.field(f), g[f])
    if "Nu" in g.files:
        for f in cases.TLS_FIELDS:
            worst[f] = rel_l2(sc.field(f), g[f])
    for c in range(setup.n_channels):
        worst[f"probe{c}"] = rel_l2(sc.probe(c), g[f"probe{c}"])
    bitwise = all(np.array_equal(sc.field(f), g[f]) for f in cases.FIELDS)
    _say("  bitwise fields: %s; " % bitwise + " ".join(f"{k}={v:.1e}" for k, v in worst.items()))
    bad = {k: v for k, v in worst.items() if not v <= 1e-12}
    assert not bad, bad


def test_fp16_storage_adaptive_scales(pkg, oracle_mod):
    """FP16 storage with device-adapted power-of-two scales stays finite and
    close to the FP64 oracle on a short run (the accuracy of the precision form
    at full length is measured by bench/precision_report)."""
    sp = pkg.scene.gain_slab_case(n_total=40, npml=6, steps=300, fine_len=100e-9, coarse=20e-9, fine=5e-9)
    setup, sc, ref = _run_pair(pkg, oracle_mod, sp, "fp16")
    lib = pkg.load()
    import ctypes as C
    scales = (C.c_double * 18)()
    assert lib.fdtd_get_fp16_scales(sc.h, scales) == 0
    assert all(s > 0 and math.frexp(s)[0] == 0.5 for s in scales)
    # Each probe is judged as the vector it records: a component that is zero
    # by symmetry at the probe point carries only rounding noise on its own.
    errs = {}
    for p, pr in enumerate(sp.probes):
        chans = [c for c, (pp, _) in enumerate(_channels(sp)) if pp == p]
        a = np.stack([sc.probe(c) for c in chans])
        b = np.stack([ref.probe(p, q) for q in range(len(pr.components))])
        errs[p] = rel_l2(a, b)
    _say("  fp16 probe rel L2: " + " ".join(f"{c}={v:.2e}" for c, v in errs.items()) + f" scales={list(scales)[:6]}")
    # FP16 keeps 11 significant bits: each store rounds by up to 2^-12, and
    # 300 steps of such rounding leave about 1e-3 relative error
    assert max(errs.values()) < 3e-3
    for f in FIELDS:
        assert np.isfinite(sc.field(f)).all()


def test_fp16_error_is_the_store_rounding(pkg, oracle_mod):
    """The device's FP16 error against the FP64 oracle equals, within a
    factor of 2, the error of the FP64 oracle with every store rounded to
    binary16 at the device's scales (oracle_set_storage): FP16 storage
    accuracy is a property of the format (tests/test_fp16_emulation.py).

    Rounding errors are pseudo-random, and the two runs round different
    values, so single probe traces (a few independent contributions each)
    scatter by a factor of 3 either way; the error of each whole field (64000
    cells) and of all probe channels together is the stable statistic."""
    sp = pkg.scene.gain_slab_case(n_total=40, npml=6, steps=300, fine_len=100e-9, coarse=20e-9, fine=5e-9)
    _say(f"[case] fp16 vs emulated binary16 stores, {sp.name}")
    setup, sc, ref = _run_pair(pkg, oracle_mod, sp, "fp16")
    emu = oracle_mod.CpuSim(sp, "oracle")
    emu.set_storage("fp16", setup.fp16_scales())
    emu.step(sp.n_steps)
    ratios = {}
    for f in FIELDS:
        r = ref.array(f)
        ratios[f] = rel_l2(sc.field(f), r) / rel_l2(emu.array(f), r)
    chans = _channels(sp)
    dev = np.concatenate([sc.probe(c) for c in range(len(chans))])
    want = np.concatenate([ref.probe(p, q) for p, q in chans])
    got = np.concatenate([emu.probe(p, q) for p, q in chans])
    ratios["probes"] = rel_l2(dev, want) / rel_l2(got, want)
    _say("  device error / emulated-store error: " + " ".join(f"{k}={v:.2f}" for k, v in ratios.items()))
    bad = {k: v for k, v in ratios.items() if not 0.5 < v < 2.0}
    assert not bad, bad
