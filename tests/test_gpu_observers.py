"""Per-step observers on the device against the oracle and the device's own
state: the DFT monitor accumulators (DftMonitor, one thread per cell and
frequency), point probes on both launch paths (more than 16 channels: the
first 16 in a probe launch, the rest inside the DFT launch), and the
stability monitor records (max |E|, max |H|, max |Pz|, max |N| as the
maxima of the fields the device holds at the recorded step).

Bars: FP64 DFT accumulators and probe traces bitwise equal to the oracle's
(same operations, same host-computed phase weights); monitor maxima equal,
bit for bit, to the maxima of the downloaded fields (FP16 storage included:
a maximum is exact and the download converts as the monitor does)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _case(pkg, n_probe_points=5):
    S = pkg.scene
    sp = S.gain_slab_case(n_total=40, npml=6, steps=60, fine_len=100e-9, coarse=20e-9, fine=5e-9)
    base = sp.probes[0]
    pts = []
    for q in range(n_probe_points):
        pts.append(S.Probe(base.x + q * 7e-9, base.y - q * 5e-9, base.z + q * 3e-9, (0, 2, 6, 11)))
    sp.probes = pts  # 20 channels
    return sp


def test_dft_and_probes_match_oracle_fp64(pkg, oracle_mod):
    sp = _case(pkg)
    setup = pkg.Setup(sp)
    assert setup.n_channels == 20 and setup.n_dfts >= 1
    _say(f"[case] {sp.name} fp64 dft+probes channels={setup.n_channels}")
    sc = pkg.Scene(setup, "fp64")
    sc.step(sp.n_steps)
    ref = oracle_mod.CpuSim(sp, "oracle")
    ref.step(sp.n_steps)
    for m in range(setup.n_dfts):
        a, b = sc.dft(m), ref.dft(m)
        assert a.shape == b.shape and np.abs(b).max() > 0
        assert np.array_equal(a, b), f"dft {m}: max diff {np.abs(a - b).max():.3e}"
    c = 0
    for p, pr in enumerate(sp.probes):
        for q in range(len(pr.components)):
            a, b = sc.probe(c), ref.probe(p, q)
            assert len(a) == sp.n_steps and np.array_equal(a, b), f"probe channel {c}"
            c += 1
    sc.close()


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_dft_and_probes_reduced_precision(pkg, oracle_mod, precision):
    from conftest import rel_l2

    sp = _case(pkg)
    setup = pkg.Setup(sp)
    _say(f"[case] {sp.name} {precision} dft+probes")
    sc = pkg.Scene(setup, precision)
    sc.step(sp.n_steps)
    ref = oracle_mod.CpuSim(sp, "oracle")
    ref.step(sp.n_steps)
    tol = 2e-4 if precision == "fp32" else 5e-2
    for m in range(setup.n_dfts):
        assert rel_l2(sc.dft(m), ref.dft(m)) <= tol
    sc.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp16"])
def test_monitor_records_are_field_maxima(pkg, precision):
    sp = _case(pkg, 1)
    setup = pkg.Setup(sp)
    _say(f"[case] {sp.name} {precision} monitor")
    sc = pkg.Scene(setup, precision)
    seen = 0
    for n in range(31):
        sc.step(1)
        if n % 10:
            continue
        recs = sc.checks()
        assert len(recs) == 1 and recs[0].step == n
        r = recs[0]
        E = max(float(np.abs(sc.field(f)).max()) for f in ("Ex", "Ey", "Ez"))
        H = max(float(np.abs(sc.field(f)).max()) for f in ("Hx", "Hy", "Hz"))
        P = float(np.abs(sc.field("Pz")).max())
        Nm = max(float(np.abs(sc.field(f)).max()) for f in ("Nu", "Ng"))
        if precision == "fp16" and n % 10 == 0:
            # storage rescaling (every 10 steps, after the record) may move
            # the downloaded values by the rounding of the new scale
            assert r.max_E == pytest.approx(E, rel=1e-3) and r.max_H == pytest.approx(H, rel=1e-3)
        else:
            assert r.max_E == E and r.max_H == H, (n, r.max_E, E, r.max_H, H)
        assert r.max_P == P and r.nan_N == 0 and r.nan_E == 0
        assert Nm > 0
        seen += 1
    assert seen == 4
    sc.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp16"])
def test_monitor_maxima_in_next_h_launch_equal_separate_pass(pkg, precision):
    """Within a multi-step call the monitor's field maxima of a monitor step
    are taken by the next step's H launch; the records (every field,
    including the population sums) equal bit for bit those of the separate
    pass over the six fields (kernel-mode bit 262144) and those of the fused
    step schedule (which keeps the separate pass) and of the dependency-
    ordered step (whose H tiles take the maxima), with a monitor step also
    as the last step of a call (where the separate pass always runs) and with
    the monitor every step."""
    sp = _case(pkg, 1)
    setup = pkg.Setup(sp)
    _say(f"[case] {sp.name} {precision} monitor maxima in the H launch")
    recs = {}
    for tag, mode, bit in (("h_launch", 0, 0), ("separate", 0, 262144), ("fused", 1, 0), ("flow", 2, 0)):
        for interval in (10, 1):
            sc = pkg.Scene(setup, precision)
            sc.set_kernel_mode(mode, bit)
            pkg.load().fdtd_set_monitor(sc.h, interval)
            out = []
            for n in (21, 10, 9):  # calls ending on and between monitor steps
                sc.step(n)
                out += [(r.step, r.max_E, r.max_H, r.max_P, r.nan_E, r.nan_H, r.nan_P, r.nan_N, r.pop_total,
                         r.pop_initial) for r in sc.checks()]
            recs[(tag, interval)] = out
            sc.close()
    for interval in (10, 1):
        a, b, c = recs[("h_launch", interval)], recs[("separate", interval)], recs[("fused", interval)]
        assert len(a) == (4 if interval == 10 else 40) and a == b, (interval, a[:3], b[:3])
        assert a == c, (interval, a[:3], c[:3])  # the fused step keeps the separate pass
        d = recs[("flow", interval)]  # the dependency-ordered step's H tiles take the maxima
        assert a == d, (interval, a[:3], d[:3])
        assert a[-1][1] > 0 and a[-1][2] > 0 and a[-1][8] > 0
