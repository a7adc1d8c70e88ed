"""Recorded stepping calls (fdtd_set_step_graphs): a call's launches are
recorded once in a HIP graph and replayed for later calls of the same shape,
with each call's DFT phase factors and probe record slots read from a device
table written before the launch.  The results equal the direct launches' bit
for bit: fields, populations, probe traces (every replay's records) and DFT
accumulators, over calls of the same shape (replays), a shorter call (a new
shape) and the asynchronous entry points the serving loop uses."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
CALLS = [50, 50, 50, 30, 50, 50, 20]  # the monitor every 10 steps: three shapes


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _state(sc, setup):
    st = {f: sc.field(f) for f in FIELDS}
    if setup.tls_on:
        st["Nu"] = sc.field("Nu")
        st["Pz"] = sc.field("Pz")
    for c in range(setup.n_channels):
        st[f"probe{c}"] = sc.probe(c)
    for m in range(setup.n_dfts):
        st[f"dft{m}"] = sc.dft(m)
    return st


@pytest.mark.parametrize("schedule", [0, 2])
@pytest.mark.parametrize("precision", ["fp32", "fp16", "fp64"])
@pytest.mark.parametrize("kind", [1, 2])
def test_step_graphs_equal_direct_launches(pkg, precision, kind, schedule):
    """schedule 2 (the dependency-ordered step) keeps its step epoch in device
    memory, so every replay of a recorded call waits on its own steps' flags."""
    sp = pkg.scene.serve_case(5 + kind, kind=kind, seed=3, lo=36, hi=52, steps=sum(CALLS))
    setup = pkg.Setup(sp)
    _say(f"[step graphs] kind={kind} {precision} schedule={schedule} shape={setup.shape} probes={setup.n_channels} "
         f"dfts={setup.n_dfts}")
    direct = pkg.Scene(setup, precision)
    for k in CALLS:
        direct.step(k)
    ref = _state(direct, setup)
    assert np.abs(ref["Ez"]).max() > 0
    direct.close()
    for mode in ("sync", "async"):
        sc = pkg.Scene(setup, precision)
        sc.set_kernel_mode(schedule, 0)
        sc.set_step_graphs(True)
        for k in CALLS:
            if mode == "sync":
                sc.step(k)
            else:
                sc.step_async(k)
                sc.step_wait()
        got = _state(sc, setup)
        sc.close()
        diff = [n for n in ref if not np.array_equal(ref[n], got[n])]
        assert not diff, f"{mode}: differs in {diff}"
