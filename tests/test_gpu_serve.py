"""Serving: continuous batching of a mixed FP32/FP16 request stream, with
both engines (requests on their own streams, or in per-precision batches).

Every request's probe traces and final fields are bitwise identical to the
same request served alone (DFT accumulators to 1e-12 relative L2), while
requests overlap on the device and admission is limited by slots and by the
device-memory budget."""
import importlib
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("engine", ["streams", "batch", "hybrid", "streams-workers"])
def test_served_results_equal_alone(pkg, engine):
    workers = 3 if engine == "streams-workers" else 0
    engine = engine.split("-")[0]
    serve = importlib.import_module("3d_fdtd_version008_amd.serve")
    reqs = serve.request_stream(10, rate=40.0, seed=2, lo=32, hi=48, steps=(60, 140))
    for r in reqs:
        r.keep_fields = True
    # a budget for about three resident requests forces queueing by memory
    est = [serve.estimate_device_bytes(pkg.Setup(r.spec), r.precision) for r in reqs]
    # hybrid: requests below the median size go to the batches
    small = int(np.median([pkg.Setup(r.spec).shape[0] ** 3 for r in reqs]))
    srv = serve.Server(budget_bytes=3.2 * max(est), max_active=8, tick=20, graph=True, setup_workers=4,
                       engine=engine, small_cells=small, stream_workers=workers)
    served = {r.rid: r for r in srv.run(reqs)}
    assert sorted(served) == [r.rid for r in reqs]
    s = serve.summarize(list(served.values()))
    _say(f"  served {s}")
    # continuous batching: some request was admitted while another was resident
    spans = sorted((x.admitted, x.finished) for x in served.values())
    overlap = any(spans[q + 1][0] < spans[q][1] for q in range(len(spans) - 1))
    assert overlap
    # the admission estimate is an upper bound of the footprint it charges
    for r, e in zip(reqs, est):
        assert served[r.rid].device_bytes <= e, (r.rid, served[r.rid].device_bytes, e)
    # the budget bounded the resident footprint
    for x in served.values():
        live = sum(y.device_bytes for y in served.values() if y.admitted <= x.admitted < y.finished)
        assert live <= srv.budget
    for r in reqs:
        alone = serve.solve_alone(r, "streams" if engine == "hybrid" else engine)
        got = served[r.rid]
        for a, b in zip(alone.probes, got.probes):
            assert np.array_equal(a, b), f"request {r.rid} probe differs"
        for f in alone.fields:
            assert np.array_equal(alone.fields[f], got.fields[f]), f"request {r.rid} field {f} differs"
        for a, b in zip(alone.dfts, got.dfts):
            assert rel_l2(b, a) <= 1e-12
        assert len(got.probes[0]) == r.steps


def test_async_steps_equal_sync_steps(pkg):
    """fdtd_step_async / fdtd_step_wait give the results of fdtd_step, also
    when a second enqueue completes the first, and when two handles' steps
    are in flight together."""
    sp = pkg.scene.gain_slab_case(n_total=40, npml=6, steps=90, fine_len=100e-9, coarse=20e-9, fine=5e-9)
    st = pkg.Setup(sp)
    ref = pkg.Scene(st, "fp32")
    ref.step(90)
    a = pkg.Scene(st, "fp32")
    b = pkg.Scene(st, "fp16")
    bref = pkg.Scene(st, "fp16")
    bref.step(90)
    a.step_async(30)
    b.step_async(45)
    a.step_async(60)  # completes the first 30 steps first
    b.step_async(45)
    a.step_wait()
    b.step_wait()
    assert a.steps_done() == 90 and b.steps_done() == 90
    for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Nu", "Pz"):
        assert np.array_equal(a.field(f), ref.field(f)), f
        assert np.array_equal(b.field(f), bref.field(f)), f
    for c in range(st.n_channels):
        assert np.array_equal(a.probe(c), ref.probe(c))
    for sc in (ref, a, b, bref):
        sc.close()


@pytest.mark.parametrize("engine", ["streams", "batch", "hybrid"])
def test_unstable_request_is_isolated(pkg, engine):
    """One request of the stream is unstable (Courant number 1.2: its fields
    grow until the reference's policy stops it).  It is retired with status
    "unstable" and its memory freed; every other request still equals its
    solo run bit for bit."""
    serve = importlib.import_module("3d_fdtd_version008_amd.serve")
    reqs = serve.request_stream(6, rate=60.0, seed=4, lo=32, hi=44, steps=(80, 160))
    bad = reqs[2]
    bad.spec = bad.spec.copy(S=1.2)
    bad.steps = bad.spec.n_steps = 400
    _say(f"[case] serve {engine} with request {bad.rid} unstable")
    small = int(np.median([pkg.Setup(r.spec).shape[0] ** 3 for r in reqs]))
    srv = serve.Server(budget_bytes=50e9, max_active=4, tick=40, graph=True, setup_workers=4, engine=engine,
                       small_cells=small)
    served = {r.rid: r for r in srv.run(reqs)}
    assert sorted(served) == [r.rid for r in reqs]
    assert served[bad.rid].status == "unstable", served[bad.rid].error
    assert served[bad.rid].steps < bad.steps
    s = serve.summarize(list(served.values()))
    assert s["failed"] == 1
    for r in reqs:
        if r.rid == bad.rid:
            continue
        got = served[r.rid]
        assert got.status == "ok"
        alone = serve.solve_alone(r, "streams" if engine == "hybrid" else engine)
        for a, b in zip(alone.probes, got.probes):
            assert np.array_equal(a, b), f"request {r.rid} probe differs"


def test_served_requests_match_oracle(pkg, oracle_mod):
    """Every request of a served mixed FP32/FP16 stream (the serving mix of
    dielectric, gain and multi-block scenes, on reduced sizes and step
    counts so the CPU oracle finishes in seconds) against the FP64 CPU oracle
    on the same request: probe traces within 2e-4 (FP32) and 5e-2 (FP16)
    relative L2, DFT spectra within 2e-4 (FP32) and 1e-2 (FP16).  FP16
    storage rounds every stored value to 11 significant bits, and over a few
    hundred steps these roundings leave errors of a few 1e-3 to 1e-2 at a
    probe; the bar is the FP16 form's stated accuracy for these requests."""
    serve = importlib.import_module("3d_fdtd_version008_amd.serve")
    reqs = serve.request_stream(8, rate=50.0, seed=7, lo=26, hi=40, steps=(150, 300))
    assert {r.precision for r in reqs} == {"fp32", "fp16"}
    srv = serve.Server(budget_bytes=50e9, max_active=4, tick=50, graph=True, setup_workers=4, engine="streams")
    served = {r.rid: r for r in srv.run(reqs)}
    bars = {"fp32": (2e-4, 2e-4), "fp16": (5e-2, 1e-2)}
    for r in reqs:
        got = served[r.rid]
        assert got.status == "ok"
        ref = oracle_mod.CpuSim(r.spec, "oracle")
        ref.step(r.steps)
        errs, c = {}, 0
        for p, pr in enumerate(r.spec.probes):
            n = len(pr.components)
            a = np.stack([got.probes[c + q] for q in range(n)])
            b = np.stack([ref.probe(p, q) for q in range(n)])
            assert a.shape == b.shape
            errs[f"probe{p}"] = rel_l2(a, b)
            c += n
        for m in range(len(r.spec.dfts)):
            errs[f"dft{m}"] = rel_l2(got.dfts[m], ref.dft(m))
        _say(f"  request {r.rid} {r.precision} {r.spec.name} steps={r.steps}: "
             + " ".join(f"{k}={v:.1e}" for k, v in errs.items()))
        bp, bd = bars[r.precision]
        bad = {k: v for k, v in errs.items() if not v <= (bp if k.startswith("probe") else bd)}
        assert not bad, (r.rid, r.precision, bad)
        ref.close()


def test_device_pool_reuse_is_bitwise(pkg):
    """With the device memory cache on, a destroyed scene's blocks (and the
    device bake's temporary buffers) are handed to the next scenes: results
    equal those of scenes on fresh allocations, the cache holds the freed
    bytes, and turning it off frees them."""
    rt = pkg.runtime
    serve = importlib.import_module("3d_fdtd_version008_amd.serve")
    reqs = serve.request_stream(4, rate=0.0, seed=3, lo=32, hi=48, steps=(60, 60))
    setups = [pkg.Setup(r.spec) for r in reqs]

    def run(st, prec):
        sc = pkg.Scene(st, prec)
        sc.step(60)
        out = {f: sc.field(f) for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}
        out["probe0"] = sc.probe(0)
        sc.close()
        return out

    rt.set_device_pool(0)
    ref = [run(st, r.precision) for st, r in zip(setups, reqs)]
    try:
        rt.set_device_pool(4 << 30)
        for rep in range(2):  # the second pass runs on blocks the first one freed
            got = [run(st, r.precision) for st, r in zip(setups, reqs)]
            for a, b in zip(ref, got):
                assert all(np.array_equal(a[k], b[k]) for k in a), f"pass {rep}"
            assert rt.device_pool_bytes() > 0
    finally:
        rt.set_device_pool(0)
    assert rt.device_pool_bytes() == 0


def test_staged_results_equal_device_reads(pkg):
    """fdtd_stage_results: the host copy enqueued after a call's steps equals
    what the device buffers hold; the next step makes reads go to the device
    again (and a new staging picks the new records up); a batch member is
    refused."""
    serve = importlib.import_module("3d_fdtd_version008_amd.serve")
    req = serve.request_stream(1, rate=0.0, seed=7, lo=36, hi=44, steps=(90, 90))[0]
    st = pkg.Setup(req.spec)
    ref = pkg.Scene(st, req.precision)
    sc = pkg.Scene(st, req.precision)
    sc.set_step_graphs(True)
    ref.step(60)
    sc.step_async(60)
    sc.stage_results()
    sc.step_wait()
    for c in range(st.n_channels):
        assert np.array_equal(sc.probe(c), ref.probe(c))
    for m in range(st.n_dfts):
        assert np.array_equal(sc.dft(m), ref.dft(m))
    ref.step(30)
    sc.step(30)  # the staged copy is stale now: reads go to the device
    for c in range(st.n_channels):
        assert len(sc.probe(c)) == len(ref.probe(c)) and np.array_equal(sc.probe(c), ref.probe(c))
    for m in range(st.n_dfts):
        assert np.array_equal(sc.dft(m), ref.dft(m))
    sc.stage_results()
    for c in range(st.n_channels):
        assert np.array_equal(sc.probe(c), ref.probe(c))
    b = pkg.Batch(req.precision)
    b.add(ref)
    with pytest.raises(pkg.runtime.FdtdError):
        ref.stage_results()
    b.close()
    ref.close()
    sc.close()
