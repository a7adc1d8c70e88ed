"""The serving loop's bookkeeping on the CPU: admission by slots and memory
budget, continuous admission while others run, per-request failure status,
deferred release of device memory, and the worker-thread variant of the
streams engine.  The device is replaced by a stand-in scene whose steps
"run" for a fixed wall time per tick (no HIP call is made), so the loop's
logic is exercised exactly as on the device, including several threads
making calls at once."""
import importlib
import threading
import time
from types import SimpleNamespace

import numpy as np
import pytest

serve = importlib.import_module("3d_fdtd_version008_amd.serve")
runtime = serve.runtime

TICK_S = 0.002  # wall time one tick of steps takes on the stand-in device


class FakeSetup:
    def __init__(self, spec):
        self.spec = spec
        self.shape = spec.shape
        self.n_channels = 2
        self.n_dfts = 1


class FakeScene:
    """A stand-in for runtime.Scene: step_async starts a timer, step_ready
    polls it, step_wait applies a 'stability policy' that fails once a
    poisoned request passes its fail step."""
    live = 0
    peak = 0
    lock = threading.Lock()

    def __init__(self, setup, precision, probe_capacity=0):
        self.setup = setup
        self.n = 0
        self.pending = 0
        self.t_done = 0.0
        self.closed = False
        self.bytes = setup.spec.bytes
        with FakeScene.lock:
            FakeScene.live += 1
            FakeScene.peak = max(FakeScene.peak, FakeScene.live)

    def stage_results(self):
        self.staged = True

    def set_step_graphs(self, on=True):
        self.graphs = bool(on)

    def step_async(self, k):
        assert not self.closed and self.pending == 0 and k > 0
        self.pending = k
        self.t_done = time.perf_counter() + TICK_S

    def step_ready(self):
        return time.perf_counter() >= self.t_done

    def step_wait(self):
        while not self.step_ready():
            time.sleep(0.0002)
        self.n += self.pending
        self.pending = 0
        fail = self.setup.spec.fail_at
        if fail is not None and self.n >= fail:
            raise runtime.FdtdError(runtime.FDTD_EUNSTABLE, f"simulation unstable at step {fail}")

    def steps_done(self):
        return self.n

    def device_bytes(self):
        return self.bytes

    def probe(self, c):
        return np.full(self.n, float(self.setup.spec.rid * 10 + c))

    def dft(self, m):
        return np.zeros(2)

    def close(self):
        if not self.closed:
            self.closed = True
            with FakeScene.lock:
                FakeScene.live -= 1


def _requests(n, seed=3, poison=None):
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    for rid in range(n):
        t += float(rng.exponential(0.002))
        steps = int(rng.integers(3, 9)) * 10
        fail = 20 if rid == poison else None
        spec = SimpleNamespace(rid=rid, shape=(8, 8, 8), bytes=int(rng.integers(1, 5)) * 100, fail_at=fail,
                               device_bake=1)
        out.append(serve.Request(rid, spec, "fp32", steps, t))
    return out


@pytest.fixture
def fake_device(monkeypatch):
    monkeypatch.setattr(runtime, "Setup", FakeSetup)
    monkeypatch.setattr(runtime, "Scene", FakeScene)
    monkeypatch.setattr(serve, "estimate_device_bytes", lambda setup, prec: setup.spec.bytes)
    FakeScene.live = 0
    FakeScene.peak = 0
    yield


@pytest.mark.parametrize("workers", [0, 3])
@pytest.mark.parametrize("defer", [False, True])
def test_streams_engine_bookkeeping(fake_device, workers, defer):
    reqs = _requests(24, poison=5)
    budget = 700
    srv = serve.Server(budget_bytes=budget, max_active=4, tick=10, setup_workers=2, engine="streams",
                       stream_workers=workers, defer_free=defer)
    res = srv.run(reqs)
    by = {r.rid: r for r in res}
    assert sorted(by) == list(range(24))
    # the poisoned request stopped at its failing tick; the others finished
    assert by[5].status == "unstable" and by[5].steps == 20
    for r in reqs:
        got = by[r.rid]
        if r.rid == 5:
            continue
        assert got.status == "ok" and got.steps == r.steps
        assert len(got.probes) == 2 and len(got.probes[0]) == r.steps
        assert np.all(got.probes[1] == r.rid * 10 + 1)
        assert got.arrival <= got.admitted <= got.finished
    # at most max_active resident and the resident footprint within budget
    # (released-but-not-freed memory still counts against it)
    assert FakeScene.peak <= 4 + (24 if defer else 0)
    for x in res:
        resident = sum(y.device_bytes for y in res if y.admitted <= x.admitted < y.finished)
        assert resident <= budget
    # continuous batching: some request was admitted while another ran
    spans = sorted((x.admitted, x.finished) for x in res)
    assert any(spans[q + 1][0] < spans[q][1] for q in range(len(spans) - 1))
    # every scene was released by the end
    assert FakeScene.live == 0
    s = serve.summarize(res)
    assert s["failed"] == 1 and s["requests"] == 23
