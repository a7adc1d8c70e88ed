"""Serving: a request loop over device-resident simulations.

Requests (a SceneSpec, a storage precision and a step count) arrive over
time.  Each is set up on the host (a thread pool overlaps this with device
work), then admitted when a slot and enough device memory are free (the
budget is checked against each scene's actual device footprint; materials
are baked on the device at admission).  Admitted requests advance together,
one tick (up to `tick` steps) at a time, so requests that arrive while others
run share the device with them (continuous batching).  Two engines:

- "streams" (default): every resident request is a plain simulation on its
  own HIP stream.  The loop polls the streams (fdtd_step_query, which never
  blocks) and gives a request its next tick as soon as its previous one has
  run, so the streams stay busy, their kernels overlap on the device and each
  request runs the plain H and E kernels; admission and retirement happen
  while the other requests keep running.
- "batch": resident requests join the batch of their precision; a tick is
  one persistent H launch and one persistent E launch per precision per step
  for all of them (replayed from a captured HIP graph).
- "hybrid": requests below `small_cells` cells go to the batches (a few
  launches per step for all of them), larger ones to their own streams; the
  batches run while the streams' steps are in flight.

Finished requests are read back and freed at the end of the tick in which
they complete, which makes room for waiting ones.

A request's results do not depend on what else is resident: it goes through
the same per-cell operations as a solo run (fields and probe traces bitwise,
DFT accumulators to 1e-12).

No reference counterpart (the reference runs one simulation per process).
"""
from __future__ import annotations

import concurrent.futures as cf
import dataclasses
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import runtime
from . import scene as scn


@dataclass
class Request:
    rid: int
    spec: "scn.SceneSpec"
    precision: str = "fp32"
    steps: int = 500
    arrival: float = 0.0          # seconds after the stream starts
    keep_fields: bool = False     # read back the final fields too


@dataclass
class Result:
    rid: int
    precision: str
    steps: int
    cells: int
    arrival: float
    admitted: float
    finished: float
    device_bytes: int
    # "ok", or "unstable" (the stability monitor stopped the solve, as the
    # reference's main returns 1) or "error" (the library refused it); a
    # failed request is retired at once, its memory freed, and the others
    # keep running
    status: str = "ok"
    error: str = ""
    probes: List[np.ndarray] = field(default_factory=list)
    dfts: List[np.ndarray] = field(default_factory=list)
    fields: Dict[str, np.ndarray] = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class _Inline:
    """Executor stand-in that runs each task at once on the calling thread."""

    def submit(self, fn, *a):
        f = cf.Future()
        try:
            f.set_result(fn(*a))
        except BaseException as e:  # handed to the caller through the future
            f.set_exception(e)
        return f

    def shutdown(self, wait=True):
        pass


@dataclass
class _Live:
    req: Request
    setup: runtime.Setup
    scene: runtime.Scene
    remaining: int
    admitted: float
    device_bytes: int
    batched: bool = False  # stepped through the batch of its precision

    @property
    def cells(self) -> int:
        return self.setup.shape[0] * self.setup.shape[1] * self.setup.shape[2]


def estimate_device_bytes(setup: runtime.Setup, precision: str) -> int:
    """Upper estimate of a scene's device footprint before it is created:
    the fields (one buffer: the two-kernel step updates in place; the fused
    schedule's second buffer is allocated only when that schedule is
    chosen), psi slabs (4 arrays per axis), material ids and axis tables,
    gain state in its bounding box, probe and DFT buffers, and 1 MiB for
    small tables."""
    s = {"fp64": 8, "fp32": 4, "fp16": 2}[precision]
    nx, ny, nz = setup.shape
    n = nx * ny * nz
    np_ = setup.npml
    slab = 2 * np_ * (ny * nz + nx * nz + nx * ny) if np_ else 0
    gain = 0
    if setup.tls_on:
        rm = setup.array("region_map")
        idx = np.argwhere(rm >= 0)
        if len(idx):
            ext = idx.max(axis=0) - idx.min(axis=0) + 1
            gain = int(np.prod(ext)) * 33
    probes = setup.n_channels * 8 * (setup.spec.n_steps + 16)
    dfts = sum(len(d.freqs) * 16 * (d.box[1] - d.box[0]) * (d.box[3] - d.box[2]) * (d.box[5] - d.box[4])
               for d in setup.spec.dfts)
    return int(6 * s * n + 2 * n + 4 * s * slab + gain + probes + dfts + (1 << 20))


class Server:
    """Continuous-batching request loop on one device.

    budget_bytes: device memory the resident requests may use together.
    max_active: resident request slots.  tick: steps per batch call (a full
    tick is replayed from a captured HIP graph when `graph`)."""

    def __init__(self, budget_bytes: float = 200e9, max_active: int = 256, tick: int = 50, graph: bool = True,
                 setup_workers: int = 8, tuning: Optional[tuple] = None, engine: str = "streams",
                 small_cells: int = 2_000_000, enqueue_workers: int = 4, stream_workers: int = 0,
                 pool_bytes: int = 24 << 30,
                 defer_free: bool = False, host_bake: bool = False, step_graphs: bool = True):
        if engine not in ("streams", "batch", "hybrid"):
            raise ValueError("engine must be 'streams', 'batch' or 'hybrid'")
        self.engine = engine
        self.small_cells = small_cells
        # device memory cache while serving: freed requests' blocks go to the
        # next admissions instead of hipFree (which waits for the whole device)
        self.pool_bytes = pool_bytes
        self.enqueue_workers = enqueue_workers
        # streams engine: threads that make the per-request device calls
        # (0: the loop thread makes them itself)
        self.stream_workers = stream_workers
        # streams engine: a finished request is read back at once but its
        # device memory is released only when admission needs the room (or
        # at the end), because a device free waits for the whole device
        self.defer_free = defer_free
        # bake materials on the host in the setup pool instead of on the
        # device at admission (the device bake frees its temporaries)
        self.host_bake = host_bake
        # requests stepped on their own streams replay each tick from a HIP
        # graph (fdtd_set_step_graphs: bitwise the direct launches), so a tick
        # costs the host one launch instead of a few per step.  Not with
        # worker threads: a tick recorded on one thread while another thread
        # makes device calls failed to record ("operation failed due to a
        # previous error during capture", 4 workers, 64 requests), so those
        # requests launch their kernels directly
        self.step_graphs = step_graphs and stream_workers == 0
        self.budget = int(budget_bytes)
        self.max_active = max_active
        self.tick = tick
        self.graph = graph
        self.setup_workers = setup_workers
        self.tuning = tuning          # (rows, wg_per_cu, chunk, planes), e.g. from Batch.autotune
        self.batches: Dict[str, runtime.Batch] = {}
        # wall seconds of the last run by activity, and its tick count
        self.stats: Dict[str, float] = {}

    def _batch(self, precision: str) -> runtime.Batch:
        b = self.batches.get(precision)
        if b is None:
            b = runtime.Batch(precision)
            b.set_graph(self.graph)
            if self.tuning:
                b.set_tuning(*self.tuning)
            self.batches[precision] = b
        return b

    def run(self, requests: List[Request]) -> List[Result]:
        """Replays the arrival times in real time and serves every request;
        returns the results in completion order."""
        # the cache is process-wide: the caller's setting comes back afterwards
        # (one server at a time per process sets it)
        before = runtime.device_pool_cap()
        runtime.set_device_pool(self.pool_bytes)
        try:
            if self.engine == "streams":
                return self._run_streams(requests)
            return self._run_ticks(requests)
        finally:
            runtime.set_device_pool(before)

    def _run_streams(self, requests: List[Request]) -> List[Result]:
        """Streams engine.  The loop polls the resident requests' streams
        (fdtd_step_query never blocks); a request whose steps have run gets
        the stability check of those steps and the enqueue of its next tick,
        or, once finished, its read-back and release.  With `stream_workers`
        > 0 these calls (and scene creation at admission) go to a pool of
        that many threads: the library's calls release the interpreter lock
        and touch only their own handle, so the host cost of kernel launches,
        uploads and frees is spread over the workers; with 0 the loop thread
        makes them in turn."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        pool = cf.ThreadPoolExecutor(max_workers=self.setup_workers)
        work = cf.ThreadPoolExecutor(max_workers=self.stream_workers) if self.stream_workers > 0 else _Inline()
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        nxt = 0
        setting_up: List[cf.Future] = []
        ready: List[tuple] = []
        admitting: List[tuple] = []    # (future, estimate)
        live: List[_Live] = []
        busy: Dict[int, cf.Future] = {}  # id(_Live) -> its tick hand-off in flight
        retiring: List[tuple] = []     # (future, _Live)
        released: List[_Live] = []     # read back, device memory not yet freed
        freeing: List[tuple] = []      # (future, bytes) of frees in flight
        used = 0
        done: List[Result] = []
        st = {"enqueue_s": 0.0, "admit_s": 0.0, "retire_s": 0.0, "free_s": 0.0, "idle_s": 0.0, "ticks": 0,
              "polls": 0}
        self.stats = st
        lock = threading.Lock()

        def count(key, v):
            with lock:
                st[key] += v

        def timed(key, fn, *a):
            ts = time.perf_counter()
            try:
                return fn(*a)
            finally:
                count(key, time.perf_counter() - ts)

        def admit(req, setup):
            ta = now()
            try:
                sc = runtime.Scene(setup, req.precision, probe_capacity=req.steps + 16)
            except runtime.FdtdError as e:
                return self._refused(req, setup, ta, e)
            lv = _Live(req, setup, sc, req.steps, now(), sc.device_bytes())
            try:
                if self.step_graphs:
                    sc.set_step_graphs(True)
                sc.step_async(min(self.tick, req.steps))
                if req.steps <= self.tick:
                    sc.stage_results()
                count("ticks", 1)
            except runtime.FdtdError as e:
                return self._retire(lv, now(), e)
            return lv

        def advance(lv):
            """The steps enqueued last have run: apply the stability policy,
            then enqueue the next tick; returns None, or the error that
            stopped the request, or True when the request has finished."""
            try:
                lv.scene.step_wait()
                left = lv.req.steps - lv.scene.steps_done()
                if left <= 0:
                    return True
                lv.scene.step_async(min(self.tick, left))
                if left <= self.tick:
                    # the last tick: its results reach host memory with it,
                    # so retirement queues no copy behind other requests
                    lv.scene.stage_results()
                count("ticks", 1)
                return None
            except runtime.FdtdError as e:
                return e

        try:
            while nxt < len(reqs) or setting_up or ready or admitting or live or retiring or freeing:
                t = now()
                progressed = False
                while nxt < len(reqs) and reqs[nxt].arrival <= t:
                    setting_up.append(pool.submit(self._prepare, reqs[nxt]))
                    nxt += 1
                still = []
                for f in setting_up:
                    if f.done():
                        ready.append(f.result())
                    else:
                        still.append(f)
                setting_up = still
                ready.sort(key=lambda x: x[0].arrival)
                waiting = []
                for req, setup, est in ready:
                    if len(live) + len(admitting) < self.max_active and used + est <= self.budget:
                        used += est  # charged at the estimate until the scene exists
                        admitting.append((work.submit(timed, "admit_s", admit, req, setup), est))
                    else:
                        waiting.append((req, setup, est))
                ready = waiting
                if released and any(used + est > self.budget for _, _, est in ready):
                    # a request waits for memory: free what finished requests hold
                    nbytes = sum(x.device_bytes for x in released)
                    freeing.append((work.submit(timed, "free_s", lambda xs: [x.scene.close() for x in xs],
                                                released), nbytes))
                    released = []
                still = []
                for f, nbytes in freeing:
                    if f.done():
                        f.result()
                        used -= nbytes
                        progressed = True
                    else:
                        still.append((f, nbytes))
                freeing = still
                still = []
                for f, est in admitting:
                    if not f.done():
                        still.append((f, est))
                        continue
                    progressed = True
                    used -= est
                    r = f.result()
                    if isinstance(r, Result):
                        done.append(r)
                    else:
                        live.append(r)
                        used += r.device_bytes
                admitting = still
                keep = []
                for lv in live:
                    f = busy.get(id(lv))
                    if f is None:
                        st["polls"] += 1
                        if lv.scene.step_ready():
                            busy[id(lv)] = work.submit(timed, "enqueue_s", advance, lv)
                            progressed = True
                        keep.append(lv)
                        continue
                    if not f.done():
                        keep.append(lv)
                        continue
                    progressed = True
                    del busy[id(lv)]
                    r = f.result()
                    if r is None:
                        keep.append(lv)
                        continue
                    err = None if r is True else r
                    retiring.append((work.submit(timed, "retire_s", lambda x, e: self._retire(
                        x, now(), e, close=not self.defer_free), lv, err), lv))
                live = keep
                still = []
                for f, lv in retiring:
                    if f.done():
                        done.append(f.result())
                        if self.defer_free:
                            released.append(lv)
                        else:
                            used -= lv.device_bytes
                        progressed = True
                    else:
                        still.append((f, lv))
                retiring = still
                if not progressed:
                    ti = now()
                    quiet = not live and not admitting and not retiring and not freeing
                    if quiet and setting_up:
                        cf.wait(setting_up, timeout=0.05, return_when=cf.FIRST_COMPLETED)
                    elif quiet and nxt < len(reqs):
                        time.sleep(max(0.0, min(0.05, reqs[nxt].arrival - now())))
                    else:
                        time.sleep(0.0002)
                    st["idle_s"] += now() - ti
        finally:
            work.shutdown(wait=True)
            pool.shutdown(wait=True)
            for x in released:
                x.scene.close()
        return done

    def _run_ticks(self, requests: List[Request]) -> List[Result]:
        reqs = sorted(requests, key=lambda r: r.arrival)
        pool = cf.ThreadPoolExecutor(max_workers=self.setup_workers)
        enq = cf.ThreadPoolExecutor(max_workers=max(1, self.enqueue_workers))
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        nxt = 0                                   # next request not yet handed to setup
        setting_up: List[cf.Future] = []
        ready: List[tuple] = []                   # (request, setup, estimate)
        live: List[_Live] = []
        used = 0
        done: List[Result] = []
        st = {"step_s": 0.0, "device_ms": 0.0, "admit_s": 0.0, "retire_s": 0.0, "idle_s": 0.0, "ticks": 0}
        self.stats = st
        try:
            while nxt < len(reqs) or setting_up or ready or live:
                t = now()
                while nxt < len(reqs) and reqs[nxt].arrival <= t:
                    r = reqs[nxt]
                    setting_up.append(pool.submit(self._prepare, r))
                    nxt += 1
                still = []
                for f in setting_up:
                    if f.done():
                        ready.append(f.result())
                    else:
                        still.append(f)
                setting_up = still
                # admission in arrival order while a slot and memory are free
                ready.sort(key=lambda x: x[0].arrival)
                waiting = []
                for req, setup, est in ready:
                    if len(live) < self.max_active and used + est <= self.budget:
                        ta = now()
                        try:
                            sc = runtime.Scene(setup, req.precision, probe_capacity=req.steps + 16)
                        except runtime.FdtdError as e:
                            done.append(self._refused(req, setup, ta, e))
                            continue
                        sc.set_stop(req.steps)
                        actual = sc.device_bytes()
                        cells = setup.shape[0] * setup.shape[1] * setup.shape[2]
                        batched = self.engine == "batch" or (self.engine == "hybrid" and cells < self.small_cells)
                        if batched:
                            self._batch(req.precision).add(sc)
                        elif self.step_graphs:
                            sc.set_step_graphs(True)
                        live.append(_Live(req, setup, sc, req.steps, now(), actual, batched))
                        used += actual
                        st["admit_s"] += now() - ta
                    else:
                        waiting.append((req, setup, est))
                ready = waiting
                if not live:
                    # idle: wait for the next arrival or a finished setup
                    ti = now()
                    if setting_up:
                        cf.wait(setting_up, timeout=0.05, return_when=cf.FIRST_COMPLETED)
                    elif nxt < len(reqs):
                        time.sleep(max(0.0, min(0.05, reqs[nxt].arrival - now())))
                    st["idle_s"] += now() - ti
                    continue
                # one tick for every resident request; a request that reaches
                # its step count mid-tick idles for the rest of it
                ts = now()
                # requests on their own streams: enqueue (largest first, so
                # their kernels start while the host enqueues the rest), run
                # the batches meanwhile, then wait
                solo = sorted((x for x in live if not x.batched), key=lambda x: -x.cells)
                failed = {}  # id(_Live) -> FdtdError

                def guarded(fn, x):
                    if id(x) in failed:
                        return
                    try:
                        fn(x)
                    except runtime.FdtdError as e:
                        failed[id(x)] = e

                if solo:
                    list(enq.map(lambda x: guarded(
                        lambda y: y.scene.step_async(min(self.tick, y.req.steps - y.scene.steps_done())), x), solo))
                for prec, b in self.batches.items():
                    if len(b):
                        try:
                            st["device_ms"] += b.step(self.tick)
                        except runtime.FdtdError as e:
                            # the batch ran its steps; the members the policy
                            # stopped are the unstable ones
                            for x in live:
                                if x.batched and x.req.precision == prec and not x.scene.stable():
                                    failed[id(x)] = e
                if solo:
                    list(enq.map(lambda x: guarded(lambda y: y.scene.step_wait(), x), solo))
                st["step_s"] += now() - ts
                st["ticks"] += 1
                for lv in live:
                    lv.remaining = lv.req.steps - lv.scene.steps_done()
                t_end = now()
                keep = []
                for lv in live:
                    if id(lv) in failed or lv.scene.steps_done() >= lv.req.steps:
                        done.append(self._retire(lv, t_end, failed.get(id(lv))))
                        used -= lv.device_bytes
                    else:
                        keep.append(lv)
                live = keep
                st["retire_s"] += now() - t_end
        finally:
            pool.shutdown(wait=True)
            enq.shutdown(wait=True)
            for b in self.batches.values():
                b.close()
            self.batches = {}
        return done

    def _prepare(self, req: Request):
        spec = dataclasses.replace(req.spec, device_bake=0) if self.host_bake else req.spec
        setup = runtime.Setup(spec)
        return req, setup, estimate_device_bytes(setup, req.precision)

    @staticmethod
    def _refused(req: Request, setup: runtime.Setup, t: float, err: Exception) -> Result:
        """A request the library refused at admission (e.g. device memory)."""
        return Result(rid=req.rid, precision=req.precision, steps=req.steps,
                      cells=setup.shape[0] * setup.shape[1] * setup.shape[2], arrival=req.arrival, admitted=t,
                      finished=t, device_bytes=0, status="error", error=str(err))

    def _retire(self, lv: _Live, t_end: float, err: Optional[Exception] = None, close: bool = True) -> Result:
        """Reads back a finished (or failed) request and (with `close`) frees
        its device state."""
        if lv.batched:
            self.batches[lv.req.precision].remove(lv.scene)
        st = lv.setup
        res = Result(rid=lv.req.rid, precision=lv.req.precision, steps=lv.req.steps,
                     cells=st.shape[0] * st.shape[1] * st.shape[2], arrival=lv.req.arrival, admitted=lv.admitted,
                     finished=t_end, device_bytes=lv.device_bytes)
        if err is not None:
            unstable = isinstance(err, runtime.FdtdError) and err.code == runtime.FDTD_EUNSTABLE
            res.status = "unstable" if unstable else "error"
            res.error = str(err)
            res.steps = lv.scene.steps_done()
        res.probes = [lv.scene.probe(c) for c in range(st.n_channels)]
        res.dfts = [lv.scene.dft(m) for m in range(st.n_dfts)]
        if lv.req.keep_fields:
            res.fields = {f: lv.scene.field(f) for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}
        if close:
            lv.scene.close()
        return res


def solve_alone(req: Request, engine: str = "streams") -> Result:
    """The same request served by itself (a plain solve, or a batch of one
    for the batch engine), for equality checks."""
    t0 = time.perf_counter()
    setup = runtime.Setup(req.spec)
    sc = runtime.Scene(setup, req.precision, probe_capacity=req.steps + 16)
    sc.set_stop(req.steps)
    b = None
    if engine == "batch":
        b = runtime.Batch(req.precision)
        b.add(sc)
        while sc.steps_done() < req.steps:
            b.step(50)
    else:
        sc.step(req.steps)
    res = Result(rid=req.rid, precision=req.precision, steps=req.steps,
                 cells=setup.shape[0] * setup.shape[1] * setup.shape[2], arrival=0.0, admitted=0.0,
                 finished=time.perf_counter() - t0, device_bytes=sc.device_bytes())
    res.probes = [sc.probe(c) for c in range(setup.n_channels)]
    res.dfts = [sc.dft(m) for m in range(setup.n_dfts)]
    if req.keep_fields:
        res.fields = {f: sc.field(f) for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}
    if b is not None:
        b.close()
    sc.close()
    return res


def request_stream(n: int, rate: float, seed: int = 1, lo: int = 48, hi: int = 224, steps=(500, 4000),
                   fp16_share: float = 0.5) -> List[Request]:
    """The serving workload: Poisson arrivals at `rate` per second; each axis
    of a request log-uniform in [lo, hi] total cells; steps uniform in
    `steps`; 40% dielectric, 40% dielectric + gain, 20% multi-block scenes;
    FP32 or FP16 storage per request."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += float(rng.exponential(1.0 / rate)) if rate > 0 else 0.0
        kind = rng.choice(3, p=[0.4, 0.4, 0.2])
        sp = scn.serve_case(rid, int(kind), seed=seed, lo=lo, hi=hi)
        nst = int(rng.integers(steps[0], steps[1] + 1))
        sp.n_steps = nst
        sp.device_bake = 1  # materials baked on the device at admission
        prec = "fp16" if rng.random() < fp16_share else "fp32"
        out.append(Request(rid, sp, prec, nst, t))
    return out


def summarize(results: List[Result]) -> dict:
    failed = sum(1 for r in results if r.status != "ok")
    results = [r for r in results if r.status == "ok"] or results
    lat = np.array([r.latency for r in results])
    work = sum(r.cells * r.steps for r in results)
    t_first = min(r.arrival for r in results)
    t_last = max(r.finished for r in results)
    return {
        "requests": len(results),
        "cell_updates": work,
        "makespan_s": t_last - t_first,
        "gcells_per_s": work / max(t_last - t_first, 1e-9) / 1e9,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_max_s": float(lat.max()),
        "failed": failed,
    }
