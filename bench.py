"""Benchmark of the per-timestep Yee update on one MI355X.

Workload (value): the ~100M-cell precision geometry, 512x512x384 total cells
(12-cell CPML, 10 nm cells in a central 400 nm z slab graded at 1.2 per cell
to 40 nm, 20 nm lateral), eps_r 3.5 substrate, 200 nm two-level gain slab, 64
seeded dielectric cuboids (eps_r U[1,12], edges 8-40 cells), Ex current sheet
source, probes and a 3-frequency DFT plane, stability monitor every 10 steps,
FP32 storage/arithmetic with FP64 populations (--workload c2: the 256^3 gain
slab without cuboids).  A "step" is one full
Yee step of the whole grid (H, CPML-H, E with J, CPML-E, gain, probes, DFT,
monitor).  Inputs are resident in HBM before the timed region.

Prints ONE JSON line (rank 0, N=1).  --gpus N > 1 is not supported: this
build runs on one GPU.

Other forms (not the default line): --form precision (512x512x384 graded grid
with 64 cuboids: FP32 and FP16-storage throughput and their error against the
FP64 device path per probe and per DFT spectrum), --form batch (256 uneven
domains in persistent batch launches against one-at-a-time solves), --form
serve (a Poisson stream of mixed FP32/FP16 requests under continuous batching:
throughput, p50/p99 latency, against one-at-a-time solves).
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
METRIC = "Gcell-updates/s & % HBM roofline; p50/p99 latency of batched solves"


def algorithmic_bytes(spec, setup, prec_bytes: int, mat_bytes: int):
    """Algorithmic HBM bytes per step: of the H kernel, of the E kernel, and
    of the whole step (the fused kernel's unit).

    Step (per cell-update): read E and H once, write E and H once: 12 s, plus
    the material id m; each CPML slab cell carries, per active axis, 4 psi
    values (2 for H, 2 for E) read and written: 8 s; each gain cell carries
    Pz, Pz_prev, Nu, Ng (FP64) read and written plus its region byte: 65 B.
    Separate kernels: H reads E, H (6 s) and writes H (3 s); E reads H, E,
    writes E: 9 s each, with the psi and gain terms split between them."""
    nx, ny, nz = setup.shape
    N = nx * ny * nz
    np_ = setup.npml
    slab = 0
    if np_ > 0:
        slab = 2 * np_ * ny * nz + nx * 2 * np_ * nz + nx * ny * 2 * np_  # (cell, axis) pairs
    s = prec_bytes
    gain_cells = int((setup.array("region_map").ravel() >= 0).sum()) if setup.tls_on else 0
    h_bytes = 9 * s * N + 4 * s * slab
    e_bytes = (9 * s + mat_bytes) * N + 4 * s * slab + 65 * gain_cells
    step_bytes = (12 * s + mat_bytes) * N + 8 * s * slab + 65 * gain_cells
    return h_bytes, e_bytes, step_bytes, gain_cells


def copy_calibration(nbytes: int) -> None:
    """One plain device-to-device copy of nbytes (reads nbytes, writes nbytes),
    so a counter run can relate FETCH_SIZE / WRITE_SIZE to known traffic."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def host_info():
    """nproc, CPU model and NUMA nodes of the host the baseline runs on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    nodes = 0
    try:
        nodes = sum(1 for d in os.listdir("/sys/devices/system/node") if d.startswith("node") and d[4:].isdigit())
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "numa_nodes": nodes}


def cpu_baseline(pkg, sample, steps: int, steps_1t: int):
    """The reference's own step (its headers compiled with OpenMP into
    oracle/_ref/libref.so, kind "reference") timed on this host's cores, on
    `sample`, a grid of the workload's structure: all threads, then one
    thread.  The reference's serial subpixel bake would take hours on a grid
    with dozens of structures, so the harness takes the MaterialGrids arrays
    of this library's bake (bitwise the reference's, tests/test_setup_parity.py)
    and runs the rest of the reference's setup itself.  Where the harness is
    absent the oracle's C restatement is timed instead (kind "port")."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # checker / CPU baseline only

    n = sample.n_cells()
    if oracle.available("ref"):
        setup = pkg.Setup(sample)
        mats = [setup.array(q) for q in pkg.runtime.MATERIAL_NAMES]
        sim = oracle.CpuSim(sample, "ref", materials=mats)
        del mats
        setup.close()
        kind, impl = "reference", "ref"
    else:
        if not oracle.available("oracle"):
            oracle.build()
        sim = oracle.CpuSim(sample, "oracle")
        kind, impl = "port", "oracle"
    threads = oracle.threads(impl)
    sim.time_steps(2)
    rate = n * steps / sim.time_steps(steps) / 1e9
    rate1 = None
    if impl == "ref" and steps_1t > 0:
        oracle.set_threads(1)
        sim.time_steps(1)
        rate1 = n * steps_1t / sim.time_steps(steps_1t) / 1e9
        oracle.set_threads(threads)
    sim.close()
    nx, ny, nz = sample.total_dims()
    what = ("oracle/_ref/libref.so: the reference's fdtd_update_H/E(_with_gain), CPMLBoundary, TLS and "
            "StabilityMonitor code in main's step order, OpenMP as the reference builds it" if impl == "ref" else
            "oracle/fdtd_oracle.c, OpenMP over i, serial CPML, J zero-fill, monitor every 10")
    out = {"value": rate, "unit": "Gcell-updates/s", "cores": threads, "kind": kind,
           "sample": f"{sample.name} ({nx}x{ny}x{nz} = {n} cells, the workload's z profile, CPML, gain slab and "
                     f"cuboids on a smaller x-y extent): {steps} FP64 steps after 2 warm-up steps ({what})"}
    if rate1 is not None:
        out["value_1_thread"] = rate1
        out["steps_1_thread"] = steps_1t
    out.update(host_info())
    return out


def _kw_param(args) -> int:
    """fdtd_set_kernel_mode's mode-0 parameter: one width, or H + 256 * E,
    plus 131072 for 64-bit element offsets."""
    o64 = 0x20000 if getattr(args, "offsets64", False) else 0
    band = (getattr(args, "band", 0) or 0) << 21
    if args.kwidth_e and args.kwidth_e != args.kwidth:
        return (args.kwidth or 32) + 256 * args.kwidth_e + o64 + band
    return (args.kwidth or (32 if o64 else 0)) + o64 + band


def plain(args):
    pkg = importlib.import_module("3d_fdtd_version008_amd")
    if args.copy_calib:
        copy_calibration(args.copy_calib << 20)
    if args.workload == "c3":
        spec = pkg.scene.precision_case(steps=5000)
    else:
        spec = pkg.scene.gain_slab_case(n_total=args.size, npml=12, steps=5000)
    t0 = time.time()
    setup = pkg.Setup(spec)
    t_setup = time.time() - t0
    sc = pkg.Scene(setup, args.precision, probe_capacity=args.steps + args.warmup + 16)
    info = sc.info()
    sc.set_kernel_mode(args.mode, _kw_param(args) if args.mode == 0 else args.ichunk)
    sc.step(args.warmup)
    sc.synchronize()
    t1 = time.perf_counter()
    # HIP events around the H and E launches of every 10th step: the kernel
    # averages are live measurements over the timed region, and the few
    # microseconds an event pair costs the stream stay out of the step time
    ms_h, ms_e, ms_call, n_timed = sc.step_timed(args.steps, args.event_every)
    sc.synchronize()
    wall = time.perf_counter() - t1
    N = setup.shape[0] * setup.shape[1] * setup.shape[2]
    value = N * args.steps / wall / 1e9
    prec_bytes = {"fp64": 8, "fp32": 4, "fp16": 2}[args.precision]
    h_bytes, e_bytes, step_bytes, gain_cells = algorithmic_bytes(spec, setup, prec_bytes, info["mat_id_bytes"])
    t_h = ms_h / n_timed / 1e3
    t_e = ms_e / n_timed / 1e3
    if args.mode == 1:
        kname, kbytes, kt = "k_step_march", step_bytes, t_h
        kernels = {kname: round(ms_h / n_timed, 4)}
    elif args.mode == 2:
        # the dependency-ordered step: one launch updates H everywhere and E
        # on every plane but the first of each of the 8 plane queues, whose E
        # tiles run in k_step_flow_edge after it; its algorithmic bytes are
        # the step's minimum on the planes it steps fully, H's on the others
        nx = setup.shape[0]
        ne = min(8, nx)
        kname, kt = "k_step_flow", t_h
        kbytes = int(step_bytes * (nx - ne) / nx + h_bytes * ne / nx)
        kernels = {"k_step_flow": round(ms_h / n_timed, 4), "k_step_flow_edge": round(ms_e / n_timed, 4)}
    else:
        kernels = {"k_update_H": round(ms_h / n_timed, 4), "k_update_E": round(ms_e / n_timed, 4)}
        if t_e >= t_h:
            kname, kbytes, kt = "k_update_E", e_bytes, t_e
        else:
            kname, kbytes, kt = "k_update_H", h_bytes, t_h
    achieved = kbytes / kt / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get(f"{spec.name}:{args.precision}:{kname}")
        except Exception:
            traffic = None
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "fp16": "f16-storage/f32"}[args.precision],
        "data": "synthetic",
        "config": {"workload": spec.name, "grid": list(setup.shape), "cells": N, "npml": setup.npml,
                   "precision": args.precision, "gain_cells": gain_cells, "materials": info["n_materials"],
                   "mat_id_bytes": info["mat_id_bytes"], "setup_s": round(t_setup, 2),
                   "kernel_mode": ["h_e_kernels", "fused_step", "flow_step"][args.mode], "fused_param": args.ichunk,
                   "wavefront_kwidth": {"H": args.kwidth or 32, "E": args.kwidth_e or args.kwidth or 32}
                   if args.mode == 0 else None},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": kbytes, "avg_launch_us": round(kt * 1e6, 2),
                     "launches_timed": int(n_timed)},
        "kernels_ms_per_step": kernels,
        # the whole step against the fused minimum: step_bytes (the bytes a
        # step must move per cell, fields read and written once, times the
        # cells) per measured step
        "step_roofline": {"bound": "hbm", "alg_bytes_per_cell": round(step_bytes / N, 2),
                          "alg_bytes_per_step": step_bytes,
                          "achieved": round(step_bytes / (wall / args.steps) / 1e9, 1), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(step_bytes / (wall / args.steps) / 1e9 / HBM_PEAK_GBS, 4)},
    }
    if traffic:
        line["roofline"]["traffic_over_algorithmic"] = round(traffic / kbytes, 3)
    sc.close()
    if not args.no_cpu:
        if args.workload == "c3":
            sample = pkg.scene.precision_case(steps=5000, shape=(128, 128, 384))
        else:
            sample = spec
        line["cpu_baseline"] = cpu_baseline(pkg, sample, args.cpu_steps, args.cpu_steps_1t)
    print(json.dumps(line), flush=True)


def _probe_errors(a, b, spec):
    """Relative L2 of each probe's recorded vector (all its components) and of
    each DFT spectrum, scene a against reference scene b."""
    out = {}
    c = 0
    for p, pr in enumerate(spec.probes):
        idx = list(range(c, c + len(pr.components)))
        c += len(pr.components)
        x = np.stack([a.probe(q) for q in idx])
        y = np.stack([b.probe(q) for q in idx])
        out[f"probe{p}"] = float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))
    for m in range(len(spec.dfts)):
        x, y = a.dft(m), b.dft(m)
        out[f"dft{m}"] = float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))
    return out


def precision_form(args):
    """FP32 and FP16-storage throughput on the ~100M-cell precision geometry
    (or, with --workload c2, the 256^3 graded gain slab), and their error against the FP64 device path (the oracle at this size,
    itself pinned to the CPU oracle at smaller sizes) after args.steps steps."""
    pkg = importlib.import_module("3d_fdtd_version008_amd")
    if args.workload == "c3":
        spec = pkg.scene.precision_case(steps=args.steps)
    else:
        spec = pkg.scene.gain_slab_case(n_total=args.size, npml=12, steps=args.steps)
    t0 = time.time()
    setup = pkg.Setup(spec)
    t_setup = time.time() - t0
    N = setup.shape[0] * setup.shape[1] * setup.shape[2]
    rates, scenes = {}, {}
    for prec in ("fp64", "fp32", "fp16"):
        sc = pkg.Scene(setup, prec, probe_capacity=args.steps + 16)
        sc.set_kernel_mode(args.mode, _kw_param(args) if args.mode == 0 else args.ichunk)
        done = 0
        if prec != "fp64":
            sc.step(args.warmup)
            sc.synchronize()
            t1 = time.perf_counter()
            sc.step(args.rate_steps)
            sc.synchronize()
            rates[prec] = N * args.rate_steps / (time.perf_counter() - t1) / 1e9
            done = args.warmup + args.rate_steps
            print(f"[precision] {prec} {rates[prec]:.2f} Gcell/s", file=sys.stderr, flush=True)
        while done < args.steps:
            k = min(500, args.steps - done)
            sc.step(k)
            done += k
            print(f"[precision] {prec} {done}/{args.steps}", file=sys.stderr, flush=True)
        scenes[prec] = sc
    err = {p: _probe_errors(scenes[p], scenes["fp64"], spec) for p in ("fp32", "fp16")}
    line = {"metric": "precision form: Gcell-updates/s and rel L2 vs FP64 per probe and per spectrum",
            "form": "precision", "config": {"workload": spec.name, "grid": list(setup.shape), "cells": N,
                                            "steps": args.steps, "setup_s": round(t_setup, 1),
                                            "kernel_mode": args.mode},
            "gcells_per_s": {k: round(v, 3) for k, v in rates.items()},
            "fp16_over_fp32": round(rates["fp16"] / rates["fp32"], 3),
            "rel_l2_vs_fp64": err}
    for sc in scenes.values():
        sc.close()
    print(json.dumps(line), flush=True)


def _solo_rate(pkg, specs, precision, steps, mode):
    """One-at-a-time plain solves: scene creation plus stepping, one request
    after another; returns (cell-updates, seconds)."""
    work, secs = 0, 0.0
    for sp in specs:
        st = pkg.Setup(sp)
        t0 = time.perf_counter()
        sc = pkg.Scene(st, precision, probe_capacity=steps + 16)
        sc.set_kernel_mode(mode)
        sc.step(steps)
        sc.synchronize()
        secs += time.perf_counter() - t0
        work += st.shape[0] * st.shape[1] * st.shape[2] * steps
        sc.close()
    return work, secs


def batch_form(args):
    """256 uneven domains advanced by persistent batch launches (HIP-graph
    blocks of 50 steps, autotuned tiling) against one-at-a-time solves."""
    pkg = importlib.import_module("3d_fdtd_version008_amd")
    specs = [pkg.scene.batch_case(i, seed=1, steps=args.steps) for i in range(args.domains)]
    t0 = time.time()
    setups = [pkg.Setup(sp) for sp in specs]
    t_setup = time.time() - t0
    cells = sum(st.shape[0] * st.shape[1] * st.shape[2] for st in setups)
    t_submit = time.perf_counter()  # every domain is submitted here
    b = pkg.Batch(args.precision)
    scenes = [pkg.Scene(st, args.precision, probe_capacity=args.steps + 200) for st in setups]
    for sc in scenes:
        b.add(sc)
    tuning, taken = b.autotune(trial_steps=2)
    b.set_graph(True)
    b.step(50)
    t1 = time.perf_counter()
    dev_ms = 0.0
    left = args.steps
    while left > 0:
        k = min(50, left)
        dev_ms += b.step(k)
        left -= k
    wall = time.perf_counter() - t1
    rate = cells * args.steps / wall / 1e9
    # solve latency of each domain: submission to its results read back
    lat = []
    for sc in scenes:
        for c in range(sc.setup.n_channels):
            sc.probe(c)
        lat.append(time.perf_counter() - t_submit)
    b.close()
    for sc in scenes:
        sc.close()
    sample = specs[:: max(1, len(specs) // args.solo_sample)][: args.solo_sample]
    w, s = _solo_rate(pkg, sample, args.precision, args.solo_steps, 0)
    solo = w / s / 1e9
    print(json.dumps({"metric": "batch form: Gcell-updates/s of 256 uneven domains", "form": "batch",
                      "value": round(rate, 3), "unit": "Gcell-updates/s", "steps": args.steps,
                      "ms_per_step": round(wall / args.steps * 1e3, 4), "device_ms_per_step": round(dev_ms / args.steps, 4),
                      "latency_p50_s": round(float(np.percentile(lat, 50)), 3),
                      "latency_p99_s": round(float(np.percentile(lat, 99)), 3),
                      "latency_note": "every domain runs the same step count in one batch, so all finish in the "
                                      "last block; latency = scene creation, autotune, warm-up block and stepping, "
                                      "then the read-back of the domains before it",
                      "config": {"domains": len(specs), "cells": cells, "precision": args.precision,
                                 "tuning_rows_wgpercu_chunk_planes": list(tuning), "autotune_steps": taken,
                                 "graph_block": 50, "setup_s": round(t_setup, 1)},
                      "one_at_a_time": {"value": round(solo, 3), "sample_domains": len(sample),
                                        "steps_each": args.solo_steps, "kernel_mode": "h_e_kernels"},
                      "speedup_vs_one_at_a_time": round(rate / solo, 2)}), flush=True)


def serve_form(args):
    """A Poisson stream of mixed FP32/FP16 requests (48^3..224^3, 500..4000
    steps, 40% dielectric, 40% gain, 20% multi-block) under continuous
    batching with a 200 GB admission budget, against one-at-a-time solves."""
    pkg = importlib.import_module("3d_fdtd_version008_amd")
    serve = importlib.import_module("3d_fdtd_version008_amd.serve")
    reqs = serve.request_stream(args.requests, rate=args.rate, seed=1)
    srv = serve.Server(budget_bytes=200e9, max_active=256, tick=args.tick, graph=True, setup_workers=8,
                       engine=args.engine, stream_workers=args.workers, defer_free=args.defer_free,
                       pool_bytes=int(args.pool_gb * (1 << 30)),
                       host_bake=args.host_bake, step_graphs=not args.no_step_graphs)
    res = srv.run(reqs)
    summ = serve.summarize(res)
    sample = reqs[:: max(1, len(reqs) // args.solo_sample)][: args.solo_sample] if args.solo_sample else []
    # the one-at-a-time baseline: the same requests solved one after another
    # on the plain path (scene creation plus stepping), each at its own
    # storage precision (matched accuracy), and all in FP32 for comparison
    solo = {}
    for key in ("matched", "fp32"):
        work, secs = 0, 0.0
        for r in sample:
            w, s = _solo_rate(pkg, [r.spec], r.precision if key == "matched" else "fp32", r.steps, 0)
            work += w
            secs += s
        solo[key] = work / secs / 1e9 if secs else None
    line = {"metric": "serve form: Gcell-updates/s, p50/p99 latency of a mixed request stream",
            "form": "serve", "value": round(summ["gcells_per_s"], 3), "unit": "Gcell-updates/s",
            "latency_p50_s": round(summ["latency_p50_s"], 3), "latency_p99_s": round(summ["latency_p99_s"], 3),
            "config": {"requests": args.requests, "arrival_rate_per_s": args.rate,
                       "fp16_share": 0.5, "budget_gb": 200, "pool_gb": args.pool_gb, "tick_steps": args.tick, "engine": args.engine,
                       "workers": args.workers, "deferred_free": args.defer_free,
                       "step_graphs": srv.step_graphs,
                       "bake": "host" if args.host_bake else "device", "makespan_s": round(summ["makespan_s"], 2),
                       "cell_updates": summ["cell_updates"], "failed": summ["failed"],
                       "gpu_max_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "runtime default")},
            "server_time_s": {k: round(v, 3) for k, v in srv.stats.items()}}
    if sample:
        line["one_at_a_time"] = {"value": round(solo["matched"], 3), "sample_requests": len(sample),
                                 "precision": "each request's own", "kernel_mode": "h_e_kernels",
                                 "value_all_fp32": round(solo["fp32"], 3)}
        line["speedup_vs_one_at_a_time"] = round(summ["gcells_per_s"] / solo["matched"], 2)
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--form", default="plain", choices=["plain", "precision", "batch", "serve"])
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--precision", default="fp32", choices=["fp64", "fp32", "fp16"])
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--workload", default="c3", choices=["c2", "c3"],
                    help="plain form: c3 = 512x512x384 graded grid with gain slab and 64 cuboids (default), "
                         "c2 = size^3 graded gain slab")
    ap.add_argument("--event-every", type=int, default=0,
                    help="plain form: HIP events around the kernels of every n-th timed step (0: 10, or fewer "
                         "for short runs so that at least 20 launches of each kernel are timed)")
    ap.add_argument("--cpu-steps", type=int, default=80, help="CPU baseline: timed FP64 steps, all threads")
    ap.add_argument("--cpu-steps-1t", type=int, default=12, help="CPU baseline: timed FP64 steps on one thread")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", type=int, default=0,
                    help="0: separate H/E kernels, 1: fused step kernel, 2: dependency-ordered step "
                         "(--ichunk: its E-tile plane lag, 0 = 2)")
    ap.add_argument("--ichunk", type=int, default=0,
                    help="fused kernel parameter (planes per work-group + 4096 * cells per lane + 65536 * rows); "
                         "0 = library default")
    ap.add_argument("--kwidth", type=int, default=0, help="H/E kernels: wavefront k-width (16, 32, 64); 0 = library default")
    ap.add_argument("--kwidth-e", type=int, default=0, help="E kernel's own wavefront k-width; 0 = same as --kwidth")
    ap.add_argument("--offsets64", action="store_true", help="H/E kernels with 64-bit element offsets")
    ap.add_argument("--band", type=int, default=0,
                    help="H/E kernels: j-tiles (of 4 rows) per band of the tile order; 0 = library default")
    ap.add_argument("--rate-steps", type=int, default=200, help="precision form: timed steps per format")
    ap.add_argument("--domains", type=int, default=256, help="batch form: domains")
    ap.add_argument("--requests", type=int, default=64, help="serve form: requests")
    ap.add_argument("--rate", type=float, default=20.0, help="serve form: arrivals per second")
    ap.add_argument("--engine", default="streams", choices=["streams", "batch", "hybrid"], help="serve form: request engine")
    ap.add_argument("--pool-gb", type=float, default=24.0,
                    help="serve form: device memory cache for freed requests' blocks (GiB; 0: hipFree at each retirement)")
    ap.add_argument("--solo-sample", type=int, default=16, help="batch/serve forms: one-at-a-time sample size")
    ap.add_argument("--tick", type=int, default=50, help="serve form: steps per tick")
    ap.add_argument("--no-step-graphs", action="store_true",
                    help="serve form: requests on streams launch every step directly instead of replaying graphs")
    ap.add_argument("--workers", type=int, default=4,
                    help="serve form, streams engine: threads for the per-request device calls (0: the loop thread; with threads the requests launch directly instead of replaying recorded ticks)")
    ap.add_argument("--defer-free", action="store_true",
                    help="serve form: free finished requests' device memory only when admission needs it")
    ap.add_argument("--host-bake", action="store_true", help="serve form: bake materials on the host during setup")
    ap.add_argument("--solo-steps", type=int, default=200, help="batch form: steps per one-at-a-time solve")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a known-size kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this build measures one GPU (N=1).",
              file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = {"plain": 1000, "precision": 4000, "batch": 2000, "serve": 0}[args.form]
    if args.event_every <= 0:
        args.event_every = max(1, min(10, args.steps // 20))
    {"plain": plain, "precision": precision_form, "batch": batch_form, "serve": serve_form}[args.form](args)


if __name__ == "__main__":
    main()
