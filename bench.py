"""Benchmark of the per-timestep Yee update on one MI355X.

Workload (value): the graded-mesh gain case, 256x256x256 total cells (12-cell
CPML, 10 nm cells in a central 400 nm z slab graded at 1.2 per cell to 40 nm,
20 nm lateral), eps_r 3.5 substrate, 200 nm two-level gain slab, Ex current
sheet source, probes and a 3-frequency DFT plane, stability monitor every 10
steps, FP32 storage/arithmetic with FP64 populations.  A "step" is one full
Yee step of the whole grid (H, CPML-H, E with J, CPML-E, gain, probes, DFT,
monitor).  Inputs are resident in HBM before the timed region.

Prints ONE JSON line (rank 0, N=1).  --gpus N > 1 is not supported: this
build runs on one GPU.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
METRIC = "Gcell-updates/s & % HBM roofline; p50/p99 latency of batched solves"


def algorithmic_bytes(spec, setup, prec_bytes: int, mat_bytes: int):
    """Algorithmic HBM bytes per step of the H kernel and of the E kernel.

    Per cell: H kernel reads E, H (6 values) and writes H (3): 9 s; E kernel
    reads H, E (6), writes E (3) and reads one material id: 9 s + m.  Each
    CPML slab cell adds, per active axis, 2 psi values read+written (4 s) in
    each kernel.  Each gain cell adds Pz, Pz_prev, Nu, Ng (FP64) read and
    written plus a region byte: 65 B, in the E kernel."""
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
    return h_bytes, e_bytes, gain_cells


def cpu_baseline(spec, steps: int):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # checker / CPU baseline only

    if not oracle.available("oracle"):
        oracle.build()
    sim = oracle.CpuSim(spec, "oracle")
    sim.time_steps(2)
    secs = sim.time_steps(steps)
    n = spec.n_cells()
    rate = n * steps / secs / 1e9
    sim.close()
    return {"value": rate, "unit": "Gcell-updates/s", "cores": oracle.threads(), "kind": "port",
            "sample": f"{spec.name}: {steps} FP64 steps of the full grid after 2 warm-up steps "
                      f"(oracle/fdtd_oracle.c, OpenMP over i, serial CPML, J zero-fill, monitor every 10)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--precision", default="fp32", choices=["fp64", "fp32", "fp16"])
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this build measures one GPU (N=1).",
              file=sys.stderr)
        sys.exit(2)

    pkg = importlib.import_module("3d_fdtd_version008_amd")
    spec = pkg.scene.gain_slab_case(n_total=args.size, npml=12, steps=5000)
    t0 = time.time()
    setup = pkg.Setup(spec)
    t_setup = time.time() - t0
    sc = pkg.Scene(setup, args.precision, probe_capacity=args.steps + args.warmup + 16)
    info = sc.info()
    sc.step(args.warmup)
    sc.synchronize()
    t1 = time.perf_counter()
    ms_h, ms_e, ms_call = sc.step_timed(args.steps)
    sc.synchronize()
    wall = time.perf_counter() - t1
    N = setup.shape[0] * setup.shape[1] * setup.shape[2]
    value = N * args.steps / wall / 1e9
    prec_bytes = {"fp64": 8, "fp32": 4, "fp16": 2}[args.precision]
    h_bytes, e_bytes, gain_cells = algorithmic_bytes(spec, setup, prec_bytes, info["mat_id_bytes"])
    t_h = ms_h / args.steps / 1e3
    t_e = ms_e / args.steps / 1e3
    if t_e >= t_h:
        kname, kbytes, kt = "k_update_E", e_bytes, t_e
    else:
        kname, kbytes, kt = "k_update_H", h_bytes, t_h
    achieved = kbytes / kt / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tfile):
        try:
            tj = json.load(open(tfile))
            traffic = tj.get(f"{spec.name}:{args.precision}:{kname}")
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
                   "mat_id_bytes": info["mat_id_bytes"], "setup_s": round(t_setup, 2)},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": kbytes, "avg_launch_us": round(kt * 1e6, 2)},
        "kernels_ms_per_step": {"k_update_H": round(ms_h / args.steps, 4), "k_update_E": round(ms_e / args.steps, 4)},
        "step_bytes_fraction": round((h_bytes + e_bytes) / (wall / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
    }
    sc.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(spec, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
