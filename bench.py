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
    ap.add_argument("--mode", type=int, default=1, help="0: separate H/E kernels, 1: fused step kernel")
    ap.add_argument("--ichunk", type=int, default=0, help="fused kernel x-planes per work-group; 0 = library default")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a known-size kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this build measures one GPU (N=1).",
              file=sys.stderr)
        sys.exit(2)

    pkg = importlib.import_module("3d_fdtd_version008_amd")
    if args.copy_calib:
        copy_calibration(args.copy_calib << 20)
    spec = pkg.scene.gain_slab_case(n_total=args.size, npml=12, steps=5000)
    t0 = time.time()
    setup = pkg.Setup(spec)
    t_setup = time.time() - t0
    sc = pkg.Scene(setup, args.precision, probe_capacity=args.steps + args.warmup + 16)
    info = sc.info()
    sc.set_kernel_mode(args.mode, args.ichunk)
    sc.step(args.warmup)
    sc.synchronize()
    t1 = time.perf_counter()
    ms_h, ms_e, ms_call = sc.step_timed(args.steps)
    sc.synchronize()
    wall = time.perf_counter() - t1
    N = setup.shape[0] * setup.shape[1] * setup.shape[2]
    value = N * args.steps / wall / 1e9
    prec_bytes = {"fp64": 8, "fp32": 4, "fp16": 2}[args.precision]
    h_bytes, e_bytes, step_bytes, gain_cells = algorithmic_bytes(spec, setup, prec_bytes, info["mat_id_bytes"])
    t_h = ms_h / args.steps / 1e3
    t_e = ms_e / args.steps / 1e3
    if args.mode == 1:
        kname, kbytes, kt = "k_step_rows", step_bytes, t_h
        kernels = {kname: round(ms_h / args.steps, 4)}
    elif t_e >= t_h:
        kname, kbytes, kt = "k_update_E", e_bytes, t_e
    else:
        kname, kbytes, kt = "k_update_H", h_bytes, t_h
    if args.mode == 0:
        kernels = {"k_update_H": round(ms_h / args.steps, 4), "k_update_E": round(ms_e / args.steps, 4)}
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
                   "mat_id_bytes": info["mat_id_bytes"], "setup_s": round(t_setup, 2),
                   "kernel_mode": ["h_e_kernels", "fused_rows"][args.mode],
                   "planes_per_chunk": args.ichunk},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": kbytes, "avg_launch_us": round(kt * 1e6, 2)},
        "kernels_ms_per_step": kernels,
        "step_alg_bytes_per_cell": round(step_bytes / N, 2),
        "step_roofline_frac": round(step_bytes / (wall / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
    }
    sc.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(spec, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
