"""FP16 storage study on the CPU (run by hand; not collected by pytest):
relative L2 of every probe vector and of the DFT spectrum against the FP64
oracle, for the emulated storage formats of oracle_set_storage
(oracle/fdtd_oracle.c), at checkpoints of a long run.

    python tests/fp16_storage_study.py [shape] [steps] [checkpoints]

Defaults: the precision geometry (graded z profile, gain slab, 64 cuboids)
on a 96x96x80 grid, 4000 steps, reported every 1000 steps, formats fp16,
fp16_residual8, fp16_psi32, fp16_stochastic (bfp16_8, bfp16_32, bfp16_64:
block floating point).  Prints one JSON line per (format, checkpoint)."""
import importlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402  (checker)


def rel(a, b):
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / d) if d > 0 else float(np.linalg.norm(a))


def main():
    pkg = importlib.import_module("3d_fdtd_version008_amd")
    shape = tuple(int(v) for v in (sys.argv[1] if len(sys.argv) > 1 else "96,96,80").split(","))
    steps = int(sys.argv[2]) if len(sys.argv) > 2 else 4000
    every = int(sys.argv[3]) if len(sys.argv) > 3 else 1000
    fmts = sys.argv[4].split(",") if len(sys.argv) > 4 else ["fp16", "fp16_residual8", "fp16_psi32", "fp16_stochastic"]
    spec = pkg.scene.precision_case(steps=steps, shape=shape)
    scales = pkg.Setup(spec).fp16_scales()
    sims = {"fp64": oracle.CpuSim(spec, "oracle")}
    for f in fmts:
        sims[f] = oracle.CpuSim(spec, "oracle")
        sims[f].set_storage(f, scales, seed=11)
    done = 0
    while done < steps:
        k = min(every, steps - done)
        for s in sims.values():
            s.step(k)
        done += k
        ref = sims["fp64"]
        for f in fmts:
            s = sims[f]
            probes = []
            for p, pr in enumerate(spec.probes):
                n = len(pr.components)
                probes.append(rel(np.stack([s.probe(p, q) for q in range(n)]),
                                  np.stack([ref.probe(p, q) for q in range(n)])))
            print(json.dumps({"shape": shape, "format": f, "steps": done, "probe_rel_l2": probes,
                              "dft_rel_l2": [rel(s.dft(m), ref.dft(m)) for m in range(len(spec.dfts))]}), flush=True)


if __name__ == "__main__":
    main()
