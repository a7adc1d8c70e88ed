"""Writes tests/golden/auto_mesh.npz: for every configuration of
mesh_cases.all_cases(), the configuration itself and the x, y, z grid lines
the reference's own AutoMeshGenerator::generate returns for it
(oracle/_ref/libref_mesh.so, compiled from the reference tree by
oracle/Makefile).  Run from the repository root:
    make -C oracle && python tests/golden/make_mesh_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import oracle  # noqa: E402
import mesh_cases  # noqa: E402

N_RANDOM, N_NEAR = 120, 60

if __name__ == "__main__":
    if not os.path.exists(oracle.REF_MESH_SO):
        sys.exit("oracle/_ref/libref_mesh.so is not built (needs the reference tree; run make -C oracle)")
    out = {}
    names = []
    for name, case in mesh_cases.all_cases(N_RANDOM, N_NEAR).items():
        cfg, dom, npml, boxes, bg = case
        lines = oracle.ref_auto_mesh(cfg, dom, npml, boxes, bg, cap=200000)
        c, b, o = mesh_cases.pack(case)
        out[f"{name}:cfg"], out[f"{name}:boxes"], out[f"{name}:overrides"] = c, b, o
        for a, ax in enumerate("xyz"):
            out[f"{name}:{ax}"] = lines[a]
        names.append(name)
    out["names"] = np.array(names)
    path = os.path.join(HERE, "auto_mesh.npz")
    np.savez_compressed(path, **out)
    print(f"auto_mesh.npz: {len(names)} cases, {os.path.getsize(path) / 1024:.0f} KiB")
