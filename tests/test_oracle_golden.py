"""The CPU oracle against the golden vectors written by the compiled reference
(tests/golden/*.npz, from oracle/_ref), and against the compiled reference
itself where it is built.  Expected agreement: bitwise (both sides FP64,
-ffp-contract=off, same operation order)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import cases  # noqa: E402


def _golden(name):
    return np.load(os.path.join(HERE, "golden", f"{name}.npz"))


@pytest.mark.parametrize("name", list(cases.CASES))
def test_oracle_matches_golden_bitwise(oracle_mod, name):
    sp = cases.CASES[name]()
    g = _golden(name)
    if "x_bounds" in g.files:  # auto-meshed: the case's lines are the reference generator's
        for ax in "xyz":
            assert np.array_equal(getattr(sp, f"{ax}_bounds"), g[f"{ax}_bounds"]), ax
    sim = oracle_mod.CpuSim(sp, "oracle")
    assert tuple(g["shape"]) == sim.shape
    assert float(g["dt"]) == sim.dt
    sim.step(int(g["steps"]))
    for f in cases.FIELDS:
        assert np.array_equal(sim.array(f), g[f]), f
    if "Nu" in g.files:
        for f in cases.TLS_FIELDS:
            assert np.array_equal(sim.array(f), g[f]), f
    c = 0
    for p in range(len(sp.probes)):
        ijk, nc = sim.probe_index(p)
        for q in range(nc):
            assert tuple(g[f"probe{c}_ijk"]) == ijk
            assert np.array_equal(sim.probe(p, q), g[f"probe{c}"]), f"probe{c}"
            c += 1
    st = sim.stability()
    assert int(g["stable"]) == int(st["stable"])
    assert float(g["max_E"]) == st["max_E"]


def test_golden_fixtures_nontrivial():
    """Every fixture carries signal (a silent all-zero run would pin nothing)."""
    for name in cases.CASES:
        g = _golden(name)
        assert np.abs(g["probe0"]).max() > 0, name
        assert max(np.abs(g[f]).max() for f in cases.FIELDS) > 0, name
        if "Nu" in g.files:
            assert np.abs(g["Pz"]).max() > 0, name


@pytest.mark.parametrize("name", ["uniform_sphere", "gain_clamp_dipole", "plane_waves", "auto_mesh_cylinder"])
def test_oracle_matches_compiled_reference(oracle_mod, name):
    if not oracle_mod.available("ref"):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    sp = cases.CASES[name]()
    a = oracle_mod.CpuSim(sp, "oracle")
    b = oracle_mod.CpuSim(sp, "ref")
    for sim in (a, b):
        sim.step(sp.n_steps + 37)
    for f in cases.FIELDS + ("psi_Ex_y", "psi_Hz_x", "psi_Ey_z", "bEx", "bEz", "bHy", "kx", "cpml_cHz"):
        try:
            ref = b.array(f)
        except KeyError:
            continue
        assert np.array_equal(a.array(f), ref), f
