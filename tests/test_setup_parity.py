"""The product's host setup (lib3dfdtd_hip.so, fdtd_setup_*: the C++ restatement
of make_context / CPML profiles / bake / TLS / sources / probes) against the
oracle's, bitwise.  CPU only: no device is touched."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import cases  # noqa: E402

SETUP_ARRAYS = ["aEx", "bEx", "aEy", "bEy", "aEz", "bEz", "aHx", "bHx", "aHy", "bHy", "aHz", "bHz",
                "dx", "dy", "dz", "inv_dx", "inv_dy", "inv_dz", "x_bounds", "y_bounds", "z_bounds",
                "kx", "ky", "kz", "cpml_bEx", "cpml_cEx", "cpml_bEy", "cpml_cEy", "cpml_bEz", "cpml_cEz",
                "cpml_bHx", "cpml_cHx", "cpml_bHy", "cpml_cHy", "cpml_bHz", "cpml_cHz",
                "Ndip", "Ng0", "Nu", "Ng"]


def _specs(pkg):
    out = [(n, f()) for n, f in cases.CASES.items()]
    out.append(("uniform_sphere_64", pkg.scene.uniform_sphere_case(core=64, npml=10, steps=10)))
    out.append(("gain_slab_64", pkg.scene.gain_slab_case(n_total=64, npml=8, steps=10, fine=5e-9, coarse=20e-9,
                                                          fine_len=100e-9, n_cuboids=8, seed=5)))
    return out


@pytest.fixture(scope="module")
def pairs(pkg, oracle_mod):
    return [(n, sp, pkg.Setup(sp), oracle_mod.CpuSim(sp, "oracle")) for n, sp in _specs(pkg)]


def test_grid_and_dt(pairs):
    for name, sp, s, o in pairs:
        assert s.shape == o.shape, name
        assert s.npml == o.npml, name
        assert s.dt == o.dt, name


def test_setup_arrays_bitwise(pairs):
    for name, sp, s, o in pairs:
        for a in SETUP_ARRAYS:
            if a in ("Ndip", "Ng0", "Nu", "Ng") and not s.tls_on:
                continue
            assert np.array_equal(s.array(a), o.array(a)), (name, a)
        if s.tls_on:
            assert np.array_equal(s.array("region_map").astype(np.int32), o.region_map()), name


def test_tls_coefficients(pairs):
    for name, sp, s, o in pairs:
        if not s.tls_on:
            continue
        c = s.array("tls_coeffs").reshape(-1, 13)
        for r in range(s.n_regions):
            ref = o.region(r)  # lambda0 omega_a gamma tau mu_z N0 hbar kapa pa2 pa3 inv_dt i0 i1 j0 j1 k0 k1
            assert c[r, 0] == ref[7] and c[r, 1] == ref[8] and c[r, 2] == ref[9] and c[r, 3] == ref[10], name
            assert c[r, 4] == ref[1] and c[r, 5] == ref[4] and c[r, 6] == ref[5], name
            assert tuple(c[r, 7:13]) == tuple(ref[11:17]), name


def test_source_values_bitwise(pairs):
    for name, sp, s, o in pairs:
        vals = s.source_values(0, 60)
        for q in range(s.n_sources):
            ref = np.array([o.source(q, n)[14] for n in range(60)])
            assert np.array_equal(vals[:, q], ref), (name, q)
        # continuation from a later step gives the same values
        later = s.source_values(40, 20)
        assert np.array_equal(later, vals[40:60]), name


def test_probe_placement(pairs):
    for name, sp, s, o in pairs:
        c = 0
        for p in range(len(sp.probes)):
            ijk, nc = o.probe_index(p)
            for q in range(nc):
                sijk, comp = s.probe(c)
                assert sijk == ijk and comp == sp.probes[p].components[q], (name, c)
                c += 1
        assert c == s.n_channels


def test_invalid_spec_rejected(pkg):
    sp = pkg.scene.uniform_sphere_case(core=8, npml=6, steps=1)
    bad = sp.copy(nx_core=0)
    with pytest.raises(pkg.FdtdError):
        pkg.Setup(bad)


def test_workload_generators_are_seeded(pkg):
    """The batching and serving workloads are pure functions of their seeds."""
    import importlib

    serve = importlib.import_module("3d_fdtd_version008_amd.serve")
    a = [pkg.scene.batch_case(i, seed=4).name for i in range(8)]
    b = [pkg.scene.batch_case(i, seed=4).name for i in range(8)]
    assert a == b and len(set(a)) == 8
    s1 = serve.request_stream(6, rate=2.0, seed=9)
    s2 = serve.request_stream(6, rate=2.0, seed=9)
    assert [(r.spec.name, r.precision, r.steps, r.arrival) for r in s1] == \
           [(r.spec.name, r.precision, r.steps, r.arrival) for r in s2]
    assert all(s1[q].arrival <= s1[q + 1].arrival for q in range(5))
    st = pkg.Setup(s1[0].spec)
    assert serve.estimate_device_bytes(st, "fp32") > serve.estimate_device_bytes(st, "fp16")


def test_reference_harness_with_prebaked_materials(pkg, oracle_mod):
    """The CPU baseline of a many-structure grid times the reference's step
    with the MaterialGrids arrays this library's bake produces, in place of
    the reference's serial bake.  On a graded grid with cuboids and gain the
    two harness states step to bitwise equal fields, TLS state and probes."""
    import oracle

    if not oracle.available("ref"):
        pytest.skip("oracle/_ref/libref.so not built")
    sp = pkg.scene.gain_slab_case(n_total=40, npml=6, steps=40, fine=5e-9, coarse=20e-9, fine_len=100e-9,
                                  n_cuboids=6, seed=4)
    setup = pkg.Setup(sp)
    mats = [setup.array(n) for n in ("aEx", "bEx", "aEy", "bEy", "aEz", "bEz", "aHx", "bHx", "aHy", "bHy", "aHz",
                                     "bHz")]
    a = oracle.CpuSim(sp, "ref")
    b = oracle.CpuSim(sp, "ref", materials=mats)
    assert a.step(sp.n_steps) == 0 and b.step(sp.n_steps) == 0
    for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Nu", "Ng", "Pz"):
        assert np.array_equal(a.array(f), b.array(f)), f
    assert np.abs(a.array("Ex")).max() > 0
    for p in range(len(sp.probes)):
        assert np.array_equal(a.probe(p, 0), b.probe(p, 0))
