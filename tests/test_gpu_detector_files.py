"""Detector files written from the device: the compiled reference harness
(oracle/_ref, with the reference's own PointFieldDetector and FieldMovie2D,
src/detectors/point_field_detector.hpp, src/detectors/field_movie_2d.hpp)
writes its files while it steps; the FP64 device path steps the same scene,
and 3d_fdtd_version008_amd/detectors.py writes the same files from the
device's probe records and field snapshots.  Every file must be identical
byte for byte (metadata.json text, *_ts.bin traces, frame_NNNN.raw slices),
on a graded mesh, for the three slice planes, derived components (|E|, S_z,
|S|), both output widths and a shifted domain origin."""
import importlib
import os
import tempfile

import numpy as np

import pytest

from conftest import assert_trees_equal, detector_case, frame_steps

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


@pytest.mark.parametrize("save_every,float64,dmin", [(1, True, (0.0, 0.0, 0.0)),
                                                     (3, False, (1.1e-7, -4.3e-8, 2.9e-8))])
def test_device_detector_files_equal_reference(pkg, oracle_mod, save_every, float64, dmin):
    if not oracle_mod.available("ref"):
        pytest.skip("compiled reference harness (oracle/_ref) not built")
    det = importlib.import_module("3d_fdtd_version008_amd.detectors")
    sp, setup, L, movies = detector_case(pkg, save_every)
    nphys = (sp.nx_core, sp.ny_core, sp.nz_core)
    p0 = sp.probes[0]
    print(f"[case] {sp.name} save_every={save_every} float64={float64}", flush=True)
    with tempfile.TemporaryDirectory() as ref_root, tempfile.TemporaryDirectory() as our_root:
        ref = oracle_mod.CpuSim(sp, impl="ref")
        ref.add_point_detector(ref_root, "probe", p0.x, p0.y, p0.z, list(p0.components), save_every, sp.n_steps,
                               float64)
        for name, comp, plane, pos in movies:
            ref.add_movie(ref_root, name, comp, plane, pos + dmin[2 - plane], save_every, sp.n_steps, float64, nphys,
                          L, dmin)
        ref.step(sp.n_steps)
        ref.close()

        sc = pkg.Scene(setup, "fp64", probe_capacity=sp.n_steps + 4)
        writers = [det.FieldMovie2DWriter(our_root, name, setup, comp, plane, pos + dmin[2 - plane], save_every,
                                          sp.n_steps, float64, nphys, L, dmin) for name, comp, plane, pos in movies]
        for k, take in frame_steps(sp.n_steps, save_every):
            sc.step(k)
            if take:
                f = {q: sc.field(q) for q in FIELDS}
                for w in writers:
                    w.frame(f)
        det.write_point_probe(our_root, "probe", setup, setup.probe(0)[0], p0.components,
                              [sc.probe(c) for c in range(len(p0.components))], save_every, sp.n_steps, float64)
        sc.close()
        files = assert_trees_equal(ref_root, our_root)
        assert len([f for f in files if f.startswith("movie_xy_Ez" + os.sep + "frame_")]) == \
            len(range(0, sp.n_steps, save_every))


def _csv_rows(path):
    with open(path) as f:
        head = f.readline()
        return head, [[float(x) for x in line.split(",")] for line in f.read().splitlines() if line]


@pytest.mark.parametrize("save_every", [1, 4])
def test_device_tls_files_equal_reference(pkg, oracle_mod, save_every):
    """main's TLS detectors (PopulationTimeSeriesDetector, PolarizationDiagnostics,
    TLSGlobalHistory, src/two_level_detectors.hpp; recorded after the
    population update, src/3D_FDTD_version008.cpp:330-337) run by the compiled
    reference, against the files written from the FP64 device run: the
    population frames and both metadata files byte for byte; the CSVs value
    by value - t, Pz and dPz/dt at the monitor cell exactly, the integrated
    totals to the reference's sequential-sum rounding (the device sum is the
    correctly rounded one), the fractions and fixed-point columns exactly
    where their digits allow."""
    if not oracle_mod.available("ref"):
        pytest.skip("compiled reference harness (oracle/_ref) not built")
    det = importlib.import_module("3d_fdtd_version008_amd.detectors")
    sp = pkg.scene.gain_slab_case(n_total=40, npml=6, steps=24, fine_len=100e-9, coarse=20e-9, fine=5e-9,
                                  n_cuboids=3, seed=3)
    setup = pkg.Setup(sp)
    region = setup.array("region_map").reshape(setup.shape) >= 0
    cells = np.argwhere(region)
    ic, jc, kc = (int(v) for v in cells[len(cells) // 2])
    steps = sp.n_steps
    print(f"[case] {sp.name} save_every={save_every} monitor {(ic, jc, kc)}", flush=True)
    with tempfile.TemporaryDirectory() as ref_root, tempfile.TemporaryDirectory() as our_root:
        ref = oracle_mod.CpuSim(sp, impl="ref")
        ref.add_tls_detectors(ref_root, ic, jc, kc, save_every, steps)
        ref.step(steps)
        ref.close()

        sc = pkg.Scene(setup, "fp64", probe_capacity=steps + 4)
        sc.add_tls_monitor(ic, jc, kc, save_every=save_every, capacity=steps + 4)
        frames = det.PopulationFrames(our_root, setup, save_every, steps)
        sc.step(1)
        frames.record(0, sc.field("Nu"), sc.field("Ng"))
        sc.step(steps - 1)
        frames.finalize(sc.field("Nu"), sc.field("Ng"))
        det.write_tls_csv(our_root, sc.tls_records(), setup.dt, save_every)
        sc.close()

        for f in ("populations/populations_Nu.bin", "populations/populations_Ng.bin",
                  "populations/metadata_populations.json"):
            a = open(os.path.join(ref_root, f), "rb").read()
            b = open(os.path.join(our_root, f), "rb").read()
            assert len(a) > 0 and a == b, f
        nrec = len(range(0, steps, save_every))
        for f, exact, sums in (("populations/inversion_time.csv", (0,), (1, 2, 3)),
                               ("polarization/polarization_time.csv", (0, 1, 2), (3,)),
                               ("tls_global_history.csv", (0, 1), (2, 3, 4))):
            ha, ra = _csv_rows(os.path.join(ref_root, f))
            hb, rb = _csv_rows(os.path.join(our_root, f))
            assert ha == hb and len(ra) == len(rb) == nrec, f
            ra, rb = np.array(ra), np.array(rb)
            for c in exact:
                assert np.array_equal(ra[:, c], rb[:, c]), (f, c)
            scale = np.abs(ra[:, sums]).max(axis=1, keepdims=True)
            assert np.all(np.abs(ra[:, sums] - rb[:, sums]) <= 1e-12 * scale), f
            np.testing.assert_allclose(rb, ra, rtol=1e-9, atol=1e-12 * float(np.abs(ra).max()))
