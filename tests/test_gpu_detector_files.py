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
