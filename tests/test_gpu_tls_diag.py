"""TLS diagnostics as device reductions (PopulationTimeSeriesDetector::record
and PolarizationDiagnostics::record, src/two_level_detectors.hpp:70-103,
230-257).  Each record must equal, on the state the device holds at that
step, the reference's quantities: every total is checked against math.fsum of
the same per-cell terms (the device rounds its double-double sum once, so it
is the correctly rounded sum: equal to within one ulp) and against the
reference's sequential idx3-order sum (within that sum's own rounding bound),
and the monitor cell's Pz and dPz/dt bitwise.  The FP64 state itself is
pinned to the oracle by the parity tests."""
import math
import os
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _case(pkg):
    return pkg.scene.gain_slab_case(n_total=40, npml=6, steps=40, fine_len=100e-9, coarse=20e-9, fine=5e-9,
                                    n_cuboids=3, seed=3)


def _expected(setup, sc):
    shape = setup.shape
    region = setup.array("region_map").reshape(shape) >= 0
    dx, dy, dz = (setup.array(n, shaped=False) for n in ("dx", "dy", "dz"))
    dV = (dx[:, None, None] * dy[None, :, None]) * dz[None, None, :]
    Nu, Ng, Pz, Ez = (sc.field(n) for n in ("Nu", "Ng", "Pz", "Ez"))
    terms = [(Nu * dV)[region], (Ng * dV)[region], (((-0.5 * Pz) * Ez) * dV)[region]]
    out = []
    for t in terms:
        seq = 0.0
        for x in t.tolist():  # the reference's sequential order (idx3 = C order)
            seq += x
        out.append((math.fsum(t.tolist()), seq, float(np.abs(t).sum()), t.size))
    return out


def _check_totals(rec, exp):
    for q, (fs, seq, mag, n) in enumerate(exp):
        got = rec[q]
        ulp = abs(np.spacing(fs)) if fs != 0 else 0.0
        assert abs(got - fs) <= ulp + 1e-28 * mag, (q, got, fs)
        assert abs(got - seq) <= 2.0 * n * np.finfo(float).eps * mag + ulp, (q, got, seq)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tls_records_match_state(pkg, precision):
    sp = _case(pkg)
    setup = pkg.Setup(sp)
    region = setup.array("region_map").reshape(setup.shape) >= 0
    cells = np.argwhere(region)
    ci, cj, ck = (int(v) for v in cells[len(cells) // 2])
    coeffs = setup.array("tls_coeffs", shaped=False)
    rmap = setup.array("region_map").reshape(setup.shape)
    inv_dt = coeffs[13 * int(rmap[ci, cj, ck]) + 3]
    _say(f"[case] {sp.name} {precision} monitor cell {(ci, cj, ck)}")
    sc = pkg.Scene(setup, precision)
    sc.add_tls_monitor(ci, cj, ck, save_every=1, capacity=64)
    for n in range(4):
        sc.step(1)
        rec = sc.tls_records()
        assert rec.shape == (n + 1, 5)
        _check_totals(rec[n], _expected(setup, sc))
        pz, pzp = sc.field("Pz")[ci, cj, ck], sc.field("Pz_prev")[ci, cj, ck]
        assert rec[n, 3] == pz
        assert rec[n, 4] == (pz - pzp) * inv_dt
    sc.close()


def test_tls_records_cadence_and_outside_cell(pkg):
    sp = _case(pkg)
    setup = pkg.Setup(sp)
    sc = pkg.Scene(setup, "fp64")
    sc.add_tls_monitor(1, 1, 1, save_every=7, capacity=64)  # a CPML cell: no gain there
    sc.step(29)  # steps 0..28: records at 0, 7, 14, 21, 28
    rec = sc.tls_records()
    assert rec.shape == (5, 5)
    _check_totals(rec[-1], _expected(setup, sc))
    assert np.all(rec[:, 3] == 0) and np.all(rec[:, 4] == 0)
    # populations stay positive and the totals move
    assert np.all(rec[:, 0] > 0) and np.all(rec[:, 1] > 0)
    d = __import__("importlib").import_module("3d_fdtd_version008_amd.detectors")
    with tempfile.TemporaryDirectory() as t:
        paths = d.write_tls_csv(t, rec, setup.dt, 7)
        lines = open(paths[0]).read().splitlines()
        assert lines[0] == "t, total_Nu, total_Ng, inversion, inv_fraction" and len(lines) == 6
        assert float(lines[-1].split(",")[1]) == rec[-1, 0]
        assert all(os.path.exists(p) for p in paths)
    sc.close()
