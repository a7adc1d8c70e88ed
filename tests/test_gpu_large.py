"""A grid of more than 2^31 cells (1320^3 = 2.3e9, 55 GB of FP32 fields):
every cell offset beyond 2^31 must be formed in 64 bits, in
the H and E kernels, the probes, the DFT monitor and the stability monitor.

Bar: the Yee update spreads a point source by one cell per step in each
direction, so after 16 steps every nonzero field lies within 17 cells of the
dipole.  On a uniform vacuum grid with the dipole 20 cells from the CPML,
that neighbourhood evolves exactly as on a 64^3 grid with the dipole at its
centre: same cell spacing, dt and source values, and zeros elsewhere.  The
probe traces (6 cells x 6 components), the DFT accumulators on an 8^3 box
round the dipole and the monitor maxima (over all 2.3e9 cells) of the two
grids must be bitwise equal.  The dipole sits at cell (1290, 1290, 1290),
whose offset 2.25e9 exceeds 2^31.

A 1640^3 grid (4.41e9 cells) also takes more than 2^32 threads per H or E
launch (17.5M tiles of 256), so its planes go into the second grid
dimension (every work-group but the first plane's 10,660 has blockIdx.y >
0); its dipole at (1610, 1610, 1610) has the offset 4.33e9, beyond 2^32.
The same bitwise bar holds."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D, NPML, STEPS = 20e-9, 10, 16
OFFSETS = [(0, 0, 0), (3, -2, 5), (-7, 4, 1), (10, 0, -9), (-12, -12, 12), (15, 1, 0)]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cone_case(S, n, src):
    c = lambda q: S.cell_center_phys(q, D, NPML)  # noqa: E731
    probes = [S.Probe(c(src[0] + a), c(src[1] + b), c(src[2] + e), (0, 1, 2, 3, 4, 5)) for (a, b, e) in OFFSETS]
    box = (src[0] - 4, src[0] + 4, src[1] - 4, src[1] + 4, src[2] - 4, src[2] + 4)
    dip = S.Dipole(c(src[0]), c(src[1]), c(src[2]), amplitude=1e-7, wavelength=1.0e-6, tau=3e-15, t0_factor=3.0,
                   waveform=S.WF_GAUSS_MOD_SINE, polarization=S.POL["Ez"])
    core = n - 2 * NPML
    return S.SceneSpec(nx_core=core, ny_core=core, nz_core=core, npml=NPML, dx_uniform=D, dy_uniform=D, dz_uniform=D,
                       S=0.99, cpml_m=3.0, cpml_rerr=1e-8, cpml_alpha0=0.05, cpml_kappa_max=5.0, dipoles=[dip],
                       probes=probes, save_every=1, dfts=[S.DFTMonitor(2, box, [2.5e14, 3.0e14])], device_bake=1,
                       n_steps=STEPS, name=f"light_cone_{n}")


def _run(pkg, n, src, precision):
    setup = pkg.Setup(_cone_case(pkg.scene, n, src))
    for p, (a, b, e) in enumerate(OFFSETS):  # the probes sit on the intended cells
        assert setup.probe(6 * p)[0] == (src[0] + a, src[1] + b, src[2] + e)
    _say(f"[case] light cone {n}^3 {precision} src={src}")
    sc = pkg.Scene(setup, precision)
    recs = []
    for _ in range(STEPS):
        sc.step(1)
        recs += [(r.step, r.max_E, r.max_H) for r in sc.checks()]
    out = {"probes": [sc.probe(c) for c in range(setup.n_channels)], "dft": sc.dft(0), "checks": recs}
    sc.close()
    setup.close()
    return out


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_grid_beyond_2_31_cells_matches_small_grid(pkg, precision):
    n, src = 1320, (1290, 1290, 1290)
    assert n ** 3 > 2 ** 31 and (src[0] * n + src[1]) * n + src[2] > 2 ** 31
    assert src[0] + 17 < n - NPML and src[0] - 17 > NPML  # the cone stays clear of the CPML
    small = _run(pkg, 64, (32, 32, 32), precision)
    big = _run(pkg, n, src, precision)
    assert len(small["checks"]) == 2 and small["checks"] == big["checks"], (small["checks"], big["checks"])
    assert small["checks"][-1][1] > 0
    for c, (a, b) in enumerate(zip(small["probes"], big["probes"])):
        assert len(a) == STEPS and np.array_equal(a, b), f"probe channel {c}"
    assert np.abs(small["dft"]).max() > 0 and np.array_equal(small["dft"], big["dft"])


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_launch_beyond_2_32_threads_matches_small_grid(pkg, precision):
    n, src = 1640, (1610, 1610, 1610)
    tiles = ((n + 63) // 64) * ((n + 3) // 4) * n
    assert tiles * 256 >= 2 ** 32 and (src[0] * n + src[1]) * n + src[2] > 2 ** 32
    assert src[0] + 17 < n - NPML and src[0] - 17 > NPML
    small = _run(pkg, 64, (32, 32, 32), precision)
    big = _run(pkg, n, src, precision)
    assert len(small["checks"]) == 2 and small["checks"] == big["checks"], (small["checks"], big["checks"])
    assert small["checks"][-1][1] > 0
    for c, (a, b) in enumerate(zip(small["probes"], big["probes"])):
        assert len(a) == STEPS and np.array_equal(a, b), f"probe channel {c}"
    assert np.abs(small["dft"]).max() > 0 and np.array_equal(small["dft"], big["dft"])
