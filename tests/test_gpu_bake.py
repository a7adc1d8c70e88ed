"""Device-side StructureScene::bake (src/structure_material.hpp:211-288):
every MaterialGrids value the device bake produces must equal the host bake's
bit for bit (the host bake is pinned to the reference's fixture arrays in
test_setup_parity), the material table must hold the same distinct tuples,
and a run on device-baked materials must give the same fields as one on
host-baked materials.  Cases cover the three shape types, overlapping
structures (later ones override earlier ones), 1, 8 and 27 subpixel samples,
a graded mesh with a gain slab, and a scene with more than 256 distinct
tuples (16-bit material ids)."""
import copy
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ["aEx", "bEx", "aEy", "bEy", "aEz", "bEz", "aHx", "bHx", "aHy", "bHy", "aHz", "bHz"]


def _say(msg):
    print(msg, flush=True)


def _shapes_case(pkg, subpixel):
    S = pkg.scene
    sp = S.uniform_sphere_case(core=40, npml=6, steps=30)
    L = 40 * 20e-9
    sp.structures = list(sp.structures) + [
        S.Structure(S.SHAPE_CYLINDER_Z, 2.0, (0.3 * L, 0.6 * L, 0.17 * L, 0.1 * L, 0.8 * L, 0)),
        S.Structure(S.SHAPE_BOX, 1.2, (0.45 * L, 0.9 * L, 0.2 * L, 0.55 * L, 0.35 * L, 0.52 * L)),
        S.Structure(S.SHAPE_SPHERE, 3.1, (0.7 * L, 0.3 * L, 0.6 * L, 0.12 * L, 0, 0)),
    ]
    sp.subpixel_samples = subpixel
    sp.name = f"shapes_s{subpixel}"
    return sp


def _many_spheres_case(pkg):
    S = pkg.scene
    sp = S.uniform_sphere_case(core=48, npml=6, steps=20)
    L = 48 * 20e-9
    rng = np.random.default_rng(7)
    st = []
    for _ in range(40):
        c = rng.uniform(0.1, 0.9, size=3) * L
        st.append(S.Structure(S.SHAPE_SPHERE, float(rng.uniform(1.0, 3.5)), (c[0], c[1], c[2], 0.08 * L, 0, 0)))
    sp.structures = st
    sp.name = "many_spheres"
    return sp


def _gain_case(pkg):
    return pkg.scene.gain_slab_case(n_total=44, npml=6, steps=40, fine_len=100e-9, coarse=20e-9, fine=5e-9,
                                    n_cuboids=6, seed=5)


def _check(pkg, sp, steps):
    host = pkg.Setup(sp)
    spd = copy.copy(sp)
    spd.device_bake = 1
    dev = pkg.Setup(spd)
    a = pkg.Scene(host, "fp64")
    b = pkg.Scene(dev, "fp64")
    try:
        for n in NAMES:
            ref = host.array(n).reshape(a.shape)
            got_b = b.material(n)
            got_a = a.material(n)
            assert np.array_equal(got_a, ref), f"{sp.name}: host-baked {n} changed on the device"
            assert np.array_equal(got_b, ref), f"{sp.name}: device-baked {n} differs from the host bake"
        ia, ib = a.info(), b.info()
        assert ia["n_materials"] == ib["n_materials"], (ia, ib)
        assert ia["mat_id_bytes"] == ib["mat_id_bytes"], (ia, ib)
        a.step(steps)
        b.step(steps)
        for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
            assert np.array_equal(a.field(f), b.field(f)), f"{sp.name}: {f} differs after {steps} steps"
        return ib
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("subpixel", [1, 8, 27])
def test_device_bake_matches_host_shapes(pkg, subpixel):
    sp = _shapes_case(pkg, subpixel)
    _say(f"[case] {sp.name}")
    _check(pkg, sp, 20)


def test_device_bake_matches_host_graded_gain(pkg):
    sp = _gain_case(pkg)
    _say(f"[case] {sp.name}")
    _check(pkg, sp, 30)


def test_device_bake_wide_ids(pkg):
    sp = _many_spheres_case(pkg)
    _say(f"[case] {sp.name}")
    info = _check(pkg, sp, 10)
    assert info["n_materials"] > 256 and info["mat_id_bytes"] == 2, info


def test_device_bake_keeps_no_host_material_arrays(pkg):
    sp = _shapes_case(pkg, 8)
    sp.device_bake = 1
    st = pkg.Setup(sp)
    with pytest.raises(Exception):
        st.array("bEx")
