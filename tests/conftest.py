import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def pkg():
    return importlib.import_module("3d_fdtd_version008_amd")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle  # noqa: E402

    if not oracle.available("oracle"):
        oracle.build()
    return oracle


def rel_l2(a, b):
    import numpy as np

    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    if den == 0.0:
        return num
    return num / den


def tree_files(root):
    """Relative paths of every file under root, sorted."""
    return sorted(os.path.relpath(os.path.join(d, f), root) for d, _, fs in os.walk(root) for f in fs)


def assert_trees_equal(ref_root, our_root):
    """The same files under both roots, each identical byte for byte."""
    ref_files, our_files = tree_files(ref_root), tree_files(our_root)
    assert ref_files == our_files
    for f in ref_files:
        a = open(os.path.join(ref_root, f), "rb").read()
        b = open(os.path.join(our_root, f), "rb").read()
        assert a == b, f"{f} differs:\n{a[:400]!r}\n{b[:400]!r}"
    return ref_files


def detector_case(pkg, save_every):
    """A small scene with one point probe (Ez, |E|, Sx, |S|) and four slice
    movies over the three planes; returns (spec, setup, physical extents,
    movie list (name, component, plane, position))."""
    S = pkg.scene
    import numpy as np

    sp = S.uniform_sphere_case(core=24, npml=6, steps=10)
    sp.save_every = save_every
    # graded spacing, so coordinates print with all 15 significant digits
    n = np.arange(sp.nx_core + 2 * sp.npml)
    sp.dx = 20e-9 * (1.0 + 0.21 * np.sin(0.37 * n + 0.1))
    sp.dy = 20e-9 * (1.0 + 0.17 * np.cos(0.29 * n))
    sp.dz = 20e-9 * (1.0 + 0.23 * np.sin(0.53 * n + 1.3))
    p0 = sp.probes[0]
    sp.probes = [S.Probe(p0.x, p0.y, p0.z, (2, 6, 8, 11))]
    setup = pkg.Setup(sp)
    npml = setup.npml
    L = tuple(float(b[len(b) - 1 - npml] - b[npml])
              for b in (setup.array(n, shaped=False) for n in ("x_bounds", "y_bounds", "z_bounds")))
    movies = [("movie_xy_Ez", 2, 0, 0.45 * L[2]), ("movie_xz_Smag", 11, 1, 0.3 * L[1]),
              ("movie_yz_Emag", 6, 2, 0.61 * L[0]), ("movie_xy_Sz", 10, 0, 0.5 * L[2])]
    return sp, setup, L, movies


def frame_steps(steps, save_every):
    """Step counts between frame points: a detector records after step n
    when n % save_every == 0, i.e. once n + 1 steps are done."""
    out, n = [], 0
    while n < steps:
        k = 1 if n % save_every == 0 else save_every - n % save_every
        k = min(k, steps - n)
        n += k
        out.append((k, (n - 1) % save_every == 0))
    return out
