"""The drop-in boundary exercised the way the reference's main() would use
it: the reference's own setup code (its headers compiled into
oracle/_ref/libref_dropin.so) builds GridSpacing, MaterialGrids, the CPML
profiles, the TLS regions and state, the sources and the probes; the adaptor
of include/fdtd_gpu_stepper.hpp (fdtd_gpu::GpuStepper, compiled with the
reference's headers) takes those objects, binds the sources (footprints, and
per-step values from each source's own waveform, with no N-sized J array)
and the point detectors, and advances the state with GpuStepper::run: blocks
of at most 10 steps (one fdtd_step call each), cut after every step a host
detector records (here every 7th step, offset 3, as a field movie would).
At each such step the downloaded fields are compared with the reference's
CPU step brought to the same step in the same process.  In the device-bake form the adaptor passes the
reference's structure list (StructureScene::items, bg) instead of its baked
MaterialGrids, and the handle's material arrays are compared with the
reference's bake too.

Bar: FP64 fields (at every host-detector step and at the end), TLS arrays
and probe traces within 1e-12 relative L2 of the reference's own step."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
NAMES = ["uniform_sphere", "gain_slab", "gain_clamp_dipole", "plane_waves", "pec_cavity", "graded_cuboids",
         "auto_mesh_cylinder"]


@pytest.mark.parametrize("device_bake", [False, True], ids=["host_bake", "device_bake"])
@pytest.mark.parametrize("name", NAMES)
def test_reference_main_with_device_step_fp64(pkg, oracle_mod, name, device_bake):
    import cases

    if not os.path.exists(oracle_mod.REF_DROPIN_SO):
        pytest.skip("oracle/_ref/libref_dropin.so is not built (needs the reference tree at build time)")
    sp = cases.CASES[name]()
    print(f"[dropin] {name} fp64 steps={sp.n_steps} device_bake={device_bake}", flush=True)
    nch, rel = oracle_mod.ref_dropin_compare(sp, sp.n_steps, 0, device_bake)
    print(f"[dropin] {name}: worst relative L2 {max(rel):.3e} over fields, TLS arrays, {nch} probe channels"
          f" and material arrays")
    assert nch == sum(len(p.components) for p in sp.probes)
    assert max(rel) <= 1e-12, dict(zip("Ex Ey Ez Hx Hy Hz Nu Ng Pz Pz_prev probes materials".split(), rel))
