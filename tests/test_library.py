"""The C-ABI library loads and exports every entry point its headers declare;
without a device the product path fails loudly (no CPU fallback)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(fdtd_[a-z_0-9]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.runtime.library_path()
    assert os.path.exists(lib)
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (fdtd_[a-z_0-9]+)", out))
    declared = _declared()
    assert declared, "no declarations parsed"
    missing = declared - exported
    assert not missing, missing
    assert set(pkg.runtime.EXPORTS) <= exported


def test_library_is_gfx950(pkg):
    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", pkg.runtime.library_path()], capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("roc-obj-ls unavailable")
    assert "gfx950" in out.stdout


def test_no_cpu_fallback_without_device(pkg):
    import ctypes

    lib = pkg.load()
    n = ctypes.c_int()
    has_dev = False
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        has_dev = hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        pass
    if has_dev:
        pytest.skip("a device is present")
    sp = pkg.scene.uniform_sphere_case(core=8, npml=4, steps=2)
    setup = pkg.Setup(sp)
    with pytest.raises(pkg.FdtdError) as e:
        pkg.Scene(setup, "fp64")
    assert e.value.code == -2
    assert lib is not None
