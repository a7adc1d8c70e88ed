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


def test_library_is_gfx950(pkg, tmp_path):
    """The library's device code is one gfx950 code object: the offload
    bundle in its .hip_fatbin section lists the gfx950 target and the host
    entry only."""
    fatbin = tmp_path / "fatbin.bin"
    subprocess.run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin", pkg.runtime.library_path(), str(fatbin)],
                   check=True)
    out = subprocess.run(["/opt/rocm/llvm/bin/clang-offload-bundler", "--list", "--type=o", f"--input={fatbin}"],
                         capture_output=True, text=True, check=True).stdout.split()
    device = [t for t in out if not t.startswith("host-")]
    assert device == ["hipv4-amdgcn-amd-amdhsa--gfx950"], out


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


def test_device_pool_setting_round_trip(pkg):
    """The device memory cache's cap is process-wide state: set, read back,
    refused when negative, and back to off; nothing is cached without blocks."""
    rt = pkg.runtime
    before = rt.device_pool_cap()
    try:
        rt.set_device_pool(1 << 30)
        assert rt.device_pool_cap() == 1 << 30
        assert rt.device_pool_bytes() == 0
        with pytest.raises(rt.FdtdError):
            rt.set_device_pool(-1)
        assert rt.device_pool_cap() == 1 << 30
        rt.set_device_pool(0)
        assert rt.device_pool_cap() == 0
    finally:
        rt.set_device_pool(before)
