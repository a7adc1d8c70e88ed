"""MI355X (gfx950) per-timestep Yee update of 3D_FDTD_version008.

Package layout:
  csrc/fdtd_kernels.hpp  HIP kernels (fused H+CPML, fused E+J+CPML+TLS, probes, DFT, monitor)
  csrc/fdtd_device.hip   C ABI of the device engine (include/fdtd_hip.h)
  csrc/fdtd_setup.cpp    host setup restating the reference's params / bake / CPML / TLS / source path
  csrc/fdtd_automesh.cpp auto nonuniform mesh generator (include/fdtd_mesh.h)
  build.py               in-tree build of lib3dfdtd_hip.so
  scene.py               scenario description (include/fdtd_scene.h) and named workloads
  runtime.py             ctypes front-end (Setup, Scene, Batch)
  automesh.py            auto-mesher front-end (AutoMeshConfig, generate, auto_mesh_spec)
  detectors.py           detector files in the reference's formats
  serve.py               continuous-batching server over the device engine

The package directory name starts with a digit, so import it with
importlib.import_module("3d_fdtd_version008_amd").
"""
from . import scene  # noqa: F401
from . import runtime  # noqa: F401
from .runtime import Batch, Setup, Scene, FdtdError, load, library_path  # noqa: F401
