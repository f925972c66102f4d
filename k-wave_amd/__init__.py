"""MI355X-native k-space pseudospectral time loop (k-Wave kspaceFirstOrder2D/3D).

Imported as ``kwave_amd`` (the directory name carries a hyphen; entry points
register it under that name).  Layout:

* ``csrc/``       HIP kernels for gfx950 + C++ host orchestration + C ABI
* ``native``      ctypes binding of include/kwave_hip.h (Simulation, run, Server)
* ``configs``     seeded synthetic workloads (C2, C4, the C5 serving stream)
* ``frontend``    k-Wave style inputs -> input-file dataset dict
* ``geometry``    makeDisc/makeBall/makeCartCircle/toneBurst/smooth
* ``h5format``    HDF5 input/output files (libkwave_io.so)
* ``build``       in-tree build recipe
"""

from . import configs, frontend, geometry  # noqa: F401
from .native import Server, Simulation, footprint_bytes, load_library, run  # noqa: F401

__all__ = ["Simulation", "Server", "run", "footprint_bytes", "load_library", "frontend", "geometry", "configs"]
