"""Build recipe for the native parts (gfx950 HIP + C++), in-tree.

Outputs (git-ignored, shipped next to the sources):
  k-wave_amd/lib/libkwave_hip.so     solver C ABI (include/kwave_hip.h)
  k-wave_amd/lib/libkwave_io.so      HDF5 file contract (include/kwave_io.h)
  k-wave_amd/bin/kspaceFirstOrder-HIP  drop-in executable for the MATLAB wrappers

HDF5 comes from the image's /opt/conda (headers + libhdf5.so), linked with a
RUNPATH because the MATLAB wrapper clears LD_LIBRARY_PATH before running the
binary (kspaceFirstOrder3DC.m:544).
"""

from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
INCLUDE = os.path.join(REPO, "include")
HDF5_ROOT = os.environ.get("KW_HDF5_ROOT", "/opt/conda")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off (they have given wrong results on these GPUs)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # multiply-adds fuse only within one source expression, so the rounding of a
    # kernel body does not depend on which kernel it is inlined into (the
    # single-simulation and batched wrappers share every body)
    "-ffp-contract=on",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]

SOLVER_SOURCES = (["kw_kernels.hip", "kw_sim.cpp", "kw_server.cpp", "kw_plane.hip", "kw_axisym.hip", "kw_noalloc.hip"]
                  + [f"kw_pencil_nt{n}.hip" for n in (64, 128, 256, 512, 1024)]
                  + [f"kw_xrow_nt{n}.hip" for n in (64, 128, 256)])
IO_SOURCES = ["kw_io.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out:
        print(out, flush=True)


def _obj(src):
    return os.path.join(LIB, "obj", os.path.splitext(src)[0] + ".o")


def hdf5_available() -> bool:
    return os.path.exists(os.path.join(HDF5_ROOT, "include", "hdf5.h")) and os.path.exists(
        os.path.join(HDF5_ROOT, "lib", "libhdf5.so"))


def build(verbose: bool = False, with_io: bool = True) -> dict:
    """Compile every native component; returns the paths that were built."""
    os.makedirs(os.path.join(LIB, "obj"), exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    inc = [f"-I{INCLUDE}", f"-I{CSRC}"]
    objs = [_obj(src) for src in SOLVER_SOURCES]
    # translation units compile in parallel (the length-specialised kernels dominate)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(len(objs), os.cpu_count() or 1)) as ex:
        futs = [ex.submit(_run, [HIPCC, *HIP_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", o], verbose)
                for src, o in zip(SOLVER_SOURCES, objs)]
        for fu in futs:
            fu.result()
    solver = os.path.join(LIB, "libkwave_hip.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", solver + ".tmp",
          "-Wl,-soname,libkwave_hip.so", "-pthread"], verbose)
    os.replace(solver + ".tmp", solver)
    built = {"solver": solver}
    if with_io and hdf5_available() and all(os.path.exists(os.path.join(CSRC, s)) for s in IO_SOURCES):
        h5inc = os.path.join(HDF5_ROOT, "include")
        h5lib = os.path.join(HDF5_ROOT, "lib")
        io = os.path.join(LIB, "libkwave_io.so")
        ioobjs = []
        for src in IO_SOURCES:
            o = _obj(src)
            _run([HIPCC, *HIP_FLAGS, *inc, f"-I{h5inc}", "-c", os.path.join(CSRC, src), "-o", o], verbose)
            ioobjs.append(o)
        rpath = f"-Wl,-rpath,$ORIGIN:/opt/rocm/lib:/usr/lib/x86_64-linux-gnu:{h5lib}"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *ioobjs, "-o", io + ".tmp",
              "-Wl,-soname,libkwave_io.so", f"-L{LIB}", "-lkwave_hip", f"-L{h5lib}", "-lhdf5", rpath], verbose)
        os.replace(io + ".tmp", io)
        built["io"] = io
        main = os.path.join(CSRC, "kw_main.cpp")
        if os.path.exists(main):
            exe = os.path.join(BIN, "kspaceFirstOrder-HIP")
            _run([HIPCC, f"--offload-arch={ARCH}", "-O2", "-std=c++17", *inc, main, "-o", exe + ".tmp",
                  f"-L{LIB}", "-lkwave_io", "-lkwave_hip",
                  f"-Wl,-rpath,$ORIGIN/../lib:/opt/rocm/lib:/usr/lib/x86_64-linux-gnu:{h5lib}"], verbose)
            os.replace(exe + ".tmp", exe)
            built["binary"] = exe
    shutil.rmtree(os.path.join(LIB, "obj"), ignore_errors=True)
    return built


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
