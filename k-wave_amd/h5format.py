"""HDF5 input/output files of the k-Wave binary contract (ctypes over libkwave_io.so).

* write_input(path, f)  -- writeMatrix.m for every dataset of the dict + writeAttributes.m
* read_file(path)       -- h5read of every root dataset (C-order arrays)
* run_file(in, out,...) -- kw_run_file in-process (what the executable does)
* run_binary(in, out,...) -- the kspaceFirstOrder-HIP executable through a process
                           boundary, as kspaceFirstOrder3DC.m:544 runs it
* kspace_first_order(kgrid, medium, source, sensor, ...) -- the
  kspaceFirstOrder2DG/3DG (and, axisymmetric=True, kspaceFirstOrderASC) flow:
  SaveToDisk input, run the binary, read back sensor_data shaped as MATLAB
  returns it (sensor_data.p is [Ns, Nt]).
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

from . import frontend
from .native import RECORD_FLAGS, output_names

PKG = os.path.dirname(os.path.abspath(__file__))
IO_LIB_PATH = os.path.join(PKG, "lib", "libkwave_io.so")
BINARY_PATH = os.path.join(PKG, "bin", "kspaceFirstOrder-HIP")

KWIO_FLOAT32, KWIO_UINT64 = 0, 1

IO_EXPORTED = ("kw_run_file", "kwio_open", "kwio_close", "kwio_write", "kwio_write_root_attributes",
               "kwio_info", "kwio_info_nd", "kwio_read", "kwio_exists", "kwio_list", "kwio_read_string_attribute",
               "kwio_last_error")


class KwRunOptions(C.Structure):
    _fields_ = [("record_flags", C.c_uint64), ("record_start_index", C.c_uint64), ("device", C.c_int32),
                ("verbose", C.c_int32), ("compression_level", C.c_int32), ("reserved", C.c_int32),
                ("t_load", C.c_double), ("t_precompute", C.c_double), ("t_loop", C.c_double),
                ("t_save", C.c_double), ("checkpoint_file", C.c_char_p), ("checkpoint_interval", C.c_double),
                ("checkpoint_timesteps", C.c_uint64), ("checkpointed", C.c_int32), ("reserved2", C.c_int32)]


_io = None


def load_io_library(path: str = IO_LIB_PATH):
    global _io
    if _io is not None:
        return _io
    if not os.path.exists(path):
        raise RuntimeError(f"HDF5 I/O library missing: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    vp, cp, u64p = C.c_void_p, C.c_char_p, C.POINTER(C.c_uint64)
    lib.kwio_open.argtypes = [cp, C.c_int, C.POINTER(vp)]
    lib.kwio_close.argtypes = [vp]
    lib.kwio_write.argtypes = [vp, cp, C.c_int, u64p, vp, C.c_int]
    lib.kwio_write_root_attributes.argtypes = [vp, cp, cp]
    lib.kwio_info.argtypes = [vp, cp, C.POINTER(C.c_int), u64p]
    lib.kwio_read.argtypes = [vp, cp, vp, C.c_size_t]
    lib.kwio_exists.argtypes = [vp, cp]
    lib.kwio_info_nd.argtypes = [vp, cp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_uint64)]
    lib.kwio_info_nd.restype = C.c_int
    lib.kwio_list.argtypes = [vp, C.c_char_p, C.c_size_t]
    lib.kwio_list.restype = C.c_size_t
    lib.kwio_read_string_attribute.argtypes = [vp, cp, cp, C.c_char_p, C.c_size_t]
    lib.kwio_last_error.restype = cp
    lib.kw_run_file.argtypes = [cp, cp, C.POINTER(KwRunOptions)]
    for n in ("kwio_open", "kwio_close", "kwio_write", "kwio_write_root_attributes", "kwio_info", "kwio_read",
              "kwio_read_string_attribute", "kw_run_file"):
        getattr(lib, n).restype = C.c_int
    _io = lib
    return lib


class H5Error(RuntimeError):
    pass


def _chk(lib, st):
    if st != 0:
        raise H5Error(f"kwio status {st}: {lib.kwio_last_error().decode(errors='replace')}")


class H5File:
    def __init__(self, path: str, mode: str = "r"):
        self.lib = load_io_library()
        h = C.c_void_p()
        _chk(self.lib, self.lib.kwio_open(path.encode(), 1 if mode == "w" else 0, C.byref(h)))
        self.h = h

    def write(self, name: str, arr, compression_level: int = 0):
        a = np.asarray(arr)
        if a.dtype == np.uint64 or np.issubdtype(a.dtype, np.integer):
            a, dt = np.ascontiguousarray(a, dtype=np.uint64), KWIO_UINT64
        else:
            a, dt = np.ascontiguousarray(a, dtype=np.float32), KWIO_FLOAT32
        shape = (1,) * (3 - a.ndim) + a.shape if a.ndim <= 3 else None
        if shape is None:
            raise ValueError(f"{name}: at most 3 dimensions")
        dims = (C.c_uint64 * 3)(*shape)
        _chk(self.lib, self.lib.kwio_write(self.h, name.encode(), dt, dims, a.ctypes.data, compression_level))

    def attributes(self, file_type: str, description: str = ""):
        _chk(self.lib, self.lib.kwio_write_root_attributes(self.h, file_type.encode(), description.encode()))

    def names(self):
        n = self.lib.kwio_list(self.h, None, 0)
        buf = C.create_string_buffer(n)
        self.lib.kwio_list(self.h, buf, n)
        return [s for s in buf.value.decode().split("\n") if s]

    def info(self, name: str):
        """(dtype, C-order dims): 3-D for the k-Wave datasets, 4-D for per-cuboid series."""
        dt, rank = C.c_int(), C.c_int()
        dims = (C.c_uint64 * 8)()
        _chk(self.lib, self.lib.kwio_info_nd(self.h, name.encode(), C.byref(dt), C.byref(rank), dims))
        d = tuple(int(dims[i]) for i in range(rank.value))
        return dt.value, (1,) * (3 - len(d)) + d if len(d) < 3 else d

    def read(self, name: str) -> np.ndarray:
        dt, dims = self.info(name)
        out = np.empty(dims, dtype=np.float32 if dt == KWIO_FLOAT32 else np.uint64)
        _chk(self.lib, self.lib.kwio_read(self.h, name.encode(), out.ctypes.data, out.nbytes))
        return out

    def attribute(self, name: str, obj: str = "/") -> str:
        buf = C.create_string_buffer(4096)
        _chk(self.lib, self.lib.kwio_read_string_attribute(self.h, obj.encode(), name.encode(), buf, 4096))
        return buf.value.decode()

    def close(self):
        if self.h:
            self.lib.kwio_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def write_input(path: str, f: dict, compression_level: int = 0,
                description: str = "Input data created by the k-wave_amd front end"):
    """Write an input-file dataset dict (saveToDisk.m:435-485)."""
    with H5File(path, "w") as h:
        for name in sorted(f):
            h.write(name, f[name], compression_level)
        h.attributes("input", description)


def read_file(path: str) -> dict:
    with H5File(path, "r") as h:
        return {n: h.read(n) for n in h.names()}


def run_file(inp: str, out: str, record=("p",), record_start_index: int = 1, device: int = 0,
             verbose: int = 0, checkpoint_file: str | None = None, checkpoint_interval: float = 0.0,
             checkpoint_timesteps: int = 0) -> KwRunOptions:
    """kw_run_file; with a checkpoint file the call may stop early (o.checkpointed == 1)."""
    lib = load_io_library()
    o = KwRunOptions()
    for r in record:
        o.record_flags |= RECORD_FLAGS[r]
    o.record_start_index = record_start_index
    o.device = device
    o.verbose = verbose
    if checkpoint_file:
        o.checkpoint_file = checkpoint_file.encode()
        o.checkpoint_interval = float(checkpoint_interval)
        o.checkpoint_timesteps = int(checkpoint_timesteps)
    _chk(lib, lib.kw_run_file(inp.encode(), out.encode(), C.byref(o)))
    return o


BINARY_FLAG = {"p": "--p_raw", "u": "--u_raw", "u_non_staggered": "--u_non_staggered_raw"}


def binary_options(record=("p",), record_start_index=None, device=None):
    """Command-line options as kspaceFirstOrder3DC.m:308-524 assembles them."""
    opts = []
    if device is not None:
        opts += ["-g", str(device)]
    for r in record:
        opts.append(BINARY_FLAG.get(r, "--" + r))
    if record_start_index is not None:
        opts += ["-s", str(record_start_index)]
    return opts


def run_binary(inp: str, out: str, record=("p",), record_start_index=None, device=None, binary=BINARY_PATH,
               timeout=None, extra=()) -> subprocess.CompletedProcess:
    """Run the executable with LD_LIBRARY_PATH cleared, from its own directory
    (kspaceFirstOrder3DC.m:544); `extra` appends options (e.g. checkpointing)."""
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = ""
    cmd = [binary, "-i", inp, "-o", out] + binary_options(record, record_start_index, device) + list(extra)
    return subprocess.run(cmd, cwd=os.path.dirname(binary), env=env, capture_output=True, text=True,
                          timeout=timeout)


def sensor_data_from_output(o: dict, record, dim: int) -> dict:
    """Output-file datasets -> MATLAB-shaped sensor_data (3DC.m:568-778)."""
    sd = {}
    for name in output_names(record, dim):
        a = o[name]
        if name in ("p", "ux", "uy", "uz") or name.endswith("_non_staggered"):
            sd[name] = a.reshape(a.shape[-2], a.shape[-1]).T          # [Ns, Nt_rec]
        elif name.endswith("_all") or name.endswith("_final"):
            sd[name] = a.transpose(2, 1, 0) if dim == 3 else a[0].T    # MATLAB [Nx, Ny(, Nz)]
        else:
            sd[name] = a.reshape(-1)                                     # [Ns]
    return sd


def kspace_first_order(kgrid, medium, source, sensor, record=("p",), record_start_index=None, pml_size=None,
                       pml_alpha=2.0, smooth_p0=True, data_path=None, device=None, delete_data=True,
                       binary=BINARY_PATH, axisymmetric=False):
    """kspaceFirstOrder2DG/3DG (kspaceFirstOrderASC with axisymmetric=True): write
    the input file, run the binary, read the output.  'I' / 'I_avg' in the record
    ask the binary for p and u_non_staggered and form the intensity on the host
    (3DC.m:508-511, 786-790)."""
    want_I = "I" in record or "I_avg" in record
    record_bin = tuple(r for r in record if r not in ("I", "I_avg"))
    if want_I:
        record_bin += tuple(r for r in ("p", "u_non_staggered") if r not in record_bin)
    f = frontend.save_to_disk_dict(kgrid, medium, source, sensor, pml_size=pml_size, pml_alpha=pml_alpha,
                                   smooth_p0=smooth_p0, axisymmetric=axisymmetric)
    d = data_path or tempfile.mkdtemp(prefix="kwave_")
    inp, out = os.path.join(d, "kwave_input.h5"), os.path.join(d, "kwave_output.h5")
    write_input(inp, f)
    r = run_binary(inp, out, record_bin, record_start_index, device, binary)
    if r.returncode != 0:
        raise H5Error(f"kspaceFirstOrder-HIP failed ({r.returncode}): {r.stderr.strip()}")
    sd = sensor_data_from_output(read_file(out), record_bin, kgrid.dim)
    if want_I:
        sd = frontend.save_intensity(sd, kgrid.dim, record)
    if delete_data:
        for p in (inp, out):
            if os.path.exists(p):
                os.remove(p)
    return sd
