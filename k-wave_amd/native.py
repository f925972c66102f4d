"""ctypes binding of the solver C ABI (include/kwave_hip.h).

``Simulation`` takes the dataset dict of an input file (the same dict
``frontend.save_to_disk_dict`` builds and ``h5format`` reads/writes), creates
a device-resident simulation, runs it and fetches outputs named like the
datasets of the output file.  There is no fallback: if libkwave_hip.so is not
built, or the HIP runtime reports no device, the calls raise.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libkwave_hip.so")

KW_ABI_VERSION = 1

RECORD_FLAGS = {
    "p": 1 << 0, "p_max": 1 << 1, "p_min": 1 << 2, "p_rms": 1 << 3, "p_max_all": 1 << 4,
    "p_min_all": 1 << 5, "p_final": 1 << 6, "u": 1 << 7, "u_max": 1 << 8, "u_min": 1 << 9,
    "u_rms": 1 << 10, "u_max_all": 1 << 11, "u_min_all": 1 << 12, "u_final": 1 << 13,
    "u_non_staggered": 1 << 14,
}

# output ids (kw_output_id)
_OUT_P = {"p": 0, "p_max": 1, "p_min": 2, "p_rms": 3, "p_max_all": 4, "p_min_all": 5, "p_final": 6}
_OUT_U_BASE = {"": 7, "_max": 10, "_min": 13, "_rms": 16, "_max_all": 19, "_min_all": 22, "_final": 25,
               "_non_staggered": 28}


def _time_series(name: str) -> bool:
    """Outputs shaped (Nt_rec, Ns)."""
    return name in ("p", "ux", "uy", "uz") or name.endswith("_non_staggered")

EXPORTED = (
    "kw_sim_create", "kw_sim_run", "kw_sim_sync", "kw_sim_steps_done", "kw_sim_output_bytes",
    "kw_sim_fetch", "kw_sim_destroy", "kw_last_error", "kw_sim_set_profiling", "kw_sim_profile_get",
    "kw_sim_footprint_bytes", "kw_diag_copy", "kw_diag_copy_ex", "kw_sim_stream", "kw_sim_memory",
    "kw_sim_launch_config", "kw_sim_run_noalloc",
    "kw_sim_checkpoint_bytes", "kw_sim_checkpoint_save", "kw_sim_checkpoint_load", "kw_sim_autotune",
    "kw_server_create", "kw_server_submit", "kw_server_poll", "kw_server_wait", "kw_server_request_info",
    "kw_server_request_error", "kw_server_output_bytes", "kw_server_fetch", "kw_server_release",
    "kw_server_stats_get", "kw_server_destroy",
)

KW_ERR_TIMEOUT = 6
REQUEST_STATES = ("queued", "running", "done", "failed", "unknown")


class KwRealArray(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("count", C.c_uint64)]


class KwSimDesc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32), ("device", C.c_int32),
        ("Nx", C.c_uint64), ("Ny", C.c_uint64), ("Nz", C.c_uint64), ("Nt", C.c_uint64),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double), ("c_ref", C.c_double),
        ("pml_x_size", C.c_uint64), ("pml_y_size", C.c_uint64), ("pml_z_size", C.c_uint64),
        ("pml_x_alpha", C.c_double), ("pml_y_alpha", C.c_double), ("pml_z_alpha", C.c_double),
        ("c0", KwRealArray), ("rho0", KwRealArray), ("rho0_sgx", KwRealArray), ("rho0_sgy", KwRealArray),
        ("rho0_sgz", KwRealArray),
        ("nonlinear_flag", C.c_uint64), ("BonA", KwRealArray),
        ("absorbing_flag", C.c_uint64), ("alpha_coeff", KwRealArray), ("alpha_power", C.c_double),
        ("p0_source_input", C.POINTER(C.c_float)),
        ("p_source_flag", C.c_uint64), ("p_source_mode", C.c_uint64), ("p_source_many", C.c_uint64),
        ("p_source_npoints", C.c_uint64), ("p_source_index", C.POINTER(C.c_uint64)),
        ("p_source_input", C.POINTER(C.c_float)),
        ("ux_source_flag", C.c_uint64), ("uy_source_flag", C.c_uint64), ("uz_source_flag", C.c_uint64),
        ("u_source_mode", C.c_uint64), ("u_source_many", C.c_uint64), ("u_source_npoints", C.c_uint64),
        ("u_source_index", C.POINTER(C.c_uint64)),
        ("ux_source_input", C.POINTER(C.c_float)), ("uy_source_input", C.POINTER(C.c_float)),
        ("uz_source_input", C.POINTER(C.c_float)),
        ("sensor_npoints", C.c_uint64), ("sensor_mask_index", C.POINTER(C.c_uint64)),
        ("record_flags", C.c_uint64), ("record_start_index", C.c_uint64),
        ("transducer_source_flag", C.c_uint64), ("transducer_source_length", C.c_uint64),
        ("transducer_source_input", C.POINTER(C.c_float)), ("delay_mask", C.POINTER(C.c_uint64)),
        ("axisymmetric_flag", C.c_uint64),
    ]


class KwLaunchConfig(C.Structure):
    _fields_ = [("row_nt", C.c_int32), ("ypass_nt", C.c_int32), ("last_nt", C.c_int32), ("plane_nt", C.c_int32),
                ("ypass_plan", C.c_int32), ("last_plan", C.c_int32), ("kx_major", C.c_int32),
                ("row_wg_target", C.c_int32)]


class KwOutputBuffer(C.Structure):
    _fields_ = [("output_id", C.c_int32), ("reserved", C.c_int32), ("host_dst", C.c_void_p),
                ("bytes", C.c_size_t)]


class KwServerConfig(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("device", C.c_int32), ("hbm_budget_bytes", C.c_uint64),
                ("max_active", C.c_uint32), ("steps_per_quantum", C.c_uint32), ("autotune", C.c_uint32),
                ("batching", C.c_uint32)]


class KwRequestInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("status", C.c_int32), ("footprint_bytes", C.c_uint64),
                ("point_updates", C.c_uint64), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double)]


class KwServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("queued", C.c_uint64), ("active", C.c_uint64), ("hbm_budget_bytes", C.c_uint64),
                ("hbm_reserved_bytes", C.c_uint64), ("hbm_peak_reserved_bytes", C.c_uint64),
                ("point_updates_done", C.c_uint64), ("elapsed_s", C.c_double)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libkwave_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"native solver library missing: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    lib.kw_sim_create.argtypes = [C.POINTER(KwSimDesc), C.POINTER(C.c_void_p)]
    lib.kw_sim_create.restype = C.c_int
    lib.kw_sim_run.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
    lib.kw_sim_run.restype = C.c_int
    lib.kw_sim_sync.argtypes = [C.c_void_p]
    lib.kw_sim_sync.restype = C.c_int
    lib.kw_sim_steps_done.argtypes = [C.c_void_p]
    lib.kw_sim_steps_done.restype = C.c_uint64
    lib.kw_sim_output_bytes.argtypes = [C.c_void_p, C.c_int]
    lib.kw_sim_output_bytes.restype = C.c_size_t
    lib.kw_sim_fetch.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]
    lib.kw_sim_fetch.restype = C.c_int
    lib.kw_sim_destroy.argtypes = [C.c_void_p]
    lib.kw_sim_destroy.restype = C.c_int
    lib.kw_last_error.argtypes = []
    lib.kw_last_error.restype = C.c_char_p
    lib.kw_sim_set_profiling.argtypes = [C.c_void_p, C.c_int]
    lib.kw_sim_set_profiling.restype = C.c_int
    lib.kw_sim_profile_get.argtypes = [C.c_void_p, C.POINTER(C.c_char_p), C.POINTER(C.c_double),
                                       C.POINTER(C.c_uint64), C.c_int]
    lib.kw_sim_profile_get.restype = C.c_int
    lib.kw_sim_footprint_bytes.argtypes = [C.POINTER(KwSimDesc)]
    lib.kw_sim_footprint_bytes.restype = C.c_uint64
    lib.kw_diag_copy.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_double)]
    lib.kw_diag_copy.restype = C.c_int
    lib.kw_sim_autotune.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int)]
    lib.kw_sim_autotune.restype = C.c_int
    lib.kw_sim_stream.argtypes = [C.c_void_p]
    lib.kw_sim_stream.restype = C.c_void_p
    lib.kw_diag_copy_ex.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    lib.kw_diag_copy_ex.restype = C.c_int
    lib.kw_sim_memory.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    lib.kw_sim_memory.restype = C.c_int
    lib.kw_sim_launch_config.argtypes = [C.c_void_p, C.POINTER(KwLaunchConfig)]
    lib.kw_sim_launch_config.restype = C.c_int
    lib.kw_sim_run_noalloc.argtypes = [C.POINTER(KwSimDesc), C.c_void_p, C.POINTER(KwOutputBuffer), C.c_int]
    lib.kw_sim_run_noalloc.restype = C.c_int
    vp, u64 = C.c_void_p, C.c_uint64
    for name, args, res in (
            ("kw_server_create", [C.POINTER(KwServerConfig), C.POINTER(vp)], C.c_int),
            ("kw_server_submit", [vp, C.POINTER(KwSimDesc), C.POINTER(u64)], C.c_int),
            ("kw_server_poll", [vp, u64], C.c_int),
            ("kw_server_wait", [vp, u64, C.c_double], C.c_int),
            ("kw_server_request_info", [vp, u64, C.POINTER(KwRequestInfo)], C.c_int),
            ("kw_server_request_error", [vp, u64], C.c_char_p),
            ("kw_server_output_bytes", [vp, u64, C.c_int], C.c_size_t),
            ("kw_server_fetch", [vp, u64, C.c_int, vp, C.c_size_t], C.c_int),
            ("kw_server_release", [vp, u64], C.c_int),
            ("kw_server_stats_get", [vp, C.POINTER(KwServerStats)], C.c_int),
            ("kw_server_destroy", [vp], C.c_int)):
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


class KWaveError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"kw status {status}: {msg}")
        self.status = status


def _check(lib, st):
    if st != 0:
        raise KWaveError(st, lib.kw_last_error().decode(errors="replace"))


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1))


def _u64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64).reshape(-1))


def _int(v, default=0):
    if v is None:
        return default
    return int(np.asarray(v).reshape(-1)[0])


def _flt(v, default=0.0):
    if v is None:
        return default
    return float(np.asarray(v, dtype=np.float64).reshape(-1)[0])


def build_desc(f: dict, record=("p",), record_start_index: int = 1, device: int = 0):
    """kw_sim_desc from an input-file dataset dict; returns (desc, keepalive)."""
    keep = []

    def ptr_f(a):
        if a is None:
            return None
        arr = _f32(a)
        keep.append(arr)
        return arr.ctypes.data_as(C.POINTER(C.c_float))

    def ptr_u(a):
        if a is None:
            return None
        arr = _u64(a)
        keep.append(arr)
        return arr.ctypes.data_as(C.POINTER(C.c_uint64))

    def real(name):
        if name not in f:
            return KwRealArray(None, 0)
        arr = _f32(f[name])
        keep.append(arr)
        return KwRealArray(arr.ctypes.data_as(C.POINTER(C.c_float)), arr.size)

    d = KwSimDesc()
    d.abi_version = KW_ABI_VERSION
    d.device = device
    d.Nx, d.Ny, d.Nz, d.Nt = _int(f["Nx"]), _int(f["Ny"]), _int(f.get("Nz"), 1), _int(f["Nt"])
    d.dx, d.dy, d.dz = _flt(f["dx"]), _flt(f["dy"]), _flt(f.get("dz"), 0.0)
    d.dt, d.c_ref = _flt(f["dt"]), _flt(f["c_ref"])
    d.pml_x_size, d.pml_y_size, d.pml_z_size = (_int(f.get("pml_x_size")), _int(f.get("pml_y_size")),
                                                _int(f.get("pml_z_size")))
    d.pml_x_alpha, d.pml_y_alpha, d.pml_z_alpha = (_flt(f.get("pml_x_alpha")), _flt(f.get("pml_y_alpha")),
                                                   _flt(f.get("pml_z_alpha")))
    d.c0, d.rho0 = real("c0"), real("rho0")
    d.rho0_sgx, d.rho0_sgy, d.rho0_sgz = real("rho0_sgx"), real("rho0_sgy"), real("rho0_sgz")
    d.nonlinear_flag = _int(f.get("nonlinear_flag"))
    d.BonA = real("BonA")
    d.absorbing_flag = _int(f.get("absorbing_flag"))
    d.alpha_coeff = real("alpha_coeff")
    d.alpha_power = _flt(f.get("alpha_power"))
    if _int(f.get("p0_source_flag")):
        d.p0_source_input = ptr_f(f["p0_source_input"])
    d.p_source_flag = _int(f.get("p_source_flag"))
    if d.p_source_flag:
        d.p_source_mode = _int(f.get("p_source_mode"))
        d.p_source_many = _int(f.get("p_source_many"))
        d.p_source_npoints = np.asarray(f["p_source_index"]).size
        d.p_source_index = ptr_u(f["p_source_index"])
        d.p_source_input = ptr_f(f["p_source_input"])
    d.ux_source_flag = _int(f.get("ux_source_flag"))
    d.uy_source_flag = _int(f.get("uy_source_flag"))
    d.uz_source_flag = _int(f.get("uz_source_flag"))
    if d.ux_source_flag or d.uy_source_flag or d.uz_source_flag:
        d.u_source_mode = _int(f.get("u_source_mode"))
        d.u_source_many = _int(f.get("u_source_many"))
        d.u_source_npoints = np.asarray(f["u_source_index"]).size
        d.u_source_index = ptr_u(f["u_source_index"])
        for k in ("ux", "uy", "uz"):
            if _int(f.get(k + "_source_flag")):
                setattr(d, k + "_source_input", ptr_f(f[k + "_source_input"]))
    d.axisymmetric_flag = _int(f.get("axisymmetric_flag"))
    d.transducer_source_flag = _int(f.get("transducer_source_flag"))
    if d.transducer_source_flag:
        sig = _f32(f["transducer_source_input"])
        d.transducer_source_length = sig.size
        keep.append(sig)
        d.transducer_source_input = sig.ctypes.data_as(C.POINTER(C.c_float))
        d.delay_mask = ptr_u(f["delay_mask"])
        d.u_source_npoints = np.asarray(f["u_source_index"]).size
        d.u_source_index = ptr_u(f["u_source_index"])
    if _int(f.get("sensor_mask_type")) == 1:
        raise ValueError("cuboid-corner sensor masks are split per cuboid by the file path (kw_run_file / "
                         "kspaceFirstOrder-HIP); in process, pass the cuboids' points as a binary mask")
    sm = f["sensor_mask_index"]
    d.sensor_npoints = np.asarray(sm).size
    d.sensor_mask_index = ptr_u(sm)
    flags = 0
    for r in record:
        if r not in RECORD_FLAGS:
            raise ValueError(f"unknown record option {r}")
        flags |= RECORD_FLAGS[r]
    d.record_flags = flags
    d.record_start_index = int(record_start_index)
    return d, keep


def output_names(record, dim: int):
    names = []
    for r in record:
        if r.startswith("u"):
            suffix = r[1:]
            names += ["u" + a + suffix for a in "xyz"[:dim]]
        else:
            names.append(r)
    return names


def output_id(name: str) -> int:
    if name in _OUT_P:
        return _OUT_P[name]
    if name[:2] in ("ux", "uy", "uz"):
        axis = "xyz".index(name[1])
        return _OUT_U_BASE[name[2:]] + axis
    raise ValueError(f"unknown output {name}")


class Simulation:
    """One device-resident simulation (kw_sim handle)."""

    def __init__(self, f: dict, record=("p",), record_start_index: int = 1, device: int = 0):
        self.lib = load_library()
        self.desc, keep = build_desc(f, record, record_start_index, device)
        self.record = tuple(record)
        self.Nx, self.Ny, self.Nz = int(self.desc.Nx), int(self.desc.Ny), int(self.desc.Nz)
        self.dim = 3 if self.Nz > 1 else 2
        self.Nt = int(self.desc.Nt)
        self.ns = int(self.desc.sensor_npoints)
        self.nrec = self.Nt - max(1, int(record_start_index)) + 1
        h = C.c_void_p()
        _check(self.lib, self.lib.kw_sim_create(C.byref(self.desc), C.byref(h)))
        del keep
        self.handle = h

    def run(self, nsteps: int | None = None, stream: int | None = None):
        n = self.Nt if nsteps is None else int(nsteps)
        _check(self.lib, self.lib.kw_sim_run(self.handle, n, C.c_void_p(stream) if stream else None))
        return self

    def sync(self):
        _check(self.lib, self.lib.kw_sim_sync(self.handle))

    @property
    def steps_done(self) -> int:
        return int(self.lib.kw_sim_steps_done(self.handle))

    def fetch(self, name: str) -> np.ndarray:
        oid = output_id(name)
        nbytes = self.lib.kw_sim_output_bytes(self.handle, oid)
        if nbytes == 0:
            raise KWaveError(1, f"output {name} was not recorded")
        out = np.empty(nbytes // 4, dtype=np.float32)
        _check(self.lib, self.lib.kw_sim_fetch(self.handle, oid, out.ctypes.data, nbytes))
        if _time_series(name):
            return out.reshape(self.nrec, self.ns)
        if name.endswith("_all") or name.endswith("_final"):
            return out.reshape(self.Nz, self.Ny, self.Nx)
        return out

    def outputs(self) -> dict:
        return {n: self.fetch(n) for n in output_names(self.record, self.dim)}

    def autotune(self, trial_steps: int = 4) -> dict:
        """Pick the fastest workgroup sizes (cached per shape); resets to step 0."""
        r, pe, pl = C.c_int(), C.c_int(), C.c_int()
        _check(self.lib, self.lib.kw_sim_autotune(self.handle, int(trial_steps), C.byref(r), C.byref(pe),
                                                  C.byref(pl)))
        return {"row_nt": r.value, "pencil_nt": pe.value, "plane_nt": pl.value}

    def launch_config(self) -> dict:
        """Workgroup sizes and pencil radix plans in use (kw_sim_launch_config)."""
        c = KwLaunchConfig()
        _check(self.lib, self.lib.kw_sim_launch_config(self.handle, C.byref(c)))
        return {k: getattr(c, k) for k, _ in KwLaunchConfig._fields_}

    def memory(self) -> dict:
        """Device bytes held now and at most since creation (kw_sim_memory)."""
        a, pk = C.c_uint64(), C.c_uint64()
        _check(self.lib, self.lib.kw_sim_memory(self.handle, C.byref(a), C.byref(pk)))
        return {"allocated": a.value, "peak": pk.value}

    def set_profiling(self, on: bool = True):
        _check(self.lib, self.lib.kw_sim_set_profiling(self.handle, 1 if on else 0))

    def profile(self) -> dict:
        cap = 16
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        n = (C.c_uint64 * cap)()
        k = self.lib.kw_sim_profile_get(self.handle, names, ms, n, cap)
        return {names[i].decode(): (ms[i], int(n[i])) for i in range(min(k, cap))}

    def close(self):
        if getattr(self, "handle", None):
            self.lib.kw_sim_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run(f: dict, record=("p",), record_start_index: int = 1, device: int = 0) -> dict:
    """Run a whole simulation and return its outputs (output-file naming)."""
    with Simulation(f, record, record_start_index, device) as sim:
        sim.run()
        return sim.outputs()


def run_noalloc(f: dict, record=("p",), record_start_index: int = 1, device: int = 0,
                stream: int | None = None) -> dict:
    """A whole simulation through kw_sim_run_noalloc (every device buffer in
    the library's static scratch region; no device allocation), returning
    the same outputs as run()."""
    lib = load_library()
    d, keep = build_desc(f, record, record_start_index, device)
    Nx, Ny, Nz = int(d.Nx), int(d.Ny), int(d.Nz)
    dim = 3 if Nz > 1 else 2
    ns = int(d.sensor_npoints)
    nrec = int(d.Nt) - max(1, int(record_start_index)) + 1
    names = output_names(record, dim)
    outs, bufs = {}, (KwOutputBuffer * max(1, len(names)))()
    for i, name in enumerate(names):
        if _time_series(name):
            a = np.empty((nrec, ns), dtype=np.float32)
        elif name.endswith("_all") or name.endswith("_final"):
            a = np.empty((Nz, Ny, Nx), dtype=np.float32)
        else:
            a = np.empty(ns, dtype=np.float32)
        outs[name] = a
        bufs[i] = KwOutputBuffer(output_id(name), 0, a.ctypes.data, a.nbytes)
    _check(lib, lib.kw_sim_run_noalloc(C.byref(d), C.c_void_p(stream) if stream else None, bufs, len(names)))
    del keep
    return outs


def footprint_bytes(f: dict, record=("p",), record_start_index: int = 1) -> int:
    lib = load_library()
    d, keep = build_desc(f, record, record_start_index)
    n = int(lib.kw_sim_footprint_bytes(C.byref(d)))
    del keep
    return n


def diag_copy(nbytes: int, device: int = 0, variant: int = 0, blocks: int = 0) -> float:
    """Time one device copy of nbytes (kw_diag_copy_ex; variant 0 is the
    counter-calibration kernel); returns ms per copy."""
    lib = load_library()
    ms = C.c_double()
    _check(lib, lib.kw_diag_copy_ex(int(nbytes), device, int(variant), int(blocks), C.byref(ms)))
    return ms.value


class Server:
    """Serving loop (kw_server handle): submit many simulations, each admitted
    by its device footprint and run concurrently with the others (batching:
    simulations of one grid shape and physics branch share each kernel
    launch, a quantum of steps replayed from a captured hipGraph)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0, max_active: int = 8,
                 steps_per_quantum: int = 16, autotune: bool = False, batching: bool = False):
        self.lib = load_library()
        cfg = KwServerConfig(KW_ABI_VERSION, device, int(hbm_budget_bytes), int(max_active),
                             int(steps_per_quantum), 1 if autotune else 0, 1 if batching else 0)
        h = C.c_void_p()
        _check(self.lib, self.lib.kw_server_create(C.byref(cfg), C.byref(h)))
        self.handle = h
        self._shape = {}

    def submit(self, f: dict, record=("p",), record_start_index: int = 1) -> int:
        d, keep = build_desc(f, record, record_start_index)
        rid = C.c_uint64()
        _check(self.lib, self.lib.kw_server_submit(self.handle, C.byref(d), C.byref(rid)))
        del keep  # the server copied every array
        dim = 3 if int(d.Nz) > 1 else 2
        nrec = int(d.Nt) - max(1, int(record_start_index)) + 1
        self._shape[rid.value] = (tuple(record), dim, int(d.sensor_npoints), nrec,
                                  (int(d.Nz), int(d.Ny), int(d.Nx)))
        return rid.value

    def state(self, rid: int) -> str:
        return REQUEST_STATES[self.lib.kw_server_poll(self.handle, rid)]

    def wait(self, rid: int, timeout: float = -1.0) -> bool:
        st = self.lib.kw_server_wait(self.handle, rid, float(timeout))
        if st == KW_ERR_TIMEOUT:
            return False
        _check(self.lib, st)
        return True

    def info(self, rid: int) -> dict:
        i = KwRequestInfo()
        _check(self.lib, self.lib.kw_server_request_info(self.handle, rid, C.byref(i)))
        return {"state": REQUEST_STATES[i.state], "status": i.status, "footprint_bytes": i.footprint_bytes,
                "point_updates": i.point_updates, "t_submit": i.t_submit, "t_admit": i.t_admit,
                "t_done": i.t_done}

    def error(self, rid: int) -> str:
        return (self.lib.kw_server_request_error(self.handle, rid) or b"").decode(errors="replace")

    def fetch(self, rid: int, name: str) -> np.ndarray:
        record, dim, ns, nrec, shape = self._shape[rid]
        oid = output_id(name)
        nbytes = self.lib.kw_server_output_bytes(self.handle, rid, oid)
        if nbytes == 0:
            raise KWaveError(1, f"output {name} of request {rid} is not available ({self.state(rid)})")
        out = np.empty(nbytes // 4, dtype=np.float32)
        _check(self.lib, self.lib.kw_server_fetch(self.handle, rid, oid, out.ctypes.data, nbytes))
        if _time_series(name):
            return out.reshape(nrec, ns)
        if name.endswith("_all") or name.endswith("_final"):
            return out.reshape(shape)
        return out

    def outputs(self, rid: int) -> dict:
        record, dim = self._shape[rid][:2]
        return {n: self.fetch(rid, n) for n in output_names(record, dim)}

    def release(self, rid: int):
        _check(self.lib, self.lib.kw_server_release(self.handle, rid))
        self._shape.pop(rid, None)

    def stats(self) -> dict:
        s = KwServerStats()
        _check(self.lib, self.lib.kw_server_stats_get(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in KwServerStats._fields_}

    def close(self):
        if getattr(self, "handle", None):
            self.lib.kw_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
