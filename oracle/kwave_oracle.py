"""FP64 CPU oracle for the k-space pseudospectral time loop of k-Wave.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP solver:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product path (``k-wave_amd``) never imports,
calls or falls back to it.

It restates, statement by statement and in double precision, what the
external ``kspaceFirstOrder-OMP/-CUDA`` binaries compute from the HDF5 input
file written by ``kspaceFirstOrder_saveToDisk.m``:

* operator precompute        k-Wave/kspaceFirstOrder3D.m:583-659,
                             k-Wave/kspaceFirstOrder2D.m:545-609
* PML profiles               k-Wave/private/getPML.m:45-85
* wavenumber vectors         k-Wave/kWaveGrid.m:417-433, 779-797 (makeDim)
* sinc                       k-Wave/private/sinc.m
* absorption operators       k-Wave/private/kspaceFirstOrder_createAbsorptionVariables.m:27-98
* dB -> Neper conversion     k-Wave/db2neper.m:50
* time loop                  k-Wave/kspaceFirstOrder3D.m:767-1206,
                             k-Wave/kspaceFirstOrder2D.m:710-1042
* sensor recording           k-Wave/private/kspaceFirstOrder_extractSensorData.m:61-297, 462-560
* axisymmetric loop          k-Wave/kspaceFirstOrderAS.m:467-571, 692-1244 (WSWA-FFT form)

Input: a dict holding the datasets of the HDF5 input file, each as a numpy
array in HDF5 C order (MATLAB [Nx,Ny,Nz] == C (Nz,Ny,Nx)); see
``k-wave_amd/h5format.py`` for the writer of the same dict.

FFT convention.  MATLAB evaluates ``real(ifftn(D .* fftn(f)))`` on the full
complex spectrum.  For real f and an operator with D(-k) = conj(D(k)) (true
for every operator on this path, including at the Nyquist bin where D is
real) this equals the half-spectrum R2C/C2R evaluation; ``full_complex=True``
evaluates the literal full-spectrum form and the tests check both agree.

Parity pins (tests/test_oracle.py): closed-form lossless solution
p_n = IFFT(FFT(p0) cos(c|k|(n-1)dt)); direct DFT of the derivative operators;
PML/kappa/absorption closed forms; p0 == two-clicks; 2D == 3D plane waves;
x/y/z rotation symmetry; additive-source scaling against a 2x finer dt; the
stability bound.  Functions without such a check say so in their docstring.
"""

from __future__ import annotations

import concurrent.futures
import math
import os
import time
from dataclasses import dataclass, field

import numpy as np
from scipy import fft as sfft

WORKERS = -1  # scipy.fft workers: all cores

# Elementwise statements on large grids are evaluated slab by slab (along the
# first, z, axis) on all cores; numpy releases the GIL inside its loops.  The
# per-element arithmetic is the statement's own, so the values are those of
# the whole-array evaluation.
PAR_MIN_ELEMS = 1 << 21
_pool = None


def _threads():
    return max(1, len(os.sched_getaffinity(0)))


def pev(fn, *xs):
    """fn(*xs) for an elementwise fn, evaluated over z-slabs in parallel.

    Operands are scalars or arrays broadcastable to the 3-D result; arrays
    spanning the result's first axis are split along it, the others (extent
    1 there, or fewer dimensions) are passed whole to every slab."""
    shapes = [np.shape(x) for x in xs]
    full = np.broadcast_shapes(*shapes)
    if len(full) != 3 or full[0] < 2 or int(np.prod(full)) < PAR_MIN_ELEMS or _threads() == 1:
        return fn(*xs)
    global _pool
    if _pool is None:
        _pool = concurrent.futures.ThreadPoolExecutor(_threads())
    n0 = full[0]
    nch = min(n0, 4 * _threads())
    bounds = [n0 * i // nch for i in range(nch + 1)]

    def args(i):
        a, b = bounds[i], bounds[i + 1]
        return [x[a:b] if (isinstance(x, np.ndarray) and x.ndim == 3 and x.shape[0] == n0) else x for x in xs]

    first = np.asarray(fn(*args(0)))
    out = np.empty(full, dtype=first.dtype)
    out[bounds[0]:bounds[1]] = first

    def part(i):
        out[bounds[i]:bounds[i + 1]] = fn(*args(i))

    list(_pool.map(part, range(1, nch)))
    return out


def psum(arrs):
    """sum(arrs) of same-shape arrays, slab-parallel (left to right, as sum())."""
    if len(arrs) == 1:
        return arrs[0]
    return pev(lambda *a: sum(a), *arrs)

RECORD_NAMES = (
    "p", "p_max", "p_min", "p_rms", "p_max_all", "p_min_all", "p_final",
    "u", "u_max", "u_min", "u_rms", "u_max_all", "u_min_all", "u_final", "u_non_staggered",
)


# ---------------------------------------------------------------------------
# scalar / grid helpers
# ---------------------------------------------------------------------------

def db2neper(alpha, y=1.0):
    """dB/(MHz^y cm) -> Np/((rad/s)^y m).  k-Wave/db2neper.m:50."""
    return 100.0 * np.asarray(alpha, dtype=np.float64) * (1e-6 / (2.0 * math.pi)) ** y / (
        20.0 * math.log10(math.e))


def make_dim(n: int, d: float) -> np.ndarray:
    """Wavenumber vector in fftshift order.  kWaveGrid.makeDim, kWaveGrid.m:779-797."""
    if n % 2 == 0:
        nx = np.arange(-n // 2, n // 2, dtype=np.float64) / n
    else:
        nx = np.arange(-(n - 1) // 2, (n - 1) // 2 + 1, dtype=np.float64) / n
    nx[n // 2] = 0.0
    return (2.0 * math.pi / d) * nx


def k_vec_unshifted(n: int, d: float) -> np.ndarray:
    """ifftshift(makeDim(n, d)): FFT-order wavenumbers (Nyquist negative for even n)."""
    return np.fft.ifftshift(make_dim(n, d))


def x_shift_neg(n: int, d: float) -> np.ndarray:
    """record.x_shift_neg = ifftshift(exp(-1i kx_vec dx / 2)), the half-cell shift
    of the u_non_staggered outputs.  kspaceFirstOrder_createStorageVariables.m:87-99."""
    return np.exp(-1j * k_vec_unshifted(n, d) * d / 2)


def sinc(x):
    """sin(x)/x with sinc(0)=1.  k-Wave/private/sinc.m."""
    x = np.asarray(x, dtype=np.float64)
    z = x == 0
    return np.sin(x + math.pi * z) / (x + z) + z


def get_pml(n: int, d: float, dt: float, c: float, size: int, alpha: float,
            staggered: bool, axisymmetric: bool = False) -> np.ndarray:
    """1-D PML multiplier.  k-Wave/private/getPML.m:50-85; axisymmetric: the
    radial axis has the layer at its outer end only (getPML.m:73-77)."""
    x = np.arange(1, size + 1, dtype=np.float64)
    if staggered:
        left = alpha * (c / d) * (((x + 0.5) - size - 1) / (0 - size)) ** 4
        right = alpha * (c / d) * ((x + 0.5) / size) ** 4
    else:
        left = alpha * (c / d) * ((x - size - 1) / (0 - size)) ** 4
        right = alpha * (c / d) * (x / size) ** 4
    left = np.exp(-left * dt / 2)
    right = np.exp(-right * dt / 2)
    pml = np.ones(n, dtype=np.float64)
    if size > 0:
        if not axisymmetric:
            pml[:size] = left
        pml[n - size:] = right
    return pml


# ---------------------------------------------------------------------------
# input-dict helpers
# ---------------------------------------------------------------------------

def _scalar(v) -> float:
    return float(np.asarray(v).reshape(-1)[0])


def _int(v) -> int:
    return int(np.asarray(v).reshape(-1)[0])


def _field_or_scalar(v, shape):
    a = np.asarray(v)
    if a.size == 1:
        return float(a.reshape(-1)[0])
    return pev(lambda x: np.asarray(x, dtype=np.float64), a.reshape(shape))


@dataclass
class Problem:
    """Datasets of one input file, decoded into FP64 working arrays."""

    Nx: int
    Ny: int
    Nz: int
    Nt: int
    dim: int
    dx: float
    dy: float
    dz: float
    dt: float
    c_ref: float
    c0: object
    rho0: object
    rho0_sg: list
    nonlinear: bool
    BonA: object
    absorbing_flag: int
    alpha_coeff: object
    alpha_power: float
    pml_size: list
    pml_alpha: list
    p0: np.ndarray | None
    p_source_flag: int
    p_source_mode: int
    p_source_index: np.ndarray | None
    p_source_input: np.ndarray | None
    u_source_flag: list
    u_source_mode: int
    u_source_index: np.ndarray | None
    u_source_input: list
    sensor_mask_index: np.ndarray
    extras: dict = field(default_factory=dict)
    axisymmetric: bool = False

    @property
    def shape(self):
        return (self.Nz, self.Ny, self.Nx)

    @property
    def N(self):
        return self.Nx * self.Ny * self.Nz


def decode(inp: dict) -> Problem:
    """Decode an input-file dict (saveToDisk.m:44-414 dataset set)."""
    Nx, Ny, Nz = _int(inp["Nx"]), _int(inp["Ny"]), _int(inp["Nz"])
    dim = 2 if Nz == 1 else 3
    shape = (Nz, Ny, Nx)
    g = lambda k, d=0.0: inp[k] if k in inp else d  # noqa: E731
    rho0_sg = [_field_or_scalar(inp["rho0_sgx"], shape), _field_or_scalar(inp["rho0_sgy"], shape)]
    if dim == 3:
        rho0_sg.append(_field_or_scalar(inp["rho0_sgz"], shape))
    absorbing = _int(g("absorbing_flag", 0))
    nonlinear = bool(_int(g("nonlinear_flag", 0)))

    def sig(name):
        if name not in inp:
            return None
        a = np.asarray(inp[name], dtype=np.float64)
        return a.reshape(a.shape[-2], a.shape[-1])  # (Nt_src, N_series)

    def idx(name):
        if name not in inp:
            return None
        return np.asarray(inp[name], dtype=np.int64).reshape(-1) - 1

    u_flags = [_int(g("ux_source_flag", 0)), _int(g("uy_source_flag", 0))]
    u_inputs = [sig("ux_source_input"), sig("uy_source_input")]
    if dim == 3:
        u_flags.append(_int(g("uz_source_flag", 0)))
        u_inputs.append(sig("uz_source_input"))
    p0 = None
    if _int(g("p0_source_flag", 0)):
        p0 = np.asarray(inp["p0_source_input"], dtype=np.float64).reshape(shape)
    extras = {}
    if _int(g("transducer_source_flag", 0)):
        # saveToDisk.m:290-293, 339-343; delay_mask is 1-based (inputChecking.m:1213)
        extras["transducer_flag"] = _int(inp["transducer_source_flag"])
        extras["transducer_signal"] = np.asarray(inp["transducer_source_input"], dtype=np.float64).reshape(-1)
        extras["transducer_delay0"] = np.asarray(inp["delay_mask"], dtype=np.int64).reshape(-1) - 1
    return Problem(
        Nx=Nx, Ny=Ny, Nz=Nz, Nt=_int(inp["Nt"]), dim=dim,
        dx=_scalar(inp["dx"]), dy=_scalar(inp["dy"]),
        dz=_scalar(inp["dz"]) if dim == 3 else 1.0,
        dt=_scalar(inp["dt"]), c_ref=_scalar(inp["c_ref"]),
        c0=_field_or_scalar(inp["c0"], shape), rho0=_field_or_scalar(inp["rho0"], shape),
        rho0_sg=rho0_sg, nonlinear=nonlinear,
        BonA=_field_or_scalar(inp["BonA"], shape) if nonlinear else 0.0,
        absorbing_flag=absorbing,
        alpha_coeff=_field_or_scalar(inp["alpha_coeff"], shape) if absorbing else 0.0,
        alpha_power=_scalar(inp["alpha_power"]) if absorbing else 0.0,
        pml_size=[_int(inp["pml_x_size"]), _int(inp["pml_y_size"])] + (
            [_int(inp["pml_z_size"])] if dim == 3 else []),
        pml_alpha=[_scalar(inp["pml_x_alpha"]), _scalar(inp["pml_y_alpha"])] + (
            [_scalar(inp["pml_z_alpha"])] if dim == 3 else []),
        p0=p0,
        p_source_flag=_int(g("p_source_flag", 0)),
        p_source_mode=_int(g("p_source_mode", 0)),
        p_source_index=idx("p_source_index"),
        p_source_input=sig("p_source_input"),
        u_source_flag=u_flags,
        u_source_mode=_int(g("u_source_mode", 0)),
        u_source_index=idx("u_source_index"),
        u_source_input=u_inputs,
        sensor_mask_index=idx("sensor_mask_index"),
        extras=extras,
        axisymmetric=bool(_int(g("axisymmetric_flag", 0))),
    )


# ---------------------------------------------------------------------------
# operators
# ---------------------------------------------------------------------------

class Operators:
    """k-space operators on the R2C half spectrum (or the full spectrum).

    kspaceFirstOrder3D.m:617-659, createAbsorptionVariables.m:27-53.
    Spectral arrays have the C-order shape of the transform output:
    (Nz, Ny, Nx//2+1) for R2C (x is the halved axis), (Nz, Ny, Nx) for full.
    """

    def __init__(self, pb: Problem, full_complex: bool = False):
        self.full = full_complex
        d = [pb.dx, pb.dy, pb.dz]
        n = [pb.Nx, pb.Ny, pb.Nz]
        kv = [k_vec_unshifted(n[a], d[a]) for a in range(3)]
        if not full_complex:
            kv[0] = kv[0][: pb.Nx // 2 + 1]
        if pb.dim == 2:
            kv[2] = np.zeros(1)
        kx = kv[0][None, None, :]
        ky = kv[1][None, :, None]
        kz = kv[2][:, None, None]
        self.k = pev(lambda kx, ky, kz: np.sqrt(kx ** 2 + ky ** 2 + kz ** 2), kx, ky, kz)
        shp = [kx, ky, kz]
        # ddx_k_shift_pos/neg (3D.m:627-634), broadcast along their axis
        self.d_pos = [1j * shp[a] * np.exp(1j * shp[a] * d[a] / 2) for a in range(pb.dim)]
        self.d_neg = [1j * shp[a] * np.exp(-1j * shp[a] * d[a] / 2) for a in range(pb.dim)]
        self.kappa = pev(lambda k: sinc(pb.c_ref * k * pb.dt / 2), self.k)              # 3D.m:652
        self.source_kappa = pev(lambda k: np.cos(pb.c_ref * k * pb.dt / 2), self.k)    # 3D.m:654
        # PML (3D.m:617-622)
        self.pml = []
        self.pml_sg = []
        for a in range(pb.dim):
            axi = pb.axisymmetric and a == 1      # radial axis (kspaceFirstOrderAS.m:470-471)
            self.pml.append(get_pml(n[a], d[a], pb.dt, pb.c_ref, pb.pml_size[a], pb.pml_alpha[a], False, axi))
            self.pml_sg.append(get_pml(n[a], d[a], pb.dt, pb.c_ref, pb.pml_size[a], pb.pml_alpha[a], True, axi))
        # absorption (createAbsorptionVariables.m:27-53, 83-97)
        self.absorb_tau = 0.0
        self.absorb_eta = 0.0
        self.absorb_nabla1 = None
        self.absorb_nabla2 = None
        if pb.absorbing_flag == 1:
            y = pb.alpha_power
            a_np = pev(lambda a: db2neper(a, y), pb.alpha_coeff)

            def kpow(k, e):
                # k^e with the Inf at k = 0 set to 0 (createAbsorptionVariables.m:33-41)
                with np.errstate(divide="ignore"):
                    n = k ** e
                n[np.isinf(n)] = 0.0
                return n

            self.absorb_nabla1 = pev(lambda k: kpow(k, y - 2), self.k)
            self.absorb_nabla2 = pev(lambda k: kpow(k, y - 1), self.k)
            c0 = pb.c0
            self.absorb_tau = pev(lambda a, c: -2.0 * a * np.power(c, y - 1), a_np, c0)
            self.absorb_eta = pev(lambda a, c: 2.0 * a * np.power(c, y) * math.tan(math.pi * y / 2), a_np, c0)
        elif pb.absorbing_flag == 2:
            a_np = db2neper(pb.alpha_coeff, 2)
            self.absorb_tau = -2.0 * a_np * pb.c0

    def bcast_pml(self, vec, axis):
        """Reshape a 1-D PML vector to broadcast along grid axis (0=x,1=y,2=z)."""
        if axis == 0:
            return vec[None, None, :]
        if axis == 1:
            return vec[None, :, None]
        return vec[:, None, None]


class Spectral:
    """Forward/inverse transforms over the active axes (x,y[,z])."""

    def __init__(self, pb: Problem, full_complex: bool):
        self.full = full_complex
        self.axes = (-1, -2) if pb.dim == 2 else (-1, -2, -3)
        self.shape = pb.shape

        self.workers = WORKERS

    def fwd(self, f):
        if self.full:
            return sfft.fftn(f, axes=self.axes, workers=self.workers)
        return sfft.rfftn(f, axes=self.axes[::-1], workers=self.workers)

    def inv_real(self, F):
        # F is always a temporary of the caller: the transform may overwrite it
        if self.full:
            return np.real(sfft.ifftn(F, axes=self.axes, workers=self.workers, overwrite_x=True))
        sh = [self.shape[a] for a in self.axes[::-1]]
        return sfft.irfftn(F, s=sh, axes=self.axes[::-1], workers=self.workers, overwrite_x=True)


class AxisymOps:
    """Axisymmetric (x axial, y radial) gradients of kspaceFirstOrderAS.m in
    the WSWA-FFT form: each field is mirrored to a 4 Ny radial period with the
    symmetry of its grid, transformed with fft2 on the expanded grid
    kWaveGrid(Nx, dx, 4 Ny, dy), and trimmed back to Ny (AS.m:488-521 operators,
    709-731 pressure gradient, 950-990 divergence, 801-827 / 876-904 / 1076-1090
    source corrections).  The input file selects WSWA (DTT) symmetry
    (inputChecking.m:1400-1411), which the reference's own test
    (testing/unit/kspaceFirstOrderAS_compare_FFT_implementations.m) holds equal
    to WSWA-FFT to 1e-14; tests/test_oracle.py repeats that comparison with
    scipy's DCT/DST.  Expanded arrays are zero-filled for every use (AS.m keeps
    one source_mat_exp across components, which WSWA does not have)."""

    def __init__(self, pb: Problem):
        if pb.dim != 2:
            raise ValueError("axisymmetric inputs are 2-D (x axial, y radial)")
        if pb.absorbing_flag == 1:
            raise ValueError("the axisymmetric code supports stokes absorption only (inputChecking.m:172-185)")
        self.Ny = pb.Ny
        kx = k_vec_unshifted(pb.Nx, pb.dx)[None, :]
        ky = k_vec_unshifted(4 * pb.Ny, pb.dy)[:, None]
        self.ddx_pos = 1j * kx * np.exp(1j * kx * pb.dx / 2)       # AS.m:476-477
        self.ddx_neg = 1j * kx * np.exp(-1j * kx * pb.dx / 2)
        self.ddy = 1j * ky                                            # AS.m:508-510
        self.y_pos = np.exp(1j * ky * pb.dy / 2)
        self.y_neg = np.exp(-1j * ky * pb.dy / 2)
        k = np.sqrt(kx ** 2 + ky ** 2)
        self.kappa = sinc(pb.c_ref * k * pb.dt / 2)                  # AS.m:514
        self.source_kappa = np.cos(pb.c_ref * k * pb.dt / 2)         # AS.m:516
        self.y_sg = (np.arange(pb.Ny) + 0.5) * pb.dy                 # y_vec_sg, AS.m:571

    def mirror(self, f, kind):
        """Radial mirror to 4 Ny.  "wswa": p, ux and their sources (AS.m:713-717);
        "hahs": uy and its source (AS.m:953-956, 884-888); "hsha": uy / y_sg (AS.m:959-962)."""
        Ny = self.Ny
        e = np.zeros(f.shape[:-2] + (4 * Ny, f.shape[-1]))
        if kind == "wswa":
            e[..., :Ny, :] = f
            e[..., Ny + 1:2 * Ny, :] = -f[..., :0:-1, :]
            e[..., 2 * Ny:3 * Ny, :] = -f
            e[..., 3 * Ny + 1:, :] = f[..., :0:-1, :]
        else:
            sgn = 1.0 if kind == "hahs" else -1.0
            e[..., :Ny, :] = f
            e[..., Ny:2 * Ny, :] = sgn * f[..., ::-1, :]
            e[..., 2 * Ny:3 * Ny, :] = -f
            e[..., 3 * Ny:, :] = -sgn * f[..., ::-1, :]
        return e

    @staticmethod
    def _f2(f):
        return sfft.fft2(f, axes=(-2, -1), workers=WORKERS)

    def _i2(self, F):
        return np.real(sfft.ifft2(F, axes=(-2, -1), workers=WORKERS))[..., :self.Ny, :]

    def grad_p(self, p):
        """[dpdx_sgx, dpdy_sgy] (AS.m:724-731)."""
        pk = self.kappa * self._f2(self.mirror(p, "wswa"))
        return [self._i2(self.ddx_pos * pk), self._i2(self.ddy * self.y_pos * pk)]

    def div(self, ux, uy):
        """[duxdx, duydy]: the radial term is (1/r) d(r uy)/dr = duy/dy + uy/y (AS.m:981-986)."""
        duxdx = self._i2(self.ddx_neg * self.kappa * self._f2(self.mirror(ux, "wswa")))
        duydy = self._i2(self.kappa * self.y_neg * (self.ddy * self._f2(self.mirror(uy, "hahs"))
                                                     + self._f2(self.mirror(uy / self.y_sg[:, None], "hsha"))))
        return [duxdx, duydy]

    def source_correction(self, sm, kind):
        """real(ifft2(source_kappa .* fft2(mirror(source_mat)))) trimmed (AS.m:806-827)."""
        return self._i2(self.source_kappa * self._f2(self.mirror(sm, kind)))


# ---------------------------------------------------------------------------
# time loop
# ---------------------------------------------------------------------------

def nonlinear_density(srho, rho0):
    """rho0_plus_rho = 2 (rhox + rhoy + rhoz) + rho0 (kspaceFirstOrder3D.m:920)."""
    return 2.0 * srho + rho0


def run(inp: dict, record=("p",), record_start_index: int = 1, full_complex: bool = False,
        nsteps: int | None = None, timings: dict | None = None, progress=None) -> dict:
    """Run the time loop; return outputs named and shaped like the output file.

    kspaceFirstOrder3D.m:767-1206 (2D.m:710-1042 for Nz == 1).
    ``record`` holds names from RECORD_NAMES; ``nsteps`` truncates the loop
    (the outputs then cover steps 1..nsteps).  ``timings`` (a dict) receives
    the wall time of every step ("step") and of its source-injection
    branches ("source"), in seconds.  ``progress(msg)``, if given, is called
    after the set-up and after every step.
    """
    pb = decode(inp)
    ops = Operators(pb, full_complex)
    if progress:
        progress("operators ready")
    sp = Spectral(pb, full_complex)
    axs = AxisymOps(pb) if pb.axisymmetric else None
    dim, dt = pb.dim, pb.dt
    shape = pb.shape
    Nt = pb.Nt if nsteps is None else nsteps
    rec = set(record)
    for r in rec:
        if r not in RECORD_NAMES:
            raise ValueError(f"unknown record option {r}")

    c0, rho0 = pb.c0, pb.rho0
    rsg_inv = [pev(lambda r: 1.0 / r, r) for r in pb.rho0_sg]
    pml = [ops.bcast_pml(ops.pml[a], a) for a in range(dim)]
    pml_sg = [ops.bcast_pml(ops.pml_sg[a], a) for a in range(dim)]

    p = np.zeros(shape)
    rho = [np.zeros(shape) for _ in range(dim)]
    u = [np.zeros(shape) for _ in range(dim)]
    p_k = sp.fwd(p)

    # sources
    p_idx = pb.p_source_index
    u_idx = pb.u_source_index

    def source_values(sig_arr, npts, t_index):
        # sig_arr: (Nt_src, N_series); N_series == 1 -> same value at every point
        row = sig_arr[t_index - 1]
        if row.size == 1:
            return np.full(npts, row[0])
        return row

    sens = pb.sensor_mask_index
    ns = sens.size
    nrec = Nt - record_start_index + 1
    out: dict = {}
    if "p" in rec:
        out["p"] = np.zeros((max(nrec, 0), ns))
    if "u" in rec:
        for a, nm in enumerate("xyz"[:dim]):
            out["u" + nm] = np.zeros((max(nrec, 0), ns))
    if "u_non_staggered" in rec and axs:
        raise ValueError("u_non_staggered is not defined for axisymmetric inputs here")
    if "u_non_staggered" in rec:
        # record.x_shift_neg = ifftshift(exp(-1i kx_vec dx / 2))
        # (kspaceFirstOrder_createStorageVariables.m:87-99); array axis of MATLAB axis a is 2 - a
        d = (pb.dx, pb.dy, pb.dz)
        shift_neg = []
        for a in range(dim):
            n = shape[2 - a]
            v = x_shift_neg(n, d[a])
            sh = [1, 1, 1]
            sh[2 - a] = n
            shift_neg.append(v.reshape(sh))
        for a, nm in enumerate("xyz"[:dim]):
            out["u" + nm + "_non_staggered"] = np.zeros((max(nrec, 0), ns))
    p_rms_acc = np.zeros(ns)
    u_rms_acc = [np.zeros(ns) for _ in range(dim)]

    if timings is not None:
        timings.setdefault("step", [])
        timings.setdefault("source", [])
    for t_index in range(1, Nt + 1):
        t_step0 = time.perf_counter()
        t_src = 0.0
        # velocity update (3D.m:790-801; AS.m:777-784)
        grads = axs.grad_p(p) if axs else None
        for a in range(dim):
            grad = grads[a] if axs else sp.inv_real(pev(lambda d, k, f: d * k * f, ops.d_pos[a], ops.kappa, p_k))
            u[a] = pev(lambda ps, ua, r, g: ps * (ps * ua - dt * r * g), pml_sg[a], u[a], rsg_inv[a], grad)
        # velocity sources (3D.m:804-884)
        t0 = time.perf_counter()
        for a in range(dim):
            if pb.u_source_flag[a] >= t_index:
                vals = source_values(pb.u_source_input[a], u_idx.size, t_index)
                uf = u[a].reshape(-1)
                if pb.u_source_mode == 0:
                    uf[u_idx] = vals
                elif pb.u_source_mode == 2:
                    sm = np.zeros(pb.N)
                    sm[u_idx] = vals
                    if axs:
                        sm = axs.source_correction(sm.reshape(shape), "wswa" if a == 0 else "hahs")
                    else:
                        sm = sp.inv_real(pev(lambda k, f: k * f, ops.source_kappa, sp.fwd(sm.reshape(shape))))
                    u[a] = pev(lambda x, y: x + y, u[a], sm)
                else:
                    uf[u_idx] = uf[u_idx] + vals
        # transducer source (3D.m:889-902): one signal, a per-point delay, x velocity only
        if pb.extras.get("transducer_flag", 0) >= t_index:
            uf = u[0].reshape(-1)
            uf[u_idx] = uf[u_idx] + pb.extras["transducer_signal"][pb.extras["transducer_delay0"] + t_index - 1]
        t_src += time.perf_counter() - t0
        # divergence (3D.m:905-907)
        du = axs.div(u[0], u[1]) if axs else [
            sp.inv_real(pev(lambda d, k, f: d * k * f, ops.d_neg[a], ops.kappa, sp.fwd(u[a]))) for a in range(dim)]
        # mass conservation (3D.m:910-925)
        if not pb.nonlinear:
            for a in range(dim):
                rho[a] = pev(lambda pm, ra, r0, d: pm * (pm * ra - dt * r0 * d), pml[a], rho[a], rho0, du[a])
        else:
            rpr = nonlinear_density(psum(rho), rho0)
            for a in range(dim):
                rho[a] = pev(lambda pm, ra, r0, d: pm * (pm * ra - dt * r0 * d), pml[a], rho[a], rpr, du[a])
        # pressure source (3D.m:928-960)
        t0 = time.perf_counter()
        if pb.p_source_flag >= t_index:
            vals = source_values(pb.p_source_input, p_idx.size, t_index)
            if pb.p_source_mode == 0:
                for a in range(dim):
                    rho[a].reshape(-1)[p_idx] = vals
            elif pb.p_source_mode == 2:
                sm = np.zeros(pb.N)
                sm[p_idx] = vals
                if axs:
                    sm = axs.source_correction(sm.reshape(shape), "wswa")
                else:
                    sm = sp.inv_real(pev(lambda k, f: k * f, ops.source_kappa, sp.fwd(sm.reshape(shape))))
                for a in range(dim):
                    rho[a] = pev(lambda x, y: x + y, rho[a], sm)
            else:
                for a in range(dim):
                    rf = rho[a].reshape(-1)
                    rf[p_idx] = rf[p_idx] + vals
        t_src += time.perf_counter() - t0
        # equation of state (3D.m:963-1017)
        srho = psum(rho)
        if pb.absorbing_flag == 1:
            sdu = psum(du)
            t1 = sp.inv_real(pev(lambda n, f: n * f, ops.absorb_nabla1, sp.fwd(pev(lambda r, d: r * d, rho0, sdu))))
            t2 = sp.inv_real(pev(lambda n, f: n * f, ops.absorb_nabla2, sp.fwd(srho)))
            p = pev(lambda c, sr, tau, a, eta, b: c ** 2 * (sr + tau * a - eta * b),
                    c0, srho, ops.absorb_tau, t1, ops.absorb_eta, t2)
            del t1, t2
        elif pb.absorbing_flag == 2:
            p = c0 ** 2 * (srho + ops.absorb_tau * rho0 * sum(du))
        else:
            p = c0 ** 2 * srho
        if pb.nonlinear:
            p = pev(lambda pp, c, b, sr, r0: pp + c ** 2 * (b * sr ** 2 / (2.0 * r0)), p, c0, pb.BonA, srho, rho0)
        if np.shape(p) != shape:
            p = np.broadcast_to(p, shape).copy()
        # initial pressure (3D.m:1021-1035)
        if t_index == 1 and pb.p0 is not None:
            p = pb.p0.copy()
            for a in range(dim):
                rho[a] = pb.p0 / (dim * c0 ** 2) * np.ones(shape)
            p0k = sp.fwd(p)
            g0 = axs.grad_p(p) if axs else None
            for a in range(dim):
                u[a] = dt * rsg_inv[a] * (g0[a] if axs else sp.inv_real(ops.d_pos[a] * ops.kappa * p0k)) / 2
        # spectrum of p (3D.m:1038)
        p_k = sp.fwd(p)
        # sensor recording (extractSensorData.m:61-297)
        if t_index >= record_start_index:
            fi = t_index - record_start_index + 1
            ps = p.reshape(-1)[sens]
            if "p" in rec:
                out["p"][fi - 1] = ps
            if "p_max" in rec:
                out["p_max"] = ps.copy() if fi == 1 else np.maximum(out["p_max"], ps)
            if "p_min" in rec:
                out["p_min"] = ps.copy() if fi == 1 else np.minimum(out["p_min"], ps)
            if "p_rms" in rec:
                p_rms_acc = np.sqrt((p_rms_acc ** 2 * (fi - 1) + ps ** 2) / fi)
                out["p_rms"] = p_rms_acc
            if "p_max_all" in rec:
                out["p_max_all"] = p.copy() if fi == 1 else np.maximum(out["p_max_all"], p)
            if "p_min_all" in rec:
                out["p_min_all"] = p.copy() if fi == 1 else np.minimum(out["p_min_all"], p)
            for a, nm in enumerate("xyz"[:dim]):
                us = u[a].reshape(-1)[sens]
                if "u" in rec:
                    out["u" + nm][fi - 1] = us
                if "u_non_staggered" in rec:
                    # real(ifft(shift_neg .* fft(u, [], axis), [], axis)) (extractSensorData.m:39-55)
                    ax = 2 - a
                    ush = np.real(np.fft.ifft(shift_neg[a] * np.fft.fft(u[a], axis=ax), axis=ax))
                    out["u" + nm + "_non_staggered"][fi - 1] = ush.reshape(-1)[sens]
                if "u_max" in rec:
                    k = "u" + nm + "_max"
                    out[k] = us.copy() if fi == 1 else np.maximum(out[k], us)
                if "u_min" in rec:
                    k = "u" + nm + "_min"
                    out[k] = us.copy() if fi == 1 else np.minimum(out[k], us)
                if "u_rms" in rec:
                    u_rms_acc[a] = np.sqrt((u_rms_acc[a] ** 2 * (fi - 1) + us ** 2) / fi)
                    out["u" + nm + "_rms"] = u_rms_acc[a]
                if "u_max_all" in rec:
                    k = "u" + nm + "_max_all"
                    out[k] = u[a].copy() if fi == 1 else np.maximum(out[k], u[a])
                if "u_min_all" in rec:
                    k = "u" + nm + "_min_all"
                    out[k] = u[a].copy() if fi == 1 else np.minimum(out[k], u[a])
        if timings is not None:
            timings["step"].append(time.perf_counter() - t_step0)
            timings["source"].append(t_src)
        if progress:
            progress(f"step {t_index} done ({time.perf_counter() - t_step0:.1f} s)")
    if "p_final" in rec:
        out["p_final"] = p.copy()
    if "u_final" in rec:
        for a, nm in enumerate("xyz"[:dim]):
            out["u" + nm + "_final"] = u[a].copy()
    return out


def exact_lossless_homogeneous(p0: np.ndarray, c: float, dt: float, d: tuple, n_steps: int):
    """Closed-form p at steps 1..n_steps for a homogeneous lossless p0 IVP without PML.

    With kappa = sinc(c k dt / 2) the leapfrog recurrence becomes
    p^{n+1} + p^{n-1} = 2 cos(c k dt) p^n, and the symmetric start of
    kspaceFirstOrder3D.m:1029-1033 gives p_n = IFFT(FFT(p0) cos(c k (n-1) dt)).
    Independent of the time loop above: used to pin it.  p0 has shape
    (Nz, Ny, Nx) with Nz == 1 for 2D; d = (dx, dy, dz).
    """
    nz, ny, nx = p0.shape
    k2 = (k_vec_unshifted(nx, d[0])[None, None, :] ** 2
          + k_vec_unshifted(ny, d[1])[None, :, None] ** 2)
    axes = (-1, -2)
    if nz > 1:
        k2 = k2 + k_vec_unshifted(nz, d[2])[:, None, None] ** 2
        axes = (-1, -2, -3)
    k = np.sqrt(k2)
    P0 = np.fft.fftn(p0, axes=axes)
    res = [np.real(np.fft.ifftn(P0 * np.cos(c * k * (n - 1) * dt), axes=axes))
           for n in range(1, n_steps + 1)]
    return np.array(res)
