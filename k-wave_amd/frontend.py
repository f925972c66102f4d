"""Front-end mirror: builds the HDF5 input-file dataset set from k-Wave style
(kgrid, medium, source, sensor) inputs, the way ``kspaceFirstOrder3D(...,
'SaveToDisk', file)`` does before the external binary runs.

Restated steps (only those that decide numbers in the file):

* grid and wavenumbers          k-Wave/kWaveGrid.m:342-433, makeTime 612-671
* reference sound speed         k-Wave/private/kspaceFirstOrder_setSoundSpeedRef.m:28-47
* p0 smoothing                  k-Wave/private/kspaceFirstOrder_inputChecking.m:1499-1553
* sensor / source indices       kspaceFirstOrder_inputChecking.m:1606-1659 (find(mask))
* source scaling                k-Wave/private/kspaceFirstOrder_scaleSourceTerms.m:27-386
* staggered density             k-Wave/kspaceFirstOrder3D.m:583-605, 2D.m:545-558
* dataset names, flags, dtypes  k-Wave/private/kspaceFirstOrder_saveToDisk.m:44-414
* PML defaults                  k-Wave/private/kspaceFirstOrder_setDefaults.m:149-168

Grid arrays are given in MATLAB index order ((Nx, Ny) or (Nx, Ny, Nz)); the
returned dict holds each dataset in HDF5 C order (MATLAB [a,b,c] -> (c,b,a)),
float32 or uint64, exactly as ``writeMatrix`` stores it.
"""

from __future__ import annotations

import math

import numpy as np

from . import geometry


class KWaveGrid:
    """Subset of kWaveGrid (kWaveGrid.m) needed to define a simulation."""

    def __init__(self, Nx, dx, Ny, dy, Nz=None, dz=None):
        self.Nx, self.dx, self.Ny, self.dy = int(Nx), float(dx), int(Ny), float(dy)
        self.Nz = int(Nz) if Nz else 1
        self.dz = float(dz) if dz else 0.0
        self.dim = 3 if Nz else 2
        self.Nt = None
        self.dt = None

    @property
    def shape(self):
        return (self.Nx, self.Ny) if self.dim == 2 else (self.Nx, self.Ny, self.Nz)

    @property
    def size(self):
        return (self.Nx * self.dx, self.Ny * self.dy, self.Nz * self.dz)

    def set_time(self, Nt, dt):
        self.Nt, self.dt = int(Nt), float(dt)

    def make_time(self, c, cfl=0.3, t_end=None):
        """kWaveGrid.makeTime (kWaveGrid.m:612-671)."""
        c = np.asarray(c, dtype=np.float64)
        c_max, c_min = float(c.max()), float(c.min())
        xs, ys, zs = self.size
        if t_end is None:
            t_end = math.sqrt(xs ** 2 + ys ** 2 + (zs ** 2 if self.dim == 3 else 0.0)) / c_min
        dmin = min(self.dx, self.dy) if self.dim == 2 else min(self.dx, self.dy, self.dz)
        self.dt = cfl * dmin / c_max
        self.Nt = int(math.floor(t_end / self.dt)) + 1
        if math.floor(t_end / self.dt) != math.ceil(t_end / self.dt) and math.fmod(t_end, self.dt) == 0:
            self.Nt += 1
        return self.dt, self.Nt


def _c_order(a: np.ndarray) -> np.ndarray:
    """MATLAB-order grid array -> HDF5 C-order (Nz, Ny, Nx) array."""
    a = np.asarray(a)
    if a.ndim == 2:
        a = a[:, :, None]
    return np.ascontiguousarray(a.transpose(2, 1, 0))


def _scalar(v, dtype=np.float32):
    return np.full((1, 1, 1), v, dtype=dtype)


def _find(mask: np.ndarray) -> np.ndarray:
    """1-based column-major indices of nonzero entries (MATLAB find)."""
    return np.flatnonzero(np.asarray(mask).reshape(-1, order="F")) + 1


def _col(idx: np.ndarray) -> np.ndarray:
    """MATLAB column vector [n,1] -> HDF5 (1,1,n) uint64."""
    return np.asarray(idx, dtype=np.uint64).reshape(1, 1, -1)


def _signal(sig: np.ndarray) -> np.ndarray:
    """MATLAB [N_series, Nt_src] -> HDF5 (1, Nt_src, N_series) float32."""
    s = np.atleast_2d(np.asarray(sig, dtype=np.float64))
    return np.ascontiguousarray(s.T[None].astype(np.float32))


def _source_labels(mask: np.ndarray, nrows: int, sig_name: str, mask_name: str):
    """Label of every source point, in find() order, for a labelled source
    mask; None for a binary mask (kspaceFirstOrder_inputChecking.m:847-884,
    964-1003).  The in-process solver then takes the signal row of each
    point's label (kspaceFirstOrder3D.m:933-957); the input file has no label
    dataset (p_source_many only says "one row per point"), so the front end
    expands the rows to one per point."""
    m = np.asarray(mask)
    u = np.unique(m)
    if u.size <= 2 and u.sum() == 1:
        return None
    if np.sum(u[1:] - u[:-1]) != u.size - 1 or not np.any(u == 1):
        raise ValueError(f"If using a labelled source.{mask_name}, the source labels must be monotonically "
                         "increasing and start from 1.")
    if nrows != u.size - 1:
        raise ValueError(f"The number of time series in source.{sig_name} must match the number of labelled "
                         f"source elements in source.{mask_name}.")
    v = m.reshape(-1, order="F")
    return v[v != 0].astype(np.int64)


def _stagger_c(rc: np.ndarray, axis: int) -> np.ndarray:
    """interpn('*linear') at +d/2 along a MATLAB axis (0 = x), last plane copied
    (3D.m:583-596), on a C-order (Nz, Ny, Nx) float64 array; returns float32."""
    ca = 2 - axis
    n = rc.shape[ca]
    a = [slice(None)] * 3
    b = [slice(None)] * 3
    a[ca] = slice(0, n - 1)
    b[ca] = slice(1, n)
    out = rc.copy()
    out[tuple(a)] += rc[tuple(b)]
    out[tuple(a)] *= 0.5
    return out.astype(np.float32)


def save_to_disk_dict(kgrid: KWaveGrid, medium: dict, source: dict, sensor: dict,
                      pml_size=None, pml_alpha=2.0, smooth_p0: bool = True,
                      sound_speed_ref=None, axisymmetric: bool = False) -> dict:
    """Dataset dict of the input file (kspaceFirstOrder_saveToDisk.m:44-414).

    axisymmetric: kspaceFirstOrderASC's file (axisymmetric_flag = 1): a 2-D grid
    with x axial and y radial from the axis (y index 0 on the axis), a radial PML
    at the outer edge only, stokes absorption (inputChecking.m:165-200) and p0
    smoothed on the WSWA-mirrored grid (inputChecking.m:1505-1536).

    medium: sound_speed, density, [alpha_coeff, alpha_power, alpha_mode='stokes'], [BonA]
    source: [p0] | [p, p_mask, p_mode] | [ux/uy/uz, u_mask, u_mode]; masks binary or labelled 1..L
    sensor: mask (binary, MATLAB order)
    """
    dim = kgrid.dim
    shape = kgrid.shape
    if kgrid.Nt is None:
        raise ValueError("kgrid time array not set")
    if pml_size is None:
        pml_size = 20 if dim == 2 else 10                       # setDefaults.m:149-168
    ps = list(pml_size) if np.ndim(pml_size) else [int(pml_size)] * dim
    pa = list(pml_alpha) if np.ndim(pml_alpha) else [float(pml_alpha)] * dim
    if axisymmetric and dim != 2:
        raise ValueError("axisymmetric simulations use a 2-D grid (x axial, y radial)")
    for a in range(dim):
        # inputChecking.m:1378-1379: the radial PML of the axisymmetric code is one-sided
        if (ps[a] > shape[a]) if (axisymmetric and a == 1) else (2 * ps[a] > shape[a]):
            raise ValueError("The size of the PML must be smaller than the size of the grid.")

    c0 = np.asarray(medium["sound_speed"], dtype=np.float64)
    rho0 = np.asarray(medium["density"], dtype=np.float64)
    hetero_c = c0.size > 1
    hetero_rho = rho0.size > 1
    c0_s = float(c0.reshape(-1)[0]) if not hetero_c else None
    if sound_speed_ref is None:
        c_ref = float(c0.max())                                    # setSoundSpeedRef.m:44
    elif sound_speed_ref == "min":
        c_ref = float(c0.min())
    elif sound_speed_ref == "mean":
        c_ref = float(c0.mean())
    elif sound_speed_ref == "max":
        c_ref = float(c0.max())
    else:
        c_ref = float(sound_speed_ref)
    dt = kgrid.dt
    d = [kgrid.dx, kgrid.dy, kgrid.dz]
    c0_flat = c0.reshape(-1, order="F") if hetero_c else None   # one reshape (a copy unless F-ordered)
    c_at = (lambda i: float(c0_flat[i - 1])) if hetero_c else (lambda i: c0_s)

    f = {}
    f["Nx"], f["Ny"], f["Nz"], f["Nt"] = (_scalar(kgrid.Nx, np.uint64), _scalar(kgrid.Ny, np.uint64),
                                          _scalar(kgrid.Nz if dim == 3 else 1, np.uint64),
                                          _scalar(kgrid.Nt, np.uint64))
    f["dx"], f["dy"] = _scalar(kgrid.dx), _scalar(kgrid.dy)
    f["pml_x_size"], f["pml_y_size"] = _scalar(ps[0], np.uint64), _scalar(ps[1], np.uint64)
    f["pml_z_size"] = _scalar(ps[2] if dim == 3 else 0, np.uint64)
    f["pml_x_alpha"], f["pml_y_alpha"] = _scalar(pa[0]), _scalar(pa[1])
    if dim == 3:
        f["dz"] = _scalar(kgrid.dz)
        f["pml_z_alpha"] = _scalar(pa[2])
    f["dt"] = _scalar(dt)
    f["c0"] = _c_order(c0).astype(np.float32) if hetero_c else _scalar(c0_s)
    f["c_ref"] = _scalar(c_ref)
    rho_c = _c_order(rho0) if hetero_rho else None
    f["rho0"] = rho_c.astype(np.float32) if hetero_rho else _scalar(float(rho0.reshape(-1)[0]))
    for a, nm in enumerate("xyz"[:dim]):
        if hetero_rho:
            f["rho0_sg" + nm] = _stagger_c(rho_c, a)
        else:
            f["rho0_sg" + nm] = _scalar(float(rho0.reshape(-1)[0]))
    del rho_c

    # flags (writeFlags.m, saveToDisk.m:88-106)
    zero = _scalar(0, np.uint64)
    for nm in ("p_source_flag", "p0_source_flag", "ux_source_flag", "uy_source_flag",
               "uz_source_flag", "sxx_source_flag", "syy_source_flag", "szz_source_flag",
               "sxy_source_flag", "sxz_source_flag", "syz_source_flag", "transducer_source_flag",
               "nonuniform_grid_flag", "nonlinear_flag", "absorbing_flag", "elastic_flag",
               "axisymmetric_flag", "sensor_mask_type"):
        f[nm] = zero.copy()

    # nonlinearity
    if "BonA" in medium and medium["BonA"] is not None:
        f["nonlinear_flag"] = _scalar(1, np.uint64)
        b = np.asarray(medium["BonA"], dtype=np.float64)
        f["BonA"] = _c_order(b).astype(np.float32) if b.size > 1 else _scalar(float(b.reshape(-1)[0]))
    # absorption (alpha_coeff stays in dB/(MHz^y cm): inputChecking.m:1666-1671)
    if "alpha_coeff" in medium and medium["alpha_coeff"] is not None:
        stokes = medium.get("alpha_mode") == "stokes" or axisymmetric
        f["absorbing_flag"] = _scalar(2 if stokes else 1, np.uint64)
        ac = np.asarray(medium["alpha_coeff"], dtype=np.float64)
        f["alpha_coeff"] = _c_order(ac).astype(np.float32) if ac.size > 1 else _scalar(float(ac.reshape(-1)[0]))
        f["alpha_power"] = _scalar(2.0 if axisymmetric else float(medium.get("alpha_power", 2.0 if stokes else 1.5)))
    if axisymmetric:
        f["axisymmetric_flag"] = _scalar(1, np.uint64)

    # initial pressure (smoothed before expansion: inputChecking.m:1499-1553)
    if source.get("p0") is not None:
        p0 = np.asarray(source["p0"], dtype=np.float64)
        if smooth_p0 and axisymmetric:
            # WSWA mirror to 4 Ny, smooth, trim (inputChecking.m:1505-1534)
            Ny = p0.shape[1]
            e = np.zeros((p0.shape[0], 4 * Ny))
            e[:, :Ny] = p0
            e[:, Ny + 1:2 * Ny] = -p0[:, :0:-1]
            e[:, 2 * Ny:3 * Ny] = -p0
            e[:, 3 * Ny + 1:] = p0[:, :0:-1]
            p0 = geometry.smooth(e, True)[:, :Ny]
        elif smooth_p0:
            p0 = geometry.smooth(p0, True)
        f["p0_source_flag"] = _scalar(1, np.uint64)
        f["p0_source_input"] = _c_order(p0).astype(np.float32)

    mode_code = {"dirichlet": 0, "additive-no-correction": 1, "additive": 2}

    def w_correction(field):
        """source.<field>_frequency_ref (inputChecking.m:818-833, 921-936): the
        signal is multiplied by cos(2 pi f_ref dt / 2) (scaleSourceTerms.m:41-42,
        270-278) and the k-space source correction is switched off."""
        f_ref = source.get(field + "_frequency_ref")
        if f_ref is None:
            return None
        f_ref = float(f_ref)
        # kgrid.k_max: the smallest of max|k| over the axes (kWaveGrid.m:410-433)
        k_max = min(2 * math.pi / dd * ((n // 2) / n if n % 2 == 0 else ((n - 1) / 2) / n)
                    for n, dd in zip(shape, d[:dim]))
        c_min = float(np.min(c0))
        if f_ref < 0 or f_ref > k_max * c_min / (2 * np.pi):
            raise ValueError(f"source.{field}_frequency_ref is higher than the maximum frequency supported by the "
                             "spatial grid.")
        return math.cos(2 * math.pi * f_ref * dt / 2)

    # pressure source (scaleSourceTerms.m:50-127)
    if source.get("p") is not None:
        sig = np.atleast_2d(np.asarray(source["p"], dtype=np.float64)).copy()
        idx = _find(source["p_mask"])
        mode = source.get("p_mode", "additive")
        wc = w_correction("p")
        if wc is not None:
            sig *= wc
            mode = "additive-no-correction"
        for r in range(sig.shape[0]):
            c_here = c_at(idx[r]) if hetero_c else c0_s
            if mode == "dirichlet":
                sig[r] = sig[r] / (dim * c_here ** 2)
            else:
                sig[r] = sig[r] * (2 * dt / (dim * c_here * kgrid.dx))
        # labelled mask: rows were scaled with the sound speed at the r-th
        # source point, as scaleSourceTerms.m:60-62,120-122 index them, then
        # each point takes its label's row
        labels = _source_labels(source["p_mask"], sig.shape[0], "p", "p_mask")
        if labels is not None:
            sig = sig[labels - 1]
        f["p_source_flag"] = _scalar(sig.shape[1], np.uint64)
        f["p_source_mode"] = _scalar(mode_code[mode], np.uint64)
        f["p_source_many"] = _scalar(int(sig.shape[0] > 1), np.uint64)
        f["p_source_index"] = _col(idx)
        f["p_source_input"] = _signal(sig)

    # velocity sources (scaleSourceTerms.m:281-353)
    u_keys = [k for k in ("ux", "uy", "uz")[:dim] if source.get(k) is not None]
    if u_keys:
        idx = _find(source["u_mask"])
        mode = source.get("u_mode", "additive")
        wcu = w_correction("u")
        if wcu is not None:
            mode = "additive-no-correction"
        labels = None
        for k in u_keys:
            labels = _source_labels(source["u_mask"], np.atleast_2d(np.asarray(source[k])).shape[0], k, "u_mask")
        many = None
        for a, k in enumerate(("ux", "uy", "uz")[:dim]):
            if source.get(k) is None:
                continue
            sig = np.atleast_2d(np.asarray(source[k], dtype=np.float64)).copy()
            if wcu is not None:
                sig *= wcu
            if mode != "dirichlet":
                for r in range(sig.shape[0]):
                    c_here = c_at(idx[r]) if hetero_c else c0_s
                    sig[r] = sig[r] * (2 * c_here * dt / d[a])
            if labels is not None:
                sig = sig[labels - 1]
            if many is None:
                many = int(sig.shape[0] > 1)
            f[k + "_source_flag"] = _scalar(sig.shape[1], np.uint64)
            f[k + "_source_input"] = _signal(sig)
        f["u_source_mode"] = _scalar(mode_code[mode], np.uint64)
        f["u_source_many"] = _scalar(many, np.uint64)
        f["u_source_index"] = _col(idx)

    # transducer source (inputChecking.m:1160-1213, scaleSourceTerms.m:376-386,
    # saveToDisk.m:290-293, 339-343): one input signal, a delay (in samples)
    # per active point, x velocity only; transmit apodization is not part of
    # the file (the binaries apply none)
    if source.get("transducer") is not None:
        if dim != 3:
            raise ValueError("Transducer inputs are only compatible with 3D simulations.")
        tr = source["transducer"]
        sig = np.asarray(tr["input_signal"], dtype=np.float64).reshape(-1)
        delay = np.asarray(tr["delay_mask"])
        active = np.asarray(tr["active_elements_mask"])
        idx = _find(active)
        flag = sig.size - int(delay.max())
        c_here = float(np.mean(c0.reshape(-1, order="F")[idx - 1])) if hetero_c else c0_s
        f["transducer_source_flag"] = _scalar(flag, np.uint64)
        f["transducer_source_input"] = _signal(sig[None] * (2 * c_here * dt / kgrid.dx))
        f["delay_mask"] = _col(delay.reshape(-1, order="F")[idx - 1].astype(np.int64) + 1)
        f["u_source_index"] = _col(idx)

    # sensor: binary mask, or cuboid corners (saveToDisk.m:35-37, 355-378)
    if sensor.get("corners") is not None:
        c = np.asarray(sensor["corners"], dtype=np.int64)
        c = c.reshape(c.shape[0], -1)
        if dim == 2:
            if c.shape[0] != 4:
                raise ValueError("2-D cuboid corners are [x1; y1; x2; y2] columns")
            c6 = np.ones((6, c.shape[1]), dtype=np.int64)       # z corners 1 (saveToDisk.m:359-366)
            c6[0], c6[1], c6[3], c6[4] = c[0], c[1], c[2], c[3]
            c = c6
        elif c.shape[0] != 6:
            raise ValueError("3-D cuboid corners are [x1; y1; z1; x2; y2; z2] columns")
        f["sensor_mask_type"] = _scalar(1, np.uint64)
        f["sensor_mask_corners"] = np.ascontiguousarray(c.T[None].astype(np.uint64))   # [6, K] -> (1, K, 6)
        return f
    mask = np.asarray(sensor["mask"])
    if mask.shape != shape:
        raise ValueError("sensor.mask must be a binary grid mask of the grid's shape")
    f["sensor_mask_index"] = _col(_find(mask))
    return f


def plane_wave_check_stable(kgrid: KWaveGrid, c_max: float) -> float:
    """Largest stable dt for a lossless homogeneous medium (checkStability.m:78-143):
    dt_limit = 2 / (c_ref * k_max) with k_max = pi * sqrt(sum 1/d^2)."""
    inv = (1 / kgrid.dx ** 2 + 1 / kgrid.dy ** 2 + (1 / kgrid.dz ** 2 if kgrid.dim == 3 else 0))
    kmax = math.pi * math.sqrt(inv)
    return 2.0 / (c_max * kmax)


class CartesianSensor:
    """Cartesian sensor points with 'linear' interpolation, for the binary-mask contract.

    kspaceFirstOrder2D interpolates p at Cartesian sensor points from the
    Delaunay triangulation of the grid points and the barycentric coordinates
    of each point in its triangle (kspaceFirstOrder_createStorageVariables.m:
    370-397 -> gridDataFast2D.m; applied each step as sum(p(tri) .* bc, 2),
    kspaceFirstOrder_extractSensorData.m:297-306).  The external binaries take
    binary masks only (kspaceFirstOrder_saveToDisk.m:35-37), so here the mask
    is the union of the triangle vertices and the weights are applied to the
    recorded vertex time series afterwards.

    Triangulation rule: the Delaunay triangulation of a regular grid is
    degenerate (every cell is cocircular), so which diagonal splits a cell is
    implementation-defined; this rule always splits cell (i, j) along the
    diagonal from grid point (i, j) to (i+1, j+1).  Interpolation of any field
    linear in x and y is exact under either choice.
    """

    def __init__(self, kgrid: KWaveGrid, cart_pos):
        if kgrid.dim != 2:
            raise NotImplementedError("Cartesian sensors are implemented for 2-D grids")
        pos = np.asarray(cart_pos, dtype=np.float64).reshape(2, -1)
        Nx, Ny = kgrid.Nx, kgrid.Ny
        # kgrid.x_vec(i) = (i - floor(N/2)) * dx, 0-based i (kWaveGrid.m:494-495, makeDim 779-797)
        gx = pos[0] / kgrid.dx + Nx // 2
        gy = pos[1] / kgrid.dy + Ny // 2
        if np.any(gx < 0) or np.any(gx > Nx - 1) or np.any(gy < 0) or np.any(gy > Ny - 1):
            raise ValueError("Cartesian points must lie within the k-space grid defined by kgrid.")
        i = np.minimum(np.floor(gx).astype(np.int64), Nx - 2)
        j = np.minimum(np.floor(gy).astype(np.int64), Ny - 2)
        ux, uy = gx - i, gy - j
        lower = ux >= uy
        v1x = np.where(lower, i + 1, i)
        v1y = np.where(lower, j, j + 1)
        vx = np.stack([i, v1x, i + 1], axis=1)
        vy = np.stack([j, v1y, j + 1], axis=1)
        self.bc = np.stack([1.0 - np.maximum(ux, uy), np.abs(ux - uy), np.minimum(ux, uy)], axis=1)
        lin = vx + Nx * vy                                  # 0-based column-major index
        uniq = np.unique(lin)                               # find() order of the mask
        flat = np.zeros(Nx * Ny)
        flat[uniq] = 1.0
        self.mask = flat.reshape((Nx, Ny), order="F")
        self.tri = np.searchsorted(uniq, lin)               # column of each vertex in the recorded data
        self.num_points = pos.shape[1]

    def apply(self, rec: np.ndarray) -> np.ndarray:
        """Interpolate recorded mask data (..., Nmask) -> (..., Ns) in float64."""
        r = np.asarray(rec, dtype=np.float64)
        return np.einsum("...sk,sk->...s", r[..., self.tri], self.bc)


def fourier_shift(data, shift: float) -> np.ndarray:
    """fourierShift.m along the last axis (the time axis of sensor data): the
    band-limited interpolant of each series advanced by `shift` samples,
    real(ifft(ifftshift(exp(i k shift)) * fft(data))), with the wavenumber
    k = 2 pi m / N, m = -floor(N/2) .. ceil(N/2) - 1, the m = 0 entry set to an
    exact 0 and the even-length Nyquist bin kept at -pi (fourierShift.m:81-112)."""
    x = np.asarray(data, dtype=np.float64)
    n = x.shape[-1]
    m = np.arange(-(n // 2), n - n // 2, dtype=np.float64)
    k = 2.0 * np.pi * m / n
    k[n // 2] = 0.0
    ph = np.fft.ifftshift(np.exp(1j * k * shift))
    return np.real(np.fft.ifft(ph * np.fft.fft(x, axis=-1), axis=-1))


def save_intensity(sensor_data, dim: int, record):
    """kspaceFirstOrder_saveIntensity.m:42-104 on the sensor data the binary
    returned (a dict, or a list of dicts, one per cuboid): u_non_staggered is
    shifted forward by dt/2 in time (fourier_shift by 1/2) so it sits on p's
    time grid, I_i = p * u_i, and I_i_avg is its mean over the time (last) axis.
    Fields the record does not ask for (the time-varying intensity,
    u_non_staggered, p) are dropped, as the MATLAB code removes them."""
    if isinstance(sensor_data, (list, tuple)):
        return [save_intensity(s, dim, record) for s in sensor_data]
    out = dict(sensor_data)
    p = np.asarray(out["p"], dtype=np.float64)
    for a in "xyz"[:dim]:
        I = p * fourier_shift(out[f"u{a}_non_staggered"], 0.5)
        out[f"I{a}"] = I
        if "I_avg" in record:
            out[f"I{a}_avg"] = I.mean(axis=-1)
    for a in "xyz"[:dim]:
        if "u_non_staggered" not in record:
            out.pop(f"u{a}_non_staggered")
        if "I" not in record:
            out.pop(f"I{a}")
    if "p" not in record:
        out.pop("p")
    return out
