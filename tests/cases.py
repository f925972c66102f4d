"""Seeded simulation inputs shared by the oracle and the GPU parity tests.

Every case is built through the front-end mirror (frontend.save_to_disk_dict)
so the oracle and the HIP solver read the identical input-file dataset dict.
Sizes are chosen so the FP64 oracle finishes in a few seconds.
"""

from __future__ import annotations

import numpy as np

import kwave_amd as kw

fe = kw.frontend
geo = kw.geometry


def smooth_field(rng, shape, sigma=2.0):
    """Unit-std Gaussian random field: white noise filtered by a Gaussian of sigma points."""
    w = rng.standard_normal(shape)
    k = [np.fft.fftfreq(n) for n in shape]
    grids = np.meshgrid(*k, indexing="ij")
    g = np.exp(-2 * (np.pi * sigma) ** 2 * sum(x ** 2 for x in grids))
    f = np.real(np.fft.ifftn(np.fft.fftn(w) * g))
    return f / f.std()


def case_2d_p0_lossless(N=64, Nt=80, pml=10, het=False, seed=1):
    rng = np.random.default_rng(seed)
    kgrid = fe.KWaveGrid(N, 1e-4, N, 1e-4)
    c = 1500.0 * (1 + (0.05 * smooth_field(rng, (N, N)) if het else 0.0))
    rho = 1000.0 * (1 + (0.05 * smooth_field(rng, (N, N)) if het else 0.0))
    kgrid.set_time(Nt, 0.3 * 1e-4 / float(np.max(c)))
    p0 = 5.0 * geo.make_disc(N, N, 0, 0, 5)
    mask = (rng.random((N, N)) < 0.02).astype(float)
    return fe.save_to_disk_dict(kgrid, {"sound_speed": c, "density": rho}, {"p0": p0}, {"mask": mask},
                                pml_size=pml)


def case_2d_absorbing_psource(N=(60, 48), Nt=120, mode="additive", many=False, nonlinear=False,
                              stokes=False, seed=2):
    rng = np.random.default_rng(seed)
    Nx, Ny = N
    kgrid = fe.KWaveGrid(Nx, 1e-4, Ny, 1.2e-4)
    c = 1500.0 * (1 + 0.04 * smooth_field(rng, (Nx, Ny)))
    rho = 1000.0 * (1 + 0.04 * smooth_field(rng, (Nx, Ny)))
    kgrid.set_time(Nt, 0.25 * 1e-4 / float(np.max(c)))
    medium = {"sound_speed": c, "density": rho,
              "alpha_coeff": 0.5 + 0.2 * rng.random((Nx, Ny)), "alpha_power": 2.0 if stokes else 1.5}
    if stokes:
        medium["alpha_mode"] = "stokes"
    if nonlinear:
        medium["BonA"] = 6.0
    pmask = np.zeros((Nx, Ny))
    pmask[15, 10:38] = 1
    npts = int(pmask.sum())
    burst = geo.tone_burst(1 / kgrid.dt, 1.5e6, 3)
    if many:
        sig = np.vstack([np.roll(burst[0], k) for k in range(npts)]) * 1e5
    else:
        sig = burst * 1e5
    mask = np.zeros((Nx, Ny))
    mask[40, :] = 1
    mask[:, 30] = 1
    return fe.save_to_disk_dict(kgrid, medium, {"p": sig, "p_mask": pmask, "p_mode": mode}, {"mask": mask},
                                pml_size=8)


def case_3d(N=(32, 28, 24), Nt=60, absorbing=True, nonlinear=False, source="ux", mode="additive", many=False,
            het=True, p0=False, seed=3, alpha_power=1.5, pml=6):
    rng = np.random.default_rng(seed)
    Nx, Ny, Nz = N
    dx = 2e-4
    kgrid = fe.KWaveGrid(Nx, dx, Ny, dx, Nz, dx)
    shape = (Nx, Ny, Nz)
    c = 1500.0 * (1 + (0.03 * smooth_field(rng, shape) if het else 0.0))
    rho = 1000.0 * (1 + (0.03 * smooth_field(rng, shape) if het else 0.0))
    kgrid.set_time(Nt, 0.3 * dx / float(np.max(c)))
    medium = {"sound_speed": c, "density": rho}
    if absorbing:
        medium["alpha_coeff"] = (0.5 + 0.25 * rng.random(shape)) if het else 0.75
        medium["alpha_power"] = alpha_power
    if nonlinear:
        medium["BonA"] = 5.0
    src = {}
    if p0:
        src["p0"] = 3.0 * geo.make_ball(Nx, Ny, Nz, 0, 0, 0, 3)
    if source:
        m = np.zeros(shape)
        m[pml + 2, Ny // 2 - 3: Ny // 2 + 3, Nz // 2 - 3: Nz // 2 + 3] = 1
        npts = int(m.sum())
        burst = geo.tone_burst(1 / kgrid.dt, 1e6, 3)[0]
        sig = np.vstack([np.roll(burst, k) for k in range(npts)]) if many else burst[None]
        if source == "p":
            src.update({"p": sig * 1e5, "p_mask": m, "p_mode": mode})
        else:
            src.update({"ux": sig * 0.5, "u_mask": m, "u_mode": mode})
            if source == "uxyz":
                src["uy"] = sig * 0.3
                src["uz"] = sig * 0.2
    mask = np.zeros(shape)
    mask[:, :, Nz // 2] = 1
    return fe.save_to_disk_dict(kgrid, medium, src, {"mask": mask}, pml_size=pml)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def case_transducer(N=(48, 40, 32), Nt=60, seed=5, het=True, steer=True, pml=4):
    """3-D transducer source: active elements on a rectangle in the plane x = pml + 3,
    beam-steering delays along y (0..7 samples), one tone-burst input signal."""
    rng = np.random.default_rng(seed)
    Nx, Ny, Nz = N
    dx = 2e-4
    kgrid = fe.KWaveGrid(Nx, dx, Ny, dx, Nz, dx)
    shape = (Nx, Ny, Nz)
    c = 1500.0 * (1 + (0.03 * smooth_field(rng, shape) if het else 0.0))
    rho = 1000.0 * (1 + (0.03 * smooth_field(rng, shape) if het else 0.0))
    kgrid.set_time(Nt, 0.3 * dx / float(np.max(c)))
    active = np.zeros(shape)
    active[pml + 2, Ny // 2 - 8: Ny // 2 + 8, Nz // 2 - 4: Nz // 2 + 4] = 1
    delay = np.zeros(shape, dtype=np.int64)
    if steer:
        yy = np.arange(Ny)[None, :, None] * np.ones(shape, dtype=np.int64)
        delay = np.where(active > 0, (yy - (Ny // 2 - 8)) // 2, 0)
    sig = geo.tone_burst(1 / kgrid.dt, 1.5e6, 3)[0] * 0.5
    sig = np.concatenate([sig, np.zeros(Nt)])
    medium = {"sound_speed": c, "density": rho, "alpha_coeff": 0.5, "alpha_power": 1.5}
    mask = np.zeros(shape)
    mask[:, :, Nz // 2] = 1
    return fe.save_to_disk_dict(kgrid, medium, {"transducer": {"input_signal": sig, "delay_mask": delay,
                                                                "active_elements_mask": active}},
                                {"mask": mask}, pml_size=pml)


def case_labelled(kind="p", mode="additive", het=False, only=None, Nt=60, seed=7):
    """Labelled source mask (kspaceFirstOrder_inputChecking.m:847-884, 964-1003):
    labels 1..3 scattered over a line of source points, one signal row per
    label.  kind "p": 2-D 48x40 pressure source; kind "u": 3-D 28x24x20 ux/uy
    source.  only=k keeps just label k's points as a binary mask driven by
    row k alone (the superposition pieces)."""
    rng = np.random.default_rng(seed)
    dim = 2 if kind == "p" else 3
    shape = (48, 40) if dim == 2 else (28, 24, 20)
    dx = 1e-4
    kgrid = fe.KWaveGrid(shape[0], dx, shape[1], dx) if dim == 2 else fe.KWaveGrid(*[v for n in shape for v in (n, dx)])
    c = 1500.0 * (1 + (0.03 * smooth_field(rng, shape) if het else 0.0))
    rho = 1000.0 * (1 + (0.03 * smooth_field(rng, shape) if het else 0.0))
    kgrid.set_time(Nt, 0.3 * dx / float(np.max(c)))
    m = np.zeros(shape)
    line = (slice(10, 11), slice(6, 34)) if dim == 2 else (slice(8, 9), slice(5, 19), slice(9, 11))
    npts = int(np.prod([s.stop - s.start for s in line]))
    labels = 1 + (np.arange(npts) * 5 + np.arange(npts) // 3) % 3
    m[line] = labels.reshape([s.stop - s.start for s in line], order="F")
    m[tuple(min(n - 6, 30) for n in shape)] = 2                          # a stray point of label 2
    t = np.arange(Nt) * kgrid.dt
    rows = np.stack([np.sin(2 * np.pi * f0 * t) * np.exp(-((t - t0) / 4e-7) ** 2)
                     for f0, t0 in ((1.0e6, 8e-7), (1.6e6, 1.1e-6), (0.7e6, 1.4e-6))])
    if only is not None:
        m = (m == only).astype(float)
        rows = rows[only - 1:only]
    src = ({"p": rows * 1e5, "p_mask": m, "p_mode": mode} if kind == "p"
           else {"ux": rows * 0.5, "uy": rows[::-1] * 0.3 if only is None else rows * 0.3,
                 "u_mask": m, "u_mode": mode})
    if kind == "u" and only is not None:
        # the uy rows are reversed in the labelled run: label k drives uy with row 4-k
        full = np.stack([np.sin(2 * np.pi * f0 * t) * np.exp(-((t - t0) / 4e-7) ** 2)
                         for f0, t0 in ((1.0e6, 8e-7), (1.6e6, 1.1e-6), (0.7e6, 1.4e-6))])
        src["uy"] = full[::-1][only - 1:only] * 0.3
    medium = {"sound_speed": c, "density": rho}
    sensor = np.zeros(shape)
    if dim == 2:
        sensor[30, :] = 1
    else:
        sensor[:, :, 12] = 1
    return fe.save_to_disk_dict(kgrid, medium, src, {"mask": sensor}, pml_size=5)


def case_axisym(Nx=64, Ny=32, Nt=50, dx=1e-4, source="p0", het=False, absorbing=False, nonlinear=False,
                mode="additive", sigma=2.5, pml=8, seed=11):
    """Axisymmetric input (kspaceFirstOrderASC): x axial, y radial from the axis.
    source "p0": a Gaussian on the axis; "p": a disc of p points across the
    axis (a piston face); "u": ux and uy on a ring; sensors on the axis and on
    a radial line."""
    rng = np.random.default_rng(seed)
    kg = fe.KWaveGrid(Nx, dx, Ny, dx)
    c = 1500.0 * (1 + (0.03 * smooth_field(rng, (Nx, Ny)) if het else 0.0))
    rho = 1000.0 * (1 + (0.03 * smooth_field(rng, (Nx, Ny)) if het else 0.0))
    kg.set_time(Nt, 0.3 * dx / float(np.max(c)))
    medium = {"sound_speed": c, "density": rho}
    if absorbing:
        medium["alpha_coeff"] = (0.5 + 0.3 * rng.random((Nx, Ny))) if het else 0.75
    if nonlinear:
        medium["BonA"] = 5.0
    src = {}
    if source == "p0":
        x = (np.arange(Nx) - Nx // 2)[:, None] * dx
        r = np.arange(Ny)[None, :] * dx
        src["p0"] = 2.0 * np.exp(-(x ** 2 + r ** 2) / (2 * (sigma * dx) ** 2))
    else:
        burst = geo.tone_burst(1 / kg.dt, 1.5e6, 3)[0]
        m = np.zeros((Nx, Ny))
        if source == "p":
            m[pml + 2, :Ny // 3] = 1
            src.update({"p": burst[None] * 1e5, "p_mask": m, "p_mode": mode})
        else:
            m[pml + 4, Ny // 4: Ny // 4 + 3] = 1
            m[pml + 5: pml + 8, Ny // 4] = 1
            src.update({"ux": burst[None] * 0.5, "uy": np.roll(burst, 5)[None] * 0.3, "u_mask": m, "u_mode": mode})
    mask = np.zeros((Nx, Ny))
    mask[:, 0] = 1
    mask[Nx // 2, :] = 1
    return fe.save_to_disk_dict(kg, medium, src, {"mask": mask}, pml_size=pml, smooth_p0=False, axisymmetric=True)


def cw_signal(t, f0, amp, ramp_periods=4):
    """createCWSignals.m: amp sin(2 pi f0 t) with a cosine up-ramp of 4 periods."""
    s = amp * np.sin(2 * np.pi * f0 * t)
    dt = t[1] - t[0]
    n = int(round(ramp_periods / f0 / dt))
    ramp = (-np.cos(np.linspace(0, np.pi, n)) + 1) * 0.5
    s[:n] *= ramp
    return s


def piston_axisym():
    """testing/unit/kspaceFirstOrderAS_piston_analytical.m: a CW piston of radius
    5 mm at 1 MHz on the axis of an axisymmetric grid (12 points per wavelength,
    CFL 0.25, 40 us).  The reference runs it with PMLInside false; here the PML
    is part of the grid (320 x 100, radices 2 and 5) and the piston sits on the
    first interior plane.  Returns (input dict, record_start_index, error_fn):
    error_fn(p) is the L-inf error (% of the peak) of the steady-state on-axis
    amplitude of the recorded p against P0 |2 sin(k (sqrt(x^2 + a^2) - x) / 2)|."""
    c0, rho0, f0, radius, mag = 1500.0, 1000.0, 1e6, 5e-3, 1e6
    ppw, cfl, t_end, periods = 12, 0.25, 40e-6, 3
    dx = c0 / (ppw * f0)
    nx_user, ny_user = 256, 80                      # roundEven(32 mm / dx), roundEven(10 mm / dx)
    pml_x, pml_y = 32, 20
    Nx, Ny = nx_user + 2 * pml_x, ny_user + pml_y
    ppp = round(ppw / cfl)
    dt = 1 / (ppp * f0)
    Nt = round(t_end / dt)
    kg = fe.KWaveGrid(Nx, dx, Ny, dx)
    kg.set_time(Nt, dt)
    rgp = round(radius / dx)
    pmask = np.zeros((Nx, Ny))
    pmask[pml_x, :rgp] = 1
    sens = np.zeros((Nx, Ny))
    sens[pml_x + 1:pml_x + nx_user, 0] = 1           # the axis, x = 2 .. end of the user grid
    sig = cw_signal(np.arange(Nt) * dt, f0, mag)
    f = fe.save_to_disk_dict(kg, {"sound_speed": c0, "density": rho0}, {"p": sig[None], "p_mask": pmask},
                             {"mask": sens}, pml_size=[pml_x, pml_y], axisymmetric=True)
    start = Nt - periods * ppp + 1

    def error_pct(p):
        n = p.shape[0]
        amp = 2 / n * np.abs(np.exp(-2j * np.pi * f0 * dt * np.arange(n)) @ p)   # 3 whole periods, rectangular
        x = np.arange(1, nx_user) * dx
        k = 2 * np.pi * f0 / c0
        a = (rgp - 0.5) * dx
        ref = mag * np.abs(2 * np.sin((k * np.sqrt(x ** 2 + a ** 2) - k * x) / 2))
        return 100 * np.max(np.abs(ref - amp)) / np.max(ref)

    return f, start, error_pct


def fourier_shift(data, shift):
    """fourierShift.m along the last axis: real(ifft(ifftshift(exp(i k shift)) .* fft(data)))."""
    n = data.shape[-1]
    k = 2 * np.pi * (np.arange(-(n // 2), n - n // 2) / n)
    k[n // 2] = 0.0
    return np.real(np.fft.ifft(np.fft.ifftshift(np.exp(1j * k * shift)) * np.fft.fft(data, axis=-1), axis=-1))


PLANE_WAVE_I0 = 10.0   # W/cm^2


def plane_wave_intensity_objects():
    """The kgrid, medium, source, sensor, PML options and record start index of
    plane_wave_intensity(), for the kspaceFirstOrder2DG-style flow."""
    Nx, Nj, dx = 64, 32, 0.05e-3
    c0, rho0, f0, I0 = 1500.0, 1000.0, 10e6, PLANE_WAVE_I0
    strength = np.sqrt(2 * I0 * 1e4 * rho0 * c0)
    ppp = 30
    dt = 1 / f0 / (2 * ppp)
    Nt = int(np.floor(3e-6 / dt + 1e-9)) + 1
    t = np.arange(Nt) * dt
    sig = strength * np.sin(2 * np.pi * f0 * t)
    ramp = (-np.cos(np.arange(200) * np.pi / 200) + 1) / 2
    sig[:200] *= ramp
    kg = fe.KWaveGrid(Nx, dx, Nj, dx)
    kg.set_time(Nt, dt)
    pm = np.zeros((Nx, Nj))
    pm[20, :] = 1                                    # source_position = PML_size + 1
    sm = np.zeros((Nx, Nj))
    sm[Nx - 22, Nj // 2 - 1] = 1                     # sensor_position = Nx - PML_size - 1, Nj/2
    pml = {"pml_size": [20, 2], "pml_alpha": [2.0, 0.0]}
    return (kg, {"sound_speed": c0, "density": rho0}, {"p": sig[None], "p_mask": pm}, {"mask": sm}, pml,
            Nt - 2 * ppp + 1)


def plane_wave_intensity():
    """testing/unit/kspaceFirstOrder_check_plane_wave_intensity.m, 2-D x case: a
    10 MHz plane wave of 10 W/cm^2 from a p source line; the time-averaged
    intensity p .* fourierShift(ux_non_staggered, 1/2) (saveIntensity.m) over the
    last two periods at a point across the grid must be 10 W/cm^2 within 1 %.
    Returns (input dict, record_start_index, error_fn(p, ux_non_staggered) -> %)."""
    kg, medium, source, sensor, pml, start = plane_wave_intensity_objects()
    f = fe.save_to_disk_dict(kg, medium, source, sensor, **pml)
    I0 = PLANE_WAVE_I0

    def error_pct(p, ux_ns):
        p = np.asarray(p, dtype=np.float64).reshape(-1)
        u = np.asarray(ux_ns, dtype=np.float64).reshape(-1)
        I_avg = np.mean(p * fourier_shift(u, 0.5)) / 1e4      # W/cm^2
        return 100 * abs(I_avg - I0) / I0

    return f, start, error_pct


def nonstaggered_velocity_order():
    """testing/unit/kspaceFirstOrder2D_check_nonstaggered_particle_velocity.m: a
    smoothed disc p0 (radius 5, magnitude 5) at the centre of a 128^2 absorbing
    grid (0.75 dB/(MHz^1.5 cm)), CFL 0.1, 6 us; sensors 40 and 41 points along
    +x.  Returns the input dict; the check orders the peaks of ux_non_staggered
    and ux (half a cell further out) at the two points."""
    Nx = Ny = 128
    kg = fe.KWaveGrid(Nx, 0.1e-3, Ny, 0.1e-3)
    medium = {"sound_speed": 1500.0, "density": 1000.0, "alpha_coeff": 0.75, "alpha_power": 1.5}
    kg.make_time(1500.0, cfl=0.1, t_end=6e-6)
    p0 = 5.0 * geo.make_disc(Nx, Ny, Nx // 2, Ny // 2, 5)
    m = np.zeros((Nx, Ny))
    m[Nx // 2 + 40 - 1, Ny // 2 - 1] = 1
    m[Nx // 2 + 41 - 1, Ny // 2 - 1] = 1
    return fe.save_to_disk_dict(kg, medium, {"p0": p0}, {"mask": m})


def velocity_order_holds(ux_ns, ux):
    """max1 >= max2 >= max3 >= max4 and ind1 <= ind2 <= ind3 <= ind4 for
    (ux_ns at point 1, ux at point 1, ux_ns at point 2, ux at point 2); series
    are (Nt, 2) as recorded."""
    s = [ux_ns[:, 0], ux[:, 0], ux_ns[:, 1], ux[:, 1]]
    mx = [float(np.max(v)) for v in s]
    ix = [int(np.argmax(v)) for v in s]
    return all(mx[i] >= mx[i + 1] for i in range(3)) and all(ix[i] <= ix[i + 1] for i in range(3)), mx, ix


def w_correction_plane_wave(kind="p", variant="k", fine=False, Nx=64, Ny=8, f0=1e6, cfl=0.5, sc=10):
    """kspaceFirstOrder*_check_source_scaling_*.m in miniature: a plane wave from a
    source line across a periodic y axis (no y PML), recorded 40 points away.
    variant "k": additive (k-space corrected); "w": additive with
    p_frequency_ref / u_frequency_ref = f0 (cos(2 pi f0 dt/2) scaling, no
    k-space correction); "none": additive-no-correction.  fine=True: dt / sc with no
    correction, the reference the coarse runs are judged against."""
    dx = 1e-4
    kg = fe.KWaveGrid(Nx, dx, Ny, dx)
    dt = cfl * dx / 1500.0 / (sc if fine else 1)
    Nt = int(round(12e-6 / dt))
    kg.set_time(Nt, dt)
    t = np.arange(Nt) * dt
    burst = np.sin(2 * np.pi * f0 * t) * np.where(t < 3 / f0, np.sin(np.pi * t * f0 / 3) ** 2, 0.0)
    m = np.zeros((Nx, Ny))
    m[12, :] = 1
    sens = np.zeros((Nx, Ny))
    sens[52, 0] = 1
    mode = "additive" if (variant == "k" and not fine) else "additive-no-correction"
    if kind == "p":
        src = {"p": burst[None] * 1e5, "p_mask": m, "p_mode": mode}
    else:
        src = {"ux": burst[None] * 0.1, "u_mask": m, "u_mode": mode}
    if variant == "w" and not fine:
        src[("p" if kind == "p" else "u") + "_frequency_ref"] = f0
    return fe.save_to_disk_dict(kg, {"sound_speed": 1500.0, "density": 1000.0}, src, {"mask": sens},
                                pml_size=[10, 0])
