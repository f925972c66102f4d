"""Seeded synthetic workloads (input-file dataset dicts).

C1  2-D 128^2 homogeneous lossless p0, Cartesian circle sensor (p)
C2  3-D 128^3 heterogeneous absorbing, 16x16 ux additive patch, mid-plane p + p_max
C3  3-D 256x256x512 nonlinear absorbing (y=2), bowl p source, all-point p_max + p_rms
C4  3-D 768^3 heterogeneous absorbing three-layer slab, 64x64 p patch, 128^2 plane p
C5  serving stream: requests on 64^3 .. 160x160x192 grids, 1% random sensor

RNG: numpy.random.default_rng(1000 + config number), draws in the order
written below.  Random media are unit-std Gaussian random fields: white noise
filtered by a Gaussian of sigma = correlation length (in points), rescaled.
"""

from __future__ import annotations

import math

import numpy as np

from . import frontend as fe
from . import geometry as geo


def gaussian_random_field(rng, shape, corr_len: float) -> np.ndarray:
    """White noise filtered by a Gaussian of std corr_len points, unit std."""
    w = rng.standard_normal(shape)
    ks = np.meshgrid(*[np.fft.fftfreq(n) for n in shape], indexing="ij", sparse=True)
    g = np.exp(-2.0 * (math.pi * corr_len) ** 2 * sum(k ** 2 for k in ks))
    f = np.fft.ifftn(np.fft.fftn(w) * g).real
    return f / f.std()


C1_DESCRIPTION = ("2-D 128^2 homogeneous lossless initial-value problem, PML 20 (alpha 2), dx 0.1 mm, disc p0 "
                  "(radius 5, 5 Pa, smoothed), 50 Cartesian sensor points on a 4 mm circle (linear "
                  "interpolation), record p, CFL 0.3, Nt 500")


def c1(Nt: int | None = None):
    """2-D homogeneous lossless p0 problem with a Cartesian circle sensor.

    Returns (input-file dataset dict, frontend.CartesianSensor): the dict's
    binary mask holds the interpolation vertices; sensor.apply() maps the
    recorded p (Nt, Nmask) to the 50 Cartesian points (Nt, 50).
    """
    N, dx = 128, 1e-4
    kgrid = fe.KWaveGrid(N, dx, N, dx)
    c0, rho0 = 1500.0, 1000.0
    kgrid.set_time(500 if Nt is None else Nt, 0.3 * dx / c0)
    p0 = 5.0 * geo.make_disc(N, N, 0, 0, 5)
    sensor = fe.CartesianSensor(kgrid, geo.make_cart_circle(4e-3, 50))
    f = fe.save_to_disk_dict(kgrid, {"sound_speed": c0, "density": rho0}, {"p0": p0}, {"mask": sensor.mask},
                             pml_size=20, pml_alpha=2.0)
    return f, sensor


def c2(Nt: int | None = None, N: int = 128) -> dict:
    """3-D heterogeneous absorbing linear medium (128^3, PML 10, dx = 0.2 mm)."""
    rng = np.random.default_rng(1002)
    dx = 0.2e-3
    shape = (N, N, N)
    kgrid = fe.KWaveGrid(N, dx, N, dx, N, dx)
    g1 = gaussian_random_field(rng, shape, 4.0)
    g2 = gaussian_random_field(rng, shape, 4.0)
    c = 1500.0 * (1 + 0.02 * g1)
    rho = 1000.0 * (1 + 0.02 * g2)
    alpha = np.full(shape, 0.0022)
    ball = geo.make_ball(N, N, N, 0, 0, 0, 20 * N // 128) > 0
    c[ball], rho[ball], alpha[ball] = 1580.0, 1050.0, 0.75
    dt = 0.3 * dx / float(c.max())
    kgrid.set_time(1200 if Nt is None else Nt, dt)
    u_mask = np.zeros(shape)
    h = 8 * N // 128
    u_mask[10, N // 2 - h: N // 2 + h, N // 2 - h: N // 2 + h] = 1       # x = 11 (1-based)
    ux = geo.tone_burst(1.0 / dt, 1e6, 3)
    sensor = np.zeros(shape)
    sensor[:, :, N // 2 - 1] = 1                                           # z = 64 (1-based)
    medium = {"sound_speed": c, "density": rho, "alpha_coeff": alpha, "alpha_power": 1.5}
    return fe.save_to_disk_dict(kgrid, medium, {"ux": ux, "u_mask": u_mask, "u_mode": "additive"},
                                {"mask": sensor}, pml_size=10)


C2_RECORD = ("p", "p_max")

C3_RECORD = ("p_max", "p_rms")
C3_DESCRIPTION = ("3-D 256x256x512 nonlinear (B/A 5) absorbing (y=2) homogeneous water, PML 10, dx 0.25 mm, "
                  "focused bowl p source (R 96, aperture 127 points) driven by a 1 MPa 0.5 MHz 5-cycle "
                  "tone burst, every grid point recording p_max and p_rms (FP64), CFL 0.2, Nt 2000")


def c3(Nt: int | None = None, shape=(256, 256, 512), radius: int = 96, diameter: int = 127,
       dx: float = 0.25e-3, pml: int = 10) -> dict:
    """3-D nonlinear focused field: makeBowl(grid, [Nx/2 Ny/2 11], R, d, [Nx/2 Ny/2 Nz/2])."""
    Nx, Ny, Nz = shape
    kgrid = fe.KWaveGrid(Nx, dx, Ny, dx, Nz, dx)
    c0 = 1500.0
    dt = 0.2 * dx / c0
    kgrid.set_time(2000 if Nt is None else Nt, dt)
    bowl = geo.make_bowl(shape, (Nx // 2, Ny // 2, 11), radius, diameter, (Nx // 2, Ny // 2, Nz // 2))
    sig = 1e6 * geo.tone_burst(1.0 / dt, 0.5e6, 5)
    medium = {"sound_speed": c0, "density": 1000.0, "alpha_coeff": 0.0022, "alpha_power": 2.0, "BonA": 5.0}
    return fe.save_to_disk_dict(kgrid, medium, {"p": sig, "p_mask": bowl, "p_mode": "additive"},
                                {"mask": np.ones(shape, dtype=np.float32)}, pml_size=pml)
C2_DESCRIPTION = ("3-D 128^3 heterogeneous absorbing (y=1.5) linear medium, PML 10, 16x16 ux additive "
                  "tone-burst patch, mid-plane binary sensor recording p and p_max, CFL 0.3, Nt 1200")


C4_RECORD = ("p",)
C4_DESCRIPTION = ("3-D 768^3 heterogeneous absorbing (y=1.1) three-layer slab (1% Gaussian perturbation), PML 20, "
                  "dx 0.2 mm, 64x64 p additive patch (1 MHz, 4 cycles), 128x128 receiver plane recording p, "
                  "CFL 0.3, Nt 1500")


def unit_random_field(rng, shape, corr_len: float) -> np.ndarray:
    """float32 white noise filtered by a Gaussian of std corr_len points, unit std (real FFTs, all cores)."""
    import scipy.fft as sfft

    w = rng.standard_normal(shape, dtype=np.float32)
    ks = [np.fft.fftfreq(n).astype(np.float32) for n in shape[:-1]] + [np.fft.rfftfreq(shape[-1]).astype(np.float32)]
    k2 = sum(k.reshape([-1 if a == i else 1 for a in range(len(shape))]) ** 2 for i, k in enumerate(ks))
    g = np.exp(np.float32(-2.0 * math.pi ** 2 * corr_len ** 2) * k2)
    spec = sfft.rfftn(w, workers=-1)
    del w
    spec *= g
    f = sfft.irfftn(spec, s=shape, workers=-1).astype(np.float32)
    del spec
    f /= np.float32(f.std())
    return f


def c4(Nt: int | None = None, N: int = 768, log: bool = False) -> dict:
    """3-D three-layer slab, heterogeneous absorbing (y=1.1), PML 20, dx = 0.2 mm."""
    import sys
    import time

    t0 = time.perf_counter()

    def say(what):
        if log:
            print(f"c4: {what} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)

    rng = np.random.default_rng(1004)
    dx = 0.2e-3
    shape = (N, N, N)
    kgrid = fe.KWaveGrid(N, dx, N, dx, N, dx)
    pml = 20
    interior = N - 2 * pml
    x = np.arange(N)
    layer = np.where(x < pml + interior // 3, 0, np.where(x < pml + 2 * interior // 3, 1, 2))
    cs, rs, al = np.array([1480.0, 1560.0, 2800.0]), np.array([1000.0, 1050.0, 1900.0]), np.array([0.0022, 0.6, 4.0])
    # fields are built in MATLAB order with Fortran memory layout, so the
    # file's (Nz, Ny, Nx) C-order datasets are views, not transposed copies
    pert = 1 + np.float32(0.01) * unit_random_field(rng, shape[::-1], 4.0).T
    c = (cs[layer].astype(np.float32)[:, None, None] * pert).astype(np.float32, order="K")
    rho = (rs[layer].astype(np.float32)[:, None, None] * pert).astype(np.float32, order="K")
    alpha = (al[layer].astype(np.float32)[:, None, None] * np.ones_like(pert)).astype(np.float32, order="K")
    del pert
    say("medium built")
    dt = 0.3 * dx / float(c.max())
    kgrid.set_time(1500 if Nt is None else Nt, dt)
    p_mask = np.zeros(shape, dtype=np.float32, order="F")
    hs = min(32, N // 2 - pml)          # 64x64 patch (smaller on reduced grids)
    p_mask[pml, N // 2 - hs: N // 2 + hs, N // 2 - hs: N // 2 + hs] = 1   # x = 21 (1-based)
    sig = geo.tone_burst(1.0 / dt, 1e6, 4)
    sensor = np.zeros(shape, dtype=np.float32, order="F")
    hr = min(64, N // 2 - pml)          # 128x128 receiver plane (smaller on reduced grids)
    sensor[N - pml - 1, N // 2 - hr: N // 2 + hr, N // 2 - hr: N // 2 + hr] = 1
    medium = {"sound_speed": c, "density": rho, "alpha_coeff": alpha, "alpha_power": 1.1}
    f = fe.save_to_disk_dict(kgrid, medium, {"p": sig, "p_mask": p_mask, "p_mode": "additive"},
                             {"mask": sensor}, pml_size=pml)
    say("input datasets built")
    return f


C5_GRIDS = ((64, 64, 64), (96, 96, 96), (128, 128, 128), (160, 160, 192))
C5_NT = (400, 1500)
C5_RECORD = ("p",)
C5_DESCRIPTION = ("serving stream: requests on 64^3, 96^3, 128^3 or 160x160x192 grids (PML 10, dx 0.1 mm), "
                  "Nt uniform in [400, 1500], smooth heterogeneous c/rho/alpha (y=1.5), planar p tone burst "
                  "at 0.5-1.5 MHz, 1% random binary sensor recording p")


def smooth_field(rng, shape, corr_len: float, lo: float, hi: float) -> np.ndarray:
    """White noise filtered by a Gaussian of std corr_len points, min-max rescaled to [lo, hi] (float32)."""
    import scipy.fft as sfft

    w = rng.standard_normal(shape, dtype=np.float32)
    ks = [np.fft.fftfreq(n).astype(np.float32) for n in shape[:-1]] + [np.fft.rfftfreq(shape[-1]).astype(np.float32)]
    k2 = sum(k.reshape([-1 if a == i else 1 for a in range(len(shape))]) ** 2 for i, k in enumerate(ks))
    g = np.exp(np.float32(-2.0 * math.pi ** 2 * corr_len ** 2) * k2)
    f = sfft.irfftn(sfft.rfftn(w, workers=-1) * g, s=shape, workers=-1).astype(np.float32)
    fmin, fmax = float(f.min()), float(f.max())
    return (lo + (hi - lo) * (f - fmin) / (fmax - fmin)).astype(np.float32)


def c5_request(i: int, grids=C5_GRIDS, nt_range=C5_NT, Nt: int | None = None) -> dict:
    """Request i of the serving stream, drawn from default_rng(5000 + i).

    Draw order: grid, Nt, source frequency, c, rho, alpha_coeff, sensor points.
    Source: p on the whole y-z plane x = 11 (1-based), additive, unit-amplitude
    3-cycle tone burst.  Sensor: round(0.01 N) distinct points, as a binary mask.
    """
    rng = np.random.default_rng(5000 + i)
    shape = tuple(int(n) for n in grids[int(rng.integers(len(grids)))])
    nt = int(rng.integers(nt_range[0], nt_range[1] + 1))
    f0 = float(rng.uniform(0.5e6, 1.5e6))
    c = smooth_field(rng, shape, 4.0, 1450.0, 1650.0)
    rho = smooth_field(rng, shape, 4.0, 950.0, 1100.0)
    alpha = smooth_field(rng, shape, 4.0, 0.002, 1.0)
    N = int(np.prod(shape))
    pts = np.sort(rng.choice(N, int(round(0.01 * N)), replace=False))
    sensor = np.zeros(shape, dtype=np.float32)
    sensor[np.unravel_index(pts, shape, order="F")] = 1
    dx = 0.1e-3
    kgrid = fe.KWaveGrid(shape[0], dx, shape[1], dx, shape[2], dx)
    dt = 0.3 * dx / float(c.max())
    kgrid.set_time(nt if Nt is None else int(Nt), dt)
    p_mask = np.zeros(shape, dtype=np.float32)
    p_mask[10, :, :] = 1                                                   # x = 11 (1-based)
    sig = geo.tone_burst(1.0 / dt, f0, 3)
    medium = {"sound_speed": c, "density": rho, "alpha_coeff": alpha, "alpha_power": 1.5}
    return fe.save_to_disk_dict(kgrid, medium, {"p": sig, "p_mask": p_mask, "p_mode": "additive"},
                                {"mask": sensor}, pml_size=10)
