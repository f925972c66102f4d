"""Host-side input generators used to build simulation inputs.

Restatements of the k-Wave helpers that decide the numbers written into an
input file (front-end side, not the time loop):

* makeDisc        k-Wave/makeDisc.m:57-92   (with private/makePixelMap.m)
* makeBall        k-Wave/makeBall.m
* makeCartCircle  k-Wave/makeCartCircle.m:54-97
* toneBurst       k-Wave/toneBurst.m:100-178 (Gaussian/Rectangular envelopes)
* getWin          k-Wave/getWin.m:196-345   (Blackman, rotation form)
* smooth          k-Wave/smooth.m:108-135

Arrays use MATLAB index order: a 2-D grid array has shape (Nx, Ny), a 3-D one
(Nx, Ny, Nz).
"""

from __future__ import annotations

import math

import numpy as np


def _pixel_dim(n: int) -> np.ndarray:
    """createPixelDim with origin 'single' and shift 0 (makePixelMap.m)."""
    if n % 2 == 0:
        return np.arange(-n // 2 + 1, n // 2 + 1, dtype=np.float64)
    return np.arange(-(n - 1) // 2, (n - 1) // 2 + 1, dtype=np.float64)


def make_disc(Nx: int, Ny: int, cx: int, cy: int, radius: float) -> np.ndarray:
    """Binary disc (1-based centre; 0 selects the grid centre).  makeDisc.m:57-92."""
    if cx == 0:
        cx = Nx // 2 + 1
    if cy == 0:
        cy = Ny // 2 + 1
    if not (1 <= cx <= Nx and 1 <= cy <= Ny):
        raise ValueError("Disc center must be within grid.")
    rx, ry = np.meshgrid(_pixel_dim(Nx), _pixel_dim(Ny), indexing="ij")
    disc = (np.sqrt(rx ** 2 + ry ** 2) <= radius).astype(np.float64)
    return np.roll(disc, (cx - math.ceil(Nx / 2), cy - math.ceil(Ny / 2)), axis=(0, 1))


def make_ball(Nx, Ny, Nz, cx, cy, cz, radius) -> np.ndarray:
    """Binary ball (1-based centre; 0 selects the grid centre).  makeBall.m."""
    cx = Nx // 2 + 1 if cx == 0 else cx
    cy = Ny // 2 + 1 if cy == 0 else cy
    cz = Nz // 2 + 1 if cz == 0 else cz
    rx, ry, rz = np.meshgrid(_pixel_dim(Nx), _pixel_dim(Ny), _pixel_dim(Nz), indexing="ij")
    ball = (np.sqrt(rx ** 2 + ry ** 2 + rz ** 2) <= radius).astype(np.float64)
    return np.roll(ball, (cx - math.ceil(Nx / 2), cy - math.ceil(Ny / 2), cz - math.ceil(Nz / 2)),
                   axis=(0, 1, 2))


def _mround(x):
    """MATLAB round (halves away from zero)."""
    x = np.asarray(x, dtype=np.float64)
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def make_bowl(grid_size, bowl_pos, radius, diameter, focus_pos) -> np.ndarray:
    """Binary bowl (spherical cap) in a 3-D grid.  makeBowl.m:80-710 (default options).

    grid_size (Nx, Ny, Nz); bowl_pos (centre of the rear surface) and
    focus_pos (any point on the axis) are 1-based grid points; radius of
    curvature and aperture diameter (odd) in grid points.  The surface is
    found in a bounding box of side round(sqrt(2) d) + 2 by a forward and a
    backward line search along x, then y, then z for the point whose distance
    from the sphere centre is closest to the radius (kept if within 0.5),
    then trimmed to the cap's half-angle asin(d / 2R) about the axis.
    """
    grid = [int(_mround(v)) for v in grid_size]
    bpos = [int(_mround(v)) for v in bowl_pos]
    fpos = [int(_mround(v)) for v in focus_pos]
    diameter = int(_mround(diameter))
    radius = float(_mround(radius))
    if diameter > 2 * radius or diameter % 2 == 0 or bpos == fpos:
        raise ValueError("invalid bowl: diameter must be odd and <= 2 radius, focus_pos != bowl_pos")
    n = int(_mround(math.sqrt(2) * diameter)) + 2
    b = [math.ceil(n / 2)] * 3                                    # bowl_pos in the box (1-based)
    f = [b[i] + (fpos[i] - bpos[i]) for i in range(3)]
    half_arc = math.asin(diameter / (2 * radius))
    dist = math.sqrt(sum((b[i] - f[i]) ** 2 for i in range(3)))
    c = [float(_mround(radius / dist * (f[i] - b[i]) + b[i])) for i in range(3)]
    ax = [np.arange(1, n + 1, dtype=np.float64) - c[i] for i in range(3)]
    pm = np.abs(np.sqrt(ax[0][:, None, None] ** 2 + ax[1][None, :, None] ** 2 + ax[2][None, None, :] ** 2)
                - radius)
    bowl = np.zeros((n, n, n), dtype=bool)
    for axis in range(3):
        if axis > 0:
            pm[bowl] = 0.0
        m = np.moveaxis(pm, axis, 0)
        for backward in (False, True):
            mm = m[::-1] if backward else m
            idx = np.argmin(mm, axis=0)                            # first minimum, as min(..., [], 1)
            val = np.take_along_axis(mm, idx[None], axis=0)[0]
            sel = np.nonzero(val < 0.5)
            pos = idx[sel]
            if backward:
                pos = n - 1 - pos
            coord = [None, None, None]
            others = [a for a in range(3) if a != axis]
            coord[axis] = pos
            coord[others[0]], coord[others[1]] = sel
            bowl[tuple(coord)] = True
    # trim to the cap
    pts = np.argwhere(bowl).astype(np.float64) + 1.0
    v1 = np.array(b, dtype=np.float64) - np.array(c)
    v2 = pts - np.array(c)
    cosang = (v2 @ v1) / (np.linalg.norm(v1) * np.linalg.norm(v2, axis=1))
    theta = np.arccos(np.clip(cosang, -1.0, 1.0))
    drop = pts[theta > half_arc].astype(np.int64) - 1
    bowl[tuple(drop.T)] = False
    # place the box in the grid (cropping what falls outside)
    out = np.zeros(grid, dtype=np.float64)
    lo = [bpos[i] - b[i] for i in range(3)]                       # 0-based grid index of box index 0
    src, dst = [], []
    for i in range(3):
        a0, a1 = max(0, -lo[i]), min(n, grid[i] - lo[i])
        src.append(slice(a0, a1))
        dst.append(slice(lo[i] + a0, lo[i] + a1))
    out[tuple(dst)] = bowl[tuple(src)]
    return out


def make_cart_circle(radius: float, num_points: int, center=(0.0, 0.0),
                     arc_angle: float = 2 * math.pi) -> np.ndarray:
    """Cartesian points on a circle, shape (2, num_points).  makeCartCircle.m:54-97."""
    full = arc_angle == 2 * math.pi
    n = num_points if full else num_points - 1
    angles = np.arange(0, n + 1) * arc_angle / n + math.pi / 2
    if full:
        angles = angles[:-1]
    circ = np.vstack([radius * np.cos(angles), radius * np.sin(-angles)])
    circ[0] += center[0]
    circ[1] += center[1]
    return circ


def _gaussian(x, magnitude=1.0, mean=0.0, variance=1.0):
    return magnitude * np.exp(-(x - mean) ** 2 / (2 * variance))


def _tukey(n: int, param: float) -> np.ndarray:
    win = np.ones(n)
    index = np.arange(0, (n - 1) * param / 2 + 1e-12)
    index = index[index <= (n - 1) * param / 2]
    p = param * n
    win[: index.size] = 0.5 * (1 + np.cos(2 * math.pi / p * (index - p / 2)))
    win[n - index.size:] = win[: index.size][::-1]
    return win


def tone_burst(sample_freq: float, signal_freq: float, num_cycles: float,
               envelope: str = "Gaussian", signal_offset: int = 0,
               signal_length: int | None = None) -> np.ndarray:
    """Windowed tone burst, shape (1, L).  toneBurst.m:100-178."""
    dt = 1.0 / sample_freq
    tone_length = num_cycles / signal_freq
    # MATLAB colon 0:dt:tone_length
    n = int(math.floor(tone_length / dt + 1e-10)) + 1
    tone_t = np.arange(n) * dt
    burst = np.sin(2 * math.pi * signal_freq * tone_t)
    if envelope == "Gaussian":
        x_lim = 3.0
        wx = np.linspace(-x_lim, x_lim, burst.size) if burst.size > 1 else np.zeros(1)
        burst = burst * _gaussian(wx, 1, 0, 1)
        burst = burst * _tukey(burst.size, 0.05)
    elif envelope == "Rectangular":
        pass
    else:
        raise ValueError(f"Unknown envelope {envelope}")
    length = signal_offset + burst.size if signal_length is None else signal_length
    sig = np.zeros((1, length))
    sig[0, signal_offset: signal_offset + burst.size] = burst
    return sig


def _cosine_series(n, N, coeffs):
    s = coeffs[0] * np.ones_like(n, dtype=np.float64)
    for i in range(1, len(coeffs)):
        s = s + (-1) ** i * coeffs[i] * np.cos(i * 2 * math.pi * n / (N - 1))
    return s


def get_win_blackman_1d(N: int, symmetric: bool = True, param: float = 0.16) -> np.ndarray:
    """1-D Blackman window (getWin.m:207-252)."""
    M = N + (0 if symmetric else 1)
    n = np.arange(M, dtype=np.float64)
    win = _cosine_series(n, M, [(1 - param) / 2, 0.5, param / 2])
    return win[:N]


def get_win_blackman_rot(sizes, symmetric) -> np.ndarray:
    """2-D/3-D Blackman window built by rotation (getWin.m:256-345)."""
    sizes = [int(s) for s in sizes]
    sym = list(symmetric)
    N = [s + (0 if y else 1) for s, y in zip(sizes, sym)]
    L = max(N)
    win_lin = get_win_blackman_1d(L, True)
    radius = (L - 1) / 2
    ll = np.linspace(-radius, radius, L)
    axes = [np.linspace(-radius, radius, n) for n in N]
    grids = np.meshgrid(*axes, indexing="ij")
    r = np.sqrt(sum(g ** 2 for g in grids))
    r = np.minimum(r, radius)
    win = np.interp(r, ll, win_lin)
    sl = tuple(slice(0, s) for s in sizes)
    return win[sl]


def smooth(mat: np.ndarray, restore_max: bool = False) -> np.ndarray:
    """Blackman k-space smoothing of a grid array.  smooth.m:108-135."""
    mat = np.asarray(mat, dtype=np.float64)
    sizes = [s for s in mat.shape if s != 1] if mat.ndim > 1 else list(mat.shape)
    sym = [bool(s % 2) for s in sizes]
    if len(sizes) == 1:
        win = get_win_blackman_1d(sizes[0], sym[0]).reshape(mat.shape)
    else:
        win = get_win_blackman_rot(sizes, sym).reshape(mat.shape)
    win = np.abs(win)
    sm = np.real(np.fft.ifftn(np.fft.fftn(mat) * np.fft.ifftshift(win)))
    if restore_max:
        sm = (np.max(np.abs(mat)) / np.max(np.abs(sm))) * sm
    return sm
