"""CPU pins for the oracle branches and front-end functions that decide the
benchmark inputs: Blackman windows against the textbook/library windows,
k-space smoothing through its eigenfunctions, tone-burst symmetry and
length, makeTime's known answer, the dirichlet source branches (the source
value is what the field holds after the step) and the full-grid `_all` /
`_final` outputs against brute force over the recorded series."""

import math

import numpy as np
import pytest
from scipy.signal import windows

import cases


# ---- getWin.m Blackman (getWin.m:207-252, 256-345) ---------------------------

@pytest.mark.parametrize("n", [7, 32, 33, 128])
def test_blackman_1d_is_the_classic_window(kw, n):
    g = kw.geometry
    w = g.get_win_blackman_1d(n, True)
    np.testing.assert_allclose(w, np.blackman(n), atol=1e-15, rtol=0)  # 0.42 - 0.5 cos + 0.08 cos
    assert abs(w[0]) < 1e-15 and abs(w[-1]) < 1e-15
    np.testing.assert_allclose(w, w[::-1], atol=1e-15)
    # periodic form (getWin's 'Symmetric', false): first n of the (n+1)-point window
    np.testing.assert_allclose(g.get_win_blackman_1d(n, False), windows.blackman(n, sym=False), atol=1e-15)


@pytest.mark.parametrize("n", [17, 33])
def test_blackman_rotation_window(kw, n):
    g = kw.geometry
    w = g.get_win_blackman_rot([n, n], [True, True])
    c = n // 2
    assert w[c, c] == pytest.approx(1.0, abs=1e-15)
    # on the axes the rotated window is the 1-D window; it depends on radius only
    np.testing.assert_allclose(w[c, :], g.get_win_blackman_1d(n, True), atol=1e-14)
    np.testing.assert_allclose(w[:, c], g.get_win_blackman_1d(n, True), atol=1e-14)
    np.testing.assert_allclose(w, w.T, atol=1e-15)
    assert abs(w[0, 0]) < 1e-15  # beyond the radius
    w3 = g.get_win_blackman_rot([n, n, n], [True] * 3)
    np.testing.assert_allclose(w3[c, c, :], g.get_win_blackman_1d(n, True), atol=1e-14)


# ---- smooth.m (smooth.m:108-135) ---------------------------------------------

@pytest.mark.parametrize("n,k", [(33, 5), (32, 5), (33, 0), (32, 11)])
def test_smooth_scales_each_frequency_by_the_window(kw, n, k):
    """A cosine at DFT bin k is an eigenfunction of the k-space smoothing:
    smooth(cos) = W(k) cos, W the window whose centre sits at k = 0."""
    g = kw.geometry
    x = np.cos(2 * math.pi * k * np.arange(n) / n)
    w = g.get_win_blackman_1d(n + (0 if n % 2 else 1), True)[:n]
    centre = n // 2
    np.testing.assert_allclose(g.smooth(x), w[centre + k] * x, atol=1e-12)


def test_smooth_preserves_constants_and_restores_max(kw):
    g = kw.geometry
    np.testing.assert_allclose(g.smooth(np.full((20, 17), 3.5)), 3.5, atol=1e-12)
    p0 = 5.0 * g.make_disc(40, 40, 0, 0, 6)
    sm = g.smooth(p0, restore_max=True)
    assert np.max(np.abs(sm)) == pytest.approx(5.0, rel=1e-12)
    assert sm.sum() == pytest.approx(np.sum(g.smooth(p0)) * 5.0 / np.max(np.abs(g.smooth(p0))), rel=1e-12)


# ---- toneBurst.m (toneBurst.m:100-178) ----------------------------------------

@pytest.mark.parametrize("fs,f0,cycles", [(50e6, 1e6, 3), (66.67e6, 0.5e6, 5), (40e6, 1.5e6, 4)])
def test_tone_burst_length_symmetry_and_frequency(kw, fs, f0, cycles):
    g = kw.geometry
    s = g.tone_burst(fs, f0, cycles)[0]
    # MATLAB colon 0:dt:T: floor(T/dt) + 1 samples
    assert s.size == math.floor(cycles / f0 * fs + 1e-10) + 1
    assert s[0] == 0.0 and s[-1] == pytest.approx(0.0, abs=1e-12)  # Tukey taper ends
    t = np.arange(s.size) / fs
    if abs((s.size - 1) / fs * f0 - cycles) < 1e-9:
        # an integer number of cycles exactly spans the burst: odd about its centre
        np.testing.assert_allclose(s, -s[::-1], atol=1e-12)
    # the envelope (analytic-signal magnitude) is a unit Gaussian spanning
    # 3 standard deviations to each end of the burst
    from scipy.signal import hilbert

    pad = 4 * s.size
    env = np.abs(hilbert(np.concatenate([np.zeros(pad), s, np.zeros(pad)])))[pad:pad + s.size]
    x = -3.0 + 6.0 * t / t[-1]
    mid = np.abs(x) < 2.0
    np.testing.assert_allclose(env[mid], np.exp(-x[mid] ** 2 / 2), atol=0.03)
    # the spectrum peaks at the carrier
    nfft = 1 << 16
    spec = np.abs(np.fft.rfft(s, nfft))
    f_peak = np.argmax(spec) * fs / nfft
    assert abs(f_peak - f0) < 0.03 * f0
    r = g.tone_burst(fs, f0, cycles, envelope="Rectangular")[0]
    np.testing.assert_allclose(r, np.sin(2 * math.pi * f0 * t), atol=1e-15)


def test_tone_burst_offset_and_length(kw):
    g = kw.geometry
    a = g.tone_burst(20e6, 1e6, 2)[0]
    b = g.tone_burst(20e6, 1e6, 2, signal_offset=7, signal_length=100)[0]
    assert b.size == 100 and np.all(b[:7] == 0) and np.all(b[7 + a.size:] == 0)
    np.testing.assert_array_equal(b[7:7 + a.size], a)


# ---- kWaveGrid.makeTime (kWaveGrid.m:612-671) ---------------------------------

def test_make_time_known_answer(kw):
    """The 2-D 128^2 configuration: dt = 0.3 * 1e-4 / 1500 = 2e-8 s, and
    t_end = the grid diagonal / c_min gives floor(t_end/dt) + 1 steps."""
    kg = kw.frontend.KWaveGrid(128, 1e-4, 128, 1e-4)
    dt, nt = kg.make_time(1500.0, 0.3)
    assert dt == pytest.approx(2e-8, rel=1e-15)
    t_end = math.hypot(0.0128, 0.0128) / 1500.0
    assert nt == 604 == math.floor(t_end / dt) + 1
    assert (nt - 1) * dt <= t_end < nt * dt
    kg3 = kw.frontend.KWaveGrid(64, 2e-4, 32, 1e-4, 16, 3e-4)
    c = np.full((64, 32, 16), 1500.0)
    c[3, 4, 5] = 1600.0
    dt3, nt3 = kg3.make_time(c, 0.2)
    assert dt3 == pytest.approx(0.2 * 1e-4 / 1600.0, rel=1e-15)       # smallest spacing, fastest speed
    t_end3 = math.sqrt(0.0128 ** 2 + 0.0032 ** 2 + 0.0048 ** 2) / 1500.0  # slowest speed
    assert nt3 == math.floor(t_end3 / dt3) + 1


# ---- dirichlet sources (kspaceFirstOrder3D.m:806-812, 930-937) ----------------

def _dirichlet_case(kw, kind, dim):
    fe, geo = kw.frontend, kw.geometry
    shape = (40, 32) if dim == 2 else (24, 20, 16)
    kg = fe.KWaveGrid(*[v for n in shape for v in (n, 1e-4)])
    kg.set_time(40, 0.3 * 1e-4 / 1500.0)
    mask = np.zeros(shape)
    mask[(12,) + tuple(slice(4, n - 4) for n in shape[1:])] = 1
    npts = int(mask.sum())
    burst = geo.tone_burst(1 / kg.dt, 2e6, 3)[0]
    sig = np.vstack([np.roll(burst, j) * (1 + 0.1 * j) for j in range(npts)])  # one series per point
    src = ({"p": sig, "p_mask": mask, "p_mode": "dirichlet"} if kind == "p"
           else {"ux": sig, "u_mask": mask, "u_mode": "dirichlet"})
    f = fe.save_to_disk_dict(kg, {"sound_speed": 1500.0, "density": 1000.0}, src, {"mask": mask}, pml_size=4)
    return f, sig


@pytest.mark.parametrize("dim", [2, 3])
def test_dirichlet_pressure_source_sets_the_pressure(kw, oracle, dim):
    """Lossless linear medium: the front end scales the signal by
    1/(dim c0^2) and the loop writes it into every density component after
    the mass update, so p = c0^2 sum(rho) equals the unscaled signal at the
    source points after every step while the signal lasts."""
    f, sig = _dirichlet_case(kw, "p", dim)
    stored = np.asarray(f["p_source_input"], dtype=np.float64).reshape(-1, sig.shape[0])  # (Nt_src, points)
    # the file holds sig / (dim c0^2), in float32
    np.testing.assert_allclose(stored * dim * 1500.0 ** 2, sig.T, rtol=2e-7, atol=1e-9)
    p = oracle.run(f, record=("p",))["p"]
    n = min(sig.shape[1], p.shape[0])
    np.testing.assert_allclose(p[:n], dim * 1500.0 ** 2 * stored[:n], rtol=1e-13, atol=1e-13 * np.max(np.abs(sig)))


@pytest.mark.parametrize("dim", [2, 3])
def test_dirichlet_velocity_source_sets_the_velocity(kw, oracle, dim):
    """The velocity signal (unscaled for dirichlet) replaces ux at the source
    points after the velocity update."""
    f, sig = _dirichlet_case(kw, "u", dim)
    stored = np.asarray(f["ux_source_input"], dtype=np.float64).reshape(-1, sig.shape[0])
    np.testing.assert_allclose(stored, sig.T, rtol=1e-7, atol=1e-12)  # float32 copy, unscaled
    ux = oracle.run(f, record=("u",))["ux"]
    n = min(sig.shape[1], ux.shape[0])
    np.testing.assert_array_equal(ux[:n], stored[:n])


# ---- full-grid outputs (extractSensorData.m:218-297, 3DC.m:589-610) -----------

@pytest.mark.parametrize("start", [1, 4])
@pytest.mark.parametrize("dim", [2, 3])
def test_all_and_final_outputs_equal_brute_force(kw, oracle, dim, start):
    if dim == 2:
        f = cases.case_2d_absorbing_psource(N=(44, 40), Nt=30, mode="additive")
    else:
        f = cases.case_3d(N=(16, 14, 12), Nt=25, pml=3, source="uxyz")
    f = dict(f)
    N = int(np.prod([int(np.ravel(f[k])[0]) for k in ("Nx", "Ny", "Nz")]))
    f["sensor_mask_index"] = np.arange(1, N + 1, dtype=np.uint64).reshape(-1, 1)  # every grid point
    rec = ("p", "p_max_all", "p_min_all", "p_final", "u", "u_max_all", "u_min_all", "u_final")
    out = oracle.run(f, record=rec, record_start_index=start)
    p = out["p"]
    assert p.shape[0] == int(np.ravel(f["Nt"])[0]) - start + 1
    # (Nz, Ny, Nx) C order flattens in the sensor (find) order
    np.testing.assert_array_equal(out["p_max_all"].reshape(-1), p.max(axis=0))
    np.testing.assert_array_equal(out["p_min_all"].reshape(-1), p.min(axis=0))
    np.testing.assert_array_equal(out["p_final"].reshape(-1), p[-1])
    for a in "xyz"[:dim]:
        u = out["u" + a]
        np.testing.assert_array_equal(out["u" + a + "_max_all"].reshape(-1), u.max(axis=0))
        np.testing.assert_array_equal(out["u" + a + "_min_all"].reshape(-1), u.min(axis=0))
        np.testing.assert_array_equal(out["u" + a + "_final"].reshape(-1), u[-1])
