"""Pins of the FP64 oracle (CPU only).

The reference ships no golden vectors for this path (its regression data is
generated at CI time and *.mat is git-ignored), so the oracle is pinned by
closed forms, brute force on tiny inputs and the invariants the reference's
own unit tests assert (k-Wave/testing/unit/...).  Each check is chosen so a
plausible slip (dropped term, wrong sign, wrong axis, wrong factor) fails it.
"""

import math

import numpy as np
import pytest

import cases


def fe_of(kw):
    return kw.frontend


# ---------------------------------------------------------------------------
# closed forms of the operator precompute
# ---------------------------------------------------------------------------

def test_make_dim_even_odd(oracle):
    """kWaveGrid.makeDim: DC exactly zero, Nyquist negative for even N."""
    k = oracle.make_dim(8, 1.0)
    np.testing.assert_allclose(k, 2 * math.pi * np.arange(-4, 4) / 8, rtol=0, atol=1e-15)
    assert k[4] == 0.0
    k = oracle.make_dim(7, 0.5)
    np.testing.assert_allclose(k, 2 * math.pi / 0.5 * np.arange(-3, 4) / 7, atol=1e-13)
    ku = oracle.k_vec_unshifted(8, 1.0)
    assert ku[0] == 0 and ku[4] == -math.pi


def test_db2neper_known_value(oracle):
    """1 dB/(MHz cm) = ln(10)/20 Np per 2*pi*1e6 rad/s per 0.01 m."""
    expect = (math.log(10) / 20) / 0.01 / (2 * math.pi * 1e6)
    assert abs(oracle.db2neper(1.0, 1.0) - expect) < 1e-15 * expect
    expect2 = (math.log(10) / 20) / 0.01 / (2 * math.pi * 1e6) ** 1.5
    assert abs(oracle.db2neper(1.0, 1.5) - expect2) < 1e-14 * expect2


def test_pml_closed_form(oracle):
    """getPML.m: exp(-alpha (c/dx) r^4 dt/2), r = 1 at the outer point, 0.5/size inside."""
    n, d, dt, c, size, a = 40, 1e-4, 2e-8, 1500.0, 10, 2.0
    p = oracle.get_pml(n, d, dt, c, size, a, False)
    assert abs(p[0] - math.exp(-a * c / d * dt / 2)) < 1e-15
    assert abs(p[-1] - math.exp(-a * c / d * dt / 2)) < 1e-15
    assert np.all(p[size:n - size] == 1.0)
    assert abs(p[n - size] - math.exp(-a * c / d * (1 / size) ** 4 * dt / 2)) < 1e-15
    ps = oracle.get_pml(n, d, dt, c, size, a, True)
    assert abs(ps[-1] - math.exp(-a * c / d * ((size + 0.5) / size) ** 4 * dt / 2)) < 1e-15
    assert abs(ps[0] - math.exp(-a * c / d * ((size - 0.5) / size) ** 4 * dt / 2)) < 1e-15
    assert np.all(np.diff(p[:size]) > 0) and np.all(np.diff(p[n - size:]) < 0)


def test_kappa_and_absorption_special_cases(kw, oracle):
    f = cases.case_2d_absorbing_psource(Nt=3)
    f["alpha_power"] = np.array([[[2.0]]], dtype=np.float32)
    ops = oracle.Operators(oracle.decode(f))
    # sinc.m evaluates sin(0 + pi)/(0 + 1) + 1 at the origin: 1 + sin(pi) in double
    assert ops.kappa[0, 0, 0] == 1.0 + math.sin(math.pi)
    assert np.all(ops.kappa[ops.k > 0] < 1.0) and np.all(ops.kappa > 0)
    # y = 2: nabla1 = k^0 == 1 everywhere (MATLAB 0^0 = 1), nabla2(0) = 0
    assert np.all(ops.absorb_nabla1 == 1.0)
    assert ops.absorb_nabla2[0, 0, 0] == 0.0
    # eta = 2 alpha c^y tan(pi y / 2) vanishes to rounding at y = 2
    assert np.max(np.abs(ops.absorb_eta)) < 1e-12 * np.max(np.abs(ops.absorb_tau)) * 1500
    f["alpha_power"] = np.array([[[1.5]]], dtype=np.float32)
    ops = oracle.Operators(oracle.decode(f))
    assert ops.absorb_nabla1[0, 0, 0] == 0.0            # k^-0.5 = Inf at DC -> 0
    k = ops.k[0, 3, 5]
    assert abs(ops.absorb_nabla1[0, 3, 5] - k ** -0.5) < 1e-12 * k ** -0.5
    pb = oracle.decode(f)
    a = oracle.db2neper(pb.alpha_coeff, 1.5)
    np.testing.assert_allclose(ops.absorb_tau, -2 * a * pb.c0 ** 0.5, rtol=1e-14)
    np.testing.assert_allclose(ops.absorb_eta, 2 * a * pb.c0 ** 1.5 * math.tan(0.75 * math.pi), rtol=1e-14)


# ---------------------------------------------------------------------------
# spectral derivative operators against a brute-force DFT
# ---------------------------------------------------------------------------

def dft_matrix(n):
    j = np.arange(n)
    return np.exp(-2j * math.pi * np.outer(j, j) / n)


@pytest.mark.parametrize("shape", [(1, 6, 8), (1, 5, 7), (4, 6, 5)])
def test_derivative_operators_match_brute_force_dft(oracle, shape):
    """real(ifftn(D kappa fftn(f))) through the R2C path == explicit DFT matrices."""
    Nz, Ny, Nx = shape
    rng = np.random.default_rng(0)
    f = rng.standard_normal(shape)
    d = (1e-3, 1.3e-3, 0.7e-3)
    inp = {"Nx": Nx, "Ny": Ny, "Nz": Nz, "Nt": 1, "dx": d[0], "dy": d[1], "dz": d[2], "dt": 1e-7,
           "c_ref": 1500.0, "c0": 1500.0, "rho0": 1000.0, "rho0_sgx": 1000.0, "rho0_sgy": 1000.0,
           "rho0_sgz": 1000.0, "pml_x_size": 0, "pml_y_size": 0, "pml_z_size": 0, "pml_x_alpha": 0,
           "pml_y_alpha": 0, "pml_z_alpha": 0, "sensor_mask_index": np.array([1])}
    pb = oracle.decode(inp)
    ops = oracle.Operators(pb)
    sp = oracle.Spectral(pb, False)
    mats = [dft_matrix(Nx), dft_matrix(Ny), dft_matrix(Nz)]
    # full DFT over the active axes
    F = np.einsum("ax,zyx->zya", mats[0], f.astype(complex))
    F = np.einsum("by,zya->zba", mats[1], F)
    if pb.dim == 3:
        F = np.einsum("cz,zba->cba", mats[2], F)
    kv = [oracle.k_vec_unshifted(n, dd) for n, dd in zip((Nx, Ny, Nz), d)]
    K = np.sqrt(kv[0][None, None, :] ** 2 + kv[1][None, :, None] ** 2
                + (kv[2][:, None, None] ** 2 if pb.dim == 3 else 0))
    kappa = np.where(K == 0, 1.0, np.sin(1500 * K * 1e-7 / 2) / np.where(K == 0, 1, 1500 * K * 1e-7 / 2))
    for axis in range(pb.dim):
        kk = [kv[0][None, None, :], kv[1][None, :, None], kv[2][:, None, None]][axis]
        for sign, dop in ((1, ops.d_pos[axis]), (-1, ops.d_neg[axis])):
            D = 1j * kk * np.exp(sign * 1j * kk * d[axis] / 2)
            G = F * D * kappa
            g = np.einsum("xa,zya->zyx", mats[0].conj(), G) / Nx
            g = np.einsum("yb,zbx->zyx", mats[1].conj(), g) / Ny
            if pb.dim == 3:
                g = np.einsum("zc,cyx->zyx", mats[2].conj(), g) / Nz
            got = sp.inv_real(dop * ops.kappa * sp.fwd(f))
            np.testing.assert_allclose(got, g.real, atol=1e-9 * np.abs(g).max(), rtol=0)


def test_half_spectrum_equals_full_complex(oracle):
    """The R2C/C2R evaluation equals MATLAB's full-spectrum real(ifftn(...))."""
    f = cases.case_3d(N=(16, 14, 12), Nt=25, nonlinear=True, source="p", many=True, pml=3)
    a = oracle.run(f, record=("p", "u"))
    b = oracle.run(f, record=("p", "u"), full_complex=True)
    for k in a:
        assert cases.rel_l2(a[k], b[k]) < 1e-12, k


# ---------------------------------------------------------------------------
# exact time-loop solution
# ---------------------------------------------------------------------------

def _homogeneous_p0(kw, shape, Nt, het_p0_seed=0):
    fe = kw.frontend
    rng = np.random.default_rng(het_p0_seed)
    d = 1e-4
    kgrid = fe.KWaveGrid(shape[0], d, shape[1], d, *((shape[2], d) if len(shape) == 3 else ()))
    kgrid.set_time(Nt, 0.3 * d / 1500.0)
    p0 = cases.smooth_field(rng, shape, 2.0)
    mask = np.ones(shape)
    return fe.save_to_disk_dict(kgrid, {"sound_speed": 1500.0, "density": 1000.0}, {"p0": p0}, {"mask": mask},
                                pml_size=2, pml_alpha=0.0, smooth_p0=False), kgrid


@pytest.mark.parametrize("shape", [(32, 24), (16, 12, 10)])
def test_lossless_homogeneous_matches_closed_form(kw, oracle, shape):
    """p_n = IFFT(FFT(p0) cos(c |k| (n-1) dt)) (kappa = sinc(c k dt/2) makes the
    leapfrog exact; kspaceFirstOrder3D.m:1029-1033 gives the symmetric start)."""
    f, kgrid = _homogeneous_p0(kw, shape, 40)
    out = oracle.run(f, record=("p",))
    p0 = np.asarray(f["p0_source_input"], dtype=np.float64).reshape(
        (kgrid.Nz if kgrid.dim == 3 else 1, kgrid.Ny, kgrid.Nx))
    val = lambda k: float(np.asarray(f[k]).reshape(-1)[0]) if k in f else 1.0  # noqa: E731  (file values)
    ex = oracle.exact_lossless_homogeneous(p0, val("c0"), val("dt"), (val("dx"), val("dy"), val("dz")), 40)
    got = out["p"]
    assert np.max(np.abs(got - ex.reshape(40, -1))) < 1e-11 * np.max(np.abs(ex))


# ---------------------------------------------------------------------------
# invariants of the reference's unit tests
# ---------------------------------------------------------------------------

def test_p0_equals_two_clicks(kw, oracle):
    """kspaceFirstOrder2D_p0_vs_two_clicks.m:36-99: a p0 source equals two
    additive-no-correction clicks of 0.5/(2 c dt/dx), PMLAlpha 0, to 1e-14."""
    fe, geo = kw.frontend, kw.geometry
    Nx = Ny = 64
    dx = 0.1e-3
    kgrid = fe.KWaveGrid(Nx, dx, Ny, dx)
    kgrid.make_time(1500.0, 0.234)
    p0 = geo.make_disc(Nx, Ny, 25, 25, 4) + geo.make_disc(Nx, Ny, 40, 30, 2)
    rr = np.hypot(*np.meshgrid(np.arange(Nx) - 32, np.arange(Ny) - 32, indexing="ij"))
    mask = (np.abs(rr - 16) < 0.5).astype(float)
    medium = {"sound_speed": 1500.0, "density": 1000.0}
    f1 = fe.save_to_disk_dict(kgrid, medium, {"p0": p0}, {"mask": mask}, pml_alpha=0.0, smooth_p0=False)
    src = {"p_mask": (p0 != 0).astype(float), "p_mode": "additive-no-correction",
           "p": np.array([[0.5, 0.5]]) / (2 * 1500.0 * kgrid.dt / kgrid.dx)}
    f2 = fe.save_to_disk_dict(kgrid, medium, src, {"mask": mask}, pml_alpha=0.0)
    # the reference runs this test in double: keep the scaled click, 0.5/(2 c dt/dx) * 2 dt/(2 c dx)
    # = 1/(4 c^2), in double rather than the file's float32
    f2["p_source_input"] = np.full((1, 2, 1), 1.0 / (4 * 1500.0 ** 2))
    a = oracle.run(f1, record=("p",))["p"]
    b = oracle.run(f2, record=("p",))["p"]
    assert np.max(np.abs(a - b)) < 1e-14 * max(1.0, np.abs(a).max()) * 10
    assert np.abs(a).max() > 0.05


def test_plane_wave_2d_equals_3d(kw, oracle):
    """kspaceFirstOrder_compare_1D_2D_3D_plane_waves.m: an x plane wave is the
    same in 2-D and 3-D (tolerance 1e-14 there)."""
    fe = kw.frontend
    Nx, Ny, Nz, dx = 64, 8, 6, 1e-4
    x = np.arange(Nx)
    prof = np.exp(-((x - 30) / 3.0) ** 2)
    medium = {"sound_speed": 1500.0, "density": 1000.0, "alpha_coeff": 0.5, "alpha_power": 1.5}
    k2 = fe.KWaveGrid(Nx, dx, Ny, dx)
    k2.set_time(60, 2e-8)
    k3 = fe.KWaveGrid(Nx, dx, Ny, dx, Nz, dx)
    k3.set_time(60, 2e-8)
    m2 = np.zeros((Nx, Ny))
    m2[:, 3] = 1
    m3 = np.zeros((Nx, Ny, Nz))
    m3[:, 3, 2] = 1
    f2 = fe.save_to_disk_dict(k2, medium, {"p0": np.tile(prof[:, None], (1, Ny))}, {"mask": m2},
                              pml_size=(8, 0), smooth_p0=False)
    f3 = fe.save_to_disk_dict(k3, medium, {"p0": np.tile(prof[:, None, None], (1, Ny, Nz))}, {"mask": m3},
                              pml_size=(8, 0, 0), smooth_p0=False)
    a = oracle.run(f2, record=("p",))["p"]
    b = oracle.run(f3, record=("p",))["p"]
    assert np.max(np.abs(a - b)) < 1e-13 * np.abs(a).max()


def test_axis_rotation_symmetry(kw, oracle):
    """The same heterogeneous absorbing problem with x and z exchanged gives
    the exchanged result (x/y/z rotation tests of the reference, 1e-14 there)."""
    fe = kw.frontend
    rng = np.random.default_rng(7)
    n, d = 12, 1e-4
    shape = (n, n, n)
    c = 1500 * (1 + 0.03 * cases.smooth_field(rng, shape))
    rho = 1000 * (1 + 0.03 * cases.smooth_field(rng, shape))
    p0 = cases.smooth_field(rng, shape)
    mask = (rng.random(shape) < 0.1).astype(float)
    kg = fe.KWaveGrid(n, d, n, d, n, d)
    kg.set_time(20, 0.25 * d / c.max())
    med = lambda T: {"sound_speed": T(c), "density": T(rho), "alpha_coeff": 0.4, "alpha_power": 1.3}  # noqa: E731
    swap = lambda a: np.ascontiguousarray(a.transpose(2, 1, 0))  # noqa: E731
    ident = lambda a: a  # noqa: E731
    fa = fe.save_to_disk_dict(kg, med(ident), {"p0": p0}, {"mask": np.ones(shape)}, pml_size=2, smooth_p0=False)
    fb = fe.save_to_disk_dict(kg, med(swap), {"p0": swap(p0)}, {"mask": np.ones(shape)}, pml_size=2,
                              smooth_p0=False)
    a = oracle.run(fa, record=("p_final",))["p_final"]
    b = oracle.run(fb, record=("p_final",))["p_final"]
    np.testing.assert_allclose(swap(b), a, atol=1e-12 * np.abs(a).max())


def test_sensor_statistics_match_brute_force(oracle):
    """p_max / p_min / p_rms recurrences (extractSensorData.m:70-91) == direct
    statistics of the recorded p time series; same for u."""
    f = cases.case_2d_absorbing_psource(Nt=50)
    o = oracle.run(f, record=("p", "p_max", "p_min", "p_rms", "u", "u_max", "u_rms"), record_start_index=7)
    p = o["p"]
    np.testing.assert_allclose(o["p_max"], p.max(0), rtol=0, atol=0)
    np.testing.assert_allclose(o["p_min"], p.min(0), rtol=0, atol=0)
    np.testing.assert_allclose(o["p_rms"], np.sqrt((p ** 2).mean(0)), rtol=1e-12)
    np.testing.assert_allclose(o["uy_max"], o["uy"].max(0), rtol=0, atol=0)
    np.testing.assert_allclose(o["ux_rms"], np.sqrt((o["ux"] ** 2).mean(0)), rtol=1e-12)
    assert p.shape[0] == 50 - 7 + 1


def test_additive_source_scaling_against_finer_dt(kw, oracle):
    """kspaceFirstOrder2D_check_source_scaling_p.m: an additive plane-wave p
    source gives the same field when dt is refined (the 2 dt/(N c dx) scaling);
    a factor error in the scaling fails by tens of percent."""
    fe, geo = kw.frontend, kw.geometry
    Nx, Ny, dx, c = 128, 4, 1e-4, 1500.0
    mask = np.zeros((Nx, Ny))
    mask[30, :] = 1
    sens = np.zeros((Nx, Ny))
    sens[80, 1] = 1
    out = []
    for refine in (1, 4):
        dt = 0.2 * dx / c / refine
        kg = fe.KWaveGrid(Nx, dx, Ny, dx)
        nt = 400 * refine
        kg.set_time(nt, dt)
        t = np.arange(nt) * dt
        sig = np.exp(-((t - 1.2e-6) / 0.3e-6) ** 2)[None]
        f = fe.save_to_disk_dict(kg, {"sound_speed": c, "density": 1000.0},
                                 {"p": sig, "p_mask": mask, "p_mode": "additive"}, {"mask": sens},
                                 pml_size=(16, 0))
        out.append(oracle.run(f, record=("p",))["p"][:, 0][::refine])
    a, b = out
    # the scaling makes a plane-wave mass source radiate its own amplitude in each
    # direction (a factor 2 or an N slip fails this), and the field is dt-independent
    assert abs(np.max(a) - 1.0) < 1e-2
    assert cases.rel_l2(a, b) < 2e-2
    del geo


def test_power_law_attenuation(kw, oracle):
    """example_na_modelling_absorption.m: the plane-wave attenuation measured
    between two sensors follows alpha0 f^y (Np/m from db2neper)."""
    fe = kw.frontend
    Nx, Ny, dx, c = 512, 4, 5e-5, 1500.0
    a0, y = 0.5, 1.5
    dt = 0.3 * dx / c
    kg = fe.KWaveGrid(Nx, dx, Ny, dx)
    nt = 1200
    kg.set_time(nt, dt)
    t = np.arange(nt) * dt
    sig = np.exp(-((t - 0.4e-6) / 0.08e-6) ** 2)[None]
    mask = np.zeros((Nx, Ny))
    mask[40, :] = 1
    sens = np.zeros((Nx, Ny))
    x1, x2 = 100, 300
    sens[x1, 1] = 1
    sens[x2, 1] = 1
    f = fe.save_to_disk_dict(kg, {"sound_speed": c, "density": 1000.0, "alpha_coeff": a0, "alpha_power": y},
                             {"p": sig, "p_mask": mask, "p_mode": "additive"}, {"mask": sens}, pml_size=(20, 0))
    p = oracle.run(f, record=("p",))["p"]
    P = np.abs(np.fft.rfft(p, n=8192, axis=0))
    freq = np.fft.rfftfreq(8192, dt)
    band = (freq > 1e6) & (freq < 5e6)
    measured = np.log(P[band, 0] / P[band, 1]) / ((x2 - x1) * dx)
    expected = oracle.db2neper(a0, y) * (2 * np.pi * freq[band]) ** y
    np.testing.assert_allclose(measured, expected, rtol=0.03)


def _plane_wave_harmonics(kw, oracle, BonA):
    fe = kw.frontend
    c, rho0, f0, P = 1500.0, 1000.0, 1e6, 5e6
    dx = c / f0 / 24
    Nx, Ny = 256, 4
    dt = 0.15 * dx / c
    kg = fe.KWaveGrid(Nx, dx, Ny, dx)
    nt = 1800
    kg.set_time(nt, dt)
    t = np.arange(nt) * dt
    r = np.clip(t / (3 / f0), 0, 1)
    sig = (P * np.sin(2 * np.pi * f0 * t) * r * r * (3 - 2 * r))[None]
    x0, xs = 40, (60, 80, 100)
    mask = np.zeros((Nx, Ny))
    mask[x0, :] = 1
    sens = np.zeros((Nx, Ny))
    for x in xs:
        sens[x, 1] = 1
    f = fe.save_to_disk_dict(kg, {"sound_speed": c, "density": rho0, "BonA": BonA},
                             {"p": sig, "p_mask": mask, "p_mode": "additive"}, {"mask": sens}, pml_size=(30, 0))
    p = oracle.run(f, record=("p",))["p"]
    n_per = int(round(1 / (f0 * dt)))
    out = []
    for j, x in enumerate(xs):
        seg = p[-4 * n_per:, j]
        S = np.abs(np.fft.rfft(seg * np.hanning(seg.size), n=1 << 16))
        fr = np.fft.rfftfreq(1 << 16, dt)
        ratio = S[np.argmin(abs(fr - 2 * f0))] / S[np.argmin(abs(fr - f0))]
        sigma = (1 + BonA / 2) * 2 * np.pi * f0 * (x - x0) * dx * P / (rho0 * c ** 3)
        out.append((ratio, sigma / 2))
    return out


@pytest.mark.parametrize("BonA,tol", [(5.0, 0.02), (0.0, 0.05)])
def test_second_harmonic_follows_fubini(kw, oracle, BonA, tol):
    """example_na_modelling_nonlinearity.m: a plane wave's second harmonic grows
    as p2/p1 = sigma/2, sigma = beta omega x P/(rho0 c^3), beta = 1 + B/2A, for
    x well inside the shock distance.  BonA = 0 isolates the convective term
    (the 2 sum(rho) of kspaceFirstOrder3D.m:920), BonA = 5 adds the B/A term."""
    for ratio, theory in _plane_wave_harmonics(kw, oracle, BonA):
        assert abs(ratio / theory - 1) < tol, (ratio, theory)


def test_stokes_attenuation(kw, oracle):
    """The 'stokes' equation of state (kspaceFirstOrder3D.m:977-983) attenuates
    as alpha0 omega^2 (alpha in Np/((rad/s)^2 m)) at low frequency."""
    fe = kw.frontend
    Nx, Ny, dx, c = 512, 4, 5e-5, 1500.0
    a0 = 0.5
    dt = 0.3 * dx / c
    kg = fe.KWaveGrid(Nx, dx, Ny, dx)
    nt = 1200
    kg.set_time(nt, dt)
    t = np.arange(nt) * dt
    sig = np.exp(-((t - 0.4e-6) / 0.08e-6) ** 2)[None]
    mask = np.zeros((Nx, Ny))
    mask[40, :] = 1
    sens = np.zeros((Nx, Ny))
    sens[100, 1] = sens[300, 1] = 1
    medium = {"sound_speed": c, "density": 1000.0, "alpha_coeff": a0, "alpha_power": 2.0, "alpha_mode": "stokes"}
    f = fe.save_to_disk_dict(kg, medium, {"p": sig, "p_mask": mask, "p_mode": "additive"}, {"mask": sens},
                             pml_size=(20, 0))
    assert int(np.asarray(f["absorbing_flag"]).reshape(-1)[0]) == 2
    p = oracle.run(f, record=("p",))["p"]
    P = np.abs(np.fft.rfft(p, n=8192, axis=0))
    freq = np.fft.rfftfreq(8192, dt)
    band = (freq > 1e6) & (freq < 3e6)
    measured = np.log(P[band, 0] / P[band, 1]) / (200 * dx)
    expected = oracle.db2neper(a0, 2) * (2 * np.pi * freq[band]) ** 2
    np.testing.assert_allclose(measured, expected, rtol=0.05)


def test_u_non_staggered_is_half_cell_shift(oracle):
    """u_non_staggered = real(ifft(exp(-i k d/2) fft(u))) moves the staggered velocity
    (at x + dx/2) back to the grid points: on a smooth field it matches the two-point
    average (u(x - dx) + u(x)) / 2 to second order, and not the opposite-side average
    (a sign mistake in the shift).  Also exact on the Nyquist-free closed form."""
    import cases
    f = cases.case_2d_p0_lossless(N=64, Nt=12, het=True)
    n = 64
    f["sensor_mask_index"] = np.arange(1, n * n + 1, dtype=np.uint64).reshape(1, 1, -1)
    out = oracle.run(f, record=("u", "u_non_staggered"))
    for a, nm in enumerate("xy"):
        u = out["u" + nm][-1].reshape(n, n)            # (y, x) C order
        ush = out["u" + nm + "_non_staggered"][-1].reshape(n, n)
        ax = 1 - a
        back = 0.5 * (np.roll(u, 1, axis=ax) + u)
        fwd = 0.5 * (u + np.roll(u, -1, axis=ax))
        e_back = np.linalg.norm(ush - back) / np.linalg.norm(ush)
        e_fwd = np.linalg.norm(ush - fwd) / np.linalg.norm(ush)
        assert e_back < 0.05 and e_fwd > 4 * e_back, (nm, e_back, e_fwd)
    # closed form: a band-limited sinusoid shifted by -dx/2 exactly
    x = np.arange(n)
    u = np.cos(2 * np.pi * 5 * (x + 0.5) / n + 0.3)    # samples of cos at the staggered points
    k = oracle.k_vec_unshifted(n, 1.0)
    ush = np.real(np.fft.ifft(np.exp(-1j * k / 2) * np.fft.fft(u)))
    assert np.allclose(ush, np.cos(2 * np.pi * 5 * x / n + 0.3), atol=1e-12)


def test_transducer_equals_delayed_many_source(oracle):
    """3D.m:889-902: point j adds signal(delay_mask(j) + t - 1) to ux.  The same
    field follows from a `many` additive-no-correction ux source whose series j is
    the signal advanced by delay_j (a one-sample or sign mistake in the delay
    indexing breaks the equality)."""
    import cases
    f = cases.case_transducer(N=(32, 28, 24), Nt=30)
    a = oracle.run(f, record=("p", "u_final"))
    flag = int(np.ravel(f["transducer_source_flag"])[0])
    sig = np.asarray(f["transducer_source_input"], dtype=np.float64).reshape(-1)
    d0 = np.asarray(f["delay_mask"], dtype=np.int64).reshape(-1) - 1
    series = np.stack([sig[d + np.arange(flag)] for d in d0])          # (Npts, flag)
    g = {k: v for k, v in f.items() if k not in ("transducer_source_flag", "transducer_source_input", "delay_mask")}
    g["ux_source_flag"] = np.array([[[flag]]], dtype=np.uint64)
    g["u_source_mode"] = np.array([[[1]]], dtype=np.uint64)
    g["u_source_many"] = np.array([[[1]]], dtype=np.uint64)
    g["ux_source_input"] = np.ascontiguousarray(series.T[None].astype(np.float32))
    b = oracle.run(g, record=("p", "u_final"))
    for k in a:
        assert np.allclose(a[k], b[k], rtol=0, atol=1e-12 * np.abs(b[k]).max()), k
    assert flag == sig.size - 7


@pytest.mark.parametrize("kind,mode", [("p", "additive"), ("u", "additive-no-correction")])
def test_labelled_source_is_sum_of_its_labels(oracle, kind, mode):
    """A labelled source mask drives each point with its label's signal row
    (kspaceFirstOrder3D.m:933-957 with p_source_sig_index, inputChecking.m:881-883).
    In a homogeneous lossless medium the field is linear in the sources, so the
    labelled run equals the sum of runs that each keep one label's points (a
    binary mask) driven by that label's row; a label/row mix-up or an
    off-by-one in the expansion breaks the sum."""
    import cases
    rec = ("p", "u_final")
    a = oracle.run(cases.case_labelled(kind, mode=mode), record=rec)
    parts = [oracle.run(cases.case_labelled(kind, mode=mode, only=k), record=rec) for k in (1, 2, 3)]
    for key in a:
        s = sum(p[key] for p in parts)
        assert np.abs(s).max() > 0
        assert np.allclose(a[key], s, rtol=0, atol=1e-11 * np.abs(s).max()), key


def _dtt_wswa(oracle, f, p, ux, uy, sm):
    """The WSWA (DTT) branch of kspaceFirstOrderAS.m (733-750 gradient, 991-1002
    divergence, 830-836 / 905-908 source corrections) with scipy's unnormalised
    DCT/DST types 2, 3, 4 (FFTW's REDFT10/01/11, RODFT11) along the radial axis."""
    from scipy import fft as sfft
    pb = oracle.decode(f)
    Ny, M = pb.Ny, 2 * pb.Ny
    kx = oracle.k_vec_unshifted(pb.Nx, pb.dx)[None, :]
    ky = (2 * np.pi * (np.arange(Ny) + 0.5) / (M * pb.dy))[:, None]      # kWaveGrid.m:843-851
    kap = oracle.sinc(pb.c_ref * np.sqrt(kx ** 2 + ky ** 2) * pb.dt / 2)
    skap = np.cos(pb.c_ref * np.sqrt(kx ** 2 + ky ** 2) * pb.dt / 2)
    dpos = 1j * kx * np.exp(1j * kx * pb.dx / 2)
    dneg = 1j * kx * np.exp(-1j * kx * pb.dx / 2)
    dct = lambda a, t: sfft.dct(a, type=t, axis=-2)                      # noqa: E731
    dst = lambda a, t: sfft.dst(a, type=t, axis=-2)                      # noqa: E731
    fx = lambda a: np.fft.fft(a, axis=-1)                                # noqa: E731
    ix = lambda a: np.real(np.fft.ifft(a, axis=-1))                      # noqa: E731
    pk = kap * fx(dct(p, 3))
    dpdx = dct(ix(dpos * pk), 2) / M
    dpdy = dst(ix(-ky * pk), 4) / M
    duxdx = dct(ix(kap * dneg * fx(dct(ux, 3))), 2) / M
    ysg = ((np.arange(Ny) + 0.5) * pb.dy)[:, None]
    duydy = dct(ix(kap * (ky * fx(dst(uy, 4)) + fx(dct(uy / ysg, 4)))), 2) / M
    s_ws = dct(ix(skap * fx(dct(sm, 3))), 2) / M
    s_ha = dst(ix(skap * fx(dst(sm, 4))), 4) / M
    return dpdx, dpdy, duxdx, duydy, s_ws, s_ha


def test_axisymmetric_wswa_fft_equals_dtt(oracle):
    """The oracle's WSWA-FFT gradients (mirror to 4 Ny, fft2, trim) equal the
    WSWA DTT form the input file selects (inputChecking.m:1400-1411) to
    machine precision, as testing/unit/kspaceFirstOrderAS_compare_FFT_implementations.m
    requires of the reference (1e-14); a wrong mirror sign, offset or wavenumber
    breaks it."""
    import cases
    f = cases.case_axisym(Nx=30, Ny=24, Nt=2, source="p")
    rng = np.random.default_rng(3)
    p, ux, uy, sm = (rng.standard_normal((1, 24, 30)) for _ in range(4))
    ax = oracle.AxisymOps(oracle.decode(f))
    got = ax.grad_p(p) + ax.div(ux, uy) + [ax.source_correction(sm, "wswa"), ax.source_correction(sm, "hahs")]
    ref = _dtt_wswa(oracle, f, p, ux, uy, sm)
    for g, r in zip(got, ref):
        assert np.abs(g - r).max() <= 1e-12 * np.abs(r).max()


def test_axisymmetric_matches_3d_cylindrical_wave(oracle):
    """An axisymmetric initial pressure (a Gaussian on the axis) run with the
    axisymmetric loop equals the 3-D loop on the same Gaussian, on the axis and
    along a radius (homogeneous, lossless, before the wave reaches a PML): the
    (1/r) d(r u_r)/dr divergence and the radial symmetries are what make the
    two agree.  Dropping the u_r / r term gives 0.6 relative error, the wrong
    symmetry for it 0.23; the two discretisations agree to 1.3e-3."""
    import cases
    fe = cases.fe
    a = oracle.run(cases.case_axisym(Nx=64, Ny=32, Nt=50), record=("p",))["p"]
    N, dx, c = 64, 1e-4, 32
    kg = fe.KWaveGrid(N, dx, N, dx, N, dx)
    kg.set_time(50, 0.3 * dx / 1500)
    g = (np.arange(N) - c) * dx
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    p0 = 2.0 * np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / (2 * (2.5 * dx) ** 2))
    m = np.zeros((N, N, N))
    m[:, c, c] = 1            # the axis
    m[c, c:, c] = 1           # a radius in the mid plane
    f3 = fe.save_to_disk_dict(kg, {"sound_speed": 1500.0, "density": 1000.0}, {"p0": p0}, {"mask": m},
                              pml_size=8, smooth_p0=False)
    b = oracle.run(f3, record=("p",))["p"]
    assert a.shape == b.shape
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 3e-3


@pytest.mark.slow
def test_axisymmetric_piston_on_axis_amplitude(oracle):
    """testing/unit/kspaceFirstOrderAS_piston_analytical.m on the oracle: the
    steady-state on-axis amplitude of a CW piston within 1 % of the closed form
    (measured 0.64 %); pins the axisymmetric additive-source scaling and
    propagation (about a minute on 8 cores)."""
    import cases
    f, start, error_pct = cases.piston_axisym()
    p = oracle.run(f, record=("p",), record_start_index=start)["p"]
    assert error_pct(p) < 1.0


def test_plane_wave_intensity_from_u_non_staggered(oracle):
    """testing/unit/kspaceFirstOrder_check_plane_wave_intensity.m (2-D x case):
    the intensity the MATLAB wrapper forms from p and ux_non_staggered
    (kspaceFirstOrder_saveIntensity.m) equals the source intensity within 1 %;
    pins u_non_staggered's amplitude and phase against p, and the p source scaling."""
    import cases
    f, start, error_pct = cases.plane_wave_intensity()
    out = oracle.run(f, record=("p", "u_non_staggered"), record_start_index=start)
    assert error_pct(out["p"], out["ux_non_staggered"]) < 1.0


def test_nonstaggered_velocity_peaks_are_ordered(oracle):
    """kspaceFirstOrder2D_check_nonstaggered_particle_velocity.m: going outward from
    a disc source, the peak of ux_non_staggered at a point, of ux (half a cell
    further), of ux_non_staggered one point on and of ux after it decrease and
    arrive in that order; a shift of the wrong sign or size breaks the order."""
    import cases
    out = oracle.run(cases.nonstaggered_velocity_order(), record=("u", "u_non_staggered"))
    ok, mx, ix = cases.velocity_order_holds(out["ux_non_staggered"], out["ux"])
    assert ok, (mx, ix)


@pytest.mark.parametrize("kind", ["p", "ux"])
def test_source_corrections_rank_against_fine_dt(oracle, kind):
    """kspaceFirstOrder2D_check_source_scaling_p/_ux.m (and the AS variants) in
    miniature: at CFL 0.5 the k-space corrected additive source matches a 10x
    finer-dt reference best, the frequency-referenced (w) correction next, no
    correction worst (err_k < err_w < err_none); the w correction is the front
    end's cos(2 pi f_ref dt/2) scaling with the k-space correction switched off."""
    import cases
    fine = oracle.run(cases.w_correction_plane_wave(kind, fine=True), record=("p",))["p"][:, 0]
    # coarse sample n against the fine series at the same time: fine sample
    # 10n+9 for a velocity source; a pressure source enters the density update,
    # which the staggered time grid places half a step off the velocity one, so
    # its comparison point is midway between fine samples 10n+4 and 10n+5
    ref = 0.5 * (fine[4::10] + fine[5::10]) if kind == "p" else fine[9::10]
    err = {}
    for v in ("k", "w", "none"):
        out = oracle.run(cases.w_correction_plane_wave(kind, variant=v), record=("p",))["p"][:, 0]
        n = min(out.size, ref.size)
        err[v] = np.max(np.abs(out[:n] - ref[:n])) / np.max(np.abs(fine))
    assert err["k"] < err["w"] < err["none"], err
    assert err["k"] < 1e-4, err
