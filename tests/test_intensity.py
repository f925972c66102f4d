"""Host-side intensity post-processing (kspaceFirstOrder_saveIntensity.m,
fourierShift.m) and the non-staggered velocity's sensor-data shape: the part of
the output contract MATLAB computes after the binary returns."""

import numpy as np
import pytest

import cases

fe = cases.fe


@pytest.mark.parametrize("n", [60, 61])
def test_fourier_shift_advances_band_limited_series(n):
    """A whole number of periods below Nyquist is interpolated exactly: shifting by
    1/2 gives the series half a sample later, by 1 a circular advance (a sign
    mistake would delay instead), and the test helper's restatement agrees."""
    t = np.arange(n)
    x = np.sin(2 * np.pi * 3 * t / n) + 0.5 * np.cos(2 * np.pi * 7 * t / n + 0.3)
    half = fe.fourier_shift(x, 0.5)
    exact = np.sin(2 * np.pi * 3 * (t + 0.5) / n) + 0.5 * np.cos(2 * np.pi * 7 * (t + 0.5) / n + 0.3)
    assert np.max(np.abs(half - exact)) < 1e-12
    assert np.max(np.abs(fe.fourier_shift(x, 1.0) - np.roll(x, -1))) < 1e-12
    assert np.max(np.abs(half - cases.fourier_shift(x, 0.5))) < 1e-13
    # along the last axis of a [Ns, Nt] block, row by row
    blk = np.stack([x, 2 * x])
    assert np.max(np.abs(fe.fourier_shift(blk, 0.5) - np.stack([half, 2 * half]))) < 1e-12


def _sd(dim, ns=3, nt=40):
    rng = np.random.default_rng(5)
    sd = {"p": rng.standard_normal((ns, nt))}
    for a in "xyz"[:dim]:
        sd[f"u{a}_non_staggered"] = rng.standard_normal((ns, nt))
    return sd


@pytest.mark.parametrize("dim", [2, 3])
def test_save_intensity_fields(dim):
    sd = _sd(dim)
    out = fe.save_intensity(sd, dim, ("I", "I_avg"))
    for a in "xyz"[:dim]:
        I = sd["p"] * fe.fourier_shift(sd[f"u{a}_non_staggered"], 0.5)
        assert np.array_equal(out[f"I{a}"], I)
        assert np.allclose(out[f"I{a}_avg"], I.mean(axis=1))
        assert f"u{a}_non_staggered" not in out
    assert "p" not in out                       # not requested: removed (saveIntensity.m:102-104)
    only_avg = fe.save_intensity(sd, dim, ("I_avg", "p", "u_non_staggered"))
    assert set(only_avg) == {"p"} | {f"u{a}_non_staggered" for a in "xyz"[:dim]} | {f"I{a}_avg" for a in "xyz"[:dim]}
    # one dict per cuboid
    two = fe.save_intensity([sd, sd], dim, ("I",))
    assert len(two) == 2 and all(np.array_equal(t["Ix"], out["Ix"]) for t in two)


def test_intensity_of_a_travelling_wave():
    """p = A sin(wt), u = p / (rho c) recorded half a step early (the staggered
    time grid): the shifted product averages to A^2 / (2 rho c), the plane-wave
    intensity, while the unshifted product is off by cos(w dt / 2)."""
    n, per, A, rc = 120, 12, 3.0, 1.5e6
    t = np.arange(n)
    w = 2 * np.pi / per
    p = A * np.sin(w * t)[None]
    u = (A / rc) * np.sin(w * (t - 0.5))[None]
    out = fe.save_intensity({"p": p, "ux_non_staggered": u, "uy_non_staggered": 0 * u}, 2, ("I_avg",))
    assert abs(out["Ix_avg"][0] - A * A / (2 * rc)) < 1e-12 * A * A / rc
    assert abs(np.mean(p * u) - A * A / (2 * rc) * np.cos(w / 2)) < 1e-12 * A * A / rc
    assert out["Iy_avg"][0] == 0.0


def test_non_staggered_sensor_data_is_a_time_series():
    """/ux_non_staggered is read like /ux: [Ns, Nt_rec] (3DC.m:706-711)."""
    from kwave_amd import h5format

    o = {"p": np.arange(12.0).reshape(1, 4, 3), "ux_non_staggered": np.arange(12.0).reshape(1, 4, 3),
         "uy_non_staggered": np.zeros((1, 4, 3))}
    sd = h5format.sensor_data_from_output(o, ("p", "u_non_staggered"), 2)
    assert sd["ux_non_staggered"].shape == (3, 4) and np.array_equal(sd["ux_non_staggered"], sd["p"])
