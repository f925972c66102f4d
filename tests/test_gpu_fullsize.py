"""GPU checks at the benchmark's full size (128^3), through properties that do
not need a full-length FP64 oracle run: the closed-form lossless solution, the
C2 workload against the oracle over its source-active opening, mirror
symmetry, and run-to-run determinism."""

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def test_closed_form_lossless_128cubed(kw, oracle):
    """p_n = IFFT(FFT(p0) cos(c|k|(n-1)dt)) at 128^3 (no PML), 300 steps, FP32 GPU."""
    print("[gpu full] closed form 128^3 lossless homogeneous, 300 steps", flush=True)
    fe = kw.frontend
    n, d, Nt = 128, 1e-4, 300
    rng = np.random.default_rng(11)
    kg = fe.KWaveGrid(n, d, n, d, n, d)
    kg.set_time(Nt, 0.3 * d / 1500.0)
    p0 = kw.geometry.smooth(3.0 * kw.geometry.make_ball(n, n, n, 60, 70, 64, 9), True)
    mask = (rng.random((n, n, n)) < 0.001).astype(float)
    f = fe.save_to_disk_dict(kg, {"sound_speed": 1500.0, "density": 1000.0}, {"p0": p0}, {"mask": mask},
                             pml_size=0, pml_alpha=0.0, smooth_p0=False)
    gpu = kw.run(f, record=("p", "p_final"))
    val = lambda k: float(np.asarray(f[k]).reshape(-1)[0])  # noqa: E731
    p0f = np.asarray(f["p0_source_input"], dtype=np.float64).reshape(n, n, n)
    idx = np.asarray(f["sensor_mask_index"]).reshape(-1).astype(np.int64) - 1
    steps = [1, 2, 50, 150, 300]
    kv = [oracle.k_vec_unshifted(n, val(a)) for a in ("dx", "dy", "dz")]
    k = np.sqrt(kv[0][None, None, :] ** 2 + kv[1][None, :, None] ** 2 + kv[2][:, None, None] ** 2)
    P0 = np.fft.fftn(p0f)
    for s in steps:
        ex = np.real(np.fft.ifftn(P0 * np.cos(val("c0") * k * (s - 1) * val("dt"))))
        e = cases.rel_l2(gpu["p"][s - 1], ex.reshape(-1)[idx])
        print(f"    step {s}: rel L2 {e:.2e}", flush=True)
        assert e < 1e-4, (s, e)
    assert cases.rel_l2(gpu["p_final"], np.real(np.fft.ifftn(P0 * np.cos(val("c0") * k * (Nt - 1) * val("dt"))))) < 1e-4


def test_c2_workload_against_oracle(kw, oracle):
    """The 128^3 benchmark workload (heterogeneous absorbing, ux additive), first 90
    steps (the source is active for the first 82) against the FP64 oracle."""
    from kwave_amd import configs

    print("[gpu full] C2 128^3, 90 steps vs oracle", flush=True)
    f = configs.c2(Nt=90)
    rec = ("p", "p_max")
    gpu = kw.run(f, record=rec)
    ref = oracle.run(f, record=rec)
    for k in rec:
        e = cases.rel_l2(gpu[k], ref[k])
        print(f"    {k}: rel L2 {e:.2e}", flush=True)
        assert e < 1e-4, (k, e)


def test_mirror_symmetry_and_determinism(kw):
    """A y-symmetric problem stays y-symmetric; two runs are bit-identical."""
    print("[gpu full] mirror symmetry + determinism 96x96x64", flush=True)
    fe = kw.frontend
    nx, ny, nz, d = 96, 96, 64, 1e-4
    rng = np.random.default_rng(3)
    c = 1500 * (1 + 0.02 * cases.smooth_field(rng, (nx, ny, nz)))
    c = 0.5 * (c + c[:, ::-1, :])
    kg = fe.KWaveGrid(nx, d, ny, d, nz, d)
    kg.set_time(120, 0.3 * d / c.max())
    p0 = kw.geometry.make_ball(nx, ny, nz, 40, 48, 32, 5) + 0.0
    p0 = 0.5 * (p0 + p0[:, ::-1, :])
    f = fe.save_to_disk_dict(kg, {"sound_speed": c, "density": 1000.0, "alpha_coeff": 0.5, "alpha_power": 1.5},
                             {"p0": p0}, {"mask": np.ones((nx, ny, nz))}, pml_size=8)
    a = kw.run(f, record=("p_final",))["p_final"]
    b = kw.run(f, record=("p_final",))["p_final"]
    assert np.array_equal(a, b)
    # p0 and the medium are mirrored about y = (ny-1)/2, which maps the staggered
    # grid onto itself, so p must be mirror-symmetric to FP32 rounding
    sym = cases.rel_l2(a, a[:, ::-1, :])
    print(f"    mirror asymmetry {sym:.2e}", flush=True)
    assert sym < 1e-4, sym


def test_c3_full_size_window_against_oracle(kw, oracle):
    """C3 at its full size (256x256x512, nonlinear, y=2 absorbing, bowl source, every
    point recording p_max/p_rms): the first 12 steps against the FP64 oracle."""
    from kwave_amd import configs

    print("[gpu full] C3 256x256x512, 12 steps vs oracle (all-point p_max, p_rms)", flush=True)
    f = configs.c3(Nt=12)
    rec = ("p_max", "p_rms")
    gpu = kw.run(f, record=rec)
    ref = oracle.run(f, record=rec)
    for k in rec:
        e = cases.rel_l2(gpu[k], ref[k])
        print(f"    {k}: rel L2 {e:.2e}", flush=True)
        assert e < 1e-4, (k, e)


def test_c3_full_size_properties(kw):
    """C3 over 400 steps: finite, p_min <= p_max, p_rms <= max|p|, and the field's
    maximum lies under the bowl aperture."""
    from kwave_amd import configs

    print("[gpu full] C3 256x256x512, 400 steps, sensor-statistics properties", flush=True)
    f = configs.c3(Nt=400)
    out = kw.run(f, record=("p_max", "p_min", "p_rms"))
    pmax, pmin, prms = out["p_max"], out["p_min"], out["p_rms"]
    assert np.all(np.isfinite(pmax)) and np.all(np.isfinite(pmin)) and np.all(np.isfinite(prms))
    assert np.all(pmin <= pmax)
    assert np.all(prms <= np.maximum(np.abs(pmax), np.abs(pmin)) * (1 + 1e-5) + 1e-3)
    j = int(np.argmax(pmax))                     # MATLAB linear index order (x fastest)
    x, y = j % 256, (j // 256) % 256
    assert abs(x - 127) <= 64 and abs(y - 127) <= 64, (x, y)


def test_c4_full_size_linearity(kw):
    """C4 at its full size (768^3, linear absorbing slab): doubling the source signal
    doubles the field (scaling by 2 commutes with rounding except where values fall
    into the subnormal range, hence a 1e-6 bound rather than bit equality)."""
    from kwave_amd import configs

    print("[gpu full] C4 768^3, 30 steps, source linearity", flush=True)
    f = configs.c4(Nt=30)
    a = kw.run(f, record=("p_final",))["p_final"]
    f2 = dict(f)
    f2["p_source_input"] = (2 * np.asarray(f["p_source_input"], dtype=np.float32)).astype(np.float32)
    b = kw.run(f2, record=("p_final",))["p_final"]
    assert np.all(np.isfinite(a)) and np.abs(a).max() > 0
    err = np.abs(b - 2 * a).max() / np.abs(2 * a).max()
    print(f"    max |p(2s) - 2 p(s)| / max |2 p(s)| = {err:.2e}", flush=True)
    assert err < 1e-6
