"""GPU parity: the HIP solver (through the C ABI) against the FP64 oracle.

Tolerance: relative L2 <= 1e-4 on every recorded output (the FP32 drop-in
bar), plus a max-abs bound relative to the output's peak.  Each case prints
its name before running so a failing GPU step is identifiable.
"""

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

TOL_L2 = 1e-4


def compare(kw, oracle, f, record, name, tol=TOL_L2, start=1):
    print(f"[gpu parity] {name} record={record}", flush=True)
    gpu = kw.run(f, record=record, record_start_index=start)
    ref = oracle.run(f, record=record, record_start_index=start)
    assert set(ref) <= set(gpu) | set(), (sorted(ref), sorted(gpu))
    for k, r in ref.items():
        g = gpu[k]
        assert g.shape == r.shape, (k, g.shape, r.shape)
        e = cases.rel_l2(g, r)
        mx = np.max(np.abs(g - r)) / max(np.max(np.abs(r)), 1e-30)
        print(f"    {k}: rel L2 {e:.2e}  max-abs/peak {mx:.2e}", flush=True)
        assert e <= tol, (name, k, e)
        assert mx <= 10 * tol, (name, k, mx)
    return gpu, ref


def test_2d_p0_lossless_homogeneous(kw, oracle):
    f = cases.case_2d_p0_lossless(N=64, Nt=80)
    compare(kw, oracle, f, ("p", "p_max", "p_min", "p_rms", "p_final"), "2d_p0_homog")


def test_2d_p0_heterogeneous_velocity_outputs(kw, oracle):
    f = cases.case_2d_p0_lossless(N=64, Nt=60, het=True)
    compare(kw, oracle, f, ("p", "u", "u_max", "u_rms", "u_final", "p_max_all", "p_min_all"), "2d_p0_het")


U_STATS = ("u", "u_max", "u_min", "u_rms", "u_max_all", "u_min_all", "u_final", "p_min")


@pytest.mark.parametrize("dim", ["2d", "3d_plane"])
def test_every_velocity_statistic(kw, oracle, dim):
    """u_min, u_max_all and u_min_all beside the other velocity outputs, with a
    record start index past 1: every statistic, the whole-grid _all fields
    included, runs over the recorded steps only (extractSensorData.m:61-297)."""
    if dim == "2d":
        f = cases.case_2d_p0_lossless(N=64, Nt=50, het=True)
    else:
        f = cases.case_3d(N=(32, 64, 64), Nt=30, source="uxyz")
    compare(kw, oracle, f, U_STATS, f"u_stats_{dim}", start=7)


@pytest.mark.parametrize("mode", ["additive", "additive-no-correction", "dirichlet"])
def test_2d_absorbing_pressure_source(kw, oracle, mode):
    f = cases.case_2d_absorbing_psource(mode=mode)
    compare(kw, oracle, f, ("p", "p_max"), f"2d_abs_p_{mode}")


def test_2d_many_additive_nonlinear(kw, oracle):
    f = cases.case_2d_absorbing_psource(many=True, nonlinear=True)
    compare(kw, oracle, f, ("p", "p_rms"), "2d_many_nonlinear")


def test_2d_stokes(kw, oracle):
    f = cases.case_2d_absorbing_psource(stokes=True)
    compare(kw, oracle, f, ("p",), "2d_stokes")


def test_3d_absorbing_heterogeneous_ux_additive(kw, oracle):
    f = cases.case_3d()
    compare(kw, oracle, f, ("p", "p_max", "u"), "3d_abs_ux_add")


@pytest.mark.parametrize("mode", ["dirichlet", "additive-no-correction"])
def test_3d_velocity_modes(kw, oracle, mode):
    f = cases.case_3d(mode=mode, source="uxyz", Nt=40)
    compare(kw, oracle, f, ("p", "u_max"), f"3d_u_{mode}")


def test_3d_nonlinear_p_source_many(kw, oracle):
    f = cases.case_3d(N=(30, 28, 36), source="p", many=True, nonlinear=True, alpha_power=2.0, Nt=50)
    compare(kw, oracle, f, ("p", "p_max", "p_rms"), "3d_nl_p_many")


def test_3d_p0_lossless_odd_grid(kw, oracle):
    f = cases.case_3d(N=(25, 21, 35), absorbing=False, source=None, p0=True, Nt=40, pml=4)
    compare(kw, oracle, f, ("p", "p_final", "u_final"), "3d_p0_odd")


def test_record_start_index(kw, oracle):
    f = cases.case_2d_p0_lossless(N=48, Nt=50)
    compare(kw, oracle, f, ("p", "p_max", "p_rms", "p_max_all"), "2d_start_index", start=20)


def test_p0_equals_two_clicks_gpu(kw, oracle):
    """kspaceFirstOrder2D_p0_vs_two_clicks.m on the GPU: both runs match the
    oracle and each other (FP32: 1e-5 of the peak instead of 1e-14)."""
    fe, geo = kw.frontend, kw.geometry
    print("[gpu parity] p0 vs two clicks", flush=True)
    Nx = Ny = 64
    kgrid = fe.KWaveGrid(Nx, 1e-4, Ny, 1e-4)
    kgrid.make_time(1500.0, 0.234)
    p0 = geo.make_disc(Nx, Ny, 25, 25, 4) + geo.make_disc(Nx, Ny, 40, 30, 2)
    rr = np.hypot(*np.meshgrid(np.arange(Nx) - 32, np.arange(Ny) - 32, indexing="ij"))
    mask = (np.abs(rr - 16) < 0.5).astype(float)
    medium = {"sound_speed": 1500.0, "density": 1000.0}
    f1 = fe.save_to_disk_dict(kgrid, medium, {"p0": p0}, {"mask": mask}, pml_alpha=0.0, smooth_p0=False)
    src = {"p_mask": (p0 != 0).astype(float), "p_mode": "additive-no-correction",
           "p": np.array([[0.5, 0.5]]) / (2 * 1500.0 * kgrid.dt / kgrid.dx)}
    f2 = fe.save_to_disk_dict(kgrid, medium, src, {"mask": mask}, pml_alpha=0.0)
    a = kw.run(f1)["p"]
    b = kw.run(f2)["p"]
    assert np.max(np.abs(a - b)) < 1e-5 * np.abs(a).max()
    assert cases.rel_l2(a, oracle.run(f1)["p"]) < 1e-4


@pytest.mark.parametrize("shape", [(64, 48, 96), (48, 64, 32), (40, 32, 192), (32, 48, 768), (48, 768, 32)])
def test_3d_specialised_pencil_lengths(kw, oracle, shape):
    """Grids whose y/z lengths have compiled pencil kernels (others use the generic ones);
    a 768-point last axis runs the 1024-thread, 16-pencil tiles, a 768-point y axis
    the 8-pencil ones."""
    f = cases.case_3d(N=shape, Nt=30, nonlinear=True, source="p", pml=4)
    compare(kw, oracle, f, ("p", "p_max", "u_final"), f"3d_pencil_{shape}")


@pytest.mark.parametrize("shape", [(96, 160), (128, 128), (64, 256), (64, 768)])
def test_2d_specialised_last_axis(kw, oracle, shape):
    f = cases.case_2d_absorbing_psource(N=shape, Nt=60)
    compare(kw, oracle, f, ("p", "p_rms"), f"2d_last_{shape}")


def _uses_plane(kw, f):
    with kw.Simulation(f, record=("p",)) as sim:
        sim.set_profiling(True)
        sim.run(2)
        sim.sync()
        return sim.profile().get("yz_plane", (0.0, 0))[1] > 0


PLANE_CASES = [
    ((48, 32, 32), dict(source="p", many=True, nonlinear=True)),     # corrected source spectra per step
    ((32, 64, 64), dict(source="uxyz")),                             # single-signal additive patterns
    ((64, 48, 48), dict(p0=True, source=None, absorbing=False)),     # p0 start chain
    ((32, 96, 96), dict(source="ux", mode="dirichlet")),
    ((32, 128, 128), dict(source="p")),
    ((48, 64, 128), dict(source="ux", alpha_power=1.2)),
    ((32, 128, 96), dict(source="p", mode="additive-no-correction")),
]


@pytest.mark.parametrize("shape,kwargs", PLANE_CASES, ids=[str(c[0]) for c in PLANE_CASES])
def test_3d_plane_pass(kw, oracle, shape, kwargs):
    """Grids whose (y, z) plane fits in LDS: kx-major spectra and the plane pass."""
    f = cases.case_3d(N=shape, Nt=30, pml=4, **kwargs)
    assert _uses_plane(kw, f), shape
    compare(kw, oracle, f, ("p", "p_max", "u_final"), f"3d_plane_{shape}")


def test_plane_pass_matches_three_pass(kw, monkeypatch):
    """The same simulation with the plane pass and with y pass / last axis / y pass."""
    f = cases.case_3d(N=(64, 64, 64), Nt=40, pml=6, source="uxyz", nonlinear=True)
    a = kw.run(f, record=("p", "u_final"))
    monkeypatch.setenv("KW_PLANE", "0")
    assert not _uses_plane(kw, f)
    b = kw.run(f, record=("p", "u_final"))
    for k in a:
        assert cases.rel_l2(a[k], b[k]) < 2e-6, k


def test_c1_cartesian_sensor(kw, oracle):
    """C1 at full size (128^2, Nt 500): GPU vertex series -> Cartesian points vs the oracle."""
    f, s = kw.configs.c1()
    print("[gpu parity] C1 128^2 Nt 500 Cartesian sensor", flush=True)
    g = s.apply(kw.run(f, record=("p",))["p"])
    r = s.apply(oracle.run(f, record=("p",))["p"])
    assert g.shape == (500, 50)
    e = cases.rel_l2(g, r)
    print(f"    p (50 Cartesian points): rel L2 {e:.2e}", flush=True)
    assert e <= TOL_L2


def test_c3_reduced(kw, oracle):
    """C3's branch (nonlinear, y=2 absorbing, bowl p source, all-point p_max/p_rms) on 64x64x128."""
    f = kw.configs.c3(Nt=100, shape=(64, 64, 128), radius=24, diameter=31, pml=6)
    compare(kw, oracle, f, ("p_max", "p_rms"), "c3_reduced")


def test_c4_reduced(kw, oracle):
    """C4's branch (three-layer slab, y=1.1, p additive patch, receiver plane) on 96^3."""
    f = kw.configs.c4(Nt=60, N=96)
    compare(kw, oracle, f, ("p",), "c4_reduced")


@pytest.mark.parametrize("case", ["2d", "3d_plane", "3d_pencil", "3d_p0_start"])
def test_u_non_staggered(kw, oracle, case):
    """--u_non_staggered_raw: u shifted to the unstaggered points, per axis."""
    if case == "2d":
        f = cases.case_2d_p0_lossless(N=64, Nt=40, het=True)
        rec = ("u_non_staggered", "u")
    elif case == "3d_plane":
        f = cases.case_3d(N=(48, 32, 32), Nt=30, pml=4, source="uxyz")
        rec = ("u_non_staggered", "p")
    elif case == "3d_pencil":
        f = cases.case_3d(N=(32, 48, 40), Nt=30, pml=4, source="uxyz")
        rec = ("u_non_staggered",)
    else:
        f = cases.case_3d(N=(32, 32, 48), Nt=20, pml=4, p0=True, source=None)
        rec = ("u_non_staggered", "u")
    compare(kw, oracle, f, rec, f"u_non_staggered_{case}", start=3 if case == "2d" else 1)


@pytest.mark.parametrize("steer", [True, False])
def test_transducer_source(kw, oracle, steer):
    """kWaveTransducer-style source (3D.m:889-902) with and without steering delays."""
    f = cases.case_transducer(steer=steer)
    compare(kw, oracle, f, ("p", "p_max", "u_final"), f"transducer_steer{steer}")


@pytest.mark.parametrize("kind,mode", [("p", "additive"), ("u", "additive"), ("p", "dirichlet")])
def test_labelled_source_mask(kw, oracle, kind, mode):
    """Labelled p_mask / u_mask (inputChecking.m:847-884, 964-1003) in a heterogeneous medium."""
    f = cases.case_labelled(kind, mode=mode, het=True)
    compare(kw, oracle, f, ("p", "p_max", "u_final"), f"labelled_{kind}_{mode}")


AXISYM_CASES = [
    ("p0", dict(source="p0")),
    ("p0_het_stokes_nonlinear", dict(source="p0", het=True, absorbing=True, nonlinear=True)),
    ("p_additive", dict(source="p")),
    ("p_dirichlet_het", dict(source="p", mode="dirichlet", het=True)),
    ("p_no_correction", dict(source="p", mode="additive-no-correction")),
    ("u_additive_het_stokes", dict(source="u", het=True, absorbing=True)),
    ("u_dirichlet", dict(source="u", mode="dirichlet")),
]


@pytest.mark.parametrize("name,kwargs", AXISYM_CASES, ids=[c[0] for c in AXISYM_CASES])
def test_axisymmetric(kw, oracle, name, kwargs):
    """kspaceFirstOrderAS (WSWA radial symmetry) through the C ABI against the oracle."""
    f = cases.case_axisym(Nx=64, Ny=48, Nt=60, **kwargs)
    compare(kw, oracle, f, ("p", "p_max", "p_rms", "u", "u_final", "p_final"), f"axisym_{name}")


def test_axisymmetric_odd_lengths(kw, oracle):
    """Generic x rows (Nx = 70) and a radial length 2 Ny = 90 with radix 3 and 5 stages."""
    f = cases.case_axisym(Nx=70, Ny=45, Nt=40, source="p", het=True, pml=6)
    compare(kw, oracle, f, ("p", "u_final"), "axisym_odd")


def test_plane_wave_intensity(kw, oracle):
    """testing/unit/kspaceFirstOrder_check_plane_wave_intensity.m through the C ABI:
    the intensity formed from p and ux_non_staggered is the source's 10 W/cm^2
    within 1 %, and both series match the oracle."""
    f, start, error_pct = cases.plane_wave_intensity()
    print("[gpu parity] plane_wave_intensity", flush=True)
    gpu = kw.run(f, record=("p", "u_non_staggered"), record_start_index=start)
    ref = oracle.run(f, record=("p", "u_non_staggered"), record_start_index=start)
    for k in ("p", "ux_non_staggered"):
        e = cases.rel_l2(gpu[k], ref[k])
        print(f"    {k}: rel L2 {e:.2e}", flush=True)
        assert e <= TOL_L2, (k, e)
    # uy is zero by symmetry (FP64 leaves ~1e-30): bounded by ux's peak instead
    assert np.abs(gpu["uy_non_staggered"]).max() <= 1e-5 * np.abs(gpu["ux_non_staggered"]).max()
    err = error_pct(gpu["p"], gpu["ux_non_staggered"])
    print(f"    intensity error {err:.3f} %", flush=True)
    assert err < 1.0


def test_nonstaggered_velocity_peaks_are_ordered(kw, oracle):
    """kspaceFirstOrder2D_check_nonstaggered_particle_velocity.m through the C ABI."""
    f = cases.nonstaggered_velocity_order()
    print("[gpu parity] nonstaggered_velocity_order", flush=True)
    rec = ("u", "u_non_staggered")
    gpu, ref = kw.run(f, record=rec), oracle.run(f, record=rec)
    for k in ("ux", "uy", "ux_non_staggered"):
        e = cases.rel_l2(gpu[k], ref[k])
        print(f"    {k}: rel L2 {e:.2e}", flush=True)
        assert e <= TOL_L2, (k, e)
    # the sensors sit on the disc's symmetry axis, where uy_non_staggered is
    # zero up to rounding in both solvers: bounded by ux's peak instead
    assert np.abs(gpu["uy_non_staggered"]).max() <= 1e-4 * np.abs(gpu["ux_non_staggered"]).max()
    ok, mx, ix = cases.velocity_order_holds(gpu["ux_non_staggered"], gpu["ux"])
    assert ok, (mx, ix)
