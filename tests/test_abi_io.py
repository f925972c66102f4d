"""CPU tests of the boundary: the C-ABI libraries load and export every
declared symbol, input validation rejects bad descriptions before any GPU
call, the HDF5 files follow writeMatrix.m/writeAttributes.m, and the
executable fails loudly (non-zero exit, message on stderr) on inputs it
cannot run.  No test here launches GPU work."""

import os
import re
import subprocess

import numpy as np
import pytest

import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(kw\w*)\s*\(", txt)))


def exported(lib_path):
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_solver_library_exports_every_declared_symbol(kw):
    kw.load_library()
    names = declared("kwave_hip.h")
    assert "kw_sim_create" in names and "kw_sim_run" in names
    missing = set(names) - exported(kw.native.LIB_PATH)
    assert not missing, missing
    assert set(kw.native.EXPORTED) == set(names)


def test_io_library_exports_every_declared_symbol(kw):
    from kwave_amd import h5format

    h5format.load_io_library()
    names = set(declared("kwave_io.h")) - set(declared("kwave_hip.h"))
    missing = names - exported(h5format.IO_LIB_PATH)
    assert not missing, missing
    assert set(h5format.IO_EXPORTED) == names


def _err(kw, f, **kwargs):
    with pytest.raises(kw.native.KWaveError) as e:
        kw.Simulation(f, **kwargs)
    return e.value


def test_create_rejects_unsupported_prime_factor(kw):
    f = cases.case_2d_p0_lossless(N=44, Nt=5)   # 44 = 4 * 11
    e = _err(kw, f)
    assert e.status == 2 and "factor" in str(e)


def test_create_rejects_bad_indices_and_pml(kw):
    f = cases.case_2d_p0_lossless(N=32, Nt=5, pml=4)
    g = dict(f)
    g["sensor_mask_index"] = np.array([[[5, 3]]], dtype=np.uint64)
    assert "strictly increasing" in str(_err(kw, g))
    g["sensor_mask_index"] = np.array([[[0]]], dtype=np.uint64)
    assert "outside" in str(_err(kw, g))
    g = dict(f)
    g["pml_x_size"] = np.array([[[17]]], dtype=np.uint64)
    assert "PML" in str(_err(kw, g))
    g = dict(f)
    g["c0"] = np.ones((1, 1, 7), dtype=np.float32)
    assert "c0" in str(_err(kw, g))


def test_record_start_beyond_nt_rejected(kw):
    f = cases.case_2d_p0_lossless(N=32, Nt=5, pml=4)
    assert "record_start_index" in str(_err(kw, f, record_start_index=6))


# ---- HDF5 format --------------------------------------------------------------

def test_dataset_roundtrip_and_attributes(kw, tmp_path):
    from kwave_amd import h5format

    p = str(tmp_path / "a.h5")
    field = np.arange(4 * 5 * 6, dtype=np.float32).reshape(4, 5, 6)
    idx = np.array([[[1, 7, 9]]], dtype=np.uint64)
    with h5format.H5File(p, "w") as h:
        h.write("field", field)
        h.write("idx", idx)
        h.write("Nx", np.array([[[6]]], dtype=np.uint64))
        h.write("dt", np.array([[[2.5e-8]]], dtype=np.float32))
        h.attributes("input", "test file")
    with h5format.H5File(p) as h:
        assert sorted(h.names()) == ["Nx", "dt", "field", "idx"]
        np.testing.assert_array_equal(h.read("field"), field)
        np.testing.assert_array_equal(h.read("idx"), idx)
        assert h.read("dt").dtype == np.float32 and h.read("Nx").dtype == np.uint64
        # writeMatrix.m:206-212 and getH5Literals.m
        assert h.attribute("data_type", "field") == "float"
        assert h.attribute("data_type", "idx") == "long"
        assert h.attribute("domain_type", "field") == "real"
        assert h.attribute("file_type") == "input"
        assert h.attribute("major_version") == "1" and h.attribute("minor_version") == "2"


def test_input_file_holds_savetodisk_dataset_set(kw, tmp_path):
    """Dataset names of a 3-D absorbing ux-source input (saveToDisk.m:44-414)."""
    from kwave_amd import h5format

    f = cases.case_3d(N=(16, 16, 16), Nt=5, pml=2)
    p = str(tmp_path / "in.h5")
    h5format.write_input(p, f)
    names = set(h5format.H5File(p).names())
    expected = {"Nx", "Ny", "Nz", "Nt", "dx", "dy", "dz", "dt", "c0", "c_ref", "rho0", "rho0_sgx", "rho0_sgy",
                "rho0_sgz", "pml_x_size", "pml_y_size", "pml_z_size", "pml_x_alpha", "pml_y_alpha", "pml_z_alpha",
                "p_source_flag", "p0_source_flag", "ux_source_flag", "uy_source_flag", "uz_source_flag",
                "sxx_source_flag", "syy_source_flag", "szz_source_flag", "sxy_source_flag", "sxz_source_flag",
                "syz_source_flag", "transducer_source_flag", "nonuniform_grid_flag", "nonlinear_flag",
                "absorbing_flag", "elastic_flag", "axisymmetric_flag", "sensor_mask_type", "alpha_coeff",
                "alpha_power", "u_source_mode", "u_source_many", "u_source_index", "ux_source_input",
                "sensor_mask_index"}
    assert names == expected, (names ^ expected)
    back = h5format.read_file(p)
    for k, v in f.items():
        np.testing.assert_array_equal(back[k], np.asarray(v).reshape(back[k].shape))


def test_2d_file_has_no_z_datasets(kw, tmp_path):
    """2-D files drop every float dataset whose name contains 'z' (saveToDisk.m:420-427)."""
    f = cases.case_2d_absorbing_psource(Nt=5)
    zs = {k for k in f if "z" in k and np.asarray(f[k]).dtype == np.float32}
    assert not zs, zs
    assert int(np.asarray(f["Nz"]).reshape(-1)[0]) == 1 and int(np.asarray(f["pml_z_size"]).reshape(-1)[0]) == 0


def test_save_in_parts_equivalence(kw, tmp_path):
    """The same input built two ways gives the same datasets (hdf5_save_in_parts.m:82-137):
    through the front end, and written by hand as example_cpp_io_in_parts.m does."""
    from kwave_amd import h5format

    fe = kw.frontend
    Nx, Ny, Nz, dx = 24, 20, 16, 1e-4
    c0, rho0, alpha, y = 1500.0, 1000.0, 0.75, 1.5
    kgrid = fe.KWaveGrid(Nx, dx, Ny, dx, Nz, dx)
    kgrid.set_time(10, 2e-8)
    pmask = np.zeros((Nx, Ny, Nz))
    pmask[4, 6:14, 5:11] = 1
    sig = np.sin(np.arange(10) * 0.3)[None]
    smask = np.zeros((Nx, Ny, Nz))
    smask[18, :, 8] = 1
    f = fe.save_to_disk_dict(kgrid, {"sound_speed": c0, "density": rho0, "alpha_coeff": alpha, "alpha_power": y},
                             {"p": sig, "p_mask": pmask, "p_mode": "additive"}, {"mask": smask}, pml_size=3)
    a = str(tmp_path / "a.h5")
    b = str(tmp_path / "b.h5")
    h5format.write_input(a, f)
    # by hand: example_cpp_io_in_parts.m:95-262
    with h5format.H5File(b, "w") as h:
        u = lambda v: np.array([[[v]]], dtype=np.uint64)  # noqa: E731
        s = lambda v: np.array([[[v]]], dtype=np.float32)  # noqa: E731
        for k, v in (("Nx", Nx), ("Ny", Ny), ("Nz", Nz), ("Nt", 10), ("pml_x_size", 3), ("pml_y_size", 3),
                     ("pml_z_size", 3), ("p_source_flag", 10), ("p_source_mode", 2), ("p_source_many", 0),
                     ("absorbing_flag", 1)):
            h.write(k, u(v))
        for k in ("p0_source_flag", "ux_source_flag", "uy_source_flag", "uz_source_flag", "sxx_source_flag",
                  "syy_source_flag", "szz_source_flag", "sxy_source_flag", "sxz_source_flag", "syz_source_flag",
                  "transducer_source_flag", "nonuniform_grid_flag", "nonlinear_flag", "elastic_flag",
                  "axisymmetric_flag", "sensor_mask_type"):
            h.write(k, u(0))
        for k, v in (("dx", dx), ("dy", dx), ("dz", dx), ("dt", 2e-8), ("c0", c0), ("c_ref", c0), ("rho0", rho0),
                     ("rho0_sgx", rho0), ("rho0_sgy", rho0), ("rho0_sgz", rho0), ("alpha_coeff", alpha),
                     ("alpha_power", y), ("pml_x_alpha", 2.0), ("pml_y_alpha", 2.0), ("pml_z_alpha", 2.0)):
            h.write(k, s(v))
        # p_source_input pre-scaled by 2 dt / (N c0 dx), N = 3 (example_cpp_io_in_parts.m:208)
        h.write("p_source_input", (sig * (2 * 2e-8 / (3 * c0 * dx))).T[None].astype(np.float32))
        h.write("p_source_index", np.flatnonzero(pmask.reshape(-1, order="F"))[None, None] + 1)
        h.write("sensor_mask_index", np.flatnonzero(smask.reshape(-1, order="F"))[None, None] + 1)
        h.attributes("input")
    ra, rb = h5format.read_file(a), h5format.read_file(b)
    assert set(ra) == set(rb), set(ra) ^ set(rb)
    diffs = [k for k in ra if ra[k].shape != rb[k].shape or not np.array_equal(ra[k], rb[k])]
    assert not diffs, diffs


# ---- executable failure handling ---------------------------------------------------

def _binary(kw):
    from kwave_amd import h5format

    if not os.path.exists(h5format.BINARY_PATH):
        pytest.fail("kspaceFirstOrder-HIP was not built")
    return h5format


def test_binary_usage_errors(kw, tmp_path):
    h5 = _binary(kw)
    r = subprocess.run([h5.BINARY_PATH], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([h5.BINARY_PATH, "-i", "a", "-o", "b", "--bogus"], capture_output=True, text=True)
    assert r.returncode == 2 and "unknown option" in r.stderr
    r = subprocess.run([h5.BINARY_PATH, "-i", "a", "-o", "b", "--benchmark"], capture_output=True, text=True)
    assert r.returncode == 3 and "not supported" in r.stderr


def test_binary_missing_input_file(kw, tmp_path):
    h5 = _binary(kw)
    r = h5.run_binary(str(tmp_path / "nope.h5"), str(tmp_path / "out.h5"))
    assert r.returncode == 1 and "cannot open" in r.stderr
    assert not os.path.exists(tmp_path / "out.h5")


@pytest.mark.parametrize("mutate,msg", [
    (lambda f: f.update(elastic_flag=np.array([[[1]]], dtype=np.uint64)), "elastic_flag"),
    (lambda f: f.update(sensor_mask_type=np.array([[[1]]], dtype=np.uint64)), "sensor_mask_corners"),
    (lambda f: f.update(sensor_mask_type=np.array([[[1]]], dtype=np.uint64),
                        sensor_mask_corners=np.array([[[2, 2, 1, 70, 5, 1]]], dtype=np.uint64)), "outside the grid"),
    (lambda f: f.update(sensor_mask_type=np.array([[[2]]], dtype=np.uint64)), "sensor_mask_type"),
    (lambda f: f.pop("c0"), "c0"),
    (lambda f: f.update(p_source_input=np.zeros((1, 3, 1), dtype=np.float32)), "p_source_input"),
    (lambda f: f.update(axisymmetric_flag=np.array([[[1]]], dtype=np.uint64)), "stokes absorption only"),
])
def test_binary_rejects_inputs_it_cannot_run(kw, tmp_path, mutate, msg):
    h5 = _binary(kw)
    f = cases.case_2d_absorbing_psource(Nt=20)
    mutate(f)
    inp = str(tmp_path / "in.h5")
    h5.write_input(inp, f)
    r = h5.run_binary(inp, str(tmp_path / "out.h5"))
    assert r.returncode == 1, r
    assert msg in r.stderr, r.stderr


def test_binary_rejects_output_file_type(kw, tmp_path):
    h5 = _binary(kw)
    f = cases.case_2d_p0_lossless(N=32, Nt=5, pml=4)
    p = str(tmp_path / "o.h5")
    with h5.H5File(p, "w") as h:
        for k, v in f.items():
            h.write(k, v)
        h.attributes("output")
    r = h5.run_binary(p, str(tmp_path / "x.h5"))
    assert r.returncode == 1 and "file_type" in r.stderr


def test_cuboid_corners_front_end_layout(kw):
    """saveToDisk.m:355-378: corners [6, K] -> (1, K, 6) uint64; 2-D [4, K] gets z = 1."""
    fe = kw.frontend
    kg = fe.KWaveGrid(32, 1e-4, 24, 1e-4)
    kg.set_time(5, 2e-8)
    f = fe.save_to_disk_dict(kg, {"sound_speed": 1500.0, "density": 1000.0}, {"p0": np.ones((32, 24))},
                             {"corners": np.array([[3, 10], [4, 2], [8, 20], [9, 7]])}, pml_size=4)
    assert int(f["sensor_mask_type"].ravel()[0]) == 1 and "sensor_mask_index" not in f
    assert f["sensor_mask_corners"].dtype == np.uint64 and f["sensor_mask_corners"].shape == (1, 2, 6)
    assert f["sensor_mask_corners"][0].tolist() == [[3, 4, 1, 8, 9, 1], [10, 2, 1, 20, 7, 1]]
    with pytest.raises(ValueError):
        kw.native.build_desc(f)


def test_labelled_source_front_end_expansion(kw):
    """inputChecking.m:847-884: a labelled p_mask becomes one row per source
    point (p_source_many = 1) taking its label's row, rows scaled first with the
    sound speed at the r-th source point (scaleSourceTerms.m:60-62); labels
    must run 1..L with one row each."""
    fe = kw.frontend
    kg = fe.KWaveGrid(24, 1e-4, 20, 1e-4)
    kg.set_time(8, 2e-8)
    c = np.full((24, 20), 1500.0)
    c[5, 3:9] = [1400, 1450, 1500, 1550, 1600, 1650]
    m = np.zeros((24, 20))
    m[5, 3:9] = [2, 1, 2, 2, 1, 2]
    sig = np.arange(16, dtype=float).reshape(2, 8) + 1
    f = fe.save_to_disk_dict(kg, {"sound_speed": c, "density": 1000.0},
                             {"p": sig, "p_mask": m, "p_mode": "additive-no-correction"}, {"mask": m > 0}, pml_size=4)
    assert int(f["p_source_many"].ravel()[0]) == 1
    rows = f["p_source_input"][0].T                     # (Npts, Nt_src)
    scale = 2 * 2e-8 / (2 * 1e-4)
    expect = [sig[L - 1] * scale / c[5, 3 + (L - 1)] for L in (2, 1, 2, 2, 1, 2)]
    assert np.allclose(rows, np.array(expect), rtol=1e-6)
    bad = m.copy()
    bad[5, 4] = 3                                        # labels 2, 3: no label 1
    with pytest.raises(ValueError, match="start from 1"):
        fe.save_to_disk_dict(kg, {"sound_speed": 1500.0, "density": 1000.0},
                             {"p": sig, "p_mask": bad * (bad != 1)}, {"mask": m > 0}, pml_size=4)
    with pytest.raises(ValueError, match="number of time series"):
        fe.save_to_disk_dict(kg, {"sound_speed": 1500.0, "density": 1000.0},
                             {"p": sig[:1].repeat(3, 0), "p_mask": m}, {"mask": m > 0}, pml_size=4)


def test_axisymmetric_front_end(kw):
    """kspaceFirstOrderASC's file: axisymmetric_flag 1, stokes absorption with
    y = 2 (inputChecking.m:165-200), a one-sided radial PML that may fill up to
    Ny (inputChecking.m:1378-1379), p0 smoothed on the WSWA-mirrored grid
    (inputChecking.m:1505-1534): smoothing a field that is already the trim of
    a mirrored band-limited one leaves its shape, a plain smooth does not."""
    fe = kw.frontend
    kg = fe.KWaveGrid(32, 1e-4, 12, 1e-4)
    kg.set_time(5, 2e-8)
    p0 = np.zeros((32, 12))
    p0[10:14, 0:3] = 1.0
    f = fe.save_to_disk_dict(kg, {"sound_speed": 1500.0, "density": 1000.0, "alpha_coeff": 0.5, "alpha_power": 1.5},
                             {"p0": p0}, {"mask": np.ones((32, 12))}, pml_size=[4, 12], axisymmetric=True)
    assert int(f["axisymmetric_flag"].ravel()[0]) == 1 and int(f["absorbing_flag"].ravel()[0]) == 2
    assert float(f["alpha_power"].ravel()[0]) == 2.0
    e = np.zeros((32, 48))
    e[:, :12] = p0
    e[:, 13:24] = -p0[:, :0:-1]
    e[:, 24:36] = -p0
    e[:, 37:] = p0[:, :0:-1]
    want = kw.geometry.smooth(e, True)[:, :12]
    assert np.allclose(f["p0_source_input"][0].T, want, atol=1e-6 * np.abs(want).max())
    plain = kw.geometry.smooth(p0, True)
    assert not np.allclose(want, plain, atol=1e-3)
    with pytest.raises(ValueError):
        fe.save_to_disk_dict(kg, {"sound_speed": 1500.0, "density": 1000.0}, {"p0": p0}, {"mask": np.ones((32, 12))},
                             pml_size=[4, 13], axisymmetric=True)


def test_frequency_ref_source_front_end(kw):
    """source.p_frequency_ref / u_frequency_ref (inputChecking.m:818-833, 921-936,
    scaleSourceTerms.m:41-42, 270-278): the signal is scaled by cos(2 pi f dt/2)
    and the mode becomes additive-no-correction; a reference frequency above
    k_max c_min / 2 pi is rejected."""
    import cases
    fk = cases.w_correction_plane_wave("p", variant="k")
    fw = cases.w_correction_plane_wave("p", variant="w")
    dt = float(fk["dt"].ravel()[0])
    assert int(fw["p_source_mode"].ravel()[0]) == 1 and int(fk["p_source_mode"].ravel()[0]) == 2
    assert np.allclose(fw["p_source_input"], fk["p_source_input"] * np.cos(np.pi * 1e6 * dt), rtol=1e-6)
    fu = cases.w_correction_plane_wave("ux", variant="w")
    assert int(fu["u_source_mode"].ravel()[0]) == 1
    with pytest.raises(ValueError, match="maximum frequency"):
        cases.w_correction_plane_wave("p", variant="w", f0=8e6)


def test_create_rejects_axisymmetric_misuse(kw):
    """axisymmetric_flag: 2-D only, stokes or no absorption, no u_non_staggered
    (checked before any device call); the radial PML may be as large as Ny."""
    f = cases.case_axisym(Nx=32, Ny=16, Nt=4, source="p0", pml=4)
    g = dict(f)
    g["pml_y_size"] = np.array([[[17]]], dtype=np.uint64)
    assert "PML" in str(_err(kw, g))
    g = dict(f)
    g["absorbing_flag"] = np.array([[[1]]], dtype=np.uint64)
    g["alpha_coeff"] = np.array([[[0.5]]], dtype=np.float32)
    g["alpha_power"] = np.array([[[1.5]]], dtype=np.float32)
    assert "stokes" in str(_err(kw, g))
    assert "u_non_staggered" in str(_err(kw, f, record=("p", "u_non_staggered")))
