"""GPU tests of the file-level drop-in: kspaceFirstOrder-HIP run as the MATLAB
wrapper runs it (process boundary, LD_LIBRARY_PATH cleared, HDF5 in/out)."""

import os

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def test_binary_matches_in_process_abi_and_oracle(kw, oracle, tmp_path):
    from kwave_amd import h5format

    print("[gpu binary] 3d absorbing ux additive via kspaceFirstOrder-HIP", flush=True)
    f = cases.case_3d(Nt=40)
    inp, out = str(tmp_path / "in.h5"), str(tmp_path / "out.h5")
    h5format.write_input(inp, f)
    rec = ("p", "p_max", "p_rms", "u_final", "p_final", "u_non_staggered")
    r = h5format.run_binary(inp, out, rec, record_start_index=5)
    assert r.returncode == 0, r.stderr
    o = h5format.read_file(out)
    for k in ("Nx", "Ny", "Nz", "pml_x_size", "pml_y_size", "pml_z_size"):
        assert o[k].dtype == np.uint64 and o[k].shape == (1, 1, 1)
    with h5format.H5File(out) as h:
        assert h.attribute("file_type") == "output"
    ns = np.asarray(f["sensor_mask_index"]).size
    assert o["p"].shape == (1, 40 - 5 + 1, ns)
    assert o["p_max"].shape == (1, 1, ns)
    assert o["p_final"].shape == (24, 28, 32)
    assert o["uz_non_staggered"].shape == (1, 40 - 5 + 1, ns)   # /u?_non_staggered (3DC.m:706-711)
    direct = kw.run(f, record=rec, record_start_index=5)
    for k in direct:
        np.testing.assert_array_equal(o[k].reshape(direct[k].shape), direct[k])  # same kernels: bit-exact
    ref = oracle.run(f, record=rec, record_start_index=5)
    for k in ref:
        assert cases.rel_l2(o[k].reshape(ref[k].shape), ref[k]) < 1e-4, k


def test_in_process_run_file(kw, tmp_path):
    from kwave_amd import h5format

    print("[gpu binary] kw_run_file 2d", flush=True)
    f = cases.case_2d_absorbing_psource(Nt=40)
    inp, out = str(tmp_path / "in.h5"), str(tmp_path / "out.h5")
    h5format.write_input(inp, f)
    opts = h5format.run_file(inp, out, ("p", "u"))
    assert opts.t_loop > 0
    o = h5format.read_file(out)
    direct = kw.run(f, record=("p", "u"))
    for k in direct:
        np.testing.assert_array_equal(o[k].reshape(direct[k].shape), direct[k])


def test_matlab_style_front_end(kw, oracle):
    """kspaceFirstOrder2DG flow: SaveToDisk -> binary -> sensor_data (MATLAB shapes)."""
    from kwave_amd import h5format

    print("[gpu binary] matlab-style 2d p0", flush=True)
    fe = kw.frontend
    kgrid = fe.KWaveGrid(64, 1e-4, 64, 1e-4)
    kgrid.set_time(50, 2e-8)
    p0 = 5.0 * kw.geometry.make_disc(64, 64, 30, 34, 4)
    mask = np.zeros((64, 64))
    mask[20, 10:50] = 1
    medium = {"sound_speed": 1500.0, "density": 1000.0}
    sd = h5format.kspace_first_order(kgrid, medium, {"p0": p0}, {"mask": mask}, record=("p", "p_final"),
                                     pml_size=10)
    assert sd["p"].shape == (40, 50)            # [Ns, Nt]
    assert sd["p_final"].shape == (64, 64)      # MATLAB [Nx, Ny]
    f = fe.save_to_disk_dict(kgrid, medium, {"p0": p0}, {"mask": mask}, pml_size=10)
    ref = oracle.run(f, record=("p", "p_final"))
    assert cases.rel_l2(sd["p"], ref["p"].T) < 1e-4
    assert cases.rel_l2(sd["p_final"], ref["p_final"][0].T) < 1e-4


@pytest.mark.parametrize("dim", [3, 2, "as"])
def test_cuboid_corners_equal_binary_mask(kw, tmp_path, dim):
    """Cuboid-corner sensors through the file path: /p/<k> etc. hold, per cuboid and in
    [cx, cy, cz, Nt] order, exactly the binary-mask run's values at those points
    (the reference's binary == cuboid invariant)."""
    from kwave_amd import h5format

    print(f"[gpu binary] cuboid corners vs binary mask, {dim}-D", flush=True)
    if dim == 3:
        f = cases.case_3d(N=(32, 28, 24), Nt=30)
        shape = (32, 28, 24)
        corners = np.array([[5, 4, 3, 12, 9, 7], [10, 8, 6, 20, 14, 10]]).T      # overlapping
    elif dim == 2:
        f = cases.case_2d_absorbing_psource(N=(60, 48), Nt=40)
        shape = (60, 48)
        corners = np.array([[10, 12, 30, 20], [25, 5, 40, 15]]).T
    else:
        # kspaceFirstOrderAS_compare_binary_and_cuboid_sensor_mask.m: the same on an
        # axisymmetric file, one cuboid touching the axis
        f = cases.case_axisym(Nx=64, Ny=48, Nt=40, source="p", het=True)
        shape = (64, 48)
        corners = np.array([[10, 1, 30, 12], [25, 5, 40, 15]]).T
        dim = 2
    rec = ("p", "p_max", "p_rms", "u", "p_final")
    fc = dict(f)
    fc.pop("sensor_mask_index")
    # the front end's corner layout (saveToDisk.m:355-378)
    c6 = corners if dim == 3 else np.vstack([corners[0], corners[1], np.ones(2, int),
                                             corners[2], corners[3], np.ones(2, int)])
    fc["sensor_mask_type"] = np.array([[[1]]], dtype=np.uint64)
    fc["sensor_mask_corners"] = np.ascontiguousarray(c6.T[None].astype(np.uint64))
    mask = np.zeros(shape)
    for k in range(corners.shape[1]):
        c = c6[:, k] - 1
        if dim == 3:
            mask[c[0]:c[3] + 1, c[1]:c[4] + 1, c[2]:c[5] + 1] = 1
        else:
            mask[c[0]:c[3] + 1, c[1]:c[4] + 1] = 1
    fb = dict(f)
    idx = np.flatnonzero(mask.reshape(-1, order="F")) + 1
    fb["sensor_mask_index"] = idx.astype(np.uint64).reshape(1, 1, -1)
    outs = {}
    for tag, ff in (("cub", fc), ("bin", fb)):
        inp, out = str(tmp_path / f"{tag}.h5"), str(tmp_path / f"{tag}_out.h5")
        h5format.write_input(inp, ff)
        r = h5format.run_binary(inp, out, rec)
        assert r.returncode == 0, r.stderr
        outs[tag] = h5format.read_file(out)
    ob, oc = outs["bin"], outs["cub"]
    nt = int(np.ravel(f["Nt"])[0])
    Nx, Ny = shape[0], shape[1]
    for k in range(corners.shape[1]):
        lo = c6[:3, k]
        hi = c6[3:, k]
        xs, ys, zs = (np.arange(lo[a], hi[a] + 1) for a in range(3))
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")               # (cz, cy, cx), x fastest
        lin = X + Nx * ((Y - 1) + Ny * (Z - 1))
        col = np.searchsorted(idx, lin.reshape(-1))
        cshape = (len(zs), len(ys), len(xs))
        for name in ("p", "ux", "uy") + (("uz",) if dim == 3 else ()):
            got = oc[f"{name}/{k + 1}"]
            assert got.shape == (nt,) + cshape, (name, got.shape)
            want = ob[name].reshape(nt, -1)[:, col].reshape((nt,) + cshape)
            np.testing.assert_array_equal(got, want)
        for name in ("p_max", "p_rms"):
            got = oc[f"{name}/{k + 1}"]
            assert got.shape == cshape
            np.testing.assert_array_equal(got, ob[name].reshape(-1)[col].reshape(cshape))
    np.testing.assert_array_equal(oc["p_final"], ob["p_final"])


def test_time_reversal_through_the_contract(kw, oracle, tmp_path):
    """Time reversal as kspaceFirstOrder3DC.m:437-460 rewrites it: the recorded sensor
    data, flipped in time, drives a dirichlet p source on the sensor mask, and the
    binary records p_final over the whole grid (the reconstruction)."""
    from kwave_amd import h5format

    print("[gpu binary] time reversal via dirichlet source + p_final", flush=True)
    fe, geo = kw.frontend, kw.geometry
    N, d = 96, 1e-4
    kg = fe.KWaveGrid(N, d, N, d)
    kg.set_time(260, 0.3 * d / 1500.0)
    p0 = 2.0 * geo.make_disc(N, N, 40, 52, 6)
    X, Y = np.meshgrid(np.arange(N) - N // 2, np.arange(N) - N // 2, indexing="ij")
    ring = (np.abs(np.sqrt(X ** 2 + Y ** 2) - 34) < 0.5).astype(float)
    med = {"sound_speed": 1500.0, "density": 1000.0}
    fwd = fe.save_to_disk_dict(kg, med, {"p0": p0}, {"mask": ring}, pml_size=10)
    rec = kw.run(fwd, record=("p",))["p"]                     # (Nt, Ns)
    tr = fe.save_to_disk_dict(kg, med, {"p": rec.T[:, ::-1], "p_mask": ring, "p_mode": "dirichlet"},
                              {"mask": np.ones((N, N))}, pml_size=10)
    inp, out = str(tmp_path / "tr.h5"), str(tmp_path / "tr_out.h5")
    h5format.write_input(inp, tr)
    r = h5format.run_binary(inp, out, ("p_final",))
    assert r.returncode == 0, r.stderr
    got = h5format.read_file(out)["p_final"].reshape(N, N)      # (y, x)
    ref = oracle.run(tr, record=("p_final",))["p_final"].reshape(N, N)
    e = cases.rel_l2(got, ref)
    print(f"    p_final rel L2 vs oracle {e:.2e}", flush=True)
    assert e < 1e-4
    # the reconstruction resembles the source distribution
    p0s = np.asarray(fwd["p0_source_input"], dtype=np.float64).reshape(N, N)
    inside = (np.sqrt(X ** 2 + Y ** 2) < 30).T
    cc = np.corrcoef(got[inside], p0s[inside])[0, 1]
    print(f"    correlation with p0 inside the ring {cc:.2f}", flush=True)
    assert cc > 0.5


def test_axisymmetric_time_reversal_through_the_contract(kw, oracle, tmp_path):
    """kspaceFirstOrderAS_time_reversal.m through the file contract: an on-axis p0
    recorded on a closed axisymmetric surface (a cylinder at r = 34 with end caps),
    then the flipped recordings as a dirichlet p source and p_final through the
    executable on an axisymmetric file; p_final matches the oracle and
    reconstructs the (smoothed) p0."""
    from kwave_amd import h5format

    print("[gpu binary] axisymmetric time reversal", flush=True)
    fe = kw.frontend
    Nx, Ny, d = 96, 48, 1e-4
    kg = fe.KWaveGrid(Nx, d, Ny, d)
    kg.set_time(300, 0.3 * d / 1500.0)
    x = (np.arange(Nx) - 40)[:, None]
    r = np.arange(Ny)[None, :]
    p0 = 2.0 * ((x ** 2 + r ** 2) <= 36).astype(float)
    sens = np.zeros((Nx, Ny))
    sens[10:Nx - 10, 34] = 1
    sens[10, :35] = 1
    sens[Nx - 11, :35] = 1
    med = {"sound_speed": 1500.0, "density": 1000.0}
    fwd = fe.save_to_disk_dict(kg, med, {"p0": p0}, {"mask": sens}, pml_size=[8, 8], axisymmetric=True)
    rec = kw.run(fwd, record=("p",))["p"]
    tr = fe.save_to_disk_dict(kg, med, {"p": rec.T[:, ::-1], "p_mask": sens, "p_mode": "dirichlet"},
                              {"mask": np.ones((Nx, Ny))}, pml_size=[8, 8], axisymmetric=True)
    inp, out = str(tmp_path / "tr.h5"), str(tmp_path / "tr_out.h5")
    h5format.write_input(inp, tr)
    res = h5format.run_binary(inp, out, ("p_final",))
    assert res.returncode == 0, res.stderr
    got = h5format.read_file(out)["p_final"].reshape(Ny, Nx)
    ref = oracle.run(tr, record=("p_final",))["p_final"].reshape(Ny, Nx)
    e = cases.rel_l2(got, ref)
    print(f"    p_final rel L2 vs oracle {e:.2e}", flush=True)
    assert e < 1e-4
    p0s = np.asarray(fwd["p0_source_input"], dtype=np.float64).reshape(Ny, Nx)
    inside = np.zeros((Ny, Nx), bool)
    inside[:30, 15:Nx - 15] = True
    cc = np.corrcoef(got[inside], p0s[inside])[0, 1]
    print(f"    correlation with p0 inside the surface {cc:.4f}", flush=True)
    assert cc > 0.99


@pytest.mark.parametrize("how,source", [("abi", "p_many"), ("binary", "p_many"), ("abi", "transducer"),
                                        ("binary", "transducer")])
def test_checkpoint_restart_is_exact(kw, tmp_path, how, source):
    """A run stopped at checkpoints and resumed (twice) gives bit-identical outputs
    to the uninterrupted run; the finished run removes the checkpoint.  The
    transducer case restarts while its delayed signal is still being applied
    (the reference binary's release notes record a wrong time-signal index on
    restart for this source, k-wave_release_notes.html:215)."""
    from kwave_amd import h5format

    print(f"[gpu binary] checkpoint-restart via {how}, {source} source", flush=True)
    if source == "transducer":
        f = cases.case_transducer(N=(32, 32, 32), Nt=50)
        assert int(np.asarray(f["transducer_source_flag"]).reshape(-1)[0]) > 40
    else:
        f = cases.case_3d(N=(32, 28, 24), Nt=50, nonlinear=True, source="p", many=True)
    rec = ("p", "p_max", "p_rms", "u_max", "p_final", "u_non_staggered")
    inp = str(tmp_path / "in.h5")
    h5format.write_input(inp, f)
    ref_out, out, ck = str(tmp_path / "ref.h5"), str(tmp_path / "out.h5"), str(tmp_path / "ck.h5")
    h5format.run_file(inp, ref_out, rec)
    stops = 0
    for _ in range(4):
        if how == "abi":
            o = h5format.run_file(inp, out, rec, checkpoint_file=ck, checkpoint_timesteps=20)
            stopped = o.checkpointed == 1
        else:
            r = h5format.run_binary(inp, out, rec, extra=["--checkpoint_file", ck, "--checkpoint_timesteps", "20"])
            assert r.returncode == 0, r.stderr
            stopped = "checkpoint saved" in r.stdout
        if not stopped:
            break
        stops += 1
        assert os.path.exists(ck) and not os.path.exists(out)
        with h5format.H5File(ck) as h:
            assert h.attribute("file_type") == "checkpoint"
    assert stops == 2                       # 50 steps in calls of 20: stop at 20, 40, finish
    assert not os.path.exists(ck)
    a, b = h5format.read_file(ref_out), h5format.read_file(out)
    assert sorted(a) == sorted(b)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_transducer_through_the_file(kw, tmp_path):
    """transducer_source_flag / transducer_source_input / delay_mask read from the
    input file give the in-process result bit for bit."""
    from kwave_amd import h5format

    print("[gpu binary] transducer source via kspaceFirstOrder-HIP", flush=True)
    f = cases.case_transducer(Nt=40)
    inp, out = str(tmp_path / "in.h5"), str(tmp_path / "out.h5")
    h5format.write_input(inp, f)
    r = h5format.run_binary(inp, out, ("p", "u_final"))
    assert r.returncode == 0, r.stderr
    o = h5format.read_file(out)
    direct = kw.run(f, record=("p", "u_final"))
    for k in direct:
        np.testing.assert_array_equal(o[k].reshape(direct[k].shape), direct[k])


def test_axisymmetric_through_the_binary(kw, oracle, tmp_path):
    """kspaceFirstOrderASC's file (axisymmetric_flag = 1) through the executable:
    bit-equal to the in-process run, within 1e-4 of the oracle; and the
    kspace_first_order(axisymmetric=True) flow returns MATLAB-shaped data."""
    from kwave_amd import h5format

    print("[gpu binary] axisymmetric u additive stokes via kspaceFirstOrder-HIP", flush=True)
    f = cases.case_axisym(Nx=64, Ny=48, Nt=50, source="u", het=True, absorbing=True)
    inp, out = str(tmp_path / "in.h5"), str(tmp_path / "out.h5")
    h5format.write_input(inp, f)
    rec = ("p", "p_max", "u_final", "p_final")
    r = h5format.run_binary(inp, out, rec)
    assert r.returncode == 0, r.stderr
    o = h5format.read_file(out)
    direct = kw.run(f, record=rec)
    for k in direct:
        np.testing.assert_array_equal(o[k].reshape(direct[k].shape), direct[k])
    ref = oracle.run(f, record=rec)
    for k in ref:
        assert cases.rel_l2(o[k].reshape(ref[k].shape), ref[k]) < 1e-4, k
    fe = cases.fe
    kg = fe.KWaveGrid(64, 1e-4, 32, 1e-4)
    kg.set_time(40, 2e-8)
    p0 = np.zeros((64, 32))
    p0[30:34, :3] = 1.0
    sensor = np.zeros((64, 32))
    sensor[:, 0] = 1
    sd = h5format.kspace_first_order(kg, {"sound_speed": 1500.0, "density": 1000.0}, {"p0": p0}, {"mask": sensor},
                                     record=("p",), pml_size=8, axisymmetric=True)
    assert sd["p"].shape == (64, 40)
    fa = fe.save_to_disk_dict(kg, {"sound_speed": 1500.0, "density": 1000.0}, {"p0": p0}, {"mask": sensor},
                              pml_size=8, axisymmetric=True)
    assert cases.rel_l2(sd["p"].T, oracle.run(fa, record=("p",))["p"]) < 1e-4


def test_wrapper_options_accepted(kw, tmp_path):
    """The options kspaceFirstOrder3DC.m passes (3DC.m:308-524: -t, -g, --verbose) and
    deflate-compressed output (-c) leave the results unchanged."""
    from kwave_amd import h5format

    print("[gpu binary] -t / -g / --verbose / -c", flush=True)
    f = cases.case_2d_absorbing_psource(N=(60, 48), Nt=30)
    inp = str(tmp_path / "in.h5")
    h5format.write_input(inp, f)
    outs = {}
    for tag, extra in (("plain", ()), ("opts", ("-t", "8", "--verbose", "2", "-c", "4"))):
        out = str(tmp_path / f"{tag}.h5")
        r = h5format.run_binary(inp, out, ("p", "p_max", "u_final"), device=0, extra=extra)
        assert r.returncode == 0, r.stderr
        outs[tag] = h5format.read_file(out)
    assert outs["opts"].keys() == outs["plain"].keys()
    for k in outs["plain"]:
        np.testing.assert_array_equal(outs["opts"][k], outs["plain"][k])


def test_intensity_through_the_front_end_flow(kw):
    """sensor.record = {'I_avg'} / {'p', 'I'} in the kspaceFirstOrder2DG-style flow:
    the binary records p and u_non_staggered (3DC.m:508-511), the host forms the
    intensity (saveIntensity.m), and the plane wave's time-averaged intensity is
    the source's 10 W/cm^2 within 1 % (kspaceFirstOrder_check_plane_wave_intensity.m)."""
    from kwave_amd import h5format

    kg, medium, source, sensor, pml, start = cases.plane_wave_intensity_objects()
    print("[gpu binary] plane-wave intensity through kspace_first_order", flush=True)
    sd = h5format.kspace_first_order(kg, medium, source, sensor, record=("I_avg",), record_start_index=start, **pml)
    assert set(sd) == {"Ix_avg", "Iy_avg"}, sorted(sd)
    err = 100 * abs(sd["Ix_avg"][0] / 1e4 - cases.PLANE_WAVE_I0) / cases.PLANE_WAVE_I0
    print(f"    Ix_avg error {err:.3f} %", flush=True)
    assert err < 1.0
    assert abs(sd["Iy_avg"][0]) <= 1e-5 * abs(sd["Ix_avg"][0])
    full = h5format.kspace_first_order(kg, medium, source, sensor, record=("p", "I"), record_start_index=start, **pml)
    assert set(full) == {"p", "Ix", "Iy"}, sorted(full)
    assert full["Ix"].shape == full["p"].shape == (1, kg.Nt - start + 1)
    assert np.isclose(full["Ix"].mean(), sd["Ix_avg"][0], rtol=1e-12)
