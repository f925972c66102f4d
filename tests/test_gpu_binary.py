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
