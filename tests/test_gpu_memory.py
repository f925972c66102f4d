"""Device-memory accounting: kw_sim_footprint_bytes (the serving loop's
admission figure) against what kw_sim_create really allocates
(kw_sim_memory), and a simulation with no sensor points."""

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,builder", [
    ("3d heterogeneous absorbing additive ux", lambda kw: cases.case_3d(N=(48, 40, 32), Nt=20, het=True)),
    ("3d homogeneous lossless p0", lambda kw: cases.case_3d(N=(32, 32, 32), Nt=20, het=False, absorbing=False,
                                                             source=None, p0=True)),
    ("3d heterogeneous many p source", lambda kw: cases.case_3d(N=(32, 28, 24), Nt=20, source="p", many=True)),
    ("c5 request (c, rho, alpha heterogeneous)",
     lambda kw: kw.configs.c5_request(2, grids=((64, 48, 40),), nt_range=(20, 20))),
])
def test_footprint_covers_allocations(kw, name, builder):
    print(f"[gpu memory] {name}", flush=True)
    f = builder(kw)
    rec = ("p", "p_max", "p_rms")
    fp = kw.native.footprint_bytes(f, rec)
    with kw.Simulation(f, record=rec) as sim:
        mem = sim.memory()
        sim.run(5)
        sim.sync()
        assert sim.memory() == mem  # the time loop allocates nothing
    print(f"    footprint {fp} B, peak {mem['peak']} B, held {mem['allocated']} B", flush=True)
    assert mem["allocated"] <= mem["peak"] <= fp
    # the formula follows the same allocation sequence: exact, not a loose bound
    assert fp == mem["peak"], (fp, mem)


def test_no_sensor_points_records_p_final(kw, oracle):
    f = cases.case_3d(N=(32, 28, 24), Nt=30)
    f = dict(f)
    f["sensor_mask_index"] = np.zeros(0, dtype=np.uint64)
    print("[gpu memory] no sensor points, p_final + p_max_all", flush=True)
    got = kw.run(f, record=("p_final", "p_max_all"))
    ref = oracle.run(f, record=("p_final", "p_max_all"))
    for k in ("p_final", "p_max_all"):
        assert cases.rel_l2(got[k], ref[k]) <= 1e-4, k


def test_time_series_start_zeroed(kw):
    """Recorded time series are zero-filled at creation, so a fetch before the
    last step returns zeros for the steps still to come -- also when the memory
    was just released by another simulation of the same size."""
    f = cases.case_3d(N=(32, 28, 24), Nt=30, source="p")
    rec = ("p", "u")
    names = ("p", "ux", "uy", "uz")
    print("[gpu memory] partial runs: rows of steps not yet run read zero", flush=True)
    with kw.Simulation(f, record=rec) as sim:
        sim.run(30)
        sim.sync()
        full = {k: sim.fetch(k) for k in names}
    for _ in range(2):
        with kw.Simulation(f, record=rec) as sim:
            sim.run(20)
            sim.sync()
            for k in names:
                got = sim.fetch(k)  # (recorded steps, sensor points)
                assert got.shape == full[k].shape and got.shape[0] == 30, (k, got.shape)
                assert np.array_equal(got[:20], full[k][:20]), k
                assert not np.any(got[20:]), k
