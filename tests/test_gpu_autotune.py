"""Launch-geometry autotuner (kw_sim_autotune).

Workgroup sizes do not change the formulas, but each instantiation may fuse
multiply-adds differently, so a tuned run matches the default one to FP32
rounding (relative L2 1e-5 here), not bit for bit.  Tuning resets the
simulation to step 0: a simulation that had already run steps and is then
tuned (cached configuration) reproduces a freshly tuned one bit for bit.
"""

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(64, 64, 64), (48, 40, 32), (64, 96)])
def test_autotune_resets_and_matches_default(kw, shape):
    if len(shape) == 3:
        f = cases.case_3d(N=shape, Nt=30, pml=4, source="uxyz", nonlinear=True)
    else:
        f = cases.case_2d_absorbing_psource(N=shape, Nt=40)
    rec = ("p", "p_max", "p_rms", "u_final")
    ref = kw.run(f, record=rec)
    with kw.Simulation(f, record=rec) as sim:
        cfg = sim.autotune(trial_steps=3)
        print(f"[gpu autotune] {shape}: {cfg}", flush=True)
        assert sim.steps_done == 0
        sim.run()
        a = sim.outputs()
    with kw.Simulation(f, record=rec) as sim:
        sim.run(7)
        sim.sync()
        assert sim.autotune(trial_steps=3) == cfg       # cached for this shape
        assert sim.steps_done == 0
        sim.run()
        b = sim.outputs()
    for k in ref:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
        assert cases.rel_l2(a[k], ref[k]) < 1e-5, k


def test_server_with_autotune(kw):
    reqs = [kw.configs.c5_request(i, grids=((32, 32, 32), (48, 32, 40)), nt_range=(20, 40)) for i in range(4)]
    with kw.Server(max_active=2, autotune=True) as srv:
        ids = [srv.submit(f) for f in reqs]
        for f, rid in zip(reqs, ids):
            assert srv.wait(rid, 120.0) and srv.state(rid) == "done", srv.error(rid)
            assert cases.rel_l2(srv.fetch(rid, "p"), kw.run(f)["p"]) < 1e-5


@pytest.mark.parametrize("shape", [(48, 160, 192), (32, 192, 160)])
def test_radix_plans_agree(kw, oracle, monkeypatch, shape):
    """Pencil radix plan 1 (radix 8 at most) against plan 0 and the oracle."""
    f = cases.case_3d(N=shape, Nt=24, pml=4, source="p")
    rec = ("p", "p_final")
    ref = oracle.run(f, record=rec)
    got = {}
    for plan in (0, 1):
        monkeypatch.setenv("KW_YPASS_PLAN", str(plan))
        monkeypatch.setenv("KW_LAST_PLAN", str(plan))
        with kw.Simulation(f, record=rec) as sim:
            cfg = sim.launch_config()
            print(f"[gpu autotune] {shape} plan {plan}: {cfg}", flush=True)
            assert cfg["kx_major"] == 0 and cfg["ypass_plan"] == plan and cfg["last_plan"] == plan
            sim.run()
            got[plan] = sim.outputs()
        for k in rec:
            assert cases.rel_l2(got[plan][k], ref[k]) <= 1e-4, (plan, k)
    for k in rec:
        assert cases.rel_l2(got[1][k], got[0][k]) < 1e-5, k


def test_autotune_reports_plans(kw):
    f = cases.case_3d(N=(32, 160, 192), Nt=12, pml=4, source="p")
    with kw.Simulation(f, record=("p",)) as sim:
        before = sim.launch_config()
        sim.autotune(trial_steps=2)
        after = sim.launch_config()
    print(f"[gpu autotune] default {before} tuned {after}", flush=True)
    assert before["ypass_plan"] == 0 and before["last_plan"] == 0
    assert after["ypass_plan"] in (0, 1) and after["last_plan"] in (0, 1)
    assert after["ypass_nt"] in (128, 256, 512) and after["last_nt"] in (128, 256, 512, 1024)
    # the row kernels' workgroup target (pairs per workgroup) is a tuned launch shape too
    assert before["row_wg_target"] == 1024 and after["row_wg_target"] in (512, 1024, 2048, 4096)


def test_row_workgroup_target_leaves_results_unchanged(kw, monkeypatch):
    """Row pairs per workgroup (through the workgroup target) change the launch
    geometry only: a row's transform is the same, so results are bit-identical."""
    f = cases.case_3d(N=(64, 48, 32), Nt=20, pml=4, source="uxyz")
    outs = []
    for wgs in ("256", "1024", "8192"):
        monkeypatch.setenv("KW_ROW_WGS", wgs)
        with kw.Simulation(f, record=("p", "u_max")) as sim:
            assert sim.launch_config()["row_wg_target"] == int(wgs)
            sim.run()
            outs.append(sim.outputs())
    for o in outs[1:]:
        for k in o:
            assert np.array_equal(o[k], outs[0][k]), k
