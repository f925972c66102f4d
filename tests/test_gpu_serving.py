"""Serving loop (kw_server_*): results in the mix equal results alone, HBM
admission, failure isolation.

Requests are drawn from the serving-stream generator (configs.c5_request)
on reduced grids and Nt so the whole file runs in seconds; each result is
compared bit for bit with the same request run alone through kw_sim_*, and
one request with the FP64 oracle (relative L2 <= 1e-4).
"""

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

SMALL_GRIDS = ((32, 32, 32), (48, 32, 40), (64, 48, 32), (40, 40, 48))


def _requests(kw, n, nt=(24, 60)):
    return [kw.configs.c5_request(i, grids=SMALL_GRIDS, nt_range=nt) for i in range(n)]


def _alone(kw, f):
    return kw.run(f, record=("p", "p_max"))


def test_serving_results_equal_alone(kw):
    reqs = _requests(kw, 8)
    with kw.Server(max_active=4, steps_per_quantum=7) as srv:
        ids = [srv.submit(f, record=("p", "p_max")) for f in reqs]
        for rid in ids:
            assert srv.wait(rid, timeout=120.0)
            assert srv.state(rid) == "done", srv.error(rid)
        for f, rid in zip(reqs, ids):
            got = srv.outputs(rid)
            ref = _alone(kw, f)
            for name in ("p", "p_max"):
                assert got[name].shape == ref[name].shape
                assert np.array_equal(got[name], ref[name]), (rid, name)
            info = srv.info(rid)
            assert info["t_submit"] <= info["t_admit"] <= info["t_done"]
            assert info["point_updates"] == int(np.prod([int(np.ravel(f[k])[0]) for k in ("Nx", "Ny", "Nz", "Nt")]))
            srv.release(rid)
            assert srv.state(rid) == "unknown"
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0 and st["active"] == 0
        assert st["hbm_reserved_bytes"] == 0


def test_serving_request_matches_oracle(kw, oracle):
    f = kw.configs.c5_request(3, grids=((48, 40, 32),), nt_range=(60, 60))
    ref = oracle.run(f, record=("p",))
    with kw.Server(max_active=2) as srv:
        rid = srv.submit(f, record=("p",))
        assert srv.wait(rid, 120.0) and srv.state(rid) == "done", srv.error(rid)
        got = srv.fetch(rid, "p")
    assert cases.rel_l2(got, ref["p"]) <= 1e-4


def test_admission_respects_hbm_budget(kw):
    reqs = _requests(kw, 5)
    fps = [kw.native.footprint_bytes(f, ("p", "p_max")) for f in reqs]
    budget = max(fps) + min(fps) // 2          # never room for the two largest together
    with kw.Server(hbm_budget_bytes=budget, max_active=8, steps_per_quantum=5) as srv:
        ids = [srv.submit(f, record=("p", "p_max")) for f in reqs]
        for rid in ids:
            assert srv.wait(rid, 120.0) and srv.state(rid) == "done", srv.error(rid)
        st = srv.stats()
        assert st["hbm_peak_reserved_bytes"] <= budget
        infos = [srv.info(r) for r in ids]
        # admission is in arrival order
        admits = [i["t_admit"] for i in infos]
        assert admits == sorted(admits)
        # any two requests resident at once fit the budget together
        for a in range(len(ids)):
            for b in range(a + 1, len(ids)):
                ia, ib = infos[a], infos[b]
                overlap = ia["t_admit"] < ib["t_done"] and ib["t_admit"] < ia["t_done"]
                if overlap:
                    assert ia["footprint_bytes"] + ib["footprint_bytes"] <= budget
        for f, rid in zip(reqs, ids):
            assert np.array_equal(srv.fetch(rid, "p"), _alone(kw, f)["p"])


def test_oversize_request_rejected_at_submit(kw):
    f = _requests(kw, 1)[0]
    with kw.Server(hbm_budget_bytes=1 << 20) as srv:
        with pytest.raises(kw.native.KWaveError) as e:
            srv.submit(f)
        assert e.value.status == 4


def test_failed_request_is_isolated(kw):
    good = _requests(kw, 2)
    bad = cases.case_2d_p0_lossless(N=44, Nt=5)    # 44 = 4 * 11: unsupported prime factor
    with kw.Server(max_active=3) as srv:
        i0 = srv.submit(good[0], record=("p",))
        ib = srv.submit(bad, record=("p",))
        i1 = srv.submit(good[1], record=("p",))
        for rid in (i0, ib, i1):
            assert srv.wait(rid, 120.0)
        assert srv.state(ib) == "failed" and "factor" in srv.error(ib)
        assert srv.info(ib)["status"] == 2
        for f, rid in ((good[0], i0), (good[1], i1)):
            assert srv.state(rid) == "done"
            assert np.array_equal(srv.fetch(rid, "p"), kw.run(f, record=("p",))["p"])
        assert srv.stats()["failed"] == 1
