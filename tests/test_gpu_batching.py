"""Batched serving (kw_server_config.batching): simulations of one grid shape
and physics branch advance through shared kernel launches (one simulation
per blockIdx.z) replayed from captured hipGraphs.  Every result must equal
the same request run alone through kw_sim_* bit for bit, whatever it shares
a launch with, however many quanta it spans and when it joins or leaves.

The reference runs one simulation per binary launch
(kspaceFirstOrder3DC.m:544; kWaveTester.m:1936-1950 loops over them), so
"alone" is the reference behaviour these results must reproduce.
"""

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

REC = ("p", "p_max", "p_min", "p_rms", "p_final", "u", "u_max")


def _alone(kw, f, rec, start):
    return kw.run(f, record=rec, record_start_index=start)


def _serve(kw, items, **cfg):
    """items: (inputs, record, start); returns each request's outputs."""
    with kw.Server(batching=True, **cfg) as srv:
        ids = [srv.submit(f, record=rec, record_start_index=start) for f, rec, start in items]
        for rid in ids:
            assert srv.wait(rid, timeout=300.0)
            assert srv.state(rid) == "done", srv.error(rid)
        outs = [srv.outputs(rid) for rid in ids]
        st = srv.stats()
        assert st["completed"] == len(items) and st["failed"] == 0 and st["hbm_reserved_bytes"] == 0
    return outs


def _check_equal(kw, items, outs):
    for n, ((f, rec, start), got) in enumerate(zip(items, outs)):
        ref = _alone(kw, f, rec, start)
        assert got.keys() == ref.keys(), n
        for k in ref:
            assert got[k].shape == ref[k].shape, (n, k)
            assert np.array_equal(got[k], ref[k]), (n, k, float(np.max(np.abs(got[k] - ref[k]))))


def test_batched_mixed_requests_equal_alone(kw):
    """Two batchable shapes (a kx-major plane-pass grid and a row-major pencil
    grid) with uneven Nt, plus requests that run unbatched in the same server
    (a last axis without specialised kernels), a p0 start, dirichlet and
    additive-no-correction sources, a three-component velocity source, a
    steered transducer, a nonlinear medium, and record start indices past 1."""
    items = []
    for i in range(10):
        f = kw.configs.c5_request(i, grids=((32, 32, 32), (64, 48, 32)), nt_range=(20, 70))
        items.append((f, ("p", "p_max", "p_rms"), 1 + (i % 3) * 5))
    items.append((kw.configs.c5_request(20, grids=((48, 32, 40),), nt_range=(30, 30)), ("p",), 1))
    items.append((cases.case_3d(N=(32, 32, 32), Nt=45, het=False, absorbing=False, source=None, p0=True), REC, 1))
    items.append((cases.case_3d(N=(32, 32, 32), Nt=40, source="p", mode="dirichlet"), REC, 3))
    items.append((cases.case_3d(N=(64, 48, 32), Nt=50, source="ux", mode="additive-no-correction"), REC, 1))
    items.append((cases.case_3d(N=(64, 48, 32), Nt=33, source="uxyz", mode="additive"), REC, 2))
    items.append((cases.case_transducer(N=(32, 32, 32), Nt=40), REC, 1))
    items.append((cases.case_3d(N=(32, 32, 32), Nt=38, nonlinear=True, source="p"), REC, 1))
    items.append((cases.case_3d(N=(32, 32, 32), Nt=38, nonlinear=True, source="p", seed=9), REC, 1))
    print(f"[batching] {len(items)} mixed requests, quantum 7", flush=True)
    outs = _serve(kw, items, max_active=32, steps_per_quantum=7)
    _check_equal(kw, items, outs)


def test_batched_joins_and_leaves(kw):
    """More requests than slots: members leave mid-quantum and new ones join
    the running batch at quantum boundaries (tables rebuilt, graph recaptured)."""
    items = [(kw.configs.c5_request(i, grids=((32, 32, 32),), nt_range=(10, 60)), ("p", "p_max"), 1)
             for i in range(12)]
    print("[batching] 12 requests of 32^3 through 5 slots, quantum 4", flush=True)
    outs = _serve(kw, items, max_active=5, steps_per_quantum=4)
    _check_equal(kw, items, outs)


@pytest.mark.parametrize("grid", [(64, 64, 64), (96, 96, 96), (128, 128, 128), (160, 160, 192)])
def test_c5_grid_batched_equals_alone_and_oracle(kw, oracle, grid):
    """The serving stream's own grid shapes (the 64^3, 96^3 and 128^3 plane
    pass, the 160x160x192 pencils), two requests per batch: bit-identical to
    alone, and within relative L2 1e-4 of the FP64 oracle."""
    reqs = [kw.configs.c5_request(i, grids=(grid,), Nt=24) for i in (100, 101)]
    items = [(f, ("p",), 1) for f in reqs]
    print(f"[batching] C5 grid {grid}: 2 requests, Nt 24", flush=True)
    outs = _serve(kw, items, max_active=4, steps_per_quantum=8)
    _check_equal(kw, items, outs)
    ref = oracle.run(reqs[0], record=("p",))["p"]
    err = cases.rel_l2(outs[0]["p"], ref)
    mx = float(np.max(np.abs(outs[0]["p"].astype(np.float64) - ref)))
    print(f"    vs FP64 oracle: rel L2 {err:.3e}, max abs {mx:.3e} (peak {float(np.max(np.abs(ref))):.3e})", flush=True)
    assert err <= 1e-4


def test_batched_768_point_last_axis_equals_alone(kw):
    """A 768-point last axis runs 1024-thread, 16-pencil tiles (the CU's whole
    LDS) in the batched kernels too: two requests per launch, bit-identical to
    each run alone."""
    items = [(cases.case_3d(N=(32, 48, 768), Nt=14, source="p", pml=4, seed=s), ("p", "p_max"), 1) for s in (3, 4)]
    print("[batching] 2 requests of 32x48x768, quantum 5", flush=True)
    outs = _serve(kw, items, max_active=4, steps_per_quantum=5)
    _check_equal(kw, items, outs)
