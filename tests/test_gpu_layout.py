"""Row-major x-spectra with padded rows (Grid::Nxs: the row stride rounded up
to a multiple of 8 complex values, so pencil tiles move whole 64-byte
sectors).  The padding changes addressing only: results must equal the
unpadded layout bit for bit, through the single-simulation path, the
batched path and the u_non_staggered / source-spectrum passes, and the
footprint formula must still match the allocations exactly.
KW_PAD_SPECTRA=1 forces the padding on grids where it is off by default."""

import os

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

REC = ("p", "p_max", "p_rms", "p_final", "u_non_staggered", "u_max")


class pad_env:
    def __init__(self, value):
        self.value = value

    def __enter__(self):
        self.old = os.environ.get("KW_PAD_SPECTRA")
        os.environ["KW_PAD_SPECTRA"] = self.value

    def __exit__(self, *exc):
        if self.old is None:
            os.environ.pop("KW_PAD_SPECTRA", None)
        else:
            os.environ["KW_PAD_SPECTRA"] = self.old


def _run(kw, f, rec, pad):
    with pad_env(pad):
        with kw.Simulation(f, record=rec) as sim:
            cfg = sim.launch_config()
            sim.run()
            return sim.outputs(), cfg


CASES = {
    "absorbing ux additive": lambda: cases.case_3d(N=(64, 48, 32), Nt=40, source="uxyz", pml=4),
    "nonlinear p many": lambda: cases.case_3d(N=(64, 48, 32), Nt=40, nonlinear=True, source="p", many=True, pml=4),
    "p0 lossless": lambda: cases.case_3d(N=(96, 64, 48), Nt=30, het=False, absorbing=False, source=None, p0=True,
                                         pml=4),
}


@pytest.mark.parametrize("name", list(CASES))
def test_padded_rows_equal_unpadded(kw, oracle, name):
    f = CASES[name]()
    print(f"[gpu layout] {name}: padded vs unpadded row-major spectra", flush=True)
    a, ca = _run(kw, f, REC, "0")
    b, cb = _run(kw, f, REC, "1")
    assert ca["kx_major"] == 0 and cb["kx_major"] == 0
    for k in a:
        assert np.array_equal(a[k], b[k]), (k, float(np.max(np.abs(a[k] - b[k]))))
    ref = oracle.run(f, record=("p",))
    assert cases.rel_l2(b["p"], ref["p"]) <= 1e-4


def test_padded_footprint_exact(kw):
    f = cases.case_3d(N=(64, 48, 32), Nt=10, source="p", pml=4)
    rec = ("p", "p_max", "u_non_staggered")
    for pad in ("0", "1"):
        with pad_env(pad):
            fp = kw.native.footprint_bytes(f, rec)
            with kw.Simulation(f, record=rec) as sim:
                mem = sim.memory()
        print(f"[gpu layout] KW_PAD_SPECTRA={pad}: footprint {fp} B, peak {mem['peak']} B", flush=True)
        assert fp == mem["peak"]


def test_padded_batched_equals_alone(kw):
    reqs = [kw.configs.c5_request(i, grids=((64, 48, 32),), nt_range=(20, 45)) for i in range(4)]
    print("[gpu layout] batched 64x48x32 requests with padded rows", flush=True)
    with pad_env("1"):
        with kw.Server(batching=True, max_active=8, steps_per_quantum=6) as srv:
            ids = [srv.submit(f, record=("p", "p_max")) for f in reqs]
            for rid in ids:
                assert srv.wait(rid, 120.0) and srv.state(rid) == "done", srv.error(rid)
            got = [srv.outputs(rid) for rid in ids]
        alone = [kw.run(f, record=("p", "p_max")) for f in reqs]
    unpadded = [kw.run(f, record=("p", "p_max")) for f in reqs]
    for g, a, u in zip(got, alone, unpadded):
        for k in ("p", "p_max"):
            assert np.array_equal(g[k], a[k]) and np.array_equal(a[k], u[k]), k
