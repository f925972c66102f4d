"""kw_sim_run_noalloc (every device buffer in the library's static scratch
region) against kw_sim_create + kw_sim_run + kw_sim_fetch: bit-identical
outputs on two calls in a row, the second on a smaller grid, plus the
size limit rejected on the host before any device work."""

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

REC = ("p", "p_max", "p_rms", "p_final", "u_max")


def _plain(kw, f, rec):
    with kw.Simulation(f, record=rec) as sim:
        sim.run()
        return sim.outputs()


def _same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        assert a[k].shape == b[k].shape, k
        assert np.array_equal(a[k], b[k]), (k, float(np.max(np.abs(a[k] - b[k]))))


def test_noalloc_equals_create_run_fetch_twice(kw):
    big = cases.case_3d(N=(48, 40, 32), Nt=40, het=True)
    small = cases.case_3d(N=(32, 28, 24), Nt=30, het=True, source="p", many=True)
    for name, f in (("48x40x32 heterogeneous absorbing ux additive", big),
                    ("32x28x24 heterogeneous absorbing p additive many", small)):
        print(f"[noalloc] {name}", flush=True)
        got = kw.native.run_noalloc(f, record=REC)
        ref = _plain(kw, f, REC)
        _same(got, ref)
        assert np.all(np.isfinite(got["p"])) and np.max(np.abs(got["p"])) > 0


def test_noalloc_2d_then_3d_p0(kw):
    for name, f in (("2d 64x64 p0", cases.case_2d_p0_lossless(N=64, Nt=50)),
                    ("3d 32^3 lossless p0", cases.case_3d(N=(32, 32, 32), Nt=25, het=False, absorbing=False,
                                                          source=None, p0=True))):
        print(f"[noalloc] {name}", flush=True)
        rec = ("p", "p_max_all")
        _same(kw.native.run_noalloc(f, record=rec), _plain(kw, f, rec))


def test_noalloc_rejects_oversize_before_device_work(kw):
    f = cases.case_3d(N=(256, 256, 128), Nt=4, het=True)
    print("[noalloc] 256x256x128 (larger than the region): host-side rejection", flush=True)
    with pytest.raises(kw.native.KWaveError, match="scratch region"):
        kw.native.run_noalloc(f, record=("p",))
    # the region is intact afterwards
    g = cases.case_3d(N=(32, 28, 24), Nt=10)
    _same(kw.native.run_noalloc(g, record=("p",)), _plain(kw, g, ("p",)))
