"""Full-length parity on the benchmark workloads: the GPU run against the FP64
oracle over the whole time loop (C2 128^3, 1200 steps; C4's branch at 96^3,
1500 steps) and over a 100-step window of C3 at its full size through the
source-active phase.  Each comparison reports the norms the reference's
parity harness prints (kWaveTester.m:2946-2983: L2, relative L2, L-infinity,
relative L-infinity) and must stay within relative L2 1e-4."""

import time

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

TOL_L2 = 1e-4


def norms(got, ref):
    g = np.asarray(got, dtype=np.float64)
    r = np.asarray(ref, dtype=np.float64)
    d = g - r
    l2, linf = float(np.linalg.norm(d)), float(np.max(np.abs(d)))
    return {"L2": l2, "rel_L2": l2 / max(float(np.linalg.norm(r)), 1e-300), "Linf": linf,
            "rel_Linf": linf / max(float(np.max(np.abs(r))), 1e-300)}


def compare(kw, oracle, f, rec, label):
    t0 = time.perf_counter()
    gpu = kw.run(f, record=rec)
    t1 = time.perf_counter()
    ref = oracle.run(f, record=rec)
    t2 = time.perf_counter()
    print(f"    {label}: GPU {t1 - t0:.1f} s, FP64 oracle {t2 - t1:.1f} s", flush=True)
    for k in rec:
        n = norms(gpu[k], ref[k])
        print(f"    {k}: " + ", ".join(f"{a} {b:.3e}" for a, b in n.items()), flush=True)
        assert np.all(np.isfinite(gpu[k]))
        assert n["rel_L2"] <= TOL_L2, (k, n)


def test_c2_full_length(kw, oracle):
    """C2 (128^3 heterogeneous absorbing, ux additive patch, mid-plane p and
    p_max): all 1200 steps."""
    print("[gpu longrun] C2 128^3, 1200 steps vs FP64 oracle", flush=True)
    compare(kw, oracle, kw.configs.c2(), ("p", "p_max"), "C2")


def test_c4_branch_full_length_96(kw, oracle):
    """C4's branch (three-layer slab, y = 1.1, p additive patch, receiver
    plane) on 96^3 for the full Nt = 1500."""
    print("[gpu longrun] C4 branch on 96^3, 1500 steps vs FP64 oracle", flush=True)
    compare(kw, oracle, kw.configs.c4(Nt=1500, N=96), ("p",), "C4@96^3")


def test_c3_full_size_100_steps(kw, oracle):
    """C3 at full size (256x256x512, nonlinear B/A = 5, y = 2, bowl source,
    every point recording p_max and FP64-accumulated p_rms): 100 steps, all
    within the bowl's source-active phase."""
    f = kw.configs.c3(Nt=100)
    nsrc = int(np.asarray(f["p_source_flag"]).reshape(-1)[0])
    print(f"[gpu longrun] C3 256x256x512, 100 steps vs FP64 oracle (source active for {nsrc})", flush=True)
    assert nsrc >= 100
    compare(kw, oracle, f, ("p_max", "p_rms"), "C3")
