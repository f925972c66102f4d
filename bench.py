"""Benchmark: grid-point-updates/s of the k-space time loop on one MI355X.

Workload (N=1): the 3-D 128^3 heterogeneous absorbing configuration (C2:
16x16 ux additive source, mid-plane sensor recording p and p_max, Nt = 1200).
A "step" is one time step of that simulation; the timed region runs K steps
of a freshly created simulation from t = 1 (K = 1200 by default = the whole
run), with inputs and operators already resident in HBM.  Warmup runs W steps
on a separate instance first.

Also reported: the roofline of the dominant kernel (algorithmic bytes per
launch / measured average launch time, per-kernel HIP events on the solver's
stream), the whole-step algorithmic bandwidth, and the FP64 CPU oracle timed
on a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
from __graft_entry__ import import_pkg  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def kernel_bytes(f: dict, kind: str, launches_per_step: dict) -> float:
    """Algorithmic HBM bytes of one launch of a kernel kind at this workload.

    Counts each distinct array the kernel must read or write once: real
    fields 4 B/point, X-domain half spectra (Nx/2+1)*Ny*Nz*8 B, spectral
    tables 4 B per half-spectrum point; scalars and 1-D vectors 0.
    """
    Nx, Ny, Nz = (int(np.asarray(f[k]).reshape(-1)[0]) for k in ("Nx", "Ny", "Nz"))
    N = Nx * Ny * Nz
    spec = (Nx // 2 + 1) * Ny * Nz * 8
    tab = (Nx // 2 + 1) * Ny * Nz * 4
    fld = 4 * N
    het = lambda k: np.asarray(f[k]).size > 1  # noqa: E731
    dim = 3 if Nz > 1 else 2
    absorbing = int(np.asarray(f.get("absorbing_flag", 0)).reshape(-1)[0])
    if kind == "x_velocity":    # per component: G in, u in/out, rho_sg^-1, G out
        per = 2 * spec + 2 * fld + sum(fld for a in "xyz"[:dim] if het("rho0_sg" + a)) / dim
        return dim * per
    if kind == "x_density":     # dim G in, dim rho in/out, rho0; absorbing: 2 G out + sum(rho)
        b = dim * spec + 2 * dim * fld + (fld if het("rho0") else 0)
        return b + (2 * spec + fld if absorbing == 1 else spec + (fld if het("c0") else 0))
    if kind == "x_pressure":    # 2 G in, sum(rho), c0, tau, eta, Pk out
        fields = 1 + (1 if het("c0") else 0) + (2 if (het("c0") or het("alpha_coeff")) else 0)
        return 2 * spec + fields * fld + spec
    if kind == "last_axis":     # averaged over the step's launches: in + out spectra + tables
        # gradient 1->dim (+kappa), divergence dim->dim (+kappa), absorption 2->2 (+nabla1, nabla2)
        launches = [(1, dim, 1), (dim, dim, 1)] + ([(2, 2, 2)] if absorbing == 1 else [])
        tot = sum((i + o) * spec + t * tab for i, o, t in launches)
        return tot / len(launches)
    if kind == "y_pass":        # averaged: read + write of the fields in each launch
        nf = [1, dim, dim, dim, dim] + ([2, 2] if absorbing == 1 else [])
        return sum(2 * n * spec for n in nf) / len(nf)
    return 0.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1200)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=120)
    ap.add_argument("--calib-copy-mib", type=int, default=1024,
                    help="also run one plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "a multi-GPU run is not part of it.", file=sys.stderr)
        sys.exit(2)

    kw = import_pkg()
    from kwave_amd import configs

    copy_gbs = None
    if args.calib_copy_mib:
        ms = kw.native.diag_copy(args.calib_copy_mib << 20)
        copy_gbs = 2 * (args.calib_copy_mib << 20) / (ms * 1e-3) / 1e9
        print(f"calibration copy: {args.calib_copy_mib} MiB in {ms:.3f} ms ({copy_gbs:.0f} GB/s read+write)",
              file=sys.stderr, flush=True)

    K, W = args.steps, args.warmup
    f = configs.c2(Nt=max(K, W, 1))
    Nx, Ny, Nz = (int(np.asarray(f[k]).reshape(-1)[0]) for k in ("Nx", "Ny", "Nz"))
    N = Nx * Ny * Nz
    rec = configs.C2_RECORD

    # warmup on a separate instance
    if W > 0:
        with kw.Simulation(f, record=rec) as sim:
            sim.run(W)
            sim.sync()

    # timed region: K steps of a fresh simulation from t = 1
    with kw.Simulation(f, record=rec) as sim:
        sim.sync()
        t0 = time.perf_counter()
        sim.run(K)
        sim.sync()
        t1 = time.perf_counter()
        p_max = sim.fetch("p_max")
        if not np.all(np.isfinite(p_max)):
            raise RuntimeError("non-finite p_max in the benchmark run")
    elapsed = t1 - t0
    value = N * K / elapsed
    ms_per_step = 1e3 * elapsed / K

    # per-kernel timing (HIP events around every launch, on the solver stream)
    prof_steps = min(K, 200)
    with kw.Simulation(f, record=rec) as sim:
        sim.set_profiling(True)
        sim.run(prof_steps)
        sim.sync()
        prof = sim.profile()
    kinds = {k: v for k, v in prof.items() if v[1] > 0}
    dom = max(kinds, key=lambda k: kinds[k][0])
    dom_ms, dom_n = kinds[dom]
    dom_avg_s = dom_ms / dom_n / 1e3
    dom_bytes = kernel_bytes(f, dom, {})
    achieved = dom_bytes / dom_avg_s / 1e9
    per_kernel = {k: {"avg_us": round(1e3 * v[0] / v[1], 2), "launches": v[1],
                      "share": round(v[0] / sum(x[0] for x in kinds.values()), 3)} for k, v in kinds.items()}

    step_bytes = 212.0 * N  # algorithmic bytes per time step of this workload (212 B/point)
    out = {
        "metric": "grid-point-updates/s",
        "value": value,
        "unit": "grid-point-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian random medium, default_rng(1002))",
        "config": {"workload": "C2 " + configs.C2_DESCRIPTION, "grid": [Nx, Ny, Nz], "Nt": K,
                   "record": list(rec)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": dom_bytes, "avg_launch_us": round(dom_avg_s * 1e6, 2)},
        "step_roofline": {"bytes_per_point": 212, "achieved_GBs": round(step_bytes / (ms_per_step / 1e3) / 1e9, 1),
                          "frac": round(step_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
        "kernels": per_kernel,
    }
    if copy_gbs:
        out["copy_GBs"] = round(copy_gbs, 1)

    if not args.no_cpu:
        from oracle import kwave_oracle  # CPU baseline only

        cs = args.cpu_steps
        t0 = time.perf_counter()
        kwave_oracle.run(f, record=rec, nsteps=cs)
        tc = time.perf_counter() - t0
        cores = len(os.sched_getaffinity(0))
        out["cpu_baseline"] = {"value": N * cs / tc, "unit": "grid-point-updates/s", "cores": cores,
                               "kind": "port",
                               "sample": f"FP64 numpy/scipy.fft oracle, first {cs} time steps of the same "
                                         f"workload (source active), scipy.fft workers={cores}, "
                                         f"{tc:.1f} s"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
