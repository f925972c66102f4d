"""Benchmark: grid-point-updates/s of the k-space time loop on one MI355X.

Workload (N=1): the 3-D 128^3 heterogeneous absorbing configuration (C2:
16x16 ux additive source, mid-plane sensor recording p and p_max, Nt = 1200).
A "step" is one time step of that simulation; the timed region runs K steps
of a freshly created simulation from t = 1 (K = 1200 by default = the whole
run), with inputs and operators already resident in HBM.  Warmup runs W steps
on a separate instance first.

Also reported: the roofline of the dominant kernel (algorithmic bytes per
launch / measured average launch time, per-kernel HIP events on the solver's
stream), the whole-step algorithmic bandwidth, the FP64 CPU oracle timed on a
bounded sample of the same workload, and the serving leg: a sample of the C5
request stream run by the plain path (one simulation at a time) and by the
serving loop (continuous batching), with throughput and p50/p99 latency for
burst and Poisson arrivals.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--workload c2|c3|c4]
(--workload c3/c4 times the large configurations, without the serving leg or the CPU baseline)
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
    if kind in ("last_axis", "yz_plane"):  # averaged over the step's launches: in + out spectra + tables
        # gradient 1->dim (+kappa), divergence dim->dim (+kappa), absorption 2->2 (+nabla1, nabla2)
        launches = [(1, dim, 1), (dim, dim, 1)] + ([(2, 2, 2)] if absorbing == 1 else [])
        tot = sum((i + o) * spec + t * tab for i, o, t in launches)
        return tot / len(launches)
    if kind == "y_pass":        # averaged: read + write of the fields in each launch
        nf = [1, dim, dim, dim, dim] + ([2, 2] if absorbing == 1 else [])
        return sum(2 * n * spec for n in nf) / len(nf)
    return 0.0


def profiled_traffic(workload: str, kind: str):
    """HBM bytes per launch of a kernel kind at this workload, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (profiles/<workload>_*.md, calibrated
    against a plain copy); None when no summary covers it."""
    import glob
    import re

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", f"{workload}_*.md"))):
        m = re.search(r"```json\s*(\{.*?\})\s*```", open(path).read(), re.S)
        if m:
            t = json.loads(m.group(1)).get("traffic_bytes_per_launch", {})
            if kind in t:
                return float(t[kind]), os.path.relpath(path, ROOT)
    return None, None


def _updates(f: dict) -> int:
    return int(np.prod([int(np.asarray(f[k]).reshape(-1)[0]) for k in ("Nx", "Ny", "Nz", "Nt")]))


def _serve_stream(kw, reqs, arrivals, max_active: int, quantum: int, autotune: bool = False):
    """Submit reqs at the given arrival offsets (s) to a fresh server; wait for all."""
    with kw.Server(max_active=max_active, steps_per_quantum=quantum, autotune=autotune) as srv:
        t0 = time.perf_counter()
        ids = []
        for f, ta in zip(reqs, arrivals):
            wait = t0 + ta - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            ids.append(srv.submit(f, record=("p",)))
        for r in ids:
            srv.wait(r)
        makespan = time.perf_counter() - t0
        infos = [srv.info(r) for r in ids]
        bad = [(r, srv.error(r)) for r, i in zip(ids, infos) if i["state"] != "done"]
        if bad:
            raise RuntimeError(f"serving requests failed: {bad[:3]}")
        outs = [srv.fetch(r, "p") for r in ids]
        peak = srv.stats()["hbm_peak_reserved_bytes"]
    lat = np.array([i["t_done"] - i["t_submit"] for i in infos])
    upd = sum(_updates(f) for f in reqs)
    res = {"max_active": max_active, "autotune": autotune, "makespan_s": round(makespan, 3),
           "updates_per_s": upd / makespan,
           "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
           "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
           "hbm_peak_reserved_GiB": round(peak / 2 ** 30, 2)}
    return res, outs


def serving_leg(kw, M: int, rates, max_active: int, quantum: int) -> dict:
    """C5 serving stream: plain (one simulation at a time) vs served (continuous
    batching over streams) on the same requests, burst and Poisson arrivals."""
    import concurrent.futures as cf
    from kwave_amd import configs

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(8) as ex:
        reqs = list(ex.map(configs.c5_request, range(M)))
    gen_s = time.perf_counter() - t0
    print(f"serving: generated {M} requests in {gen_s:.1f} s", file=sys.stderr, flush=True)
    kw.run(configs.c5_request(0, grids=((32, 32, 32),), nt_range=(8, 8)))  # code objects loaded
    burst = [0.0] * M
    plain, plain_out = _serve_stream(kw, reqs, burst, 1, quantum)
    print(f"serving: plain burst {plain}", file=sys.stderr, flush=True)
    served, served_out = _serve_stream(kw, reqs, burst, max_active, quantum, autotune=True)
    print(f"serving: served burst {served}", file=sys.stderr, flush=True)
    # served requests run autotuned launch shapes: equal to the plain path to FP32 rounding
    rel = max(float(np.linalg.norm(b.astype(np.float64) - a) / max(np.linalg.norm(a.astype(np.float64)), 1e-30))
              for a, b in zip(plain_out, served_out))
    del plain_out, served_out
    out = {"requests": M, "sample": f"c5_request(0..{M - 1}) of the seeded serving stream",
           "updates": sum(_updates(f) for f in reqs), "steps_per_quantum": quantum,
           "burst": {"plain": plain, "served": served,
                     "speedup": round(served["updates_per_s"] / plain["updates_per_s"], 3)},
           "max_rel_l2_vs_plain": rel, "matches_plain": bool(rel <= 1e-5), "poisson": []}
    capacity = M / plain["makespan_s"]  # plain-path requests per second
    rng = np.random.default_rng(5999)
    gaps = rng.exponential(1.0, M)
    for r in rates:
        lam = r * capacity
        arr = list(np.cumsum(gaps / lam) - gaps[0] / lam)
        p, _ = _serve_stream(kw, reqs, arr, 1, quantum)
        s, _ = _serve_stream(kw, reqs, arr, max_active, quantum, autotune=True)
        out["poisson"].append({"rate_per_s": round(lam, 3), "load_vs_plain_capacity": r, "plain": p, "served": s})
        print(f"serving: poisson load {r}: plain p50/p99 {p['p50_latency_s']}/{p['p99_latency_s']} s, "
              f"served {s['p50_latency_s']}/{s['p99_latency_s']} s", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("c2", "c3", "c4"), default="c2",
                    help="c2 (default): the 128^3 headline workload; c3/c4: the large configurations")
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: 1200 for c2, 100 otherwise)")
    ap.add_argument("--warmup", type=int, default=None, help="warmup steps (default: 100 for c2, 10 otherwise)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=120)
    ap.add_argument("--calib-copy-mib", type=int, default=1024,
                    help="also run one plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--serving-requests", type=int, default=24,
                    help="C5 serving-stream sample size (0 skips the serving leg)")
    ap.add_argument("--serving-rates", default="0.9",
                    help="Poisson arrival rates as fractions of the plain path's capacity")
    ap.add_argument("--serving-max-active", type=int, default=8)
    ap.add_argument("--serving-quantum", type=int, default=16)
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

    wl = args.workload
    K = args.steps if args.steps is not None else (1200 if wl == "c2" else 100)
    W = args.warmup if args.warmup is not None else (100 if wl == "c2" else 10)
    prof_steps = min(K, 200 if wl == "c2" else 50)
    t_gen = time.perf_counter()
    if wl == "c2":
        f, rec, desc, bpp = configs.c2(Nt=max(K, W, 1)), configs.C2_RECORD, configs.C2_DESCRIPTION, 212.0
        data = "synthetic (seeded Gaussian random medium, default_rng(1002))"
    elif wl == "c3":
        f, rec, desc, bpp = (configs.c3(Nt=W + K + prof_steps), configs.C3_RECORD, configs.C3_DESCRIPTION,
                             208.0)
        data = "synthetic (homogeneous medium, makeBowl source)"
    else:
        f, rec, desc, bpp = (configs.c4(Nt=W + K + prof_steps, log=True), configs.C4_RECORD,
                             configs.C4_DESCRIPTION, 212.0)
        data = "synthetic (seeded three-layer slab with Gaussian perturbation, default_rng(1004))"
    print(f"{wl}: inputs generated in {time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)
    Nx, Ny, Nz = (int(np.asarray(f[k]).reshape(-1)[0]) for k in ("Nx", "Ny", "Nz"))
    N = Nx * Ny * Nz

    def timed(sim, steps):
        sim.sync()
        t0 = time.perf_counter()
        sim.run(steps)
        sim.sync()
        return time.perf_counter() - t0

    if wl == "c2":
        # warmup on a separate instance; timed region: K steps of a fresh simulation from t = 1
        if W > 0:
            with kw.Simulation(f, record=rec) as sim:
                sim.run(W)
                sim.sync()
        with kw.Simulation(f, record=rec) as sim:
            elapsed = timed(sim, K)
            chk = sim.fetch(rec[-1])
        # per-kernel timing (HIP events around every launch, on the solver stream)
        with kw.Simulation(f, record=rec) as sim:
            sim.set_profiling(True)
            sim.run(prof_steps)
            sim.sync()
            prof = sim.profile()
    else:
        # one resident simulation (tens of GB): W warmup steps, K timed, then profiled steps
        with kw.Simulation(f, record=rec) as sim:
            print(f"{wl}: simulation created ({kw.native.footprint_bytes(f, rec) / 2**30:.1f} GiB)",
                  file=sys.stderr, flush=True)
            if W > 0:
                timed(sim, W)
            elapsed = timed(sim, K)
            print(f"{wl}: {K} timed steps in {elapsed:.2f} s", file=sys.stderr, flush=True)
            sim.set_profiling(True)
            sim.run(prof_steps)
            sim.sync()
            prof = sim.profile()
            chk = sim.fetch(rec[-1])
    if not np.all(np.isfinite(chk)):
        raise RuntimeError(f"non-finite {rec[-1]} in the benchmark run")
    del chk
    value = N * K / elapsed
    ms_per_step = 1e3 * elapsed / K
    kinds = {k: v for k, v in prof.items() if v[1] > 0}
    dom = max(kinds, key=lambda k: kinds[k][0])
    dom_ms, dom_n = kinds[dom]
    dom_avg_s = dom_ms / dom_n / 1e3
    dom_bytes = kernel_bytes(f, dom, {})
    achieved = dom_bytes / dom_avg_s / 1e9
    per_kernel = {k: {"avg_us": round(1e3 * v[0] / v[1], 2), "launches": v[1],
                      "share": round(v[0] / sum(x[0] for x in kinds.values()), 3)} for k, v in kinds.items()}

    traffic, traffic_src = profiled_traffic(wl, dom)
    step_bytes = bpp * N  # algorithmic bytes per time step of this workload
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
        "data": data,
        "config": {"workload": wl.upper() + " " + desc, "grid": [Nx, Ny, Nz], "Nt": K,
                   "record": list(rec)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src, "bytes_per_launch": dom_bytes,
                     "avg_launch_us": round(dom_avg_s * 1e6, 2)},
        "step_roofline": {"bytes_per_point": bpp, "achieved_GBs": round(step_bytes / (ms_per_step / 1e3) / 1e9, 1),
                          "frac": round(step_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
        "kernels": per_kernel,
    }
    if copy_gbs:
        out["copy_GBs"] = round(copy_gbs, 1)

    if args.serving_requests > 0 and wl == "c2":
        rates = [float(x) for x in args.serving_rates.split(",") if x]
        out["serving"] = serving_leg(kw, args.serving_requests, rates, args.serving_max_active,
                                     args.serving_quantum)

    if not args.no_cpu and wl == "c2":
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
