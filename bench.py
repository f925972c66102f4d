"""Benchmark: grid-point-updates/s of the k-space time loop on one MI355X.

Workload (N=1, default): the largest single-GPU configuration, 768^3
heterogeneous absorbing three-layer slab (PML 20, 64x64 p additive patch,
128x128 receiver plane recording p, CFL 0.3, Nt = 1500).  A "step" is one
time step of that simulation; the timed region runs K steps of one resident
simulation from t = 1 after W untimed warmup steps of a separate instance,
with inputs and operators already resident in HBM.

Also reported, on the same line:
  roofline       the dominant kernel (largest share of the timed region):
                 algorithmic bytes per launch / its average launch duration,
                 measured with HIP events on the solver's stream over the
                 timed region; `traffic` = its counter-measured HBM bytes per
                 launch from the committed rocprofv3 summary
  step_roofline  algorithmic bytes per grid point and step (212 B, +20 B per
                 active additive source component) / step time
  hbm_counters   counter-measured FETCH+WRITE bytes per step (committed
                 summary, per kernel kind x launches per step) / step time
  copy_ceiling   the best of several streaming device copies (>= 1 GiB)
  cpu_baseline   the FP64 oracle on the host cores, a bounded sample of the
                 same workload (one step), extrapolated to Nt
  c2             the 128^3 configuration (plane-pass layout) as a second line
  serving        a sample of the serving stream: plain path (one simulation at
                 a time) vs the serving loop with concurrent streams and with
                 batched launches (hipGraph quanta), burst and Poisson
                 arrivals at several loads, p50/p99 latency, and each mode's
                 relative L2 / max-abs error against the FP64 oracle

Usage: python bench.py [--gpus 1] [--workload c4|c2|c3] [--steps K] [--warmup W]
                       [--no-cpu] [--serving-requests M] [--serving-rates 0.5,0.9,1.5]
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
BASE_BYTES_PER_POINT = {"c2": 212.0, "c3": 208.0, "c4": 212.0}  # algorithmic model, per step
SOURCE_BYTES_PER_POINT = 20.0  # per additive source component while its signal lasts


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _dims(f):
    return tuple(int(np.asarray(f[k]).reshape(-1)[0]) for k in ("Nx", "Ny", "Nz"))


def kernel_bytes(f: dict, kind: str) -> float:
    """Algorithmic HBM bytes of one launch of a kernel kind at this workload.

    Counts each distinct array the kernel must read or write once: real
    fields 4 B/point, X-domain half spectra (Nx/2+1)*Ny*Nz*8 B, spectral
    tables 4 B per half-spectrum point; scalars and 1-D vectors 0.  The
    spectral passes are averaged over one step's launches of that kind.
    """
    Nx, Ny, Nz = _dims(f)
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
    if kind in ("last_axis", "yz_plane"):  # in + out spectra + tables, averaged over the step's launches
        # gradient 1->dim (+kappa), divergence dim->dim (+kappa), absorption 2->2 (+nabla1, nabla2)
        launches = [(1, dim, 1), (dim, dim, 1)] + ([(2, 2, 2)] if absorbing == 1 else [])
        return sum((i + o) * spec + t * tab for i, o, t in launches) / len(launches)
    if kind == "y_pass":        # read + write of the fields of each launch: the inputs and the
        # outputs of each spectral operator (gradient 1, dim; divergence dim, dim; absorption 2, 2)
        nf = [1, dim, dim, dim] + ([2, 2] if absorbing == 1 else [])
        return sum(2 * n * spec for n in nf) / len(nf)
    return 0.0


def step_bytes_per_point(f: dict, wl: str, t_first: int, t_last: int) -> float:
    """Mean algorithmic bytes per grid point and step over steps t_first..t_last:
    the configuration's base figure plus 20 B per additive source component
    whose signal is still running."""
    comps = []
    for k in ("p", "ux", "uy", "uz"):
        flag = int(np.asarray(f.get(k + "_source_flag", 0)).reshape(-1)[0])
        mode = int(np.asarray(f.get(("p" if k == "p" else "u") + "_source_mode", 0)).reshape(-1)[0])
        if flag and mode == 2:
            comps.append(flag)
    steps = np.arange(t_first, t_last + 1)
    active = sum((steps <= fl).astype(float) for fl in comps) if comps else np.zeros(len(steps))
    return float(np.mean(BASE_BYTES_PER_POINT[wl] + SOURCE_BYTES_PER_POINT * active))


def profiled_traffic(workload: str):
    """Counter-measured HBM bytes per launch by kernel kind at this workload,
    from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    (profiles/<workload>_*.md, calibrated against a plain copy); ({}, None)
    when no summary covers it."""
    import glob
    import re

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", f"{workload}_*.md"))):
        m = re.search(r"```json\s*(\{.*?\})\s*```", open(path).read(), re.S)
        if m:
            t = json.loads(m.group(1)).get("traffic_bytes_per_launch", {})
            if t:
                return {k: float(v) for k, v in t.items()}, os.path.relpath(path, ROOT)
    return {}, None


def _updates(f: dict) -> int:
    return int(np.prod([int(np.asarray(f[k]).reshape(-1)[0]) for k in ("Nx", "Ny", "Nz", "Nt")]))


def cpu_info() -> dict:
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "cores": len(os.sched_getaffinity(0))}


def mem_available() -> int:
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemAvailable:"):
                return int(line.split()[1]) * 1024
    except OSError:
        pass
    return 0


# ---------------------------------------------------------------------------
# device copy: counter calibration kernel and the measured copy ceiling
# ---------------------------------------------------------------------------

def copy_legs(kw, calib_mib: int, ceiling_gib: int) -> dict:
    out = {}
    if calib_mib:
        ms = kw.native.diag_copy(calib_mib << 20, variant=0)
        out["calibration_copy_GBs"] = round(2 * (calib_mib << 20) / (ms * 1e-3) / 1e9, 1)
        log(f"calibration copy (variant 0): {calib_mib} MiB in {ms:.3f} ms ({out['calibration_copy_GBs']} GB/s)")
    if ceiling_gib:
        nbytes = ceiling_gib << 30
        best, sweep = None, []
        for variant in (0, 1, 2, 3, 4):
            for blocks in (2048, 8192, 32768):
                ms = kw.native.diag_copy(nbytes, variant=variant, blocks=blocks)
                gbs = 2 * nbytes / (ms * 1e-3) / 1e9
                sweep.append((variant, blocks, round(gbs, 1)))
                if best is None or gbs > best[2]:
                    best = (variant, blocks, gbs)
        out["copy_ceiling"] = {"GBs": round(best[2], 1), "frac_of_peak": round(best[2] / HBM_PEAK_GBS, 4),
                               "variant": best[0], "blocks": best[1], "bytes": nbytes,
                               "sweep_variant_blocks_GBs": sweep}
        log(f"copy ceiling: {best[2]:.0f} GB/s (variant {best[0]}, {best[1]} blocks) over {ceiling_gib} GiB")
    return out


# ---------------------------------------------------------------------------
# one configuration on one simulation
# ---------------------------------------------------------------------------

def measure_workload(kw, wl: str, K: int, W: int, keep_inputs: bool = False) -> tuple[dict, dict | None]:
    """Times K steps of the workload from t = 1; returns (line fields, inputs)."""
    from kwave_amd import configs

    t_gen = time.perf_counter()
    if wl == "c2":
        f, rec, desc = configs.c2(Nt=max(K, W, 1)), configs.C2_RECORD, configs.C2_DESCRIPTION
        data = "synthetic (seeded Gaussian random medium, default_rng(1002))"
        nt_full = 1200
    elif wl == "c3":
        f, rec, desc = configs.c3(Nt=max(K, W, 1)), configs.C3_RECORD, configs.C3_DESCRIPTION
        data = "synthetic (homogeneous medium, makeBowl source)"
        nt_full = 2000
    else:
        f, rec, desc = configs.c4(Nt=max(K, W, 1), log=True), configs.C4_RECORD, configs.C4_DESCRIPTION
        data = "synthetic (seeded three-layer slab with Gaussian perturbation, default_rng(1004))"
        nt_full = 1500
    log(f"{wl}: inputs generated in {time.perf_counter() - t_gen:.1f} s")
    Nx, Ny, Nz = _dims(f)
    N = Nx * Ny * Nz

    def timed(sim, steps):
        sim.sync()
        t0 = time.perf_counter()
        sim.run(steps)
        sim.sync()
        return time.perf_counter() - t0

    if W > 0:  # warmup on a separate instance (code objects, caches, allocator pools)
        with kw.Simulation(f, record=rec) as sim:
            timed(sim, W)
        log(f"{wl}: {W} warmup steps done")
    with kw.Simulation(f, record=rec) as sim:
        mem = sim.memory()
        log(f"{wl}: simulation created ({mem['allocated'] / 2**30:.2f} GiB on the device)")
        # per-kernel HIP events on the solver stream over the timed region; on
        # small grids (launch-bound) the event pairs perturb the step, so they
        # time a second run of the same steps instead
        in_region = N >= 64 * 2**20
        sim.set_profiling(in_region)
        elapsed = timed(sim, K)
        prof = sim.profile() if in_region else None
        chk = sim.fetch(rec[-1])
    if not np.all(np.isfinite(chk)):
        raise RuntimeError(f"non-finite {rec[-1]} in the benchmark run")
    del chk
    if prof is None:
        with kw.Simulation(f, record=rec) as sim:
            sim.set_profiling(True)
            sim.run(K)
            sim.sync()
            prof = sim.profile()
    log(f"{wl}: {K} timed steps in {elapsed:.3f} s ({1e3 * elapsed / K:.3f} ms/step)")

    ms_per_step = 1e3 * elapsed / K
    kinds = {k: v for k, v in prof.items() if v[1] > 0}
    total_ms = sum(v[0] for v in kinds.values())
    dom = max(kinds, key=lambda k: kinds[k][0])
    dom_ms, dom_n = kinds[dom]
    dom_avg_s = dom_ms / dom_n / 1e3
    dom_bytes = kernel_bytes(f, dom)
    achieved = dom_bytes / dom_avg_s / 1e9
    per_kernel = {k: {"avg_us": round(1e3 * v[0] / v[1], 2), "launches": v[1],
                      "share": round(v[0] / total_ms, 3),
                      "algorithmic_bytes_per_launch": kernel_bytes(f, k),
                      "algorithmic_GBs": round(kernel_bytes(f, k) / (v[0] / v[1] / 1e3) / 1e9, 1)}
                  for k, v in kinds.items()}
    traffic, traffic_src = profiled_traffic(wl)
    bpp = step_bytes_per_point(f, wl, 1, K)
    step_GBs = bpp * N / (ms_per_step / 1e3) / 1e9
    line = {
        "value": N * K / elapsed,
        "ms_per_step": ms_per_step,
        "data": data,
        "config": {"workload": wl.upper() + " " + desc, "grid": [Nx, Ny, Nz], "Nt": nt_full,
                   "timed_steps": [1, K], "record": list(rec)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic.get(dom), "traffic_source": traffic_src,
                     "bytes_per_launch": dom_bytes, "avg_launch_us": round(dom_avg_s * 1e6, 2),
                     "timing": "HIP events on the solver stream over the timed region" if in_region
                               else "HIP events on the solver stream, a second run of the same steps"},
        "step_roofline": {"bytes_per_point": round(bpp, 2), "achieved_GBs": round(step_GBs, 1),
                          "frac": round(step_GBs / HBM_PEAK_GBS, 4)},
        "kernels": per_kernel,
    }
    if traffic and all(k in traffic for k in kinds):
        step_traffic = sum(traffic[k] * kinds[k][1] / K for k in kinds)
        cnt_GBs = step_traffic / (ms_per_step / 1e3) / 1e9
        line["hbm_counters"] = {"bytes_per_step": step_traffic, "achieved_GBs": round(cnt_GBs, 1),
                                "frac": round(cnt_GBs / HBM_PEAK_GBS, 4),
                                "vs_algorithmic": round(step_traffic / (bpp * N), 3), "source": traffic_src}
    return line, (f if keep_inputs else None)


# ---------------------------------------------------------------------------
# CPU baseline: the FP64 oracle on a bounded sample of the same workload
# ---------------------------------------------------------------------------

def cpu_baseline(f: dict, rec, steps: int, nt_full: int) -> dict:
    import scipy

    from oracle import kwave_oracle  # CPU baseline only

    Nx, Ny, Nz = _dims(f)
    N = Nx * Ny * Nz
    info = cpu_info()
    need = 320 * N  # FP64 oracle working set, measured ~300 B per grid point at 256^3
    avail = mem_available()
    if avail and need > 0.85 * avail:
        return {"value": None, "unit": "grid-point-updates/s", "cores": info["cores"], "kind": "port",
                "sample": f"skipped: the FP64 oracle needs ~{need / 2**30:.0f} GiB, {avail / 2**30:.0f} GiB available"}
    tm = {}
    t0 = time.perf_counter()
    kwave_oracle.run(f, record=rec, nsteps=steps, timings=tm, progress=lambda m: log(f"cpu baseline: {m}"))
    wall = time.perf_counter() - t0
    t_act = float(np.mean(tm["step"]))
    t_src = float(np.mean(tm["source"]))
    t_steady = max(t_act - t_src, 1e-9)
    flags = [int(np.asarray(f.get(k + "_source_flag", 0)).reshape(-1)[0]) for k in ("p", "ux", "uy", "uz")]
    n_src = min(nt_full, max(flags))
    T = n_src * t_act + (nt_full - n_src) * t_steady
    return {"value": N * nt_full / T, "unit": "grid-point-updates/s", "cores": info["cores"], "kind": "port",
            "extrapolated": True,
            "sample": (f"FP64 numpy {np.__version__} / scipy {scipy.__version__} (scipy.fft pocketfft, "
                       f"workers={info['cores']}, slab-parallel elementwise) oracle on {info['cpu_model']}: "
                       f"{steps} time step(s) of the same {Nx}x{Ny}x{Nz} workload with the source active, "
                       f"{t_act:.1f} s/step of which {t_src:.2f} s in the source branch; extrapolated to "
                       f"Nt={nt_full} as {n_src} source-active steps + {nt_full - n_src} steady-state steps "
                       f"of {t_steady:.1f} s (oracle call {wall:.0f} s incl. operator set-up)"),
            "step_s": round(t_act, 3), "steady_step_s": round(t_steady, 3), "cpu_model": info["cpu_model"]}


# ---------------------------------------------------------------------------
# serving leg
# ---------------------------------------------------------------------------

SERVING_MODES = {
    # one simulation at a time: the reference's one-process-per-simulation loop, on the GPU
    "plain": {"max_active": 1},
    # concurrent simulations, each on its own stream (autotuned launch shapes)
    "streams": {"max_active": 8, "autotune": True},
    # simulations of one grid shape share each kernel launch; quanta replayed from hipGraphs
    "batched": {"max_active": 32, "batching": True},
}


def _serve_stream(kw, reqs, arrivals, quantum: int, mode: str):
    """Submit reqs at the given arrival offsets (s) to a fresh server; wait for all."""
    cfg = dict(SERVING_MODES[mode])
    with kw.Server(steps_per_quantum=quantum, **cfg) as srv:
        t0 = time.perf_counter()
        ids = []
        for k, (f, ta) in enumerate(zip(reqs, arrivals)):
            wait = t0 + ta - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            ids.append(srv.submit(f, record=("p",)))
            if (k + 1) % 128 == 0 and k + 1 < len(reqs):
                log(f"serving: {mode}: {k + 1}/{len(reqs)} requests submitted")
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
    res = {"mode": mode, **{k: v for k, v in cfg.items()}, "makespan_s": round(makespan, 3),
           "updates_per_s": upd / makespan,
           "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
           "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
           "hbm_peak_reserved_GiB": round(peak / 2 ** 30, 2)}
    return res, outs


def _err(got, ref):
    g = np.asarray(got, dtype=np.float64)
    r = np.asarray(ref, dtype=np.float64)
    nr = np.linalg.norm(r)
    return float(np.linalg.norm(g - r) / (nr if nr > 0 else 1.0)), float(np.max(np.abs(g - r)))


def serving_leg(kw, M: int, rates, quantum: int, accuracy_requests: int, small_requests: int = 0) -> dict:
    """Serving stream sample: plain (one simulation at a time), streams
    (concurrent simulations on their own streams) and batched (shared
    launches, hipGraph quanta) on the same requests, burst and Poisson
    arrivals; each mode's error against the FP64 oracle on the smallest
    requests."""
    import concurrent.futures as cf

    from kwave_amd import configs

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(8) as ex:
        reqs = list(ex.map(configs.c5_request, range(M)))
    log(f"serving: generated {M} requests in {time.perf_counter() - t0:.1f} s")
    kw.run(configs.c5_request(0, grids=((32, 32, 32),), nt_range=(8, 8)))  # code objects loaded
    burst = [0.0] * M
    runs, outs = {}, {}
    for mode in SERVING_MODES:
        runs[mode], outs[mode] = _serve_stream(kw, reqs, burst, quantum, mode)
        log(f"serving: {mode} burst {runs[mode]}")
    plain = runs["plain"]
    grids = {}
    for f in reqs:
        g = "x".join(str(n) for n in _dims(f))
        grids[g] = grids.get(g, 0) + 1
    out = {"requests": M, "sample": f"c5_request(0..{M - 1}) of the seeded serving stream", "grids": grids,
           "updates": sum(_updates(f) for f in reqs), "steps_per_quantum": quantum,
           "burst": {**runs, **{f"speedup_{m}": round(runs[m]["updates_per_s"] / plain["updates_per_s"], 3)
                                for m in SERVING_MODES if m != "plain"}},
           # streams run autotuned launch shapes (equal to plain to FP32 rounding); batched runs
           # the plain launch shapes and must equal plain bit for bit
           "max_rel_l2_streams_vs_plain": max(_err(b, a)[0] for a, b in zip(outs["plain"], outs["streams"])),
           "batched_bit_identical_to_plain": bool(all(np.array_equal(a, b)
                                                      for a, b in zip(outs["plain"], outs["batched"]))),
           "poisson": []}
    if accuracy_requests > 0:
        from oracle import kwave_oracle  # accuracy check only

        order = sorted(range(M), key=lambda i: (int(np.prod(_dims(reqs[i]))), _updates(reqs[i])))[:accuracy_requests]
        acc = {m: {"max_rel_l2": 0.0, "max_abs": 0.0} for m in SERVING_MODES}
        acc["requests"] = []
        for i in order:
            t1 = time.perf_counter()
            ref = kwave_oracle.run(reqs[i], record=("p",))["p"]
            entry = {"request": i, "grid": list(_dims(reqs[i])), "Nt": int(np.asarray(reqs[i]["Nt"]).reshape(-1)[0]),
                     "peak_abs_p": float(np.max(np.abs(ref)))}
            for m in SERVING_MODES:
                e = _err(outs[m][i], ref)
                acc[m]["max_rel_l2"] = max(acc[m]["max_rel_l2"], e[0])
                acc[m]["max_abs"] = max(acc[m]["max_abs"], e[1])
                entry[f"{m}_rel_l2"], entry[f"{m}_max_abs"] = e
            acc["requests"].append(entry)
            log(f"serving: request {i} vs FP64 oracle ({time.perf_counter() - t1:.1f} s): {entry}")
        acc["tolerance_rel_l2"] = 1e-4
        acc["within_tolerance"] = bool(max(acc[m]["max_rel_l2"] for m in SERVING_MODES) <= 1e-4)
        out["accuracy_vs_oracle"] = acc
    del outs
    if small_requests > 0:
        # the serving stream's two smaller grid shapes only: grids that leave most of
        # the GPU idle one at a time, where shared launches pay most
        with cf.ThreadPoolExecutor(8) as ex:
            sreqs = list(ex.map(lambda i: configs.c5_request(i, grids=configs.C5_GRIDS[:2]), range(small_requests)))
        sb = {m: _serve_stream(kw, sreqs, [0.0] * small_requests, quantum, m)[0] for m in SERVING_MODES}
        out["small_grid_burst"] = {"requests": small_requests,
                                   "sample": f"c5_request(0..{small_requests - 1}) on the 64^3 and 96^3 grids",
                                   "updates": sum(_updates(f) for f in sreqs), **sb,
                                   **{f"speedup_{m}": round(sb[m]["updates_per_s"] / sb["plain"]["updates_per_s"], 3)
                                      for m in SERVING_MODES if m != "plain"}}
        log(f"serving: small-grid burst " + ", ".join(f"{m} {sb[m]['updates_per_s']:.3g}/s" for m in SERVING_MODES))
    capacity = M / plain["makespan_s"]  # plain-path requests per second
    rng = np.random.default_rng(5999)
    gaps = rng.exponential(1.0, M)
    for r in rates:
        lam = r * capacity
        arr = list(np.cumsum(gaps / lam) - gaps[0] / lam)
        res = {}
        for m in SERVING_MODES:
            res[m] = _serve_stream(kw, reqs, arr, quantum, m)[0]
            log(f"serving: poisson load {r}, {m}: {res[m]}")
        out["poisson"].append({"rate_per_s": round(lam, 3), "load_vs_plain_capacity": r, **res,
                               **{f"throughput_ratio_{m}": round(res[m]["updates_per_s"] / res["plain"]["updates_per_s"], 3)
                                  for m in SERVING_MODES if m != "plain"}})
        log(f"serving: poisson load {r}: " + ", ".join(
            f"{m} p50/p99 {res[m]['p50_latency_s']}/{res[m]['p99_latency_s']} s" for m in SERVING_MODES))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("c2", "c3", "c4"), default="c4",
                    help="c4 (default): 768^3, the largest single-GPU configuration; c2 128^3; c3 256x256x512")
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: 100; 1200 for c2)")
    ap.add_argument("--warmup", type=int, default=None, help="warmup steps (default: 10; 100 for c2)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="oracle steps of the CPU baseline sample (default: 1; 120 for c2)")
    ap.add_argument("--calib-copy-mib", type=int, default=1024,
                    help="one plain device copy of this many MiB (counter calibration); 0 skips")
    ap.add_argument("--ceiling-copy-gib", type=int, default=4,
                    help="streaming-copy sweep over this many GiB (the copy ceiling); 0 skips")
    ap.add_argument("--no-c2", action="store_true", help="skip the 128^3 second line")
    ap.add_argument("--serving-requests", type=int, default=48,
                    help="serving-stream sample size (0 skips the serving leg)")
    ap.add_argument("--serving-rates", default="0.5,0.9,1.5",
                    help="Poisson arrival rates as fractions of the plain path's capacity")
    ap.add_argument("--serving-quantum", type=int, default=16)
    ap.add_argument("--serving-small-requests", type=int, default=48,
                    help="a second burst on the 64^3 and 96^3 serving grids only (0 skips)")
    ap.add_argument("--accuracy-requests", type=int, default=2,
                    help="smallest serving requests checked against the FP64 oracle (both modes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "a multi-GPU run is not part of it.", file=sys.stderr)
        sys.exit(2)

    kw = import_pkg()
    wl = args.workload
    K = args.steps if args.steps is not None else (1200 if wl == "c2" else 100)
    W = args.warmup if args.warmup is not None else (100 if wl == "c2" else 10)

    copies = copy_legs(kw, args.calib_copy_mib, args.ceiling_copy_gib)
    line, f = measure_workload(kw, wl, K, W, keep_inputs=not args.no_cpu)
    out = {
        "metric": "grid-point-updates/s",
        "value": line["value"],
        "unit": "grid-point-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": line["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": line["data"],
        "config": line["config"],
        "roofline": line["roofline"],
        "step_roofline": line["step_roofline"],
    }
    for k in ("hbm_counters", "kernels"):
        if k in line:
            out[k] = line[k]
    out.update(copies)
    if "copy_ceiling" in copies:
        c = copies["copy_ceiling"]["GBs"]
        out["step_roofline"]["frac_of_copy_ceiling"] = round(out["step_roofline"]["achieved_GBs"] / c, 4)
        if "hbm_counters" in out:
            out["hbm_counters"]["frac_of_copy_ceiling"] = round(out["hbm_counters"]["achieved_GBs"] / c, 4)

    if wl != "c2" and not args.no_c2:
        c2, _ = measure_workload(kw, "c2", 1200, 100)
        out["c2"] = {k: c2[k] for k in ("value", "ms_per_step", "roofline", "step_roofline", "kernels",
                                        "hbm_counters") if k in c2}
        out["c2"]["workload"] = c2["config"]["workload"]

    if args.serving_requests > 0:
        rates = [float(x) for x in args.serving_rates.split(",") if x]
        out["serving"] = serving_leg(kw, args.serving_requests, rates, args.serving_quantum,
                                     args.accuracy_requests, args.serving_small_requests)

    if not args.no_cpu:
        from kwave_amd import configs

        cs = args.cpu_steps if args.cpu_steps is not None else (120 if wl == "c2" else 1)
        rec = {"c2": configs.C2_RECORD, "c3": configs.C3_RECORD, "c4": configs.C4_RECORD}[wl]
        log(f"cpu baseline: FP64 oracle, {cs} step(s) of {wl}")
        out["cpu_baseline"] = cpu_baseline(f, rec, cs, out["config"]["Nt"])
        log(f"cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
