#!/usr/bin/env python
"""Benchmark of the DIT/LDM spectral-synthesis path on one MI355X.

Default (no flags): every configuration in one JSON line.  The headline
(``value``) is the largest single-GPU configuration, ``h2o_100m``: 100 M
synthetic H2O/CH4-like lines, 1000-4000 cm-1 at wstep 0.005 cm-1 (600,001
grid points), T=2500 K, p=5 atm, x=0.1, RADIS's default dxG/dxL.  A step is
one full spectrum from the device-resident line list: axis reduction, axis
set-up, image-correction table, LDM fill, forward FFT pass, fused
lineshape/reduction FFT pass and inverse FFT pass.  Inputs are resident in
HBM before timing starts.  The line also carries ``co2_5m`` (5 M lines,
2000-2500 cm-1 at 0.001), ``batch_256`` (256 requests in one hipGraph
replay, every one checked against the CPU oracle at RADIS's FFT length) and
``serve`` (Poisson arrivals over two resident lists, p50/p99).

  --workload co_5k | co2_5m | h2o_100m      one single-spectrum line
  --workload batch_256                      the batched form alone
  --workload serve                          the served form alone
  --workload serve_sweep                    serving capacity: offered load
                                            from 200 req/s to saturation
  --window-inset CM                         with a single-spectrum workload:
                                            the window narrowed by CM cm-1 at
                                            both ends, strictly inside the
                                            list (the per-block axis reduction
                                            runs instead of the whole-list
                                            Pareto sets)
"""
import argparse
import dataclasses
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "lines/s (LDM fill+FFT) at ≤1e-4 rel err; batched spectra/s + p99 latency"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="all")
    ap.add_argument("--lines", type=int, default=0, help="override the line count")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-lines", type=int, default=0, help="lines in the CPU baseline sample (0 = auto)")
    ap.add_argument("--graph", action="store_true", help="replay a captured hipGraph per step")
    ap.add_argument("--serve-seconds", type=float, default=15.0, help="length of the Poisson request stream")
    ap.add_argument("--serve-rates", default="200,1000,3000,6000,10000,15000,20000",
                    help="offered loads (req/s) of --workload serve_sweep")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the fill tile width and the FFT schedule per size class before timing "
                         "(cache in a fresh temporary file)")
    ap.add_argument("--window-inset", type=float, default=0.0, metavar="CM",
                    help="narrow a single-spectrum workload's window by CM cm-1 at both ends")
    ap.add_argument("--copy-ref", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (HBM counter calibration)")
    return ap.parse_args()


def log(msg):
    """Progress on stderr (stdout carries only the JSON line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def stage_bytes(n_lines, info, db):
    """Algorithmic HBM bytes per launch of each kernel of one request.

    axis        : 12 B per scanned line (g_air, g_self, n_air f32) [+4 n_self] [+9 nu0 f64 + iso u8 with several
                  isotopologues]; the lines of whole blocks are replaced by their width candidates, 16 B each
                  (+ 2 nu0 per block and isotopologue with several)
    tile_bounds : 0 (a few dozen probes of nu0 per tile: latency, not bytes)
    wrap        : 0 (4 K bytes per L-axis node of the image-correction table)
    fill        : 28 B/line (nu0 f64; S_ref, El, g_air, g_self, n_air f32) [+4 n_self] [+4 delta] [+1 iso]
                  + 4*M*(N_v+1) LDM write
    fft_a       : 4*M*(N_v+1) LDM read + 8*P*L packed-row spectra write
    fft_b       : 8*P*L read + 8*L write
    fft_c       : 8*L read + 4*N_v abscoeff write
    """
    M = int(info.n_g * info.n_l)
    P = (M + 1) // 2
    L = int(info.fft_len)
    V = int(info.n_v) + 1
    ns = 4 if db.has_nself else 0
    per_fill = 28 + ns + (4 if db.has_delta else 0) + (1 if db.has_iso else 0)
    return {
        "axis": int(info.axis_scanned) * (12 + ns + (9 if db.has_iso else 0)) + 16 * int(info.axis_candidates),
        "finalize": 0,
        "tile_bounds": 0,
        "fill": n_lines * per_fill + 4 * M * V,
        "fft_a": 4 * M * V + 8 * P * L,
        "fft_b": 8 * P * L + 8 * L,
        "fft_c": 8 * L + 4 * int(info.n_v),
        "wrap": 0,
    }


def step_alg_bytes(n_lines, info, db):
    """Algorithmic bytes of one spectrum as the path's scope defines them:
    the line columns once, the (2 N_v x M) float32 LDM written by the fill
    and read by the forward FFT, the (N_v + 1)-bin spectrum written and read
    once, abscoeff written:
        b_line N_w + 4 M L + 4 M L + 8 (L/2+1) + 8 (L/2+1) + 4 N_v,  L = 2 N_v.
    The four-step intermediate this build writes and reads again, and the
    axis pass, count as overhead against it."""
    M = int(info.n_g * info.n_l)
    nv = int(info.n_v)
    L = 2 * nv
    b_line = 28 + (4 if db.has_nself else 0) + (4 if db.has_delta else 0) + (1 if db.has_iso else 0)
    return b_line * n_lines + 8 * M * L + 16 * (L // 2 + 1) + 4 * nv


def read_traffic(kernel, workload):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 --pmc
    summary, if present (profiles/<kernel>_<workload>_traffic.txt)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def radis_cost_model_s(n_g, n_l, n_v):
    """RADIS's own run-time estimate for broadening_method='fft',
    optimization='simple' at equilibrium (lbl/factory.py:2139-2146)."""
    import math

    return 4.675e-08 * (1 + n_l * n_g) * n_v * math.log(n_v)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(workload, lines_all, cond, n_sample, dev):
    """FP64 NumPy restatement of RADIS's CPU LDM-fft path (the oracle) timed on
    a bounded sample of the workload (``n_sample`` lines by a seeded choice,
    same grid), on one core and with all-core FFTs; on the same sample the
    GPU result is checked against it: relative error to the peak against
    RADIS's result (L = 2 N_v), the RADIS_LDM_PERIOD_L result against the
    oracle at the GPU's transform length, and the number of grid / log-width
    index mismatches."""
    import numpy as np

    from oracle import ldm_oracle as oracle
    from radis_amd import engine

    n_sample = min(n_sample, len(lines_all))
    if n_sample == len(lines_all):
        sub = lines_all
    else:
        idx = np.sort(np.random.default_rng(0).choice(len(lines_all), n_sample, replace=False))
        sub = lines_all.subset(idx)
    d = sub.as_dict()
    args = (d, cond.wmin, cond.wmax, cond.wstep, cond.Tgas, cond.pressure_bar, cond.mole_fraction, cond.dxG,
            cond.dxL)
    t0 = time.perf_counter()
    ref, det = oracle.eq_abscoeff(*args, Tref=cond.Tref, return_details=True)
    dt1 = time.perf_counter() - t0
    ncpu = os.cpu_count() or 1
    t0 = time.perf_counter()
    oracle.eq_abscoeff(*args, Tref=cond.Tref, fft_workers=ncpu)
    dtn = time.perf_counter() - t0
    # the same sample on the GPU: RADIS's result, and the circular result at
    # the transform length
    db = engine.LineDatabase(dev, sub)
    b = engine.SpectrumBatch(db, [engine.request_from_conditions(db, cond),
                                  engine.request_from_conditions(db, cond, period_fft_len=True)])
    gpu, gpu_period = b.results(b.run())
    fft_len = int(b.info(1).fft_len)
    ref_L = oracle.eq_abscoeff(*args, Tref=cond.Tref, L=fft_len)
    t = b.line_terms(0)
    mism = int(np.count_nonzero(t["k0"] != det["ki0"]) + np.count_nonzero(t["l0"] != det["li0"])
               + np.count_nonzero(t["m0"] != det["mi0"])) if t["k0"].size == det["ki0"].size else -1
    b.close()
    db.close()
    n_win = int(det["i1"] - det["i0"])
    return {
        "value": round(n_win / dt1, 1),
        "unit": "lines/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{n_sample} of the {len(lines_all)} {workload} lines ({n_win} in the window), full "
                  f"{cond.wmin:g}-{cond.wmax:g} cm-1 grid at wstep {cond.wstep:g}, L=2*N_v, one spectrum: "
                  f"{dt1:.2f} s on one core of {cpu_model()}",
        "all_cores": {"value": round(n_win / dtn, 1), "cores": ncpu,
                      "note": "same, with scipy.fft workers=os.cpu_count() (the scatter stays serial)"},
        "radis_cost_model_s": round(radis_cost_model_s(det["NG"], det["NL"], det["Nv"]), 3),
        "parity": {"sample_lines": n_win,
                   "rel_err_vs_radis_L_2Nv": float(oracle.rel_peak_error(gpu, ref)),
                   "rel_err_period_L_vs_oracle_same_L": float(oracle.rel_peak_error(gpu_period, ref_L)),
                   "fft_len": fft_len, "index_mismatches": mism, "tolerance": 1e-4},
    }


def run_autotune(db, reqs):
    """Tune every size class of ``reqs`` into a fresh temporary cache (which
    later ``make_request`` calls of this process read)."""
    import tempfile

    from radis_amd import autotune

    os.environ["RADIS_AMD_TUNE_CACHE"] = os.path.join(tempfile.mkdtemp(prefix="radis_amd_tune_"), "tune.json")
    t0 = time.perf_counter()
    entries = autotune.tune_many(db, reqs)
    dt = time.perf_counter() - t0
    if len(entries) == 1:
        (k, e), = entries.items()
        return {"key": k, **{kk: (round(v, 4) if isinstance(v, float) else v) for kk, v in e.items()},
                "tuning_s": round(dt, 2)}
    return {"size_classes": len(entries), "tuning_s": round(dt, 2),
            "compiled_fft_classes": sum(1 for e in entries.values() if not e["generic"])}


COMMON = {"metric": METRIC, "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
          "dtype": "f32 (f64 indices, Boltzmann, Q ratio)", "data": "synthetic"}


def bench_single(args, workload):
    import gc

    import numpy as np
    import torch

    from radis_amd import engine, workloads

    log(f"{workload}: generating lines")
    lines, cond = workloads.workload(workload, n=args.lines or None)
    if args.window_inset:
        cond = dataclasses.replace(cond, wmin=cond.wmin + args.window_inset, wmax=cond.wmax - args.window_inset)
    n_lines = len(lines)
    dev = engine.Device(0)
    log(f"{workload}: uploading {n_lines} lines")
    db = engine.LineDatabase(dev, lines)
    req = engine.request_from_conditions(db, cond)
    tuned = None
    if args.autotune:
        tuned = run_autotune(db, [req])
        req = engine.request_from_conditions(db, cond)
    batch = engine.SpectrumBatch(db, [req])
    outs = batch.alloc_outputs()
    stream = torch.cuda.current_stream()
    eager = lambda: batch.run(outs, stream)  # noqa: E731

    def timed_loop(step, n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n

    for _ in range(args.warmup):
        eager()
    torch.cuda.synchronize()
    inf = batch.info(0)
    if inf.status:
        raise RuntimeError(f"device status {inf.status}")
    n_win = int(inf.line_end - inf.line_begin)
    # 1) the timed steps: eager launches (or graph replays with --graph), no
    #    per-kernel events
    log(f"{workload}: timing {args.steps} steps")
    if args.graph:
        batch.capture(outs)
        replay = lambda: batch.replay(stream)  # noqa: E731
        for _ in range(args.warmup):
            replay()
        dt = timed_loop(replay, args.steps) * args.steps
    else:
        dt = timed_loop(eager, args.steps) * args.steps
    ms = dt / args.steps * 1e3
    value = n_win * args.steps / dt
    # 2) per-kernel device times from start/stop events of 20 separate runs;
    #    the device time of a run is the sum of its stages (median reported)
    batch.enable_timing(True)
    per_run = []
    for _ in range(20):
        eager()
        per_run.append(batch.stage_ms())
    batch.enable_timing(False)
    stages = {k: float(np.mean([r[k] for r in per_run])) for k in per_run[0]}
    device_ms = sorted(sum(r.values()) for r in per_run)
    device_ms_median = 0.5 * (device_ms[9] + device_ms[10])
    # 3) the other launch mode, for reference
    if args.graph:
        other = {"eager_ms_per_step": round(timed_loop(eager, args.steps) * 1e3, 4)}
    else:
        batch.capture(outs)
        replay = lambda: batch.replay(stream)  # noqa: E731
        for _ in range(args.warmup):
            replay()
        other = {"graph_ms_per_step": round(timed_loop(replay, args.steps) * 1e3, 4)}
    bytes_ = stage_bytes(n_win, inf, db)
    alg = step_alg_bytes(n_win, inf, db)
    dom = max(stages, key=stages.get)
    achieved = bytes_[dom] / (stages[dom] * 1e-3) / 1e9
    moved = sum(bytes_.values())
    roof = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": read_traffic(dom, workload),
        # whole step: the scope's algorithmic bytes per spectrum over the
        # timed step, and the bytes the kernels move (four-step intermediate
        # included) over the same time
        "step_alg_bytes": alg,
        "step_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "step_moved_bytes": moved,
        "step_moved_frac": round(moved / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "stage_ms": {k: round(v, 4) for k, v in stages.items()},
        "stage_GBps": {k: round(bytes_[k] / (v * 1e-3) / 1e9, 1) for k, v in stages.items() if v > 0},
        "fill_atomics_per_s": round(8 * n_win / (stages["fill"] * 1e-3), 1),
    }
    out = {
        **COMMON,
        "value": round(value, 1),
        "unit": "lines/s",
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "config": {
            "workload": workload,
            "lines": n_lines,
            "lines_in_window": n_win,
            "grid_points": int(inf.n_v),
            "fft_len": int(inf.fft_len),
            "n1_n2": [int(inf.n1), int(inf.n2)],
            "N_G": int(inf.n_g),
            "N_L": int(inf.n_l),
            "T_K": cond.Tgas,
            "p_bar": cond.pressure_bar,
            "x": cond.mole_fraction,
            "wstep": cond.wstep,
            "window": [cond.wmin, cond.wmax],
            # a window holding every line of a list without an isotopologue
            # column takes the list's Pareto sets in k_finalize; any other
            # runs k_axis over the window's blocks
            "axis_reduction": ("whole-list Pareto sets" if n_win == n_lines and lines.iso is None
                               else "k_axis over the window's blocks"),
            "axis_candidates": int(inf.axis_candidates),
            "axis_scanned_lines": int(inf.axis_scanned),
            "graph": bool(args.graph),
            "compiled_fft": bool(inf.compiled_fft),
            "tile_w": int(inf.tile_w),
            "autotune": tuned,
            "parallelism": "single-gpu",
        },
        "roofline": roof,
        "device_ms_median": round(device_ms_median, 4),
        "device_lines_per_s_median": round(n_win / (device_ms_median * 1e-3), 1),
        **other,
    }
    batch.close()
    if not args.no_cpu_baseline:
        n_cpu = args.cpu_lines or {"co_5k": 5000, "co2_5m": 5_000_000, "h2o_100m": 2_000_000}.get(workload, 500_000)
        log(f"{workload}: CPU baseline on {n_cpu} lines")
        out["cpu_baseline"] = cpu_baseline(workload, lines, cond, n_cpu, dev)
    db.close()
    dev.close()
    del lines, db
    gc.collect()
    log(f"{workload}: {value:.4g} lines/s, {ms:.4f} ms/step")
    return out


def copy_reference(mib):
    """Plain device copy of known size: the reference that FETCH_SIZE and
    WRITE_SIZE are scaled by when HBM traffic per launch is derived."""
    import torch

    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_reference_bytes_each_way": n * 4, "copies": 5}), flush=True)


# ---- oracle parity of every batched request, on the host's cores ----------
# Worker processes are started fresh (spawn) and build the line list from its
# seed themselves; they import neither torch nor the HIP library.
_W_LINES = None


def _parity_init(n_lines, seed):
    global _W_LINES
    from radis_amd import workloads

    _W_LINES = workloads.co2_like(n_lines, seed=seed).as_dict()


def _parity_task(task):
    from oracle import ldm_oracle as oracle

    i, c, got = task
    ref = oracle.eq_abscoeff(_W_LINES, *c)
    return i, float(oracle.rel_peak_error(got, ref))


def batch_parity(conds, results, n_lines, seed=1001, workers=0):
    """Every request against the oracle at RADIS's FFT length 2 N_v (on up to
    16 host processes): errors relative to each request's peak."""
    import multiprocessing as mp

    workers = workers or max(1, min(16, os.cpu_count() or 1))
    tasks = [(i, (c.wmin, c.wmax, c.wstep, c.Tgas, c.pressure_bar, c.mole_fraction, c.dxG, c.dxL), results[i])
             for i, c in enumerate(conds)]
    # biggest windows first, so the pool drains evenly
    tasks.sort(key=lambda t: -t[2].size)
    errs = [None] * len(conds)
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(workers, initializer=_parity_init, initargs=(n_lines, seed)) as pool:
        for i, e in pool.imap_unordered(_parity_task, tasks, chunksize=1):
            errs[i] = e
    return errs, time.perf_counter() - t0, workers


def bench_batch(args, parity=True):
    """256 requests over the 5 M-line list in one launch set replayed as a
    hipGraph, against 256 sequential single-request runs; every request
    checked against the oracle at L = 2 N_v."""
    import numpy as np
    import torch

    from radis_amd import engine, workloads

    n_list = args.lines or 5_000_000
    log("batch_256: generating lines")
    lines = workloads.co2_like(n_list)
    conds = workloads.random_requests(256, 2000.0, 2500.0, seed=1003, wsteps=(0.001,))
    dev = engine.Device(0)
    db = engine.LineDatabase(dev, lines)
    del lines
    reqs = [engine.request_from_conditions(db, c) for c in conds]
    tuned = None
    if args.autotune:
        tuned = run_autotune(db, reqs)
        reqs = [engine.request_from_conditions(db, c) for c in conds]
    batch = engine.SpectrumBatch(db, reqs)
    infos = [batch.info(i) for i in range(batch.n)]
    n_lines = sum(int(i.line_end - i.line_begin) for i in infos)
    stream = torch.cuda.Stream()
    # per-stage device times from eager runs
    outs = batch.alloc_outputs()
    batch.enable_timing(True)
    for _ in range(3):
        batch.run(outs, stream)
    stages = batch.stage_ms()
    batch.enable_timing(False)
    gouts = batch.capture(stream=stream)
    for _ in range(args.warmup):
        batch.replay(stream)
    torch.cuda.synchronize()
    log(f"batch_256: timing {args.steps} graph replays")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.replay(stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    res = batch.results(gouts)
    # sequential plain calls
    singles = [engine.SpectrumBatch(db, [r]) for r in reqs]
    souts = [b.alloc_outputs() for b in singles]
    for b, o in zip(singles, souts):
        b.run(o, stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for b, o in zip(singles, souts):
        b.run(o, stream)
    torch.cuda.synchronize()
    dt_seq = time.perf_counter() - t1
    mx = max(float(torch.max(torch.abs(a - c)).item() / max(torch.max(torch.abs(c)).item(), 1e-300))
             for a, c in zip(gouts, [o[0] for o in souts]))
    ws = batch.workspace_bytes()
    ratios = [int(i.fft_len) / (2 * int(i.n_v)) for i in infos]
    for b in singles:
        b.close()
    batch.close()
    db.close()
    dev.close()
    out = {
        **COMMON, "value": round(256 * args.steps / dt, 2), "unit": "spectra/s", "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3),
        "config": {"workload": "batch_256", "lines": n_list, "requests": 256, "window_lines_total": n_lines,
                   "wstep": 0.001, "graph": True, "workspace_bytes": ws, "autotune": tuned,
                   "fft_len_over_2Nv": [round(min(ratios), 4), round(float(np.median(ratios)), 4),
                                        round(max(ratios), 4)],
                   "parallelism": "single-gpu"},
        "lines_per_s": round(n_lines * args.steps / dt, 1),
        "sequential_256": {"spectra_per_s": round(256 / dt_seq, 2), "ms": round(dt_seq * 1e3, 3),
                           "speedup_batched": round((dt_seq / (dt / args.steps)), 2)},
        "batched_vs_sequential_max_rel_diff": mx,
        "stage_ms": {k: round(v, 4) for k, v in stages.items()},
    }
    if parity:
        log("batch_256: oracle parity of all 256 requests")
        errs, t_par, nw = batch_parity(conds, res, n_list)
        worst = int(np.argmax(errs))
        out["parity"] = {"requests": len(errs), "reference": "oracle at L = 2 N_v (RADIS CPU LDM)",
                         "max_rel_err": max(errs), "median_rel_err": float(np.median(errs)),
                         "worst_request": worst, "over_1e-4": int(sum(e > 1e-4 for e in errs)),
                         "over_1e-5": int(sum(e > 1e-5 for e in errs)), "tolerance": 1e-4,
                         "host_s": round(t_par, 1), "host_processes": nw}
    log(f"batch_256: {out['value']:.4g} spectra/s")
    return out


def _serve_setup(args):
    from radis_amd import engine, workloads

    dev = engine.Device(0)
    log("serve: generating and uploading the two lists")
    dbs = {"co2": engine.LineDatabase(dev, workloads.co2_like(5_000_000)),
           "h2o": engine.LineDatabase(dev, workloads.h2o_like(20_000_000, seed=1005))}
    return dev, dbs


def _serve_stream(dev, dbs, rate, dur, seed=1004):
    from radis_amd import serve, workloads

    t = serve.poisson_arrivals(rate, dur, seed=seed)
    ca = workloads.random_requests(len(t), 2000.0, 2500.0, seed=seed + 2, wsteps=(0.001, 0.002, 0.005))
    cb = workloads.random_requests(len(t), 1000.0, 4000.0, seed=seed + 3, wsteps=(0.001, 0.002, 0.005))
    arrivals = [(ti, "co2", ca[i]) if i % 2 == 0 else (ti, "h2o", cb[i]) for i, ti in enumerate(t)]
    srv = serve.SpectrumServer(dev, dbs, budget_bytes=96 << 30, max_batch=32, n_streams=4)
    # warm-up: a few requests through the same path
    srv.run_stream([(0.0, a[1], a[2]) for a in arrivals[:8]], timeout=120)
    srv.done.clear()
    t0 = time.perf_counter()
    done = srv.run_stream(arrivals, timeout=dur * 5 + 60)
    wall = time.perf_counter() - t0
    lat = serve.latency_summary(done)
    span = max(r.t_done for r in done) - min(r.t_arrival for r in done)
    return {"offered_per_s": rate, "duration_s": dur, "requests": len(done),
            "spectra_per_s": round(len(done) / span, 2), "lines_per_s": round(sum(r.n_lines for r in done) / span, 1),
            "latency": lat, "errors": sum(1 for r in done if r.error), "wall_s": round(wall, 2),
            "max_batches_in_flight": srv.max_in_flight}


def bench_serve(args):
    """Poisson arrivals at 200 req/s against two resident lists."""
    dev, dbs = _serve_setup(args)
    dur = args.serve_seconds
    log(f"serve: {dur:g} s at 200 req/s")
    r = _serve_stream(dev, dbs, 200.0, dur)
    for d in dbs.values():
        d.close()
    dev.close()
    return {
        **COMMON, "value": r["spectra_per_s"], "unit": "spectra/s", "steps": r["requests"], "warmup": 8,
        "ms_per_step": round(1e3 / max(r["spectra_per_s"], 1e-9), 3),
        "config": {"workload": "serve", "arrival_rate_per_s": 200, "duration_s": dur,
                   "dbs": {"co2": 5_000_000, "h2o": 20_000_000}, "wsteps": [0.001, 0.002, 0.005],
                   "max_batch": 32, "streams": 4, "parallelism": "single-gpu"},
        "latency": r["latency"], "errors": r["errors"], "lines_per_s": r["lines_per_s"], "wall_s": r["wall_s"],
        "max_batches_in_flight": r["max_batches_in_flight"],
    }


def bench_serve_sweep(args):
    """Serving capacity: one Poisson stream per offered load, from 200 req/s
    up, each --serve-seconds long; stops after the first rate whose served
    rate falls below 90 % of the offered one (saturation)."""
    dev, dbs = _serve_setup(args)
    rates = [float(x) for x in args.serve_rates.split(",")]
    rows = []
    for rate in rates:
        log(f"serve sweep: {rate:g} req/s for {args.serve_seconds:g} s")
        r = _serve_stream(dev, dbs, rate, args.serve_seconds, seed=1004 + len(rows))
        rows.append(r)
        log(f"serve sweep: {rate:g} offered -> {r['spectra_per_s']:g} served, p50 {r['latency']['p50_ms']} ms, "
            f"p99 {r['latency']['p99_ms']} ms, {r['max_batches_in_flight']} batches in flight at most")
        if r["errors"] or r["spectra_per_s"] < 0.9 * rate:
            break
    for d in dbs.values():
        d.close()
    dev.close()
    knee = max(rows, key=lambda r: r["spectra_per_s"])
    return {
        **COMMON, "value": knee["spectra_per_s"], "unit": "spectra/s", "steps": sum(r["requests"] for r in rows),
        "warmup": 8, "ms_per_step": round(1e3 / max(knee["spectra_per_s"], 1e-9), 3),
        "config": {"workload": "serve_sweep", "duration_s_per_rate": args.serve_seconds,
                   "dbs": {"co2": 5_000_000, "h2o": 20_000_000}, "wsteps": [0.001, 0.002, 0.005],
                   "max_batch": 32, "streams": 4, "parallelism": "single-gpu"},
        "rates": rows,
    }


SUMMARY_DROP = ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "dtype", "data")


def main():
    args = parse()
    if args.copy_ref:
        copy_reference(args.copy_ref)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the synthesis runs on one "
              f"MI355X (sharding across GPUs is designed, not implemented)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "all":
        # headline: the largest single-GPU configuration; the others ride along
        out = bench_single(args, "h2o_100m")
        co2 = bench_single(args, "co2_5m")
        bt = bench_batch(args)
        sv = bench_serve(args)
        for k, v in (("co2_5m", co2), ("batch_256", bt), ("serve", sv)):
            out[k] = {kk: vv for kk, vv in v.items() if kk not in SUMMARY_DROP}
    elif args.workload == "batch_256":
        out = bench_batch(args)
    elif args.workload == "serve":
        out = bench_serve(args)
    elif args.workload == "serve_sweep":
        out = bench_serve_sweep(args)
    else:
        out = bench_single(args, args.workload)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
