#!/usr/bin/env python
"""Benchmark of the DIT/LDM spectral-synthesis path on one MI355X.

Default workload (``co2_5m``): 5 M synthetic CO2-like lines, 2000-2500 cm-1
at wstep 0.001 cm-1 (500,002 grid points), T=1500 K, p=1 atm, x=0.1.  A step
is one full spectrum from the device-resident line list: axis reduction,
axis set-up, LDM fill, forward FFT pass, fused lineshape/reduction FFT pass
and inverse FFT pass.  Inputs are resident in HBM before timing starts.

Prints one JSON line (rank 0).  Other workloads:
  --workload co_5k | co2_5m | h2o_100m      single spectra (lines/s)
  --workload batch_256                      256 requests in one launch set
                                            (spectra/s), hipGraph replay
  --workload serve                          Poisson 200 req/s over two
                                            resident lists (spectra/s, p50/p99)
"""
import argparse
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
    ap.add_argument("--workload", default="co2_5m")
    ap.add_argument("--lines", type=int, default=0, help="override the line count")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-lines", type=int, default=0, help="lines in the CPU baseline sample (0 = auto)")
    ap.add_argument("--graph", action="store_true", help="replay a captured hipGraph per step")
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="length of the Poisson request stream")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the fill tile width and the FFT schedule per size class before timing "
                         "(cache in a fresh temporary file)")
    ap.add_argument("--copy-ref", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (HBM counter calibration)")
    return ap.parse_args()


def stage_bytes(n_lines, info, db):
    """Algorithmic HBM bytes per launch of each kernel of one request.

    axis        : 12 B per scanned line (g_air, g_self, n_air f32) [+4 n_self] [+9 nu0 f64 + iso u8 with several
                  isotopologues]; the lines of whole blocks are replaced by their width candidates, 16 B each
                  (+ 2 nu0 per block and isotopologue with several)
    tile_bounds : 0 (a few dozen probes of nu0 per tile: latency, not bytes)
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
    }


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


def cpu_baseline(workload, lines_all, cond, n_sample, dev, fft_len):
    """FP64 NumPy restatement of RADIS's CPU LDM-fft path (the oracle) timed on
    a bounded sample of the workload (the first ``n_sample`` lines by a seeded
    choice, same grid), on one core and with all-core FFTs; on the same sample
    the GPU result is checked against it: relative error to the peak at
    L = 2 N_v (RADIS CPU semantics) and at the GPU's own L, and the number of
    grid / log-width index mismatches."""
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
    ref_L = oracle.eq_abscoeff(*args, Tref=cond.Tref, L=int(fft_len))
    # the same sample on the GPU
    db = engine.LineDatabase(dev, sub)
    req = engine.request_from_conditions(db, cond, fft_len=int(fft_len))
    b = engine.SpectrumBatch(db, [req])
    gpu = b.results(b.run())[0]
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
                   "rel_err_L_2Nv": float(oracle.rel_peak_error(gpu, ref)),
                   "rel_err_same_L": float(oracle.rel_peak_error(gpu, ref_L)),
                   "fft_len": int(fft_len), "index_mismatches": mism, "tolerance": 1e-4},
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


def bench_single(args):
    import numpy as np
    import torch

    from radis_amd import engine, workloads

    lines, cond = workloads.workload(args.workload, n=args.lines or None)
    n_lines = len(lines)
    dev = engine.Device(0)
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
    roof = None
    if stages:
        dom = max(stages, key=stages.get)
        achieved = bytes_[dom] / (stages[dom] * 1e-3) / 1e9
        roof = {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": read_traffic(dom, args.workload),
            "stage_ms": {k: round(v, 4) for k, v in stages.items()},
            "stage_GBps": {k: round(bytes_[k] / (v * 1e-3) / 1e9, 1) for k, v in stages.items() if v > 0},
            "fill_atomics_per_s": round(8 * n_win / (stages["fill"] * 1e-3), 1),
        }
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "lines/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 indices, Boltzmann, Q ratio)",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
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
    if not args.no_cpu_baseline:
        n_cpu = args.cpu_lines or {"co_5k": 5000, "co2_5m": 5_000_000, "h2o_100m": 2_000_000}.get(args.workload,
                                                                                                 500_000)
        out["cpu_baseline"] = cpu_baseline(args.workload, lines, cond, n_cpu, dev, inf.fft_len)
    batch.close()
    db.close()
    dev.close()
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


def bench_batch(args):
    """256 requests over the 5 M-line list in one launch set replayed as a
    hipGraph, against 256 sequential single-request runs."""
    import torch

    from radis_amd import engine, workloads

    lines = workloads.co2_like(args.lines or 5_000_000)
    conds = workloads.random_requests(256, 2000.0, 2500.0, seed=1003, wsteps=(0.001,))
    dev = engine.Device(0)
    db = engine.LineDatabase(dev, lines)
    reqs = [engine.request_from_conditions(db, c) for c in conds]
    tuned = None
    if args.autotune:
        tuned = run_autotune(db, reqs)
        reqs = [engine.request_from_conditions(db, c) for c in conds]
    batch = engine.SpectrumBatch(db, reqs)
    n_lines = sum(int(batch.infos[i].line_end - batch.infos[i].line_begin) for i in range(batch.n))
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
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.replay(stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = [batch.info(i).status for i in range(batch.n)]
    if any(st):
        raise RuntimeError(f"device status {st}")
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
             for a, c in zip(gouts[:16], [o[0] for o in souts[:16]]))
    ws = batch.workspace_bytes()
    for b in singles:
        b.close()
    batch.close()
    db.close()
    dev.close()
    value = 256 * args.steps / dt
    return {
        "metric": METRIC, "value": round(value, 2), "unit": "spectra/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 indices, Boltzmann, Q ratio)",
        "data": "synthetic",
        "config": {"workload": "batch_256", "lines": len(lines), "requests": 256, "window_lines_total": n_lines,
                   "wstep": 0.001, "graph": True, "workspace_bytes": ws, "autotune": tuned,
                   "parallelism": "single-gpu"},
        "lines_per_s": round(n_lines * args.steps / dt, 1),
        "sequential_256": {"spectra_per_s": round(256 / dt_seq, 2), "ms": round(dt_seq * 1e3, 3),
                           "speedup_batched": round((dt_seq / (dt / args.steps)), 2)},
        "batched_vs_sequential_max_rel_diff": mx,
        "stage_ms": {k: round(v, 4) for k, v in stages.items()},
    }


def bench_serve(args):
    """Poisson arrivals at 200 req/s against two resident lists."""
    from radis_amd import engine, serve, workloads

    dev = engine.Device(0)
    dbs = {"co2": engine.LineDatabase(dev, workloads.co2_like(5_000_000)),
           "h2o": engine.LineDatabase(dev, workloads.h2o_like(20_000_000, seed=1005))}
    dur = args.serve_seconds
    t = serve.poisson_arrivals(200.0, dur, seed=1004)
    ca = workloads.random_requests(len(t), 2000.0, 2500.0, seed=1006, wsteps=(0.001, 0.002, 0.005))
    cb = workloads.random_requests(len(t), 1000.0, 4000.0, seed=1007, wsteps=(0.001, 0.002, 0.005))
    arrivals = [(ti, "co2", ca[i]) if i % 2 == 0 else (ti, "h2o", cb[i]) for i, ti in enumerate(t)]
    srv = serve.SpectrumServer(dev, dbs, budget_bytes=96 << 30, max_batch=32, n_streams=4)
    # warm-up: a few requests through the same path
    srv.run_stream([(0.0, a[1], a[2]) for a in arrivals[:8]], timeout=120)
    srv.done.clear()
    t0 = time.perf_counter()
    done = srv.run_stream(arrivals, timeout=dur * 5 + 60)
    wall = time.perf_counter() - t0
    errors = [r for r in done if r.error]
    lat = serve.latency_summary(done)
    n_lines = sum(r.n_lines for r in done)
    span = max(r.t_done for r in done) - min(r.t_arrival for r in done)
    for d in dbs.values():
        d.close()
    dev.close()
    return {
        "metric": METRIC, "value": round(len(done) / span, 2), "unit": "spectra/s", "n_gpus": 1,
        "steps": len(done), "warmup": 8, "ms_per_step": round(span / max(len(done), 1) * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (f64 indices, Boltzmann, Q ratio)", "data": "synthetic",
        "config": {"workload": "serve", "arrival_rate_per_s": 200, "duration_s": dur,
                   "dbs": {"co2": 5_000_000, "h2o": 20_000_000}, "wsteps": [0.001, 0.002, 0.005],
                   "max_batch": 32, "streams": 4, "parallelism": "single-gpu"},
        "latency": lat, "errors": len(errors), "lines_per_s": round(n_lines / span, 1), "wall_s": round(wall, 2),
    }


def main():
    args = parse()
    if args.copy_ref:
        copy_reference(args.copy_ref)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the synthesis runs on one "
              f"MI355X (sharding across GPUs is designed, not implemented)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch_256":
        out = bench_batch(args)
    elif args.workload == "serve":
        out = bench_serve(args)
    else:
        out = bench_single(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
