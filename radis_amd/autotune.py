"""Autotuner of a request's fill tile width and FFT schedule.

The FFT schedule is the split L = n1 * n2 together with the pass kernels:
compile-time-sized passes (fixed radix plans, for n1 in 2^a*{1,3,5} and n2 in
2^a) or run-time-sized ones (any 7-smooth split).  The fill kernel and the
FFT passes are separate kernels, so the two choices are tuned one after the
other on their own stage times: the tile width on the fill time with the
planner's FFT, then the FFT schedule on the time of the three FFT passes with
the chosen tile.  Each candidate's time is the median over a few timed runs
(per-kernel HIP events).

Results are cached in a JSON file keyed by the planner's default FFT length
and LDM row capacity (and the LDM steps), so requests of the same size class
share one tuning: ``$RADIS_AMD_TUNE_CACHE`` or
``~/.cache/radis_amd/autotune.json``.  ``engine.make_request`` applies a
cached entry unless the caller fixes the schedule.
"""
from __future__ import annotations

import copy
import json
import os

_CACHE = None
_CACHE_PATH = None
# sizes with compile-time-sized passes (mirrors FC_N1 / FC_N2 of the library;
# a candidate is kept only if the planned request reports compiled_fft)
FC_N1 = (256, 320, 384, 512, 640, 768, 1024)
FC_N2 = (64, 128, 256, 512, 1024, 2048)
TILE_WIDTHS = (16, 32, 64, 128, 256, 512)
FILL_LDS_MAX = 60 * 1024  # bytes of FP64 tile cells per fill workgroup
FFT_SLACK = 1.6           # compiled splits considered up to this factor above the smallest generic L
# a transform longer than 2 N_v needs room for the library's circular-image
# correction: L >= wrap_floor(2 N_v) (mirrors LDM_WRAP_* of the library)
WRAP_KMAX = 256
WRAP_C = 4.0 * 10.0 / (2.0 * 3.141592653589793)


def _wrap_bins(L, need):
    import math

    w = min(L - need, 0.25 * need)
    return math.ceil(WRAP_C * L / w)


def wrap_floor(need):
    """Smallest transform length above ``need`` = 2 N_v whose circular
    images the library can correct."""
    import math

    L = math.ceil(need * WRAP_KMAX / (WRAP_KMAX - WRAP_C))
    while L <= 40 * need and _wrap_bins(L, need) > WRAP_KMAX:
        L += 1
    return L


FILL_STAGES = ("tile_bounds", "fill")
FFT_STAGES = ("fft_a", "fft_b", "fft_c")


def cache_path():
    p = os.environ.get("RADIS_AMD_TUNE_CACHE")
    if p:
        return p
    home = os.path.expanduser("~")
    base = os.path.join(home, ".cache", "radis_amd") if os.access(home, os.W_OK) else "/tmp/radis_amd"
    return os.path.join(base, "autotune.json")


def _load():
    global _CACHE, _CACHE_PATH
    path = cache_path()
    if _CACHE is None or _CACHE_PATH != path:
        try:
            with open(path) as f:
                _CACHE = json.load(f)
        except (OSError, ValueError):
            _CACHE = {}
        _CACHE_PATH = path
    return _CACHE


def _save():
    path = cache_path()
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(_CACHE, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def clear():
    """Forget the cached entries (memory and file)."""
    global _CACHE
    _load()
    _CACHE = {}
    _save()


def _generic_flag():
    from . import _native

    return _native.GENERIC_FFT


def _untuned(req):
    r = copy.copy(req)
    r.fft_len, r.fft_n1, r.tile_w = 0, 0, 0
    r.flags = int(r.flags) & ~_generic_flag()
    return r


def _key_info(db, req):
    _, info = db.request_bytes(_untuned(req))
    return f"L={int(info.fft_len)}/mcap={int(info.m_cap)}/dxG={req.dxG:.6g}/dxL={req.dxL:.6g}", info


def key_of(db, req):
    """Size class of a request: the planner's FFT length and LDM row
    capacity for it (no overrides), and the LDM steps."""
    return _key_info(db, req)[0]


def lookup(db, req):
    """Tuned (fft_len, fft_n1, generic_fft, tile_w) for the request's size
    class, or None."""
    c = _load()
    if not c:
        return None
    e = c.get(key_of(db, req))
    if not e:
        return None
    return int(e["fft_len"]), int(e["fft_n1"]), bool(e["generic"]), int(e["tile_w"])


def apply(db, req):
    """Set a cached schedule on ``req`` in place; returns whether one was
    found.  Requests of one size class can differ slightly in N_v: the FFT
    schedule is applied only if its length is admissible for this request's
    2 N_v (2 N_v itself, or at least wrap_floor(2 N_v)); the tile width
    always."""
    c = _load()
    if not c:
        return False
    k, info = _key_info(db, req)
    e = c.get(k)
    if not e:
        return False
    req.tile_w = int(e["tile_w"])
    need, L = 2 * int(info.n_v), int(e["fft_len"])
    if L == need or L >= wrap_floor(need):
        req.fft_len, req.fft_n1 = int(e["fft_len"]), int(e["fft_n1"])
        req.flags = (int(req.flags) | _generic_flag()) if e["generic"] else (int(req.flags) & ~_generic_flag())
    return True


def _smooth16(limit):
    out = []
    for a in range(1, limit // 16 + 1):
        x = a
        for p in (2, 3, 5, 7):
            while x % p == 0:
                x //= p
        if x == 1:
            out.append(16 * a)
    return out


def fft_candidates(nv, n_generic=4):
    """(fft_len, n1, generic) candidates for a grid of nv points: the smallest
    admissible 7-smooth length (2 nv itself, else at least wrap_floor(2 nv))
    with up to ``n_generic`` splits (run-time-sized passes), and every
    admissible compiled split up to FFT_SLACK times that length."""
    need = 2 * nv
    lo = wrap_floor(need)
    c2 = _smooth16(8192)
    gen = []
    for n1 in _smooth16(1024):
        if need % n1 == 0 and need // n1 in c2:
            gen.append((need, n1))
            continue
        n2min = -(-lo // n1)
        n2 = next((v for v in c2 if v >= n2min), None)
        if n2 is not None:
            gen.append((n1 * n2, n1))
    if not gen:
        return []
    Lmin = min(L for L, _ in gen)
    best = sorted({n1 for L, n1 in gen if L == Lmin}, key=lambda n1: abs(n1 - 512))[:n_generic]
    out = [(Lmin, n1, True) for n1 in best]
    for n1 in FC_N1:
        for n2 in FC_N2:
            L = n1 * n2
            if (L == need or lo <= L) and L <= FFT_SLACK * Lmin:
                out.append((L, n1, False))
    return out


def fill_pad(w):
    """LDS row stride of a fill tile beyond its width (mirrors ldm_fill_pad)."""
    return 3 if w >= 64 else 2


def tile_candidates(m_cap):
    return [w for w in TILE_WIDTHS if m_cap * (w + fill_pad(w)) * 8 <= FILL_LDS_MAX]


def _time(db, req, stages, reps):
    """Median over ``reps`` timed runs of the summed ``stages`` (ms) and the
    planned info, or (None, None) if the schedule does not plan."""
    import statistics

    from . import _native, engine

    try:
        b = engine.SpectrumBatch(db, [req])
    except _native.LdmError:
        return None, None
    try:
        outs = b.alloc_outputs()
        b.run(outs)
        b.run(outs)
        b.enable_timing(True)
        ts = []
        for _ in range(reps):
            b.run(outs)
            st = b.stage_ms()
            ts.append(sum(st[s] for s in stages))
        return statistics.median(ts), b.info(0)
    finally:
        b.close()


def tune(db, req, reps=5, verbose=False):
    """Time the candidates for ``req``; store and return the best entry."""
    base = _untuned(req)
    _, info = db.request_bytes(base)
    nv, m_cap = int(info.n_v), int(info.m_cap)
    # 1) fill tile width with the planner's FFT
    fills = []
    for w in tile_candidates(m_cap):
        r = copy.copy(base)
        r.tile_w = w
        ms, _ = _time(db, r, FILL_STAGES, reps)
        if ms is not None:
            fills.append((ms, w))
            if verbose:
                print(f"  tile_w={w:4d}: fill {ms:.4f} ms", flush=True)
    fill_ms, tile_w = min(fills) if fills else (float("nan"), int(info.tile_w))
    # 2) FFT schedule with that tile
    ffts = []
    for L, n1, generic in fft_candidates(nv):
        r = copy.copy(base)
        r.tile_w, r.fft_len, r.fft_n1 = tile_w, L, n1
        if generic:
            r.flags = int(r.flags) | _generic_flag()
        ms, inf = _time(db, r, FFT_STAGES, reps)
        if ms is None or (not generic and not inf.compiled_fft):
            continue
        ffts.append((ms, L, n1, generic))
        if verbose:
            print(f"  L={L:8d} n1={n1:5d} {'generic ' if generic else 'compiled'}: fft {ms:.4f} ms", flush=True)
    if ffts:
        fft_ms, L, n1, generic = min(ffts)
    else:
        fft_ms, L, n1, generic = float("nan"), int(info.fft_len), int(info.n1), not bool(info.compiled_fft)
    entry = {"fft_len": int(L), "fft_n1": int(n1), "generic": bool(generic), "tile_w": int(tile_w),
             "fill_ms": fill_ms, "fft_ms": fft_ms, "candidates": len(fills) + len(ffts)}
    c = _load()
    c[key_of(db, req)] = entry
    _save()
    return entry


def tune_many(db, reqs, reps=5, verbose=False):
    """Tune one representative request per size class (the first request of
    each class without a cached entry); returns {key: entry}."""
    done = {}
    c = _load()
    for r in reqs:
        k = key_of(db, r)
        if k in done:
            continue
        if k in c:
            done[k] = c[k]
            continue
        done[k] = tune(db, r, reps=reps, verbose=verbose)
        if verbose:
            print(k, done[k], flush=True)
    return done
