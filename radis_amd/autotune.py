"""Autotuner of the four-step FFT split and the fill tile width.

For a request, candidate FFT splits ``L = n1 * n2`` whose length is within
4 % of the smallest supported ``L >= 2 N_v``, and fill tile widths that fit
LDS, are timed on the device (sum of the per-stage HIP-event times over a
few runs); the fastest pair is stored in a JSON cache keyed by the grid
length and LDM step sizes and used by later requests with the same key.

Cache file: ``$RADIS_AMD_TUNE_CACHE`` or ``~/.cache/radis_amd/autotune.json``.
"""
from __future__ import annotations

import json
import math
import os

_CACHE = None
_MAX_N1, _MAX_N2 = 1024, 8192
_FILL_LDS = 60 * 1024


def cache_path():
    p = os.environ.get("RADIS_AMD_TUNE_CACHE")
    if p:
        return p
    home = os.path.expanduser("~")
    base = os.path.join(home, ".cache", "radis_amd") if os.access(home, os.W_OK) else "/tmp/radis_amd"
    return os.path.join(base, "autotune.json")


def _load():
    global _CACHE
    if _CACHE is None:
        try:
            with open(cache_path()) as f:
                _CACHE = json.load(f)
        except (OSError, ValueError):
            _CACHE = {}
    return _CACHE


def _save():
    path = cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(_CACHE, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def key_of(nv, dxG, dxL):
    return f"nv={int(nv)}/dxG={dxG:.6g}/dxL={dxL:.6g}"


def _grid_len(req):
    if req.n_v:
        return int(req.n_v)
    return int(math.ceil(((req.wmax + req.wstep) - req.wmin) / req.wstep))


def lookup(req):
    """(fft_n1, tile_w) tuned for this request's key, else (0, 0)."""
    c = _load()
    if not c:
        return 0, 0
    e = c.get(key_of(_grid_len(req), req.dxG, req.dxL))
    return (int(e["fft_n1"]), int(e["tile_w"])) if e else (0, 0)


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


def fft_candidates(nv, slack=1.04, max_n=6):
    """n1 values whose L = n1 * n2 (smallest valid n2) is within ``slack`` of
    the smallest, spread over the n1 range."""
    need = 2 * nv
    c2 = _smooth16(_MAX_N2)
    cands = []
    for n1 in _smooth16(_MAX_N1):
        n2min = -(-need // n1)
        n2 = next((v for v in c2 if v >= n2min), None)
        if n2 is not None:
            cands.append((n1 * n2, n1))
    if not cands:
        return []
    Lmin = min(L for L, _ in cands)
    good = sorted({n1 for L, n1 in cands if L <= slack * Lmin})
    if len(good) > max_n:
        idx = [round(i * (len(good) - 1) / (max_n - 1)) for i in range(max_n)]
        good = [good[i] for i in sorted(set(idx))]
    return good


def tune(db, req, reps=5, tile_widths=(64, 128, 256, 512, 1024), verbose=False):
    """Time the candidates for ``req`` on the device; store and return the
    best (fft_n1, tile_w, ms)."""
    import copy

    from . import _native, engine

    nv = _grid_len(req)
    _, info = db.request_bytes(req)
    m_cap = int(info.m_cap)
    widths = [w for w in tile_widths if m_cap * (w + 2) * 8 <= _FILL_LDS] or [0]
    results = []
    for n1 in fft_candidates(nv) or [0]:
        for w in widths:
            r = copy.copy(req)
            r.fft_n1, r.tile_w = n1, w
            try:
                b = engine.SpectrumBatch(db, [r])
            except _native.LdmError:
                continue
            outs = b.alloc_outputs()
            for _ in range(2):
                b.run(outs)
            b.enable_timing(True)
            for _ in range(reps):
                b.run(outs)
            ms = sum(b.stage_ms().values())
            b.close()
            results.append((ms, n1, w))
            if verbose:
                print(f"  n1={n1:5d} tile_w={w:5d}: {ms:.4f} ms", flush=True)
    if not results:
        return 0, 0, float("nan")
    ms, n1, w = min(results)
    c = _load()
    c[key_of(nv, req.dxG, req.dxL)] = {"fft_n1": n1, "tile_w": w, "ms": ms, "candidates": len(results)}
    _save()
    return n1, w, ms
