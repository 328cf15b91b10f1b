"""Reader of HITRAN 160-character ``.par`` line records into a line list.

Fixed-width fields (HITRAN 2004 format, as read by RADIS's ``hit2df``,
radis/api/hitranapi.py): molecule (2), isotopologue (1), nu (12), S (10),
A (10), gamma_air (5), gamma_self (5), E'' (10), n_air (4), delta_air (8),
and, where present, the global quanta (15 + 15, columns 68-97) and the
statistical weights g', g'' (7 + 7, columns 147-160).
"""
from __future__ import annotations

import numpy as np

from .workloads import LineList

# HITRAN molparam molar masses (g/mol) of the first CO isotopologues,
# radis/db/molparam.txt:51-56
CO_MOLAR_MASS = [0.0, 27.994915, 28.998270, 29.999161, 28.999130, 31.002516, 30.002485]

_FIELDS = [
    ("mol", 0, 2, int),
    ("iso", 2, 3, int),
    ("nu0", 3, 15, float),
    ("S_ref", 15, 25, float),
    ("A", 25, 35, float),
    ("g_air", 35, 40, float),
    ("g_self", 40, 45, float),
    ("El", 45, 55, float),
    ("n_air", 55, 59, float),
    ("delta", 59, 67, float),
]
# read when present: upper/lower global quanta (columns 68-82, 83-97) and
# upper/lower statistical weights g', g'' (columns 147-153, 154-160)
_GQ_U, _GQ_L = (67, 82), (82, 97)
_GU, _GL = (146, 153), (153, 160)


def _open_text(path):
    """A .par file, or one compressed with bzip2 or gzip (HITEMP distributes
    its .par files as .bz2)."""
    p = str(path)
    if p.endswith(".bz2"):
        import bz2

        return bz2.open(p, "rt")
    if p.endswith(".gz"):
        import gzip

        return gzip.open(p, "rt")
    return open(p)


def read_par(path):
    """Columns of a .par file (plain, .bz2 or .gz) as a dict of numpy arrays
    (file order)."""
    cols = {k: [] for k, *_ in _FIELDS}
    extra = {"gu": [], "gl": [], "gq_u": [], "gq_l": []}

    def field(rec, ab, conv, missing):
        a, b = ab
        txt = rec[a:b] if len(rec) >= b else ""
        return conv(txt) if txt.strip() else missing

    with _open_text(path) as f:
        for line in f:
            rec = line.rstrip("\n")
            if len(rec) < 67:
                continue
            for k, a, b, conv in _FIELDS:
                cols[k].append(conv(rec[a:b]))
            extra["gu"].append(field(rec, _GU, float, np.nan))
            extra["gl"].append(field(rec, _GL, float, np.nan))
            extra["gq_u"].append(field(rec, _GQ_U, str.strip, ""))
            extra["gq_l"].append(field(rec, _GQ_L, str.strip, ""))
    out = {k: np.asarray(v) for k, v in cols.items()}
    out["gu"] = np.asarray(extra["gu"], dtype=np.float64)
    out["gl"] = np.asarray(extra["gl"], dtype=np.float64)
    out["gq_u"] = np.asarray(extra["gq_u"])
    out["gq_l"] = np.asarray(extra["gq_l"])
    out["iso"] = out["iso"].astype(np.int64)
    return out


def line_list(path, molar_mass=CO_MOLAR_MASS, q_coef=(0.36, 1.0), with_shift=True):
    """LineList sorted by nu0, float32 storage columns, optional pressure
    shift, isotopologue slots as in the file."""
    c = read_par(path)
    order = np.argsort(c["nu0"], kind="stable")
    f32 = lambda a: np.ascontiguousarray(a[order], dtype=np.float32)
    return LineList(
        nu0=np.ascontiguousarray(c["nu0"][order]),
        S_ref=f32(c["S_ref"]), El=f32(c["El"]), g_air=f32(c["g_air"]), g_self=f32(c["g_self"]),
        n_air=f32(c["n_air"]), molar_mass=np.asarray(molar_mass, dtype=np.float64), q_coef=q_coef,
        delta=f32(c["delta"]) if with_shift else None,
        iso=np.ascontiguousarray(c["iso"][order], dtype=np.uint8),
        name="hitran_par",
    )
