"""Reader of HITRAN 160-character ``.par`` line records into a line list.

Fixed-width fields (HITRAN 2004 format, as read by RADIS's ``hit2df``,
radis/api/hitranapi.py): molecule (2), isotopologue (1), nu (12), S (10),
A (10), gamma_air (5), gamma_self (5), E'' (10), n_air (4), delta_air (8),
and the upper-state weight g' (7) at columns 147-153.
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
# upper-state statistical weight g' (HITRAN 2004 columns 147-153), read when present
_GU = ("gu", 146, 153, float)


def read_par(path):
    """Columns of a .par file as a dict of numpy arrays (file order)."""
    cols = {k: [] for k, *_ in _FIELDS}
    gu = []
    with open(path) as f:
        for line in f:
            rec = line.rstrip("\n")
            if len(rec) < 67:
                continue
            for k, a, b, conv in _FIELDS:
                cols[k].append(conv(rec[a:b]))
            gu.append(float(rec[_GU[1]:_GU[2]]) if len(rec) >= _GU[2] and rec[_GU[1]:_GU[2]].strip() else np.nan)
    out = {k: np.asarray(v) for k, v in cols.items()}
    out["gu"] = np.asarray(gu, dtype=np.float64)
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
