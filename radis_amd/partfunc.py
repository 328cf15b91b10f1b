"""Tabulated partition functions (radis/levels/partfunc_cdsd.py:44-117).

The reference reads a whitespace table of Q(T) per isotopologue (comment
lines start with ``#``, a ``T(K)`` column, one column per isotopologue
named as in ``Itable``), interpolates each column with a cubic
``scipy.interpolate.splrep`` and evaluates it with ``splev`` inside 0.95
min(T) .. 1.05 max(T).  Here the same spline is handed to the device
(``LineDatabase.set_partition_table``), which evaluates Q(Tref)/Q(T) per
request itself.
"""
from __future__ import annotations

import numpy as np

# CO2 isotopologue columns of the CDSD table, radis/levels/partfunc_cdsd.py:74
CDSD_CO2_COLUMNS = {1: "12C16O2", 2: "13C16O2", 3: "16O12C18O", 4: "16O12C17O"}


def read_table(path):
    """(T, {column name: Q}) from a whitespace Q(T) table with a ``T(K)``
    header column; ``#`` starts a comment."""
    header, rows = None, []
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            if header is None:
                header = line.split()
                continue
            rows.append([float(v) for v in line.split()])
    if header is None or "T(K)" not in header:
        raise KeyError(f"Missing columns (T(K)) in {path}")
    a = np.asarray(rows, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] != len(header):
        raise ValueError(f"{path}: every row needs {len(header)} values")
    cols = {h: a[:, i] for i, h in enumerate(header)}
    return cols.pop("T(K)"), cols


def set_cdsd_co2_tables(db, path, isotopes=(1, 2, 3, 4)):
    """Give ``db`` the CDSD CO2 partition functions of ``path`` for the
    isotopologue slots ``isotopes`` (PartFuncCO2_CDSDtab's table)."""
    T, cols = read_table(path)
    for iso in isotopes:
        if iso not in CDSD_CO2_COLUMNS:
            raise KeyError(f"Isotope {iso} not defined in CDSD tabulated partition functions: "
                           f"{list(CDSD_CO2_COLUMNS)}")
        db.set_partition_table(iso, T, cols[CDSD_CO2_COLUMNS[iso]])
