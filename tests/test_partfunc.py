"""Tabulated partition functions on the CPU: the oracle's FITPACK splev
restatement and the CDSD table reader (radis/levels/partfunc_cdsd.py:44-117).

The fixture ``cdsd_co2_partition_functions.txt`` is the reference's own data
file ``radis/db/CO2/partition_functions.txt`` (CDSD-4000 Q(T), 296-6000 K).
"""
import numpy as np
import pytest
from scipy.interpolate import splev, splrep

from oracle import ldm_oracle as oracle
from radis_amd import partfunc

FIX = "tests/golden/cdsd_co2_partition_functions.txt"


def _tables():
    T, cols = partfunc.read_table(FIX)
    return T, {iso: cols[name] for iso, name in partfunc.CDSD_CO2_COLUMNS.items()}


def test_reader_columns():
    T, cols = partfunc.read_table(FIX)
    assert T[0] == 296.0 and T[-1] == 6000.0 and np.all(np.diff(T) > 0)
    assert set(cols) == set(partfunc.CDSD_CO2_COLUMNS.values())
    assert cols["12C16O2"][0] == 291.045 and cols["16O12C17O"][-1] == 28399312.0


def test_reader_needs_temperature_column(tmp_path):
    p = tmp_path / "q.txt"
    p.write_text("# no T column\n X Y\n 1 2\n")
    with pytest.raises(KeyError):
        partfunc.read_table(str(p))


def test_golden_values_of_the_reference_test():
    """radis/test/levels/test_partfunc.py:151-152: Q of 12C16O2 from the
    CDSD table is 291.0447781984652 at 300 K and 114689.88454184022 at
    3000 K, with rtol 1e-3 there: the table's own 3000 K entry is 114762,
    0.06 % above, so that tolerance is the golden values' own."""
    T, q = _tables()
    t, c, k = splrep(T, q[1])
    got = oracle.splev_fitpack(t, c, k, np.array([300.0, 3000.0]))
    np.testing.assert_allclose(got, [291.0447781984652, 114689.88454184022], rtol=1e-3)


@pytest.mark.parametrize("iso", [1, 2, 3, 4])
def test_oracle_splev_matches_scipy(iso):
    """The restatement against scipy's own splev, inside the table, on the
    knots, and over the 5 % extrapolation margins, for cubic and linear
    splines (a wrong interval, index or recursion term fails here)."""
    T, q = _tables()
    x = np.concatenate([np.linspace(0.95 * T[0], 1.05 * T[-1], 397), T, [T[0] * 0.95, T[-1] * 1.05]])
    for k in (3, 1):
        t, c, kk = splrep(T, q[iso], k=k)
        np.testing.assert_allclose(oracle.splev_fitpack(t, c, kk, x), splev(x, (t, c, kk)), rtol=1e-13)


def test_interpolating_spline_reproduces_the_table():
    T, q = _tables()
    for iso in (1, 2, 3, 4):
        t, c, k = splrep(T, q[iso])
        np.testing.assert_allclose(oracle.splev_fitpack(t, c, k, T), q[iso], rtol=1e-12)
