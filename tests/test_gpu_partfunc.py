"""Tabulated partition functions evaluated on the device (PartFuncCO2_CDSDtab,
radis/levels/partfunc_cdsd.py:84-117, feeding calc_linestrength_eq's
Qref/Q, radis/lbl/base.py:2605-2697) against the oracle's FITPACK splev, on
the reference's CDSD CO2 table (fixture), with a synthetic four-isotopologue
CO2 line list.
"""
import numpy as np
import pytest
from scipy.interpolate import splrep

from oracle import ldm_oracle as oracle
from radis_amd import _native, engine, partfunc

pytestmark = pytest.mark.gpu

FIX = "tests/golden/cdsd_co2_partition_functions.txt"
# 12C16O2, 13C16O2, 16O12C18O, 16O12C17O molar masses (g/mol), radis/db/molparam.txt:26-29
MM = [0.0, 43.98983, 44.993185, 45.994076, 44.994045]


def _splines():
    T, cols = partfunc.read_table(FIX)
    return {iso: splrep(T, cols[name]) for iso, name in partfunc.CDSD_CO2_COLUMNS.items()}


def _lines(n=40_000, seed=7):
    rng = np.random.default_rng(seed)
    nu = np.sort(rng.uniform(2200.0, 2300.0, n))
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    sp = _splines()
    return dict(nu0=nu, S_ref=f32(10.0 ** rng.uniform(-26.0, -19.0, n)), El=f32(rng.exponential(2500.0, n)),
                g_air=f32(rng.uniform(0.05, 0.08, n)), g_self=f32(rng.uniform(0.07, 0.1, n)),
                n_air=f32(rng.uniform(0.6, 0.8, n)), iso=rng.integers(1, 5, n).astype(np.uint8),
                molar_mass=np.array(MM), Q=lambda s, T: float(oracle.splev_fitpack(*sp[s], T)))


@pytest.fixture(scope="module")
def dev():
    d = engine.Device(0)
    yield d
    d.close()


def test_device_evaluation_matches_oracle(dev):
    db = engine.LineDatabase(dev, _lines(2000))
    partfunc.set_cdsd_co2_tables(db, FIX)
    sp = _splines()
    T = np.concatenate([np.linspace(0.95 * 296.0, 1.05 * 6000.0, 501), [296.0, 300.0, 3000.0, 6000.0]])
    for iso in (1, 2, 3, 4):
        got = db.partition_function(iso, T)
        ref = oracle.splev_fitpack(*sp[iso], T)
        np.testing.assert_allclose(got, ref, rtol=1e-12)
    # the reference test's golden values (radis/test/levels/test_partfunc.py:151-152, rtol 1e-3)
    np.testing.assert_allclose(db.partition_function(1, [300.0, 3000.0]), [291.0447781984652, 114689.88454184022],
                               rtol=1e-3)
    db.close()


@pytest.mark.parametrize("Tgas", [300.0, 1500.0, 5500.0])
def test_device_q_spectrum(dev, Tgas):
    lines = _lines()
    db = engine.LineDatabase(dev, {k: v for k, v in lines.items() if k != "Q"})
    partfunc.set_cdsd_co2_tables(db, FIX)
    wmin, wmax, wstep, p, x = 2200.0, 2300.0, 0.002, 1.0, 0.1
    q_ref = np.ones(16)
    for s in (1, 2, 3, 4):
        q_ref[s] = lines["Q"](s, 296.0) / lines["Q"](s, Tgas)
    r_dev = engine.make_request(db, wmin, wmax, wstep, Tgas, p, x)  # tables cover every slot: device Q
    assert r_dev.flags & _native.DEVICE_Q
    r_host = engine.make_request(db, wmin, wmax, wstep, Tgas, p, x, q_ratio=q_ref)
    assert not r_host.flags & _native.DEVICE_Q
    b = engine.SpectrumBatch(db, [r_dev, r_host])
    a_dev, a_host = b.results(b.run())
    assert oracle.rel_peak_error(a_dev, a_host) <= 1e-6
    ref = oracle.eq_abscoeff(lines, wmin, wmax, wstep, Tgas, p, x)
    assert oracle.rel_peak_error(a_dev, ref) <= 1e-4
    b.close()
    db.close()


def test_device_q_errors(dev):
    lines = _lines(5000)
    db = engine.LineDatabase(dev, {k: v for k, v in lines.items() if k != "Q"})
    T, cols = partfunc.read_table(FIX)
    db.set_partition_table(1, T, cols["12C16O2"])
    # slots 2-4 have lines but no table
    req = engine.make_request(db, 2200.0, 2300.0, 0.01, 1000.0, 1.0, 0.1, device_q=True)
    with pytest.raises(_native.LdmError):
        engine.SpectrumBatch(db, [req])
    partfunc.set_cdsd_co2_tables(db, FIX)
    # below the table's 5 % margin (0.95 * 296 K)
    req = engine.make_request(db, 2200.0, 2300.0, 0.01, 250.0, 1.0, 0.1, device_q=True)
    with pytest.raises(_native.LdmError):
        engine.SpectrumBatch(db, [req])
    with pytest.raises(ValueError):
        engine.make_request(db, 2200.0, 2300.0, 0.01, 1000.0, 1.0, 0.1, device_q=True, Tvib=2000.0, Trot=1000.0)
    with pytest.raises(_native.LdmError):
        db.partition_function(7, [300.0])
    db.close()


def test_factory_parfunc_file():
    """SpectrumFactory with a tabulated partition-function file (the
    load_databank ``parfunc`` argument, radis/lbl/loader.py:1023-1060, 1188):
    Q comes from the table, evaluated on the device."""
    from radis_amd.factory import SpectrumFactory

    lines = _lines(20_000, seed=11)
    cols = {k: v for k, v in lines.items() if k != "Q"}
    sf = SpectrumFactory(2200.0, 2300.0, wstep=0.002, pressure=1.0, mole_fraction=0.1)
    sf.load_lines(cols, parfunc=FIX, parfuncfmt="cdsd")
    s = sf.eq_spectrum_gpu(Tgas=1800.0)
    _, k = s.get("abscoeff")
    ref = oracle.eq_abscoeff(lines, 2200.0, 2300.0, 0.002, 1800.0, 1.0, 0.1)
    assert oracle.rel_peak_error(k, ref) <= 1e-4
    k2 = s.recalc_gpu("abscoeff", Tgas=900.0)
    ref2 = oracle.eq_abscoeff(lines, 2200.0, 2300.0, 0.002, 900.0, 1.0, 0.1)
    assert oracle.rel_peak_error(k2, ref2) <= 1e-4
    s.exit_gpu()
    with pytest.raises(NotImplementedError):
        SpectrumFactory(2200.0, 2300.0).load_lines(cols, parfunc=FIX, parfuncfmt="hapi")
