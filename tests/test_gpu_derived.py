"""Derived quantities written by the last GPU pass (absorbance, emissivity,
transmittance, radiance; radis/lbl/factory.py:1289-1297) and the
``SpectrumFactory`` facade that reads them (radis/test/gpu/test_gpu.py).

Bars: the derived arrays against the oracle's restatement applied to the
GPU's own abscoeff to FP32 rounding (the epilogue alone), and against the
oracle's full chain within 1e-4 of each quantity's peak (transmittance: 1e-4
absolute).
"""
import numpy as np
import pytest

from oracle import ldm_oracle as oracle
from radis_amd import _native, engine, factory, hitran, workloads

pytestmark = pytest.mark.gpu

HITRAN = "tests/golden/hitran_co_3iso_2000_2300cm.par"
REL_TOL = 1e-4
KEYS = ("absorbance", "emissivity", "transmittance", "radiance")


@pytest.fixture(scope="module")
def dev():
    d = engine.Device(0)
    yield d
    d.close()


def _err(key, got, ref):
    if key == "transmittance":
        return float(np.max(np.abs(got - ref)))
    return oracle.rel_peak_error(got, ref)


def _check_epilogue(c, abscoeff, derived, path_length, tol=2e-6):
    """Derived arrays vs the oracle formulas on the GPU's own abscoeff."""
    w = oracle.wavenumber_grid(c.wmin, c.wmax, c.wstep)
    ref = oracle.derived_quantities(w, abscoeff, path_length, c.Tgas)
    for k in KEYS:
        e = _err(k, derived[k], ref[k])
        assert e <= tol, (k, e)


def test_derived_match_oracle_compiled_and_generic(dev):
    lines = hitran.line_list(HITRAN)
    c = workloads.Conditions(2000.0, 2300.0, 0.002, 1500.0, 2.0, 0.5)
    db = engine.LineDatabase(dev, lines)
    rc = engine.request_from_conditions(db, c)
    rg = engine.make_request(db, c.wmin, c.wmax, c.wstep, c.Tgas, c.pressure_bar, c.mole_fraction, c.dxG, c.dxL,
                             generic_fft=True)
    b = engine.SpectrumBatch(db, [rc, rg])
    plain = b.results(b.run())
    assert b.info(0).compiled_fft == 1 and b.info(1).compiled_fft == 0
    L_path = 12.5
    b.set_derived(0, L_path)
    b.set_derived(1, L_path)
    res = b.results(b.run())
    for i in range(2):
        np.testing.assert_array_equal(res[i], plain[i])  # abscoeff untouched by the epilogue
        d = b.derived_results(i)
        _check_epilogue(c, res[i], d, L_path)
        inf = b.info(i)
        ref_a = oracle.eq_abscoeff(lines.as_dict(), c.wmin, c.wmax, c.wstep, c.Tgas, c.pressure_bar,
                                   c.mole_fraction, c.dxG, c.dxL)
        ref = oracle.derived_quantities(oracle.wavenumber_grid(c.wmin, c.wmax, c.wstep), ref_a, L_path, c.Tgas)
        for k in KEYS:
            e = _err(k, d[k], ref[k])
            print(f"req {i} {k}: {e:.2e}")
            assert e <= REL_TOL, (i, k, e)
        assert d["emissivity"].max() > 0.5  # optically thick peaks exercise the expm1 range
    b.close()
    db.close()


def test_derived_selection_graph_replay_and_off(dev):
    lines = workloads.co2_like(100_000, seed=31)
    c = workloads.Conditions(2200.0, 2240.0, 0.001, 800.0, 1.0, 0.2)
    db = engine.LineDatabase(dev, lines)
    b = engine.SpectrumBatch(db, [engine.request_from_conditions(db, c)])
    tens = b.set_derived(0, 3.0, which=("transmittance", "radiance"))
    assert sorted(tens) == ["radiance", "transmittance"]
    eager = b.results(b.run())[0]
    d0 = b.derived_results(0)
    assert sorted(d0) == ["radiance", "transmittance"]
    _check_epilogue(c, eager, {**oracle.derived_quantities(oracle.wavenumber_grid(c.wmin, c.wmax, c.wstep),
                                                           eager, 3.0, c.Tgas), **d0}, 3.0)
    outs = b.capture()
    for rep in range(3):
        for t in tens.values():
            t.fill_(-1.0)
        b.replay()
        got = b.results(outs)[0]
        dg = b.derived_results(0)
        assert oracle.rel_peak_error(got, eager) <= 1e-6
        for k in dg:
            assert _err(k, dg[k], d0[k]) <= 1e-6, (rep, k)
        _ = b.results(b.run())  # eager work between replays
    # a new path length on the next replay
    tens = b.set_derived(0, 6.0, which=("absorbance",))
    b.replay()
    got = b.results(outs)[0]
    a6 = b.derived_results(0)["absorbance"]
    np.testing.assert_allclose(a6, 6.0 * got, rtol=1e-6, atol=1e-12 * got.max())
    # switched off: the former tensors are left alone
    t_abs = tens["absorbance"]
    t_abs.fill_(-2.0)
    b.set_derived(0, 0.0, which=())
    b.results(b.run())
    assert b.derived_results(0) == {}
    assert float(t_abs.min()) == -2.0 and float(t_abs.max()) == -2.0
    b.close()
    db.close()


def test_derived_with_power_of_two_scale(dev):
    """S(T) beyond the FP32 range: abscoeff and absorbance are stored times
    2**-scale_log2, emissivity/transmittance/radiance are true values."""
    lines = workloads.h2o_like(200_000, seed=8, chunk=200_000)
    c = workloads.Conditions(2000.0, 2100.0, 0.005, 2500.0, 5.0 * workloads.ATM, 0.1)
    db = engine.LineDatabase(dev, lines)
    b = engine.SpectrumBatch(db, [engine.request_from_conditions(db, c)])
    b.set_derived(0, 1e-3)
    res = b.results(b.run())[0]
    assert b.info(0).scale_log2 != 0
    d = b.derived_results(0)
    _check_epilogue(c, res, d, 1e-3)
    assert d["emissivity"].max() == 1.0 and d["transmittance"].min() == 0.0
    b.close()
    db.close()


def test_set_derived_rejects_bad_arguments(dev):
    lines = workloads.co_like(1000, seed=4)
    db = engine.LineDatabase(dev, lines)
    b = engine.SpectrumBatch(db, [engine.make_request(db, 2100.0, 2110.0, 0.01, 1000.0, 1.0, 0.1)])
    with pytest.raises(_native.LdmError):
        b.set_derived(0, -1.0)
    with pytest.raises(_native.LdmError):
        b.set_derived(3, 1.0)
    with pytest.raises(ValueError):
        b.set_derived(0, 1.0, which=("absorbance", "opacity"))
    b.close()
    db.close()


# --- facade: radis/test/gpu/test_gpu.py -------------------------------------------
def _oracle_spectrum(lines, wmin, wmax, wstep, T, p, x, path_length):
    a = oracle.eq_abscoeff(lines.as_dict(), wmin, wmax, wstep, T, p, x, engine.DXG_DEFAULT, engine.DXL_DEFAULT)
    d = oracle.derived_quantities(oracle.wavenumber_grid(wmin, wmax, wstep), a, path_length, T)
    return dict(abscoeff=a, absorbance=d["absorbance"], emissivity_noslit=d["emissivity"],
                transmittance_noslit=d["transmittance"], radiance_noslit=d["radiance"])


def test_eq_spectrum_gpu_matches_oracle():
    """test_gpu.py:62-112: every quantity of eq_spectrum_gpu against the
    CPU chain (here the oracle, at RADIS's FFT length 2 N_v)."""
    lines = hitran.line_list(HITRAN)
    sf = factory.SpectrumFactory(wavenum_min=2100.0, wavenum_max=2200.0, wstep=0.001, pressure=0.1,
                                 mole_fraction=0.01, path_length=1, lines=lines)
    s = sf.eq_spectrum_gpu(Tgas=1000.0, name="GPU")
    ref = _oracle_spectrum(lines, 2100.0, 2200.0, 0.001, 1000.0, 0.1, 0.01, 1.0)
    for var in ("abscoeff", "absorbance", "emissivity_noslit", "transmittance_noslit", "radiance_noslit"):
        w, y = s.get(var)
        assert y.shape == w.shape == ref[var].shape
        e = (float(np.max(np.abs(y - ref[var]))) if var == "transmittance_noslit"
             else oracle.rel_peak_error(y, ref[var]))
        print(f"{var}: {e:.2e}")
        assert e <= REL_TOL, (var, e)
    assert s.conditions["NwG"] > 0 and s.conditions["NwL"] > 0
    s.exit_gpu()
    assert s.gpu_app is None


def test_multiple_gpu_calls_and_recalc_path_length():
    """test_gpu.py:139-194: recalc_gpu with 3x the path gives 3x the
    absorbance; a second factory call after exit_gpu gives 2x."""
    lines = hitran.line_list(HITRAN)
    sf = factory.SpectrumFactory(wavenum_min=2000.0, wavenum_max=2010.0, wstep=0.001, pressure=0.1,
                                 mole_fraction=0.01, lines=lines)
    s = sf.eq_spectrum_gpu(Tgas=300.0, path_length=1.0)
    _, A1 = s.get("absorbance")
    A2 = s.recalc_gpu("absorbance", path_length=3.0)
    T3 = s.recalc_gpu("transmittance_noslit", Tgas=700.0)
    s.exit_gpu()
    assert np.allclose(3 * A1, A2, rtol=1e-4) and np.allclose(3 * A1, A2, atol=1e-6)
    assert s.conditions["Tgas"] == 700.0 and np.all((T3 >= 0) & (T3 <= 1))
    s3 = sf.eq_spectrum_gpu(Tgas=300.0, path_length=2.0, exit_gpu=True)
    _, A3 = s3.get("absorbance")
    assert np.allclose(2 * A1, A3, rtol=1e-4)
    with pytest.raises(RuntimeError):
        s3.recalc_gpu("absorbance", Tgas=400.0)


def test_calc_spectrum_gpu_mode():
    lines = hitran.line_list(HITRAN)
    s = factory.calc_spectrum(2150.0, 2160.0, Tgas=1200.0, pressure=1.0, mole_fraction=0.1, path_length=5.0,
                              wstep=0.002, lines=lines, mode="gpu")
    w, rad = s.get("radiance_noslit")
    _, a = s.get("abscoeff")
    ref = oracle.derived_quantities(w, a, 5.0, 1200.0)
    assert oracle.rel_peak_error(rad, ref["radiance"]) <= 2e-6
    assert s.gpu_app is None


def test_factory_min_rms_option():
    lines = hitran.line_list(HITRAN)
    sf = factory.SpectrumFactory(wavenum_min=2150.0, wavenum_max=2200.0, wstep=0.001, pressure=0.05,
                                 mole_fraction=0.1, optimization="min-RMS", lines=lines)
    s = sf.eq_spectrum_gpu(Tgas=2500.0)
    ref = oracle.eq_abscoeff(lines.as_dict(), 2150.0, 2200.0, 0.001, 2500.0, 0.05, 0.1, engine.DXG_DEFAULT,
                             engine.DXL_DEFAULT, optimization="min-RMS")
    assert oracle.rel_peak_error(s.get("abscoeff")[1], ref) <= REL_TOL
    assert s.conditions["optimization"] == "min-RMS"
    s.exit_gpu()
    with pytest.raises(NotImplementedError):
        factory.SpectrumFactory(wavenum_min=2150.0, wavenum_max=2200.0, optimization=None, lines=lines)


def test_apply_slit_on_gpu_matches_oracle():
    """Spectrum.apply_slit (radis/spectrum/spectrum.py:3127) on the GPU:
    radiance, transmittance and emissivity against the oracle's convolution
    of the spectrum's own *_noslit arrays with the same sampled slit; the
    'valid' edges are NaN; energy is conserved (test/tools/test_slit.py:422-455)."""
    from radis_amd import slit

    lines = hitran.line_list(HITRAN)
    sf = factory.SpectrumFactory(wavenum_min=2100.0, wavenum_max=2200.0, wstep=0.002, pressure=1.0,
                                 mole_fraction=0.1, path_length=10.0, lines=lines)
    s = sf.eq_spectrum_gpu(Tgas=1500.0, exit_gpu=True)
    for shape, fwhm in (("triangular", 0.5), ("gaussian", 0.3), ("trapezoidal", (0.4, 0.6))):
        s.apply_slit(fwhm, unit="cm-1", shape=shape)
        ws, Is = s.get_slit()
        dw = float(np.abs(np.diff(s.get("abscoeff")[0])).min())  # the step apply_slit works with
        g = slit.slit_on_grid(ws, Is, dw)
        for q in ("radiance", "transmittance", "emissivity"):
            w, y = s.get(q)
            _, y0 = s.get(q + "_noslit")
            ref = oracle.slit_convolve(y0.astype(np.float32).astype(np.float64), g.astype(np.float32), dw)
            np.testing.assert_array_equal(np.isnan(y), np.isnan(ref))
            ok = ~np.isnan(ref)
            err = np.max(np.abs(y[ok] - ref[ok])) / np.max(np.abs(ref[ok]))
            print(f"{shape} {q}: {err:.1e}")
            assert err <= 3e-7, (shape, q, err)
        w, r = s.get("radiance")
        _, r0 = s.get("radiance_noslit")
        ok = ~np.isnan(r)
        assert np.isclose(np.trapezoid(r[ok], w[ok]), np.trapezoid(r0[ok], w[ok]), rtol=3e-2)
    with pytest.raises(AssertionError):
        s.apply_slit(500.0)  # wider than the spectral range in 'valid' mode


def test_eq_spectrum_gpu_interactive_sliders():
    """radis/lbl/factory.py:1393-1523 (docstring example): sliders for Tgas,
    path_length and the slit; moving one recomputes the plotted quantity on
    the resident GPU application, equal to a fresh computation at the new
    conditions."""
    from radis_amd.plot_tools import ParamRange

    lines = hitran.line_list(HITRAN)
    sf = factory.SpectrumFactory(wavenum_min=2100.0, wavenum_max=2200.0, wstep=0.002, lines=lines)
    s = sf.eq_spectrum_gpu_interactive(var="transmittance", Tgas=ParamRange(300.0, 2000.0, 1200.0),
                                       pressure=0.2, mole_fraction=0.1, path_length=ParamRange(0, 10, 2.0),
                                       slit_function=ParamRange(0, 1, 0), mpl_backend="Agg")
    assert set(sf.interactive_params) == {"Tgas", "path_length", "slit_function"}
    line = sf.interactive_params["Tgas"].widget.ax.figure.axes[0].lines[0]  # the plotted quantity
    sf.interactive_params["Tgas"].widget.set_val(1500.0)
    sf.interactive_params["path_length"].widget.set_val(5.0)
    assert s.conditions["Tgas"] == 1500.0 and s.conditions["path_length"] == 5.0
    ref = sf.eq_spectrum_gpu(Tgas=1500.0, pressure=0.2, mole_fraction=0.1, path_length=5.0, exit_gpu=True)
    np.testing.assert_allclose(s.get("transmittance")[1], ref.get("transmittance_noslit")[1], atol=1e-6)
    sf.interactive_params["slit_function"].widget.set_val(0.5)
    ref.apply_slit(0.5, unit="cm-1")
    np.testing.assert_allclose(s.get("transmittance")[1], ref.get("transmittance")[1], atol=1e-6)
    np.testing.assert_array_equal(line.get_ydata(), s.get("transmittance")[1])
    s.exit_gpu()
    assert s.gpu_app is None
