"""GPU parity of the HIP LDM path against the CPU oracle (FP64 NumPy
restatement of RADIS's LDM-fft path), through the C ABI.

Bar: integer grid indices bit-exact; abscoeff within 1e-4 of peak
(max |gpu - oracle| / max |oracle|) against RADIS's result, the oracle at
FFT length L = 2 N_v, whatever transform length the GPU plans (its circular
images are corrected); RADIS_LDM_PERIOD_L requests against the oracle at
the GPU's own length.
"""
import numpy as np
import pytest

from oracle import ldm_oracle as oracle
from radis_amd import _native, engine, hitran, workloads

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4  # of peak abscoeff


@pytest.fixture(scope="module")
def dev():
    d = engine.Device(0)
    yield d
    d.close()


def _run(dev, lines, conds, fft_len=0):
    db = engine.LineDatabase(dev, lines)
    reqs = [engine.request_from_conditions(db, c, fft_len=fft_len) for c in conds]
    b = engine.SpectrumBatch(db, reqs)
    outs = b.run()
    res = b.results(outs)
    return db, b, res


def _oracle(lines, c, L=None, details=False):
    d = lines.as_dict() if hasattr(lines, "as_dict") else lines
    return oracle.eq_abscoeff(d, c.wmin, c.wmax, c.wstep, c.Tgas, c.pressure_bar, c.mole_fraction,
                              c.dxG, c.dxL, Tref=c.Tref, L=L, return_details=details)


def _check_indices(b, det, i=0, frac_atol=3e-5):
    """Integer indices bit-exact; grid fraction to FP32 rounding; log-axis
    fractions to the FP32 fast-path bound (FP64 when requested)."""
    t = b.line_terms(i)
    assert t["k0"].size == det["ki0"].size
    np.testing.assert_array_equal(t["k0"], det["ki0"])
    np.testing.assert_array_equal(t["l0"], det["li0"])
    np.testing.assert_array_equal(t["m0"], det["mi0"])
    np.testing.assert_allclose(t["tv"], det["tvi"], atol=1e-6)
    np.testing.assert_allclose(t["tg"], det["tGi"], atol=frac_atol)
    np.testing.assert_allclose(t["tl"], det["tLi"], atol=frac_atol)
    return t


def test_co5k_full(dev):
    """5,000 CO-like lines, 200,001-point grid (the CPU-runnable workload)."""
    lines, c = workloads.workload("co_5k")
    db, b, res = _run(dev, lines, [c])
    inf = b.info(0)
    assert inf.n_v == 200001
    ref, det = _oracle(lines, c, details=True)
    assert (inf.n_g, inf.n_l) == (det["NG"], det["NL"])
    t = _check_indices(b, det)
    S_dens = det["S"] * oracle.number_density(c.mole_fraction, c.pressure_bar, c.Tgas)
    np.testing.assert_allclose(np.ldexp(t["I"].astype(np.float64), inf.scale_log2), S_dens, rtol=2e-6)
    err = oracle.rel_peak_error(res[0], ref)
    print(f"co_5k rel err {err:.3e} L={inf.fft_len} NG={inf.n_g} NL={inf.n_l}")
    assert err <= REL_TOL


@pytest.mark.parametrize("fp64_widths,tol", [(True, 1e-6), (False, 6e-5)])
def test_co5k_ldm_fill_matches_add_at(dev, fp64_widths, tol):
    """The LDS-tiled fill equals np.add.at into the float64 LDM (to FP32
    rounding with FP64 widths; to the fast path's fraction bound, 3e-5 of
    an axis step per weight, otherwise)."""
    lines, c = workloads.workload("co_5k")
    db = engine.LineDatabase(dev, lines)
    b = engine.SpectrumBatch(db, [engine.request_from_conditions(db, c, fp64_widths=fp64_widths)])
    b.results(b.run())
    inf = b.info(0)
    ref, det = _oracle(lines, c, details=True)
    S = det["S"] * oracle.number_density(c.mole_fraction, c.pressure_bar, c.Tgas)
    LDM = oracle.ldm_scatter(inf.n_v + 1, det["NG"], det["NL"], det["ki0"], det["ki0"] + 1, det["tvi"],
                             det["li0"], det["li0"] + 1, det["tGi"], det["mi0"], det["mi0"] + 1, det["tLi"], S)
    want = LDM.reshape(inf.n_v + 1, -1).T  # rows r = l*N_L + m
    got = np.ldexp(b.ldm(0).astype(np.float64), inf.scale_log2)
    assert got.shape == want.shape
    scale = np.abs(want).max()
    assert np.max(np.abs(got - want)) <= tol * scale


def test_co2_subset_full_grid(dev):
    """200k CO2-like lines on the 500,002-point 2000-2500 cm-1 grid at 1500 K."""
    lines = workloads.co2_like(200_000, seed=11)
    c = workloads.CONDITIONS["co2_5m"]
    db, b, res = _run(dev, lines, [c])
    inf = b.info(0)
    assert inf.n_v == 500002
    ref, det = _oracle(lines, c, details=True)
    _check_indices(b, det)
    err = oracle.rel_peak_error(res[0], ref)
    print(f"co2 subset rel err {err:.3e}")
    assert err <= REL_TOL


def test_windows_batch_equals_oracle(dev):
    """A ragged batch of windows and conditions; each request vs the oracle."""
    lines = workloads.co2_like(300_000, seed=12)
    conds = workloads.random_requests(6, 2000.0, 2500.0, seed=5, median_width=20.0, wmax_width=60.0)
    db, b, res = _run(dev, lines, conds)
    for i, c in enumerate(conds):
        inf = b.info(i)
        ref, det = _oracle(lines, c, details=True)
        _check_indices(b, det, i)
        err = oracle.rel_peak_error(res[i], ref)
        print(f"req {i}: W={c.wmax - c.wmin:.1f} T={c.Tgas:.0f} p={c.pressure_bar:.2f} err {err:.3e}")
        assert err <= REL_TOL


def test_batch_equals_single(dev):
    lines = workloads.co2_like(100_000, seed=13)
    conds = workloads.random_requests(4, 2000.0, 2500.0, seed=9, median_width=10.0, wmax_width=30.0)
    db, b, res = _run(dev, lines, conds)
    for i, c in enumerate(conds):
        b1 = engine.SpectrumBatch(db, [engine.request_from_conditions(db, c)])
        r1 = b1.results(b1.run())[0]
        assert oracle.rel_peak_error(r1, res[i]) <= 1e-6


def test_graph_replay_matches_eager(dev):
    lines = workloads.co2_like(100_000, seed=14)
    conds = workloads.random_requests(3, 2000.0, 2500.0, seed=21, median_width=10.0, wmax_width=30.0)
    db = engine.LineDatabase(dev, lines)
    b = engine.SpectrumBatch(db, [engine.request_from_conditions(db, c) for c in conds])
    eager = b.results(b.run())
    outs = b.capture()
    for rep in range(3):
        b.replay()
        got = b.results(outs)
        for i in range(len(conds)):
            assert oracle.rel_peak_error(got[i], eager[i]) <= 1e-6, (rep, i)
        _ = b.results(b.run())  # eager work between replays
    # change conditions, replay picks them up
    c2 = workloads.Conditions(conds[0].wmin, conds[0].wmax, conds[0].wstep, 900.0, 3.0, 0.3)
    b.set_request(0, engine.request_from_conditions(db, c2))
    b.replay()
    got = b.results(outs)[0]
    ref = _oracle(lines, c2)
    assert oracle.rel_peak_error(got, ref) <= REL_TOL


def test_co5k_fp64_widths(dev):
    """Full-FP64 width path: fractions to FP32 rounding of the FP64 values."""
    lines, c = workloads.workload("co_5k")
    db = engine.LineDatabase(dev, lines)
    b = engine.SpectrumBatch(db, [engine.request_from_conditions(db, c, fp64_widths=True)])
    res = b.results(b.run())
    inf = b.info(0)
    ref, det = _oracle(lines, c, details=True)
    _check_indices(b, det, frac_atol=1e-6)
    assert oracle.rel_peak_error(res[0], ref) <= REL_TOL


def test_hitran_fixture_isotopes_and_shift(dev):
    """Real HITRAN CO lines (3 isotopologues, pressure shifts), the
    reference's test fixture test/files/hitran_co_3iso_2000_2300cm.par."""
    lines = hitran.line_list("tests/golden/hitran_co_3iso_2000_2300cm.par")
    conds = [workloads.Conditions(2000.0, 2300.0, 0.002, 1500.0, 2.0, 0.5),
             workloads.Conditions(2100.0, 2200.0, 0.001, 300.0, 1.0, 0.95, dxG=0.2, dxL=0.2),
             workloads.Conditions(2000.0, 2300.0, 0.01, 3000.0, 0.01, 0.1)]
    db, b, res = _run(dev, lines, conds)
    for i, c in enumerate(conds):
        inf = b.info(i)
        ref, det = _oracle(lines, c, details=True)
        _check_indices(b, det, i)
        err = oracle.rel_peak_error(res[i], ref)
        print(f"hitran cond {i}: err {err:.3e} NG={inf.n_g} NL={inf.n_l}")
        assert err <= REL_TOL


def test_lines_outside_grid_are_clamped(dev):
    """Lines beyond the grid ends land on the edge points (np.interp clamp)."""
    lines = workloads.co_like(4000, seed=3)
    c = workloads.Conditions(2000.0, 2050.0, 0.002, 800.0, 1.0, 0.3, dxG=0.2, dxL=0.2)
    db = engine.LineDatabase(dev, lines)
    req = engine.make_request(db, c.wmin, c.wmax, c.wstep, c.Tgas, c.pressure_bar, c.mole_fraction,
                              c.dxG, c.dxL, line_window=(1990.0, 2060.0))
    b = engine.SpectrumBatch(db, [req])
    got = b.results(b.run())[0]
    inf = b.info(0)
    sub = lines.subset(slice(*db.window(1990.0, 2060.0)))
    w = oracle.wavenumber_grid(c.wmin, c.wmax, c.wstep)
    S, nu, wG, wL = oracle.line_parameters(sub.as_dict(), c.Tgas, c.pressure_bar, c.mole_fraction)
    wLa, wGa = oracle.init_w_axis(wL, c.dxL), oracle.init_w_axis(wG, c.dxG)
    k0, k1, tv = oracle.get_indices(nu, w)
    l0, l1, tg = oracle.get_indices(np.log(wG), np.log(wGa))
    m0, m1, tl = oracle.get_indices(np.log(wL), np.log(wLa))
    assert np.any(k0 == 0) and np.any(k0 == w.size - 1)
    L = 2 * w.size
    LDM = oracle.ldm_scatter(L, wGa.size, wLa.size, k0, k1, tv, l0, l1, tg, m0, m1, tl, S)
    ref = oracle.ldm_apply_fft(LDM, oracle.lineshape_templates(L, c.wstep, wGa, wLa), w.size, c.wstep)
    ref = ref * oracle.number_density(c.mole_fraction, c.pressure_bar, c.Tgas)
    assert oracle.rel_peak_error(got, ref) <= REL_TOL


def test_empty_window_gives_zeros(dev):
    lines = workloads.co_like(1000, seed=4)
    db = engine.LineDatabase(dev, lines)
    req = engine.make_request(db, 2400.0, 2410.0, 0.01, 1000.0, 1.0, 0.1, dxG=0.2, dxL=0.2)
    b = engine.SpectrumBatch(db, [req])
    got = b.results(b.run())[0]
    assert got.shape == (np.arange(2400.0, 2410.0 + 0.01, 0.01).size,) and np.all(got == 0.0)
    assert b.info(0).n_g == 0


def test_extreme_intensities_use_power_of_two_scale(dev):
    """E'' spread of the 100M-line H2O-like list at 2500 K makes S(T) exceed
    the FP32 range; the library scales by 2**-s and reports s."""
    lines = workloads.h2o_like(200_000, seed=8, chunk=200_000)
    c = workloads.Conditions(2000.0, 2100.0, 0.005, 2500.0, 5.0 * workloads.ATM, 0.1)
    db, b, res = _run(dev, lines, [c])
    inf = b.info(0)
    assert inf.scale_log2 != 0
    ref = _oracle(lines, c)
    assert not np.isfinite(np.float32(ref.max())) or ref.max() > 1e30
    assert oracle.rel_peak_error(res[0], ref) <= REL_TOL


@pytest.mark.parametrize("label,cond", [
    ("pressure", workloads.Conditions(2000.0, 2300.0, 0.002, 1500.0, 2.0, 0.5)),
    ("doppler", workloads.Conditions(2100.0, 2200.0, 0.0005, 3000.0, 0.01, 0.1, dxG=0.2, dxL=0.2)),
])
def test_min_rms_weights_match_oracle(dev, label, cond):
    """optimization='min-RMS' (radis/lbl/broadening.py:2265-2303): indices
    bit-exact and abscoeff within 1e-4 of peak, on the HITRAN fixture in a
    pressure- and a Doppler-broadened regime.  The corrected fractions match
    to FP32 rounding with FP64 log-width positions; on the FP32 fast path
    the corrections amplify its ~3e-5 fraction error (R_LG grows as the
    Doppler width dominates), so there only abscoeff is held to the bar."""
    lines = hitran.line_list("tests/golden/hitran_co_3iso_2000_2300cm.par")
    db = engine.LineDatabase(dev, lines)
    c = cond
    reqs = [engine.make_request(db, c.wmin, c.wmax, c.wstep, c.Tgas, c.pressure_bar, c.mole_fraction, c.dxG, c.dxL,
                                optimization="min-RMS", fp64_widths=f64) for f64 in (False, True)]
    b = engine.SpectrumBatch(db, reqs)
    res = b.results(b.run())
    inf = b.info(0)
    ref, det = oracle.eq_abscoeff(lines.as_dict(), c.wmin, c.wmax, c.wstep, c.Tgas, c.pressure_bar, c.mole_fraction,
                                  c.dxG, c.dxL, return_details=True, optimization="min-RMS")
    assert np.max(np.abs(det["aGi"] - det["tGi"])) > 1e-3  # the corrections are not negligible here
    simple = oracle.eq_abscoeff(lines.as_dict(), c.wmin, c.wmax, c.wstep, c.Tgas, c.pressure_bar, c.mole_fraction,
                                c.dxG, c.dxL)
    for i in range(2):
        assert int(b.info(i).fft_len) == int(inf.fft_len)
        t = b.line_terms(i)
        np.testing.assert_array_equal(t["k0"], det["ki0"])
        np.testing.assert_array_equal(t["l0"], det["li0"])
        np.testing.assert_array_equal(t["m0"], det["mi0"])
        if i == 1:
            np.testing.assert_allclose(t["tg"], det["aGi"], atol=2e-6, rtol=1e-5)
            np.testing.assert_allclose(t["tl"], det["aLi"], atol=2e-6, rtol=1e-5)
        err = oracle.rel_peak_error(res[i], ref)
        dw = max(np.max(np.abs(t["tg"] - det["aGi"])), np.max(np.abs(t["tl"] - det["aLi"])))
        print(f"min-RMS {label} fp64_widths={i}: err {err:.2e}, max fraction diff {dw:.1e}, "
              f"min-RMS vs simple {oracle.rel_peak_error(simple, ref):.2e}")
        assert err <= REL_TOL
    b.close()
    db.close()


def _axis_lines(n, seed, multi_iso, with_nself):
    import dataclasses

    lines = workloads.co2_like(n, seed=seed)
    rng = np.random.default_rng(seed + 1)
    kw = {}
    if multi_iso:
        kw["iso"] = rng.integers(1, 4, n).astype(np.uint8)
        kw["molar_mass"] = np.array([0.0, 44.0, 45.0, 46.0])
    if with_nself:
        kw["n_self"] = rng.uniform(0.4, 0.9, n).astype(np.float32)
    return dataclasses.replace(lines, **kw)


@pytest.mark.parametrize("multi_iso,with_nself", [(False, False), (True, True)])
def test_axis_candidates_equal_full_scan(dev, multi_iso, with_nself):
    """The width extremes from the per-block candidates are bit-identical to
    a scan of every window line (RADIS_LDM_SCAN_AXIS), for T above, at and
    below Tref and mole fractions 0, 0.3 and 1; both agree with NumPy's
    FP64 widths to rounding."""
    lines = _axis_lines(400_000, 40 + int(multi_iso), multi_iso, with_nself)
    db = engine.LineDatabase(dev, lines)
    for T, x in [(1500.0, 0.3), (296.0, 0.3), (200.0, 0.0), (900.0, 1.0)]:
        c = workloads.Conditions(2003.7, 2371.2, 0.002, T, 1.3, x)
        r0 = engine.request_from_conditions(db, c)
        r1 = engine.request_from_conditions(db, c)
        r1.flags = int(r1.flags) | _native.SCAN_AXIS
        b = engine.SpectrumBatch(db, [r0, r1])
        res = b.results(b.run())
        i0, i1 = b.info(0), b.info(1)
        for k in ("wg_min", "wg_max", "wl_min", "wl_max"):
            assert getattr(i0, k) == getattr(i1, k), (T, x, k)
        assert (i0.n_g, i0.n_l) == (i1.n_g, i1.n_l)
        np.testing.assert_array_equal(res[0], res[1])
        d = lines.as_dict()
        sel = slice(np.searchsorted(d["nu0"], c.wmin, side="left"), np.searchsorted(d["nu0"], c.wmax, side="right"))
        S, _, wG, wL = oracle.line_parameters({k: (v[sel] if isinstance(v, np.ndarray) and v.shape == d["nu0"].shape
                                                   else v) for k, v in d.items()}, T, 1.3, x)
        np.testing.assert_allclose([i0.wl_min, i0.wl_max], [wL.min(), wL.max()], rtol=1e-15)
        np.testing.assert_allclose([i0.wg_min, i0.wg_max], [wG.min(), wG.max()], rtol=1e-15)
        b.close()
    db.close()


@pytest.mark.parametrize("with_nself", [False, True])
def test_axis_whole_list_from_pareto_sets(dev, with_nself):
    """A window holding every line of a single-isotopologue list takes its
    Lorentzian width extremes from the list's Pareto sets in k_finalize (no
    k_axis items): bit-identical to the full scan and to NumPy's FP64
    widths to rounding, for T above, at and below Tref."""
    lines = _axis_lines(400_000, 44, False, with_nself)
    d = lines.as_dict()
    db = engine.LineDatabase(dev, lines)
    for T, x in [(1500.0, 0.3), (296.0, 0.3), (200.0, 0.0), (900.0, 1.0)]:
        c = workloads.Conditions(float(d["nu0"][0]) - 1.0, float(d["nu0"][-1]) + 1.0, 0.002, T, 1.3, x)
        r0 = engine.request_from_conditions(db, c)
        r1 = engine.request_from_conditions(db, c)
        r1.flags = int(r1.flags) | _native.SCAN_AXIS
        b = engine.SpectrumBatch(db, [r0, r1])
        res = b.results(b.run())
        i0, i1 = b.info(0), b.info(1)
        assert i0.axis_scanned == 0 and i0.axis_candidates > 0, (T, x)
        assert i1.axis_scanned == len(d["nu0"])
        for k in ("wg_min", "wg_max", "wl_min", "wl_max"):
            assert getattr(i0, k) == getattr(i1, k), (T, x, k)
        np.testing.assert_array_equal(res[0], res[1])
        _, _, wG, wL = oracle.line_parameters(d, T, 1.3, x)
        np.testing.assert_allclose([i0.wl_min, i0.wl_max], [wL.min(), wL.max()], rtol=1e-15)
        np.testing.assert_allclose([i0.wg_min, i0.wg_max], [wG.min(), wG.max()], rtol=1e-15)
        b.close()
    db.close()


def test_grid_index_at_and_just_below_grid_points(dev):
    """Lines on grid points, and 1..64 ulp and 1e-12..1e-6 steps below the
    next one, where NumPy's (1/dw)*(x - w[j]) + j rounds up to j + 1: the
    index is NumPy's on every line (the fill's fast fraction path hands
    these to the exact FP64 division)."""
    wmin, wmax, wstep = 2000.0, 2400.0, 0.001
    w = oracle.wavenumber_grid(wmin, wmax, wstep)
    rng = np.random.default_rng(77)
    j = rng.integers(1, w.size - 2, 600)
    nu = [w[j]]
    for k in (1, 2, 8, 64):
        x = w[j + 1].copy()
        for _ in range(k):
            x = np.nextafter(x, -np.inf)
        nu.append(x)
    for frac in (1e-12, 1e-9, 1e-7, 1e-6):
        nu.append(w[j + 1] - frac * wstep)
    nu = np.sort(np.concatenate(nu))
    n = nu.size
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    lines = workloads.LineList(nu0=nu, S_ref=f32(np.full(n, 1e-20)), El=f32(np.zeros(n)),
                               g_air=f32(rng.uniform(0.03, 0.08, n)), g_self=f32(rng.uniform(0.05, 0.1, n)),
                               n_air=f32(np.full(n, 0.7)), molar_mass=np.array([0.0, 28.0]), q_coef=(0.36, 1.0),
                               name="grid_edges")
    c = workloads.Conditions(wmin, wmax, wstep, 1000.0, 1.0, 0.1)
    db, b, res = _run(dev, lines, [c])
    inf = b.info(0)
    ref, det = _oracle(lines, c, details=True)
    _check_indices(b, det)
    assert oracle.rel_peak_error(res[0], ref) <= REL_TOL
    # NumPy's sum rounds up to j + 1 only where x's spacing is finer than
    # half an ulp of j relative to the step (a grid starting near zero):
    # w = 0.001 + 0.01 i, x one ulp below w[1] gets index 1
    w2 = oracle.wavenumber_grid(0.001, 0.5, 0.01)
    x = np.nextafter(w2[1], -np.inf)
    assert oracle.get_indices(np.array([x]), w2)[0][0] == 1
    nu2 = np.array([w2[0], x, w2[3] + 0.004])
    lines2 = workloads.LineList(nu0=nu2, S_ref=f32(np.full(3, 1e-20)), El=f32(np.zeros(3)), g_air=f32(np.full(3, 0.05)),
                                g_self=f32(np.full(3, 0.07)), n_air=f32(np.full(3, 0.7)),
                                molar_mass=np.array([0.0, 28.0]), q_coef=(0.36, 1.0), name="round_up")
    c2 = workloads.Conditions(0.001, 0.5, 0.01, 1000.0, 1.0, 0.1, dxG=0.5, dxL=0.5)
    db2, b2, res2 = _run(dev, lines2, [c2])
    ref2, det2 = _oracle(lines2, c2, details=True)
    t2 = _check_indices(b2, det2)
    assert t2["k0"][1] == 1 and t2["tv"][1] == 0.0
    assert oracle.rel_peak_error(res2[0], ref2) <= REL_TOL
