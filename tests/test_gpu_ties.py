"""Log-width axis indices next to axis nodes, and the FP64 width extremes
against NumPy's, bit for bit where the arithmetic allows it.

The LDM axes are w_min exp(dx i) (radis/lbl/broadening.py:1823-1837) and a
line's log-width position is np.interp's on log(axis) (2151-2154).  The
device evaluates exp and log with the HIP device library and NumPy with the
C library, so a node's log can differ by an ulp or two: a line within a few
ulp of a node can take the neighbouring index.  Its weights are then 1 and 0
on the node it shares with NumPy's pair, so the LDM does not change.  The
tests place lines at node * (1 + delta): bit-exact indices for |delta| >=
1e-14 (about ten ulp of the log width), the LDM equal to np.add.at for every
delta, exact ties included.

Widths: the Gaussian FWHM is the same sequence of correctly rounded FP64
operations on both sides (bit-identical).  The Lorentzian extremes evaluate
(Tref/T)^n with the device's pow, NumPy with the C library's; the two agree
to within an ulp or two (exactly when T = Tref).
"""
import numpy as np
import pytest

from oracle import ldm_oracle as oracle
from radis_amd import engine, workloads

pytestmark = pytest.mark.gpu

DELTAS = (-1e-12, -1e-14, 0.0, 1e-14, 1e-12)


@pytest.fixture(scope="module")
def dev():
    d = engine.Device(0)
    yield d
    d.close()


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _lines(nu, g_air, n_air=0.7):
    n = nu.size
    return workloads.LineList(nu0=np.ascontiguousarray(nu, dtype=np.float64), S_ref=_f32(np.full(n, 1e-20)),
                              El=_f32(np.zeros(n)), g_air=_f32(g_air), g_self=_f32(np.full(n, 0.07)),
                              n_air=_f32(np.full(n, n_air)), molar_mass=np.array([0.0, 28.0]), q_coef=(0.36, 1.0),
                              name="ties")


def _ldm_check(b, i, det, c, rel=1e-5):
    """The device LDM against np.add.at with NumPy's indices and fractions."""
    inf = b.info(i)
    nv = int(inf.n_v)
    S = det["S"] * oracle.number_density(c.mole_fraction, c.pressure_bar, c.Tgas)
    want = oracle.ldm_scatter(nv + 1, det["NG"], det["NL"], det["ki0"], det["ki0"] + 1, det["tvi"], det["li0"],
                              det["li0"] + 1, det["tGi"], det["mi0"], det["mi0"] + 1, det["tLi"], S)
    want = want.reshape(nv + 1, -1).T
    got = np.ldexp(b.ldm(i).astype(np.float64), inf.scale_log2)
    assert got.shape == want.shape
    assert np.max(np.abs(got - want)) <= rel * np.abs(want).max()


def test_gaussian_axis_ties(dev):
    """Lines whose Gaussian FWHM (nu0 / c * const, exact in FP64) sits at
    G-axis nodes times (1 + delta); T = Tref, x = 0.1, p = 1 atm."""
    dxG, nu_min = 0.01, 2000.0
    nodes = nu_min * np.exp(dxG * np.arange(1, 19))  # wG is linear in nu0: the nodes' line centres
    pairs = [(nu_min, 1.0)] + [(v * (1 + d), d) for v in nodes for d in DELTAS]
    pairs.sort()
    nu = np.array([p[0] for p in pairs])
    delta = np.array([p[1] for p in pairs])
    lines = _lines(nu, np.full(nu.size, 0.05))
    c = workloads.Conditions(1999.0, 2401.0, 0.01, 296.0, 1.01325, 0.1, dxG=dxG, dxL=0.2)
    db = engine.LineDatabase(dev, lines)
    ref, det = oracle.eq_abscoeff(lines.as_dict(), c.wmin, c.wmax, c.wstep, c.Tgas, c.pressure_bar,
                                  c.mole_fraction, c.dxG, c.dxL, return_details=True)
    for fp64 in (False, True):
        b = engine.SpectrumBatch(db, [engine.request_from_conditions(db, c, fp64_widths=fp64)])
        b.results(b.run())
        inf = b.info(0)
        assert inf.wg_min == det["wG_dat"].min() and inf.wg_max == det["wG_dat"].max()  # bit for bit
        assert (inf.n_g, inf.n_l) == (det["NG"], det["NL"])
        t = b.line_terms(0)
        off = np.abs(delta) >= 1e-14
        np.testing.assert_array_equal(t["l0"][off], det["li0"][off])
        flips = int(np.count_nonzero(t["l0"][~off] != det["li0"][~off]))
        print(f"G axis, fp64_widths={fp64}: {int(off.sum())} lines at |delta| >= 1e-14 equal, "
              f"{flips} of {int((~off).sum())} exact ties take the neighbouring index")
        _ldm_check(b, 0, det, c)
        b.close()
    db.close()


def test_lorentzian_axis_ties(dev):
    """At T = Tref the Lorentzian FWHM is exact in FP64 on both sides
    ((Tref/T)^n = 1).  Each request's dxL puts line 3 at node 3 times
    (1 + delta) above the narrowest line."""
    g = np.array([0.04, 0.047, 0.0555, 0.061, 0.0712], dtype=np.float32)
    nu = np.linspace(2100.0, 2200.0, g.size)
    lines = _lines(nu, g)
    p_bar, x = 2.0, 0.1
    _, _, _, wL = oracle.line_parameters(lines.as_dict(), 296.0, p_bar, x)
    conds = []
    for d in DELTAS:
        dxL = float(np.log((wL[3] / wL[0]) / (1 + d)) / 3.0)  # node 3 = wL_min exp(3 dxL) = wL[3] / (1 + delta)
        conds.append(workloads.Conditions(2050.0, 2250.0, 0.01, 296.0, p_bar, x, dxG=0.2, dxL=dxL))
    db = engine.LineDatabase(dev, lines)
    b = engine.SpectrumBatch(db, [engine.request_from_conditions(db, c) for c in conds])
    b.results(b.run())
    for i, (c, d) in enumerate(zip(conds, DELTAS)):
        inf = b.info(i)
        ref, det = oracle.eq_abscoeff(lines.as_dict(), c.wmin, c.wmax, c.wstep, c.Tgas, c.pressure_bar,
                                      c.mole_fraction, c.dxG, c.dxL, return_details=True)
        assert inf.wl_min == det["wL_dat"].min() and inf.wl_max == det["wL_dat"].max()  # bit for bit at T = Tref
        t = b.line_terms(i)
        print(f"L axis delta={d:g}: device m0 {t['m0'].tolist()}, NumPy {det['mi0'].tolist()}")
        if d != 0.0:
            np.testing.assert_array_equal(t["m0"], det["mi0"])
        _ldm_check(b, i, det, c)
    b.close()
    db.close()


def test_lorentzian_width_extremes_against_numpy_pow(dev):
    """wl_min, wl_max (FP64, device pow) against NumPy's (Tref/T)**n form
    over many conditions: equal to within 2 ulp; the count of exact matches
    is reported."""
    lines = workloads.co2_like(100_000, seed=71)
    db = engine.LineDatabase(dev, lines)
    rng = np.random.default_rng(72)
    conds = [workloads.Conditions(2000.0, 2500.0, 0.01, float(rng.uniform(200.0, 3000.0)),
                                  float(rng.uniform(0.1, 20.0)), float(rng.uniform(0.0, 1.0))) for _ in range(24)]
    b = engine.SpectrumBatch(db, [engine.request_from_conditions(db, c) for c in conds])
    b.results(b.run())
    d = lines.as_dict()
    exact = 0
    worst = 0
    for i, c in enumerate(conds):
        inf = b.info(i)
        _, _, wG, wL = oracle.line_parameters(d, c.Tgas, c.pressure_bar, c.mole_fraction)
        for got, want in ((inf.wl_min, wL.min()), (inf.wl_max, wL.max())):
            ulp = abs(int(np.float64(got).view(np.int64)) - int(np.float64(want).view(np.int64)))
            worst = max(worst, ulp)
            exact += ulp == 0
        assert inf.wg_min == wG.min() and inf.wg_max == wG.max()
    print(f"Lorentzian width extremes: {exact} of {2 * len(conds)} bit-identical to NumPy's pow form, "
          f"largest difference {worst} ulp")
    assert worst <= 2
    b.close()
    db.close()
