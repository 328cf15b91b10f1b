"""Every compile-time-sized FFT split n1 x n2 against the run-time-sized
(generic) passes on the same request, and against the FP64 oracle at the
same FFT length (RADIS_LDM_PERIOD_L: the circular result at L itself).  The grid fills the transform (2 N_v just below L) so every
bin of both passes carries signal through the lineshape product."""
import numpy as np
import pytest

from oracle import ldm_oracle as oracle
from radis_amd import engine, workloads

FC_N1 = [256, 320, 384, 512, 640, 768, 1024]
FC_N2 = [64, 128, 256, 512, 1024, 2048]
WSTEP = 0.002
WMIN = 2000.0


def _lines(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    nu = np.sort(rng.uniform(lo, hi, n))
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    return workloads.LineList(
        nu0=nu, S_ref=f32(10.0 ** rng.uniform(-24.0, -19.0, n)), El=f32(rng.uniform(0.0, 5000.0, n)),
        g_air=f32(rng.uniform(0.03, 0.08, n)), g_self=f32(rng.uniform(0.05, 0.10, n)),
        n_air=f32(rng.uniform(0.5, 0.8, n)), molar_mass=np.array([0.0, 44.0]), q_coef=(0.94, 1.1),
        name="fft_sizes")


def _grid(L):
    """wmax such that 2 * len(arange(WMIN, wmax + WSTEP, WSTEP)) <= L, close to L."""
    nv = L // 2 - 2
    while True:
        wmax = WMIN + (nv - 1) * WSTEP
        size = np.arange(WMIN, wmax + WSTEP, WSTEP).size
        if 2 * size <= L:
            return wmax, size
        nv -= 1


@pytest.fixture(scope="module")
def dev():
    d = engine.Device(0)
    yield d
    d.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n1", FC_N1)
def test_compiled_splits_match_generic_and_oracle(dev, n1):
    for n2 in FC_N2:
        L = n1 * n2
        wmax, nv = _grid(L)
        # lines over the first ~300 cm-1 (a few N_G rows), the rest of the grid
        # holds the profile wings and zeros
        lines = _lines(3000, WMIN + 1.0, min(WMIN + 300.0, wmax - 1.0), seed=n1 * 7 + n2)
        db = engine.LineDatabase(dev, lines)
        cond = dict(wmin=WMIN, wmax=wmax, wstep=WSTEP, Tgas=1200.0, pressure_bar=1.0, mole_fraction=0.2,
                    dxG=0.2, dxL=0.2)
        # the transform's own period (2 N_v is just below L)
        rc = engine.make_request(db, **cond, fft_len=L, fft_n1=n1, period_fft_len=True)
        rg = engine.make_request(db, **cond, fft_len=L, fft_n1=n1, generic_fft=True, period_fft_len=True)
        b = engine.SpectrumBatch(db, [rc, rg])
        out = b.results(b.run())
        ic, ig = b.info(0), b.info(1)
        assert (ic.n1, ic.n2, ic.compiled_fft) == (n1, n2, 1), (n1, n2)
        assert (ig.n1, ig.n2, ig.compiled_fft) == (n1, n2, 0), (n1, n2)
        assert ic.n_v == nv and ic.fft_len == L
        ref = oracle.eq_abscoeff(lines.as_dict(), WMIN, wmax, WSTEP, cond["Tgas"], cond["pressure_bar"],
                                 cond["mole_fraction"], 0.2, 0.2, L=L)
        e_cg = oracle.rel_peak_error(out[0], out[1])
        e_ref = oracle.rel_peak_error(out[0], ref)
        print(f"n1={n1} n2={n2} L={L}: compiled vs generic {e_cg:.2e}, vs oracle {e_ref:.2e}")
        assert e_cg <= 2e-5, (n1, n2, e_cg)
        assert e_ref <= 1e-4, (n1, n2, e_ref)
        b.close()
        db.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n1", [512, 640])
def test_grid_filling_half_the_transform(dev, n1):
    """N_v = L/2 exactly: the LDM's length V = N_v + 1 reaches into the
    upper half of pass A's columns, so pass A takes its full-load path
    (with V <= L/2 it skips that half).  Lines crowd the grid's end; the
    point n = N_v itself only ever gets zero weight (a line on the last
    grid point has position N_v - 1 exactly)."""
    L = n1 * 2048
    nv = L // 2
    wmax = WMIN + (nv - 1) * WSTEP
    while np.arange(WMIN, wmax + WSTEP, WSTEP).size > nv:
        wmax -= WSTEP / 4
    while np.arange(WMIN, wmax + WSTEP, WSTEP).size < nv:
        wmax += WSTEP / 4
    assert np.arange(WMIN, wmax + WSTEP, WSTEP).size == nv
    lines = _lines(2000, wmax - 2.0, wmax + 0.5 * WSTEP, seed=n1)
    db = engine.LineDatabase(dev, lines)
    cond = dict(wmin=WMIN, wmax=wmax, wstep=WSTEP, Tgas=1200.0, pressure_bar=1.0, mole_fraction=0.2,
                dxG=0.2, dxL=0.2)
    b = engine.SpectrumBatch(db, [engine.make_request(db, **cond, fft_len=L, fft_n1=n1)])
    out = b.results(b.run())[0]
    inf = b.info(0)
    assert inf.n_v == nv and inf.fft_len == L and inf.compiled_fft == 1
    t = b.line_terms(0)
    assert np.max(t["k0"]) == nv - 2  # lines in the last interval (the window ends at wmax)
    ref = oracle.eq_abscoeff(lines.as_dict(), WMIN, wmax, WSTEP, cond["Tgas"], cond["pressure_bar"],
                             cond["mole_fraction"], 0.2, 0.2, L=L)
    assert oracle.rel_peak_error(out, ref) <= 1e-4
    b.close()
    db.close()


@pytest.mark.gpu
def test_grid_beyond_the_compiled_and_short_splits(dev):
    """A 6,000,001-point grid (2 N_v = 12 M, beyond 1024 x 8192): the planner
    takes a run-time-sized split with n1 above 1024 (passes A and C hold one
    or a few columns per workgroup); against the oracle at L = 2 N_v."""
    rng = np.random.default_rng(81)
    n = 2000
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    lines = workloads.LineList(
        nu0=np.sort(rng.uniform(2001.0, 2599.0, n)), S_ref=f32(10.0 ** rng.uniform(-22.0, -20.0, n)),
        El=f32(rng.uniform(0.0, 2000.0, n)), g_air=f32(rng.uniform(0.050, 0.051, n)),
        g_self=f32(np.full(n, 0.07)), n_air=f32(np.full(n, 0.7)), molar_mass=np.array([0.0, 28.0]),
        q_coef=(0.94, 1.1), name="long_grid")
    wmax = 2000.0 + 6_000_000 * 1e-4
    db = engine.LineDatabase(dev, lines)
    cond = dict(wmin=2000.0, wmax=wmax, wstep=1e-4, Tgas=500.0, pressure_bar=0.2, mole_fraction=0.1)
    b = engine.SpectrumBatch(db, [engine.make_request(db, **cond)])
    out = b.results(b.run())[0]
    inf = b.info(0)
    print(f"N_v={inf.n_v} L={inf.fft_len} split {inf.n1} x {inf.n2} compiled={inf.compiled_fft} "
          f"N_G={inf.n_g} N_L={inf.n_l}", flush=True)
    assert 2 * inf.n_v > 1024 * 8192 and inf.n1 > 1024 and inf.compiled_fft == 0
    ref = oracle.eq_abscoeff(lines.as_dict(), cond["wmin"], wmax, 1e-4, cond["Tgas"], cond["pressure_bar"],
                             cond["mole_fraction"], workloads.DXG_DEFAULT, workloads.DXL_DEFAULT)
    err = oracle.rel_peak_error(out, ref)
    print(f"long grid: rel err {err:.3e}", flush=True)
    assert err <= 1e-4
    b.close()
    db.close()
