"""An LDM taller than one fill tile's LDS (fine LDM axes: several hundred
rows N_G x N_L) is filled in row groups: each (tile, row group) workgroup
reads the tile's lines and keeps the corners that fall in its rows.  Against
the oracle: the LDM itself (np.add.at), the line indices and the spectrum;
and a batch mixing such a request with an ordinary one (the launch then runs
the row-group fill kernel for both)."""
import numpy as np
import pytest

from oracle import ldm_oracle as oracle
from radis_amd import engine, workloads

pytestmark = pytest.mark.gpu

# rows one 16-point tile holds in the fill's largest LDS budget (60 KB of
# FP64 cells, stride 16 + 2): taller LDMs were refused before row groups
ONE_TILE_ROWS = 60 * 1024 // ((16 + 2) * 8)


@pytest.fixture(scope="module")
def dev():
    d = engine.Device(0)
    yield d
    d.close()


def _lines(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    nu = np.sort(rng.uniform(lo, hi, n))
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    return workloads.LineList(
        nu0=nu, S_ref=f32(10.0 ** rng.uniform(-23.0, -19.0, n)), El=f32(rng.uniform(0.0, 4000.0, n)),
        g_air=f32(rng.uniform(0.03, 0.08, n)), g_self=f32(rng.uniform(0.05, 0.10, n)),
        n_air=f32(rng.uniform(0.5, 0.8, n)), molar_mass=np.array([0.0, 44.0]), q_coef=(0.94, 1.1),
        name="row_groups")


# fine axes: N_G ~ 20 (2000-2200 cm-1, dxG 0.005), N_L ~ 45 (dxL 0.03)
TALL = workloads.Conditions(2000.0, 2200.0, 0.01, 1000.0, 2.0, 0.3, dxG=0.005, dxL=0.03)
PLAIN = workloads.Conditions(2050.0, 2100.0, 0.01, 800.0, 1.0, 0.1)


def _oracle(lines, c):
    return oracle.eq_abscoeff(lines.as_dict(), c.wmin, c.wmax, c.wstep, c.Tgas, c.pressure_bar, c.mole_fraction,
                              c.dxG, c.dxL, return_details=True)


@pytest.mark.parametrize("fp64_widths", [True, False])
def test_tall_ldm_fills_in_row_groups(dev, fp64_widths):
    """Indices bit-exact and the spectrum within 1e-4 of peak; with FP64
    widths (fractions as the oracle's) the LDM of every row group equals
    np.add.at into the float64 LDM to FP32 rounding.  (The FP32 fast path's
    fractions carry its position error, which grows as 1/dx on fine axes:
    its LDM is compared through the spectrum only.)"""
    lines = _lines(20_000, 2000.5, 2199.5, seed=71)
    db = engine.LineDatabase(dev, lines)
    b = engine.SpectrumBatch(db, [engine.request_from_conditions(db, TALL, fp64_widths=fp64_widths)])
    out = b.results(b.run())[0]
    inf = b.info(0)
    ref, det = _oracle(lines, TALL)
    M = det["NG"] * det["NL"]
    print(f"N_G={det['NG']} N_L={det['NL']} M={M} m_cap={inf.m_cap} tile_w={inf.tile_w}", flush=True)
    assert (inf.n_g, inf.n_l) == (det["NG"], det["NL"])
    assert M > ONE_TILE_ROWS
    t = b.line_terms(0)
    np.testing.assert_array_equal(t["k0"], det["ki0"])
    np.testing.assert_array_equal(t["l0"], det["li0"])
    np.testing.assert_array_equal(t["m0"], det["mi0"])
    if fp64_widths:
        S = det["S"] * oracle.number_density(TALL.mole_fraction, TALL.pressure_bar, TALL.Tgas)
        LDM = oracle.ldm_scatter(inf.n_v + 1, det["NG"], det["NL"], det["ki0"], det["ki0"] + 1, det["tvi"],
                                 det["li0"], det["li0"] + 1, det["tGi"], det["mi0"], det["mi0"] + 1, det["tLi"], S)
        want = LDM.reshape(inf.n_v + 1, -1).T  # rows r = l*N_L + m
        got = np.ldexp(b.ldm(0).astype(np.float64), inf.scale_log2)
        assert got.shape == want.shape
        d = np.max(np.abs(got - want)) / np.abs(want).max()
        print(f"row groups: LDM max diff {d:.3e} of its peak", flush=True)
        assert d <= 1e-6
    err = oracle.rel_peak_error(out, ref)
    print(f"row groups: rel err {err:.3e}", flush=True)
    assert err <= 1e-4
    b.close()
    db.close()


def test_row_group_launch_with_an_ordinary_request(dev):
    lines = _lines(20_000, 2000.5, 2199.5, seed=72)
    db = engine.LineDatabase(dev, lines)
    b = engine.SpectrumBatch(db, [engine.request_from_conditions(db, c) for c in (PLAIN, TALL)])
    outs = b.results(b.run())
    for k, c in enumerate((PLAIN, TALL)):
        ref, _ = _oracle(lines, c)
        err = oracle.rel_peak_error(outs[k], ref)
        print(f"request {k}: rel err {err:.3e}", flush=True)
        assert err <= 1e-4
    b.close()
    db.close()


def test_lorentzian_axis_beyond_64_nodes(dev):
    """N_L between 64 and 128 (dxL 0.015): the axis tables, the image
    correction of every L node (pass B stages it for the block's rows in
    launch-sized LDS) and the spectrum against the oracle."""
    lines = _lines(20_000, 2000.5, 2199.5, seed=73)
    c = workloads.Conditions(2000.0, 2200.0, 0.01, 1000.0, 2.0, 0.3, dxL=0.015)
    db = engine.LineDatabase(dev, lines)
    b = engine.SpectrumBatch(db, [engine.request_from_conditions(db, c)])
    out = b.results(b.run())[0]
    inf = b.info(0)
    ref, det = _oracle(lines, c)
    print(f"N_G={det['NG']} N_L={det['NL']} L={inf.fft_len} 2N_v={2 * inf.n_v}", flush=True)
    assert 64 < det["NL"] <= 128 and (inf.n_g, inf.n_l) == (det["NG"], det["NL"])
    assert inf.fft_len != 2 * inf.n_v  # the image correction runs
    t = b.line_terms(0)
    np.testing.assert_array_equal(t["l0"], det["li0"])
    np.testing.assert_array_equal(t["m0"], det["mi0"])
    err = oracle.rel_peak_error(out, ref)
    print(f"N_L > 64: rel err {err:.3e}", flush=True)
    assert err <= 1e-4
    b.close()
    db.close()
