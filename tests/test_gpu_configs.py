"""The benchmark configurations at their real shapes against the oracle (the
pattern of radis/test/gpu/test_gpu.py:60-121: the GPU spectrum against the
CPU one on the configured case), at RADIS's FFT length 2 N_v.

* h2o_100m: the full 1000-4000 cm-1 grid at 0.005 cm-1 (600,001 points),
  2500 K, 5 atm, RADIS's default dxG/dxL: N_G = N_L = 12, so M = 144 LDM
  rows, 72 pass-B pairs and the 640 x 2048 split, on a 600k-line sample of
  the list (same distributions and seed).
* batch_256: the 256 requests over the 5 M-line list in one launch set
  replayed as a hipGraph, 16 of them against the oracle (the six with the
  largest circular-image differences among them).
"""
import numpy as np
import pytest

from oracle import ldm_oracle as oracle
from radis_amd import engine, workloads

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4


@pytest.fixture(scope="module")
def dev():
    d = engine.Device(0)
    yield d
    d.close()


def test_h2o_full_grid_m144(dev):
    lines = workloads.h2o_like(600_000, seed=1002, chunk=600_000)
    c = workloads.CONDITIONS["h2o_100m"]
    db = engine.LineDatabase(dev, lines)
    b = engine.SpectrumBatch(db, [engine.request_from_conditions(db, c)])
    got = b.results(b.run())[0]
    inf = b.info(0)
    print(f"h2o full grid: N_v={inf.n_v} L={inf.fft_len} split {inf.n1}x{inf.n2} compiled={inf.compiled_fft} "
          f"N_G={inf.n_g} N_L={inf.n_l} scale_log2={inf.scale_log2}", flush=True)
    assert inf.n_v == 600_001
    assert (inf.n_g, inf.n_l) == (12, 12)
    assert (inf.n1, inf.n2, inf.compiled_fft) == (640, 2048, 1)
    ref, det = oracle.eq_abscoeff(lines.as_dict(), c.wmin, c.wmax, c.wstep, c.Tgas, c.pressure_bar, c.mole_fraction,
                                  c.dxG, c.dxL, return_details=True)
    assert (det["NG"], det["NL"]) == (12, 12)
    t = b.line_terms(0)
    np.testing.assert_array_equal(t["k0"], det["ki0"])
    np.testing.assert_array_equal(t["l0"], det["li0"])
    np.testing.assert_array_equal(t["m0"], det["mi0"])
    err = oracle.rel_peak_error(got, ref)
    print(f"h2o full grid: rel err {err:.2e} over {t['k0'].size} lines, indices equal", flush=True)
    assert err <= REL_TOL
    b.close()
    db.close()


def test_batch256_graph_replay_against_oracle(dev):
    lines = workloads.co2_like(5_000_000)
    conds = workloads.random_requests(256, 2000.0, 2500.0, seed=1003)
    db = engine.LineDatabase(dev, lines)
    b = engine.SpectrumBatch(db, [engine.request_from_conditions(db, c) for c in conds])
    outs = b.capture()
    for _ in range(2):
        b.replay()
    res = b.results(outs)
    pick = [75, 48, 157, 44, 215, 181] + [i for i in range(256) if i not in (75, 48, 157, 44, 215, 181)][:10]
    d = lines.as_dict()
    worst = 0.0
    for i in pick:
        c = conds[i]
        ref = oracle.eq_abscoeff(d, c.wmin, c.wmax, c.wstep, c.Tgas, c.pressure_bar, c.mole_fraction, c.dxG, c.dxL)
        err = oracle.rel_peak_error(res[i], ref)
        worst = max(worst, err)
        inf = b.info(i)
        print(f"batch req {i}: W={c.wmax - c.wmin:.1f} p={c.pressure_bar:.2f} L/2Nv={inf.fft_len / (2 * inf.n_v):.3f}"
              f" err {err:.2e}", flush=True)
        assert err <= REL_TOL, i
    print(f"batch_256 graph replay: 16 requests, worst {worst:.2e}")
    b.close()
    db.close()
