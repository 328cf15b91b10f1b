"""Autotuner: candidate generation (CPU) and a tuned request on the device
(schedule applied through make_request, result within tolerance)."""
import numpy as np
import pytest

from radis_amd import autotune


@pytest.mark.parametrize("nv", [1000, 20001, 200001, 500002, 600001])
def test_fft_candidates_cover_the_grid(nv):
    c = autotune.fft_candidates(nv)
    assert c, nv
    gen = [x for x in c if x[2]]
    comp = [x for x in c if not x[2]]
    Lmin = gen[0][0]
    for L, n1, generic in c:
        assert L >= 2 * nv and L % n1 == 0
        n2 = L // n1
        if generic:
            assert L == Lmin and n1 % 16 == 0 and n2 % 16 == 0 and n1 <= 1024 and n2 <= 8192
        else:
            assert n1 in autotune.FC_N1 and n2 in autotune.FC_N2 and L <= autotune.FFT_SLACK * Lmin
    # 7-smooth lengths (multiples of 256) are dense
    assert Lmin <= 1.06 * 2 * nv or Lmin <= 4096
    if nv == 500002:
        assert (1048576, 512, False) in comp


def test_tile_candidates_fit_lds():
    for m_cap in (4, 27, 40, 144, 300):
        ws = autotune.tile_candidates(m_cap)
        assert ws and all(m_cap * (w + autotune.fill_pad(w)) * 8 <= autotune.FILL_LDS_MAX for w in ws)
    assert autotune.tile_candidates(500) == []


@pytest.mark.gpu
def test_tuned_schedule_applies_and_matches_oracle(tmp_path, monkeypatch):
    from oracle import ldm_oracle as oracle
    from radis_amd import engine, workloads

    monkeypatch.setenv("RADIS_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    lines = workloads.co2_like(200_000, seed=77)
    c = workloads.Conditions(2100.0, 2160.0, 0.001, 1100.0, 2.0, 0.3)
    dev = engine.Device(0)
    db = engine.LineDatabase(dev, lines)
    r0 = engine.request_from_conditions(db, c)
    assert (r0.fft_len, r0.fft_n1, r0.tile_w) == (0, 0, 0)  # empty cache: planner defaults
    entry = autotune.tune(db, r0, reps=3)
    assert entry["candidates"] >= 2 and entry["tile_w"] in autotune.TILE_WIDTHS
    r1 = engine.request_from_conditions(db, c)
    assert (r1.tile_w, r1.fft_n1) == (entry["tile_w"], entry["fft_n1"])
    assert r1.fft_len == entry["fft_len"]
    assert bool(r1.flags & 2) == entry["generic"]
    b = engine.SpectrumBatch(db, [r1])
    out = b.results(b.run())[0]
    inf = b.info(0)
    assert int(inf.fft_len) == entry["fft_len"] and int(inf.tile_w) == entry["tile_w"]
    ref = oracle.eq_abscoeff(lines.as_dict(), c.wmin, c.wmax, c.wstep, c.Tgas, c.pressure_bar, c.mole_fraction,
                             c.dxG, c.dxL)
    assert oracle.rel_peak_error(out, ref) <= 1e-4
    # an explicit schedule is never overridden by the cache
    r2 = engine.make_request(db, c.wmin, c.wmax, c.wstep, c.Tgas, c.pressure_bar, c.mole_fraction, c.dxG, c.dxL,
                             tile_w=32)
    assert r2.tile_w == 32 and r2.fft_len == 0
    b.close()
    db.close()
    dev.close()
    assert np.isfinite(entry["fill_ms"]) and np.isfinite(entry["fft_ms"])
