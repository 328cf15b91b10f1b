"""The no-allocation entry point (radis_ldm_run_noalloc) against the batch
entry point (radis_ldm_batch_create + radis_ldm_batch_run) on the same
requests: two calls in a row, the second with a smaller request, the device's
free memory unchanged across them, and a request beyond the scratch array
refused."""
import numpy as np
import pytest

from oracle import ldm_oracle as oracle
from radis_amd import _native, engine, workloads

pytestmark = pytest.mark.gpu


def test_noalloc_matches_batch_entry_point():
    import torch

    dev = engine.Device(0)
    lines = workloads.co2_like(400_000, seed=61)
    db = engine.LineDatabase(dev, lines)
    conds = [workloads.Conditions(2050.0, 2250.0, 0.001, 1300.0, 2.0, 0.2),
             workloads.Conditions(2300.0, 2325.0, 0.001, 600.0, 15.0, 0.5)]
    reqs = [engine.request_from_conditions(db, c) for c in conds]
    # the existing entry point, one batch per request
    want, winfo = [], []
    for r in reqs:
        b = engine.SpectrumBatch(db, [r])
        want.append(b.results(b.run())[0])
        winfo.append(b.info(0))
        b.close()
    print(f"requests: N_v {[i.n_v for i in winfo]}, L {[i.fft_len for i in winfo]}", flush=True)
    outs = [torch.full((int(i.n_v),), float("nan"), dtype=torch.float32, device="cuda") for i in winfo]
    results = [torch.zeros(6, dtype=torch.float64, device="cuda") for _ in reqs]
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    infos = [engine.run_noalloc(db, r, o, res) for r, o, res in zip(reqs, outs, results)]
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert free1 == free0
    for k, (c, inf, o, res) in enumerate(zip(conds, infos, outs, results)):
        r = engine.read_result(res)
        w = winfo[k]
        assert (r.n_g, r.n_l, r.status) == (w.n_g, w.n_l, 0)
        assert (inf.n_v, inf.fft_len, inf.scale_log2) == (w.n_v, w.fft_len, w.scale_log2)
        got = np.ldexp(o.double().cpu().numpy(), int(inf.scale_log2))
        err = oracle.rel_peak_error(got, want[k])
        ref = oracle.eq_abscoeff(lines.as_dict(), c.wmin, c.wmax, c.wstep, c.Tgas, c.pressure_bar,
                                 c.mole_fraction, c.dxG, c.dxL)
        print(f"call {k}: N_G={r.n_g} N_L={r.n_l} no-alloc vs batch {err:.2e}, vs oracle "
              f"{oracle.rel_peak_error(got, ref):.2e}", flush=True)
        assert err <= 1e-6
        assert oracle.rel_peak_error(got, ref) <= 1e-4
    db.close()
    dev.close()


def test_noalloc_refuses_a_request_beyond_its_scratch():
    import torch

    dev = engine.Device(0)
    lines = workloads.h2o_like(200_000, seed=62, chunk=200_000)
    db = engine.LineDatabase(dev, lines)
    c = workloads.CONDITIONS["h2o_100m"]
    req = engine.request_from_conditions(db, c)
    out = torch.empty(600_001, dtype=torch.float32, device="cuda")
    with pytest.raises(_native.LdmError, match="scratch"):
        engine.run_noalloc(db, req, out)
    db.close()
    dev.close()
