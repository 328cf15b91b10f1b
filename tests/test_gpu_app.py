"""``gpuApp`` (radis/gpu/gpu.py:35-380) driven the way the reference factory
builds it (radis/lbl/factory.py:1186-1258): float32 line columns, Einstein-A
intensities ``S0``, ``gamma_arr = [self, air]``, per-iso molar masses and
partition-function callables.
"""
import math

import numpy as np
import pytest

from oracle import ldm_oracle as oracle
from radis_amd import gpu_app, hitran

FIX = "tests/golden/hitran_co_3iso_2000_2300cm.par"
# CO isotopologues 1-3: abundance and Q(296 K), radis/db/molparam.txt:51-53
IA = {1: 0.986544, 2: 1.10836e-2, 3: 1.97822e-3}
Q296 = {1: 107.12, 2: 224.08, 3: 112.47}
C_CGS = 2.99792458e10


def _s0(c):
    """Einstein-A form the factory hands over, radis/lbl/base.py:2437-2441."""
    ia = np.array([IA[int(i)] for i in c["iso"]])
    return ia * c["gu"] * c["A"] / (8 * math.pi * C_CGS * c["nu0"] ** 2)


def test_s0_conversion_reproduces_tabulated_intensities():
    """S(296 K) recovered from A, g', abundance and Q(296) agrees with the
    HITRAN-tabulated intensity of every line of the fixture to the
    precision of the tabulated values (a wrong Boltzmann or
    stimulated-emission term is off by orders of magnitude)."""
    c = hitran.read_par(FIX)
    assert np.all(np.isfinite(c["gu"]))
    q = np.array([Q296[int(i)] for i in c["iso"]])
    S = gpu_app.s_ref_from_s0(_s0(c), c["El"], c["nu0"], q)
    np.testing.assert_allclose(S, c["S_ref"], rtol=5e-3)


def _Q(iso, T):
    return Q296[iso] * (T / 296.0) ** 1.05


def _factory_inputs(vmin, vmax):
    c = hitran.read_par(FIX)
    order = np.argsort(c["nu0"], kind="stable")
    c = {k: v[order] for k, v in c.items()}
    keep = (c["nu0"] >= vmin) & (c["nu0"] <= vmax)
    c = {k: v[keep] for k, v in c.items()}
    f32 = lambda a: np.asarray(a, dtype=np.float32)  # noqa: E731  (factory.py:1247-1251)
    gamma = np.zeros((2, c["nu0"].size), dtype=np.float32)
    gamma[0], gamma[1] = c["g_self"], c["g_air"]
    Mm = np.array(hitran.CO_MOLAR_MASS[:4])
    Qlist = [lambda T: 1.0] + [(lambda T, i=i: _Q(i, T)) for i in (1, 2, 3)]
    args = (f32(c["nu0"]), f32(c["delta"]), f32(c["n_air"]), f32(_s0(c)), f32(c["El"]), gamma,
            c["iso"].astype(np.uint32), Mm, Qlist)
    return c, args


def _oracle_lines(args):
    v0, da, na, S0, El, gamma, iso, Mm, _ = args
    q = np.array([Q296[int(i)] for i in iso])
    return dict(nu0=v0.astype(np.float64), S_ref=gpu_app.s_ref_from_s0(S0, El, v0, q), El=El, g_air=gamma[1],
                g_self=gamma[0], n_air=na, delta=da, iso=iso.astype(np.int64), molar_mass=Mm, Q=_Q)


@pytest.mark.gpu
def test_gpuapp_iterate_matches_oracle():
    vmin, dv, Nv = 2050.0, 0.002, 100_001
    vmax = vmin + (Nv - 1) * dv
    assert oracle.wavenumber_grid(vmin, vmax, dv).size == Nv
    _, args = _factory_inputs(vmin, vmax)
    app = gpu_app.gpuApp(vmin, Nv, dv, 0.14, 0.2, *args)
    lines = _oracle_lines(args)
    for (p, T, x) in [(1.0, 1000.0, 0.1), (0.1, 2500.0, 0.01), (5.0, 300.0, 0.5)]:
        a, times = app.iterate(p, T, x)
        assert a.dtype == np.float32 and a.shape == (Nv,)
        assert set(times) >= {"fill", "fft_a", "fft_b", "fft_c", "total"}
        ref = oracle.eq_abscoeff(lines, vmin, vmax, dv, T, p, x, 0.14, 0.2)
        err = oracle.rel_peak_error(a, ref)
        NwL, NwG = app.get_griddims()
        print(f"p={p} T={T} x={x}: err {err:.2e} NwL={NwL} NwG={NwG}")
        assert err <= 1e-4
        assert NwL > 0 and NwG > 0
    # derived quantities from the last pass, path of 20 cm
    a, _ = app.iterate(1.0, 1500.0, 0.2, l=20.0, derived=True)
    w = vmin + np.arange(Nv) * ((vmin + dv) - vmin)
    ref = oracle.derived_quantities(w, a.astype(np.float64), 20.0, 1500.0)
    assert oracle.rel_peak_error(app.derived["radiance"], ref["radiance"]) <= 2e-6
    assert np.max(np.abs(app.derived["transmittance"] - ref["transmittance"])) <= 2e-6
    a2, _ = app.iterate(1.0, 1500.0, 0.2)
    assert app.derived == {}
    np.testing.assert_array_equal(a, a2)
    app.__del__()


@pytest.mark.gpu
def test_two_live_gpuapps_do_not_share_state():
    """The reference keeps its parameters in module globals, so two live
    gpuApps clobber each other (radis/gpu/params.py:66-86); here each owns
    its database and batch."""
    vmin, dv, Nv = 2050.0, 0.002, 50_001
    vmax = vmin + (Nv - 1) * dv
    _, args = _factory_inputs(vmin, vmax)
    lines = _oracle_lines(args)
    n = args[0].size
    half = [a[::2] if isinstance(a, np.ndarray) and a.shape == (n,) else a for a in args]
    half[5] = args[5][:, ::2]  # gamma_arr
    app1 = gpu_app.gpuApp(vmin, Nv, dv, 0.14, 0.2, *args)
    app2 = gpu_app.gpuApp(vmin, Nv, dv, 0.14, 0.2, *half)
    a1, _ = app1.iterate(1.0, 1000.0, 0.1)
    a2, _ = app2.iterate(0.5, 2000.0, 0.3)
    a1b, _ = app1.iterate(1.0, 1000.0, 0.1)
    np.testing.assert_array_equal(a1, a1b)
    ref1 = oracle.eq_abscoeff(lines, vmin, vmax, dv, 1000.0, 1.0, 0.1, 0.14, 0.2)
    ref2 = oracle.eq_abscoeff(_oracle_lines(tuple(half)), vmin, vmax, dv, 2000.0, 0.5, 0.3, 0.14, 0.2)
    assert oracle.rel_peak_error(a1, ref1) <= 1e-4
    assert oracle.rel_peak_error(a2, ref2) <= 1e-4
    app1.__del__()
    app2.__del__()
