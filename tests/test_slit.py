"""Slit functions (host, radis/tools/slit.py) and the GPU slit convolution
(radis_ldm_slit_convolve) against the oracle's direct sum, which is itself
pinned to scipy.signal.oaconvolve, the routine RADIS calls."""
import numpy as np
import pytest

from oracle import ldm_oracle as oracle
from radis_amd import slit


def _fwhm(w, I):
    """Full width at half maximum by linear interpolation of the two
    crossings (radis/tools/slit.py:758-814 get_FWHM)."""
    h = I.max() / 2
    up = np.flatnonzero(I >= h)
    i0, i1 = up.min(), up.max()
    x0 = w[i0 - 1] + (h - I[i0 - 1]) * (w[i0] - w[i0 - 1]) / (I[i0] - I[i0 - 1])
    x1 = w[i1] + (h - I[i1]) * (w[i1 + 1] - w[i1]) / (I[i1 + 1] - I[i1])
    return x1 - x0


@pytest.mark.parametrize("shape,arg", [("triangular", 0.4), ("gaussian", 0.4), ("trapezoidal", (0.36, 0.44))])
@pytest.mark.parametrize("wstep", [0.01, 0.0073, 0.002])
def test_slit_shapes_fwhm_and_area(shape, arg, wstep):
    """test/tools/test_slit.py:67-117: FWHM within 2 wstep of the request;
    area 1 with norm_by='area', max 1 with 'max'."""
    w, I = slit.get_slit_function(arg, shape=shape, center_wavespace=2100.0, wstep=wstep)
    assert np.isclose(_fwhm(w, I), 0.4, atol=2 * wstep)
    assert np.isclose(np.trapezoid(I, w) if hasattr(np, "trapezoid") else np.trapz(I, w), 1.0, rtol=1e-12)
    assert np.isclose(np.sum(w * I) / np.sum(I), 2100.0, atol=wstep)  # centred
    w2, I2 = slit.get_slit_function(arg, shape=shape, center_wavespace=2100.0, wstep=wstep, norm_by="max")
    assert I2.max() == 1.0
    g = slit.slit_on_grid(w, I, wstep)
    assert abs(g.sum() * wstep - 1.0) < 2e-2  # resampled on the grid, the area stays ~1


def test_slit_rejections():
    with pytest.raises(NotImplementedError):
        slit.get_slit_function(0.4, unit="nm", center_wavespace=2100.0, wstep=0.01)
    with pytest.raises(TypeError):
        slit.get_slit_function(0.4, shape="trapezoidal", center_wavespace=2100.0, wstep=0.01)
    with pytest.raises(ValueError):
        slit.slit_on_grid(np.array([0.0, 0.01, 0.02]), np.array([0.1, -0.2, 0.1]), 0.01)


@pytest.mark.parametrize("k", [1, 2, 7, 101, 2048, 2049, 5000])
def test_oracle_slit_convolve_is_oaconvolve(k):
    from scipy.signal import oaconvolve

    rng = np.random.default_rng(k)
    I = rng.random(9000) ** 4
    s = rng.random(k)
    for mode in ("same", "valid"):
        ref = oaconvolve(I, s, mode="same") * 0.003
        got = oracle.slit_convolve(I, s, 0.003, mode=mode)
        ok = ~np.isnan(got)
        np.testing.assert_allclose(got[ok], ref[ok], rtol=1e-10, atol=1e-12 * ref.max())
        if mode == "valid":
            assert np.sum(~ok) == (min(9000, k) - 1) // 2 + min(9000, k) // 2


def test_oracle_slit_convolve_delta_and_constant():
    """A delta returns the slit times wstep; a constant returns the slit's
    discrete area away from the edges (odd slits, as the generators make;
    with an even one RADIS's 'valid' edge keeps one partial point)."""
    s = np.array([1.0, 3.0, 2.0, 0.5, 0.25])
    I = np.zeros(21)
    I[10] = 1.0
    out = oracle.slit_convolve(I, s, 0.1, mode="same")
    c = (s.size - 1) // 2
    np.testing.assert_allclose(out[10 - c:10 - c + s.size], 0.1 * s)
    out = oracle.slit_convolve(np.full(50, 2.0), s, 0.1, mode="valid")
    np.testing.assert_allclose(out[~np.isnan(out)], 2.0 * 0.1 * s.sum())


@pytest.mark.gpu
@pytest.mark.parametrize("n,k", [(1000, 1), (1000, 2), (100003, 101), (100003, 2049), (50000, 5000), (300, 999)])
def test_gpu_slit_convolve_matches_oracle(n, k):
    import torch

    rng = np.random.default_rng(n + k)
    I = (rng.random(n) ** 6).astype(np.float32)
    s = rng.random(k)
    x = torch.as_tensor(I, device="cuda")
    for mode in ("same", "valid") if k <= n else ("same",):
        got = slit.convolve_device(x, s, 0.002, mode=mode).double().cpu().numpy()
        ref = oracle.slit_convolve(I.astype(np.float64), s.astype(np.float32).astype(np.float64), 0.002, mode=mode)
        np.testing.assert_array_equal(np.isnan(got), np.isnan(ref))
        ok = ~np.isnan(ref)
        err = np.max(np.abs(got[ok] - ref[ok])) / np.max(np.abs(ref[ok]))
        assert err <= 3e-7, (n, k, mode, err)


@pytest.mark.gpu
def test_gpu_slit_convolve_rejects_bad_arguments():
    import torch

    from radis_amd import _native

    x = torch.zeros(100, device="cuda")
    lib = _native.load()
    from ctypes import c_void_p
    with pytest.raises(_native.LdmError):
        _native.check(lib.radis_ldm_slit_convolve(c_void_p(x.data_ptr()), 100, c_void_p(x.data_ptr()), 3, 1.0,
                                                  c_void_p(x.data_ptr()), None))  # in == out
    with pytest.raises(_native.LdmError):
        _native.check(lib.radis_ldm_slit_convolve(c_void_p(x.data_ptr()), 0, c_void_p(x.data_ptr()), 3, 1.0,
                                                  c_void_p(x.data_ptr() + 400), None))
    with pytest.raises(ValueError):
        slit.convolve_device(x, np.ones(150), 1.0, mode="valid")
