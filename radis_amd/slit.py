"""Instrumental slit functions and their convolution on the GPU
(radis/tools/slit.py, Spectrum.apply_slit in radis/spectrum/spectrum.py:3127).

The slit is generated on the host, as RADIS does, because it is a few
hundred to a few thousand samples. It is then resampled on the spectrum
grid exactly as ``convolve_with_slit`` does (a linear spline evaluated at
``arange(w_slit[0], w_slit[-1] + wstep, wstep)``) and convolved with the
spectral quantities on the device by ``radis_ldm_slit_convolve``. The
convolution is the ``'same'`` one of ``scipy.signal.oaconvolve`` times
``wstep``. In ``'valid'`` mode the edges that the slit reaches past the
computed range become NaN, as ``remove_boundary`` does.

Supported: slits given by FWHM (triangular, gaussian) or by (top, base)
(trapezoidal), or as sampled (w, I) arrays, in cm-1; normalisation by area
or by max. Slits in nm on cm-1 spectra (RADIS resamples the spectrum in
wavelength for that), slit dispersion and experimental-slit file import are
not part of this package.
"""
from __future__ import annotations

import math
from ctypes import c_void_p

import numpy as np

from . import _native
from ._native import check

SLIT_SHAPES = ("triangular", "gaussian", "trapezoidal")


def _trapezoid(y, x):
    from scipy.integrate import trapezoid

    return trapezoid(y, x=x)


def _normalize(w, I, norm_by):
    if norm_by == "area":
        return I / _trapezoid(I, w)
    if norm_by == "max":
        return I / np.max(I)
    if norm_by is None:
        return I
    raise ValueError(f"Unknown normalization type: `norm_by` = {norm_by}")


def triangular_slit(FWHM, wstep, center=0.0, norm_by="area", footerspacing=0):
    """Triangle of full width at half maximum FWHM sampled at ~wstep, odd
    number of points, centred on its maximum; radis/tools/slit.py:1522-1650."""
    slope = 1 / (FWHM - wstep)
    a = int(FWHM / wstep)
    half = 1 - np.arange(0, a) * slope * wstep
    if FWHM % wstep:  # one more zero when wstep does not divide FWHM
        footerspacing += 1
    half = np.hstack((half, np.zeros(int(footerspacing))))
    wh = np.linspace(0, len(half) * wstep, len(half))
    I = np.hstack((half[1:][::-1], half))
    w = np.hstack((-wh[1:][::-1], wh)) + center
    return w, _normalize(w, I, norm_by)


def gaussian_slit(FWHM, wstep, center=0.0, norm_by="area", calc_range=4, footerspacing=0):
    """Gaussian of full width at half maximum FWHM over +-calc_range sigma;
    radis/tools/slit.py:1786-1880."""
    f = int(footerspacing)
    sigma = FWHM / 2 / math.sqrt(2 * math.log(2))
    a = 2 * int(calc_range * sigma // wstep / 2)
    w0 = wstep * np.linspace(-a, a, 2 * a + 1)
    I = np.hstack((np.zeros(f), np.exp(-(w0**2) / (2 * sigma**2)), np.zeros(f)))
    w = wstep * np.linspace(-(a + f), a + f, 2 * a + 2 * f + 1) + center
    return w, _normalize(w, I, norm_by)


def trapezoidal_slit(top, base, wstep, center=0.0, norm_by="area", footerspacing=0):
    """Trapezoid of plateau width ``top`` and base width ``base`` (FWHM =
    (top + base)/2); radis/tools/slit.py:1651-1785."""
    if top > base:
        top, base = base, top
    FWHM = (base + top) / 2
    b = 2 * int(top / wstep // 2) + 1  # points on the plateau (odd)
    slope = 1 / (FWHM - b * wstep)
    a = int(FWHM / wstep) - b  # points on a slope
    side = 1 - np.arange(0, a + 1) * slope * wstep
    if len(side) == 0:
        side = np.ones(1)
    if side[-1] < 0:
        side[-1] = 0
    elif side[-1] > 0:
        footerspacing += 1
    side = np.hstack((side, np.zeros(int(footerspacing))))
    I = np.hstack((side[1:][::-1], np.ones(b), side[1:]))
    w = wstep * np.linspace(-len(I) / 2, len(I) / 2, len(I)) + center
    return w, _normalize(w, I, norm_by)


def get_slit_function(slit_function, unit="cm-1", norm_by="area", shape="triangular", center_wavespace=None,
                      return_unit="cm-1", wstep=None):
    """(w_slit, I_slit) in ``return_unit`` (cm-1 here), as
    radis/tools/slit.py:65-475 builds it for a FWHM, a (top, base) tuple or a
    sampled (w, I) pair."""
    if return_unit != "cm-1":
        raise NotImplementedError("spectra of this package are in cm-1")
    if unit != "cm-1":
        # RADIS applies an nm slit to a cm-1 spectrum by resampling the
        # spectrum in wavelength (Spectrum._apply_slit_nm_on_cm1)
        raise NotImplementedError("slits in cm-1 only")
    if isinstance(slit_function, (int, float)):
        FWHM = float(slit_function)
        if shape == "triangular":
            return triangular_slit(FWHM, wstep, center=center_wavespace, norm_by=norm_by)
        if shape == "gaussian":
            return gaussian_slit(FWHM, wstep, center=center_wavespace, norm_by=norm_by)
        if shape == "trapezoidal":
            raise TypeError("A (top, base) tuple must be given with a trapezoidal slit")
        raise TypeError(f"Slit function ({shape}) not in known slit shapes: {SLIT_SHAPES}")
    if isinstance(slit_function, tuple):
        top, base = slit_function
        if shape not in ("trapezoidal", "triangular"):
            raise TypeError("A (top, base) tuple must be used with a trapezoidal slit")
        return trapezoidal_slit(top, base, wstep, center=center_wavespace, norm_by=norm_by)
    if isinstance(slit_function, (list, np.ndarray)) and np.shape(slit_function)[0] == 2:
        w, I = (np.asarray(a, dtype=np.float64) for a in slit_function)
        return w, _normalize(w, I, norm_by)
    raise TypeError(f"Wrong format for slit function: {type(slit_function)}")


def slit_on_grid(w_slit, I_slit, wstep):
    """The slit sampled at the spectrum step, as convolve_with_slit does it
    (radis/tools/slit.py:690-707): a linear spline through (w_slit, I_slit)
    at arange(w_slit[0], w_slit[-1] + wstep, wstep) unless the slit already
    has that step."""
    from scipy.interpolate import splev, splrep

    w_slit = np.asarray(w_slit, dtype=np.float64)
    I_slit = np.asarray(I_slit, dtype=np.float64)
    if w_slit[-1] < w_slit[0]:
        w_slit, I_slit = w_slit[::-1], I_slit[::-1]
    if np.allclose(np.diff(w_slit), wstep):
        I = I_slit
    else:
        I = splev(np.arange(w_slit[0], w_slit[-1] + wstep, wstep), splrep(w_slit, I_slit, k=1))
    if np.any(np.isnan(I)):
        raise ValueError("Slit has nan value")
    if not (I_slit >= 0).all():
        raise ValueError("Slit is partially negative")
    return np.asarray(I, dtype=np.float64)


def convolve_device(x, slit, wstep, mode="valid", out=None, stream=None):
    """Convolve the float32 device tensor ``x`` with the sampled slit (numpy,
    on the spectrum grid) on the GPU; returns a new float32 device tensor
    (``'valid'``: NaN over the edges the slit reaches past the grid)."""
    import torch

    if mode not in ("valid", "same"):
        raise ValueError(f"mode must be 'valid' or 'same', not {mode!r}")
    if x.dtype != torch.float32 or not x.is_cuda or not x.is_contiguous():
        raise ValueError("x must be a contiguous float32 device tensor")
    n, k = x.numel(), len(slit)
    if mode == "valid" and k > n:
        raise ValueError(f"slit of {k} points is longer than the spectrum ({n} points): use mode='same'")
    ks = torch.as_tensor(np.ascontiguousarray(slit, dtype=np.float32), device=x.device)
    y = torch.empty_like(x) if out is None else out
    lib = _native.load()
    s = 0 if stream is None else stream.cuda_stream
    check(lib.radis_ldm_slit_convolve(c_void_p(x.data_ptr()), n, c_void_p(ks.data_ptr()), k, float(wstep),
                                      c_void_p(y.data_ptr()), c_void_p(s)))
    if mode == "valid":
        la = min(n, k)
        a, b = int((la - 1) / 2), int(la / 2)  # radis/tools/slit.py:1064-1071
        if a:
            y[:a] = float("nan")
        if b:
            y[-b:] = float("nan")
    return y
