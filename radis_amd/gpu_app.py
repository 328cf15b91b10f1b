"""``gpuApp``: the reference's GPU application interface, backed by the HIP
library (radis/gpu/gpu.py:35-380).

Same constructor arguments, ``iterate``, ``get_griddims`` and ``__del__``.
The arithmetic follows RADIS's CPU LDM-fft path (the semantics its GPU tests
compare against):

* pressure in bar is converted to atm for the Lorentzian widths and the
  pressure shift (radis/lbl/broadening.py:1063, lbl/base.py:2351);
* LDM axes span the exact min/max widths of the lines in the window
  (radis/lbl/broadening.py:1823-1852);
* grid indices follow ``np.interp`` (radis/lbl/broadening.py:2151-2154),
  clamping out-of-range lines onto the grid edges;
* the Voigt FT templates include alias folding (1915-1926).

The intensities passed as ``S0`` follow the reference's Einstein-A
convention (radis/lbl/base.py:2437-2441, cmdFillLDM.comp:126); they are
converted once to the tabulated S(Tref) the CPU formula starts from:
``S_ref = S0 * exp(-c2 El/Tref) * (1 - exp(-c2 v0/Tref)) / Q(Tref)``,
which is the identity that makes the two formulas agree.
"""
from __future__ import annotations

import time

import numpy as np

from . import _native
from .engine import (DXG_DEFAULT, DXL_DEFAULT, FOLDING_THRESH_DEFAULT, Device, LineDatabase, SpectrumBatch,
                     describe_status, make_request)

C2 = 6.62607015e-34 * 2.99792458e8 / 1.3806490e-23 * 100  # hc/k in cm K
TREF = 296.0  # radis/gpu/params.py:146


def s_ref_from_s0(S0, El, v0, Q_Tref, Tref=TREF):
    """Tabulated-form intensity S(Tref) from the Einstein-A form
    ``S0 = Ia g' A / (8 pi c v0^2)`` (radis/lbl/base.py:2437-2441): times the
    Boltzmann factor and stimulated-emission term at Tref over Q(Tref)."""
    El = np.asarray(El, dtype=np.float64)
    v0 = np.asarray(v0, dtype=np.float64)
    return np.asarray(S0, dtype=np.float64) * np.exp(-C2 * El / Tref) * (1 - np.exp(-C2 * v0 / Tref)) / Q_Tref


def next_fast_len_even(n):
    """FFT length the planner uses for a grid needing at least ``n`` points
    (counterpart of radis/gpu/gpu.py:22-29, for this library's radices)."""
    from .engine import plan_grid

    return plan_grid(0.0, (n // 2 - 1) * 1.0, 1.0)[1]


class gpuApp:
    """Drop-in for ``radis.gpu.gpu.gpuApp`` (radis/gpu/gpu.py:35-54)."""

    def __init__(self, vmin, Nv, dv, dxG, dxL, v0, da, na, S0, El, gamma_arr, iso, Mm_arr, Q_intp_list,
                 verbose=0, device_id=0, folding_thresh=FOLDING_THRESH_DEFAULT):
        self.backend = "hip"
        self.verbose = verbose
        if isinstance(device_id, str):
            device_id = 0
        self.vmin, self.Nv, self.dv = float(vmin), int(Nv), float(dv)
        self.dxG, self.dxL = float(dxG), float(dxL)
        self.folding_thresh = folding_thresh
        self.Q_intp_list = list(Q_intp_list)
        v0 = np.asarray(v0, dtype=np.float64)
        iso = np.asarray(iso).astype(np.int64)
        Mm = np.asarray(Mm_arr, dtype=np.float64)
        gamma_arr = np.asarray(gamma_arr, dtype=np.float64)
        El64 = np.asarray(El, dtype=np.float64)
        Q = np.array([self.Q_intp_list[int(i)](TREF) for i in range(len(self.Q_intp_list))], dtype=np.float64)
        S_ref = s_ref_from_s0(S0, El64, v0, Q[iso])
        # grid: vmin + i*dv for i < Nv; every line handed in is used (lines
        # outside the grid are clamped onto its edges, as np.interp does)
        self.wmax = self.vmin + (self.Nv - 1) * self.dv
        lines = dict(
            nu0=v0, S_ref=S_ref, El=El64, g_air=gamma_arr[-1], g_self=gamma_arr[0],
            n_air=np.asarray(na, dtype=np.float64),
            delta=np.asarray(da, dtype=np.float64) if np.any(np.asarray(da) != 0) else None,
            iso=iso.astype(np.uint8) if np.unique(iso).size > 1 or iso[0] != 1 else None,
            molar_mass=Mm, Q=lambda s, T: self.Q_intp_list[s](T),
        )
        t0 = time.perf_counter()
        self._dev = Device(int(device_id))
        self._db = LineDatabase(self._dev, lines)
        self._batch = None
        self._out = None
        self._griddims = (0, 0)
        if verbose >= 2:
            print(f"gpuApp: {self._db.n} lines uploaded in {time.perf_counter() - t0:.3f} s")

    def _request(self, p, T, x):
        return make_request(self._db, self.vmin, self.wmax, self.dv, T, p, x, dxG=self.dxG, dxL=self.dxL,
                            Tref=TREF, folding_thresh=self.folding_thresh, n_v=self.Nv,
                            line_window=(-np.inf, np.inf))

    def iterate(self, p, T, mole_fraction, verbose=0, l=1.0, slit_FWHM=0.0, derived=False):
        """abscoeff (cm-1, float32[Nv]) and per-stage device times (ms) for
        pressure ``p`` (bar), temperature ``T`` (K) and mole fraction
        (radis/gpu/gpu.py:302-370).  ``slit_FWHM`` is accepted and ignored,
        as in the reference; so is ``l`` unless ``derived`` is set: then the
        last GPU pass also writes absorbance, emissivity, transmittance and
        radiance for a path of ``l`` cm (radis/lbl/factory.py:1289-1297),
        left as float64 arrays in ``self.derived``."""
        req = self._request(p, T, mole_fraction)
        t0 = time.perf_counter()
        if self._batch is None:
            self._batch = SpectrumBatch(self._db, [req])
            if self._batch.n_v(0) != self.Nv:
                raise ValueError(f"grid has {self._batch.n_v(0)} points, expected Nv={self.Nv}")
            self._batch.enable_timing(True)
            self._out = self._batch.alloc_outputs()
        else:
            self._batch.set_request(0, req)
        want = float(l) if derived else None
        if want != getattr(self, "_derived_l", None):
            self._batch.set_derived(0, want or 0.0, which=SpectrumBatch.DERIVED if derived else ())
            self._derived_l = want
        self._batch.run(self._out)
        inf = self._batch.info(0)
        if inf.status:
            raise _native.LdmError(_native.EDEVICE, describe_status(inf.status))
        self._griddims = (int(inf.n_l), int(inf.n_g))
        out = self._out[0].cpu().numpy()
        if inf.scale_log2:
            with np.errstate(over="ignore"):
                out = np.ldexp(out.astype(np.float64), int(inf.scale_log2)).astype(np.float32)
        self.derived = self._batch.derived_results(0) if derived else {}
        times = self._batch.stage_ms()
        times["total"] = (time.perf_counter() - t0) * 1e3
        if verbose == 1:
            print("Finished calculating spectrum!")
        return out.astype(np.float32, copy=False), times

    def get_griddims(self):
        """(N_L, N_G) of the last iteration (radis/gpu/gpu.py:372-374)."""
        return self._griddims

    def __del__(self):
        for name in ("_batch", "_db", "_dev"):
            obj = getattr(self, name, None)
            if obj is not None:
                try:
                    obj.close()
                except Exception:
                    pass
                setattr(self, name, None)
