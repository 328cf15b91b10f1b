"""Caller-side facade of the GPU path: ``SpectrumFactory.eq_spectrum_gpu``,
``Spectrum.recalc_gpu`` / ``exit_gpu`` and ``calc_spectrum(mode='gpu')``
(radis/lbl/factory.py:1027-1391, radis/spectrum/spectrum.py:6438-6577,
radis/lbl/calc.py:37, 1014-1028).

Only the parts of those callers that belong to the GPU path are reproduced:
unit handling (K, bar, cm), the calculation grid, the synthesis itself, the
derived quantities computed right after it (absorbance, emissivity,
transmittance, Planck radiance; radis/lbl/factory.py:1284-1297, written here
by the last GPU pass next to abscoeff) and the ``NwL``/``NwG`` conditions.  Line lists are handed in directly (``lines=``);
databank download and file parsing are not part of this package.
"""
from __future__ import annotations

import time

import numpy as np

from . import engine, partfunc
from .engine import DXG_DEFAULT, DXL_DEFAULT, FOLDING_THRESH_DEFAULT

class Spectrum:
    """Result of ``eq_spectrum_gpu`` holding the GPU application until
    ``exit_gpu`` (radis/spectrum/spectrum.py:176-179, 366, 528)."""

    DERIVED_NAMES = {"absorbance": "absorbance", "emissivity": "emissivity_noslit",
                     "transmittance": "transmittance_noslit", "radiance": "radiance_noslit"}
    UNITS = {"abscoeff": "cm-1", "absorbance": "", "emissivity_noslit": "", "transmittance_noslit": "",
             "radiance_noslit": "mW/cm2/sr/nm", "wavenumber": "cm-1", "radiance": "mW/cm2/sr/nm",
             "transmittance": "", "emissivity": ""}

    def __init__(self, wavenumber, abscoeff, derived, conditions, gpu_app=None, name=None):
        self._w = wavenumber
        self.conditions = dict(conditions)
        self.gpu_app = gpu_app
        self.name = name
        self._q = {}
        self._set_quantities(abscoeff, derived)

    def _set_quantities(self, abscoeff, derived):
        """abscoeff and the quantities the last GPU pass derived from it
        (radis/lbl/factory.py:1284-1297), cropped to ``woutrange``."""
        a, b = self.conditions["woutrange"]
        self._q = {"abscoeff": abscoeff[a:b]}
        for k, name in self.DERIVED_NAMES.items():
            self._q[name] = derived[k][a:b]

    def get_vars(self):
        return list(self._q)

    def apply_slit(self, slit_function, unit="cm-1", shape="triangular", center_wavespace=None, norm_by="area",
                   mode="valid", store=True):
        """Convolve every ``*_noslit`` quantity with the slit on the GPU
        (radis/spectrum/spectrum.py:3127-3500, cm-1 slits): radiance,
        transmittance and emissivity.  Returns the spectrum."""
        import torch

        from . import slit as _slit

        w = self._w
        if isinstance(slit_function, (int, float)) and float(slit_function) == 0.0:
            # no slit (the interactive view's default): the convolved
            # quantities are the unconvolved ones
            if store:
                for qns in [k for k in self._q if k.endswith("_noslit")]:
                    self._q[qns[: -len("_noslit")]] = self._q[qns].copy()
                self._slit = (np.array([0.0]), np.array([1.0]))
                self.conditions.update(slit_function=0.0, slit_unit=unit, slit_shape=shape, norm_by=norm_by)
            return self
        wstep = float(np.abs(np.diff(w)).min())
        if center_wavespace is None:
            center_wavespace = w[len(w) // 2]
        wslit, Islit = _slit.get_slit_function(slit_function, unit=unit, norm_by=norm_by, shape=shape,
                                               center_wavespace=center_wavespace, return_unit="cm-1", wstep=wstep)
        if mode == "valid" and abs(w[-1] - w[0]) < abs(wslit[-1] - wslit[0]):
            raise AssertionError("slit wider than the spectral range: nothing would be left in mode='valid'")
        on_grid = _slit.slit_on_grid(wslit, Islit, wstep)
        conv = {}
        for qns in [k for k in self._q if k.endswith("_noslit")]:
            x = torch.as_tensor(np.ascontiguousarray(self._q[qns], dtype=np.float32), device="cuda")
            conv[qns[: -len("_noslit")]] = _slit.convolve_device(x, on_grid, wstep, mode=mode).double().cpu().numpy()
        if store:
            self._q.update(conv)
            self._slit = (wslit, Islit)
            self.conditions.update(slit_function=slit_function, slit_unit=unit, slit_shape=shape, norm_by=norm_by)
        return self

    def get_slit(self):
        """(w, I) of the last slit applied (radis/spectrum/spectrum.py get_slit)."""
        if getattr(self, "_slit", None) is None:
            raise KeyError("no slit applied")
        return self._slit

    def get(self, var, wunit="cm-1", Iunit="default", copy=True):
        """(wavenumber, quantity); only cm-1 and the default units."""
        if wunit not in ("cm-1", "default"):
            raise NotImplementedError("only wunit='cm-1'")
        if Iunit not in ("default", self.UNITS.get(var)):
            raise NotImplementedError("only default units")
        y = self._q[var]
        return (self._w.copy(), y.copy()) if copy else (self._w, y)

    def recalc_gpu(self, var="abscoeff", wunit="default", Iunit="default", pressure=None, Tgas=None,
                   mole_fraction=None, path_length=None, slit_function=None):
        """Recompute with new conditions on the resident GPU application;
        radis/spectrum/spectrum.py:6438-6572."""
        if self.gpu_app is None:
            raise RuntimeError("GPU not initialized: recalc_gpu() needs a spectrum from eq_spectrum_gpu() "
                               "that has not called exit_gpu()")
        t0 = time.perf_counter()
        for k, v in (("Tgas", Tgas), ("pressure", pressure), ("mole_fraction", mole_fraction),
                     ("path_length", path_length), ("slit_function", slit_function)):
            if v is not None:
                self.conditions[k] = v
        self.conditions["Tvib"] = self.conditions["Trot"] = self.conditions["Tgas"]
        abscoeff, _ = self.gpu_app.iterate(self.conditions["pressure"], self.conditions["Tgas"],
                                           self.conditions["mole_fraction"],
                                           path_length=self.conditions["path_length"])
        NwL, NwG = self.gpu_app.get_griddims()
        self.conditions["NwL"], self.conditions["NwG"] = NwL, NwG
        self._set_quantities(abscoeff, self.gpu_app.derived)
        if getattr(self, "_slit", None) is not None:
            # the slit-convolved quantities follow the new conditions (RADIS
            # recomputes them from the conditions when they are asked for)
            c = self.conditions
            self.apply_slit(c.get("slit_function", 0.0), unit=c.get("slit_unit", "cm-1"),
                            shape=c.get("slit_shape", "triangular"), norm_by=c.get("norm_by", "area"))
        self.conditions["calculation_time"] = time.perf_counter() - t0
        return self.get(var, wunit=wunit, Iunit=Iunit, copy=False)[1]

    def plot(self, var=None, wunit="default", Iunit="default", show=False, **kwargs):
        """Plot ``var`` against wavenumber with matplotlib; returns the
        Line2D (radis/spectrum/spectrum.py plot, cm-1 and default units)."""
        import matplotlib.pyplot as plt

        var = var or self.get_vars()[0]
        w, y = self.get(var, wunit=wunit, Iunit=Iunit)
        fig = plt.figure()
        ax = fig.add_subplot(111)
        (line,) = ax.plot(w, y, **kwargs)
        ax.set_xlabel("Wavenumber (cm-1)")
        unit = self.UNITS.get(var, "")
        ax.set_ylabel(f"{var} ({unit})" if unit else var)
        if show:
            plt.show()
        return line

    def exit_gpu(self, event=None):
        """radis/spectrum/spectrum.py:6574-6577."""
        if self.gpu_app is not None:
            self.gpu_app.__del__()
            self.gpu_app = None


class SpectrumFactory:
    """The parts of ``radis.lbl.factory.SpectrumFactory`` that drive the GPU
    path (radis/lbl/factory.py:415-452 defaults)."""

    def __init__(self, wavenum_min=None, wavenum_max=None, wstep=0.01, Tref=296, pressure=1.01325,
                 mole_fraction=1, path_length=1, dxG=DXG_DEFAULT, dxL=DXL_DEFAULT,
                 folding_thresh=FOLDING_THRESH_DEFAULT, optimization="simple", lines=None, verbose=0, **kwargs):
        if wavenum_min is None or wavenum_max is None:
            raise ValueError("wavenum_min and wavenum_max (cm-1) are required")
        if kwargs.get("neighbour_lines", 0):
            raise NotImplementedError("neighbour_lines > 0")
        self.wavenum_min, self.wavenum_max, self.wstep = float(wavenum_min), float(wavenum_max), float(wstep)
        self.Tref = float(Tref)
        self.pressure, self.mole_fraction, self.path_length = pressure, mole_fraction, path_length
        self.dxG, self.dxL, self.folding_thresh = dxG, dxL, folding_thresh
        if optimization not in ("simple", "min-RMS"):
            raise NotImplementedError(f"optimization={optimization!r} (this package: 'simple' or 'min-RMS')")
        self.optimization = optimization
        self.verbose = verbose
        self.lines = None
        self.parfunc = None
        if lines is not None:
            self.load_lines(lines, parfunc=kwargs.get("parfunc"), parfuncfmt=kwargs.get("parfuncfmt", "cdsd"))

    def load_lines(self, lines, parfunc=None, parfuncfmt="cdsd"):
        """Use ``lines`` (a LineList or a mapping of columns) as the line
        database; replaces load_databank/fetch_databank.  ``parfunc`` is a
        tabulated partition-function file (``parfuncfmt='cdsd'``: the CDSD
        CO2 table format, radis/lbl/loader.py:162, 1023-1060, 1188;
        radis/levels/partfunc_cdsd.py:44-117), interpolated and evaluated
        on the device for every isotopologue of the lines."""
        if parfunc is not None and parfuncfmt != "cdsd":
            raise NotImplementedError(f"parfuncfmt={parfuncfmt!r} (this package reads 'cdsd' tables; "
                                      "'hapi' needs the HITRAN Python interface)")
        self.lines = lines
        self.parfunc = parfunc
        return self

    @property
    def wavenumber(self):
        # radis/lbl/factory.py:2590-2592
        return np.arange(self.wavenum_min, self.wavenum_max + self.wstep, self.wstep)

    def _line_columns(self):
        L = self.lines
        get = (lambda k: getattr(L, k, None)) if not isinstance(L, dict) else L.get
        nu0 = np.asarray(get("nu0"), dtype=np.float64)
        keep = (nu0 >= self.wavenum_min) & (nu0 <= self.wavenum_max)  # radis/lbl/loader.py:2707
        col = lambda k: None if get(k) is None else np.asarray(get(k))[keep]
        return dict(nu0=nu0[keep], S_ref=col("S_ref"), El=col("El"), g_air=col("g_air"), g_self=col("g_self"),
                    n_air=col("n_air"), n_self=col("n_self"), delta=col("delta"), iso=col("iso"),
                    molar_mass=get("molar_mass"), Q=get("Q"))

    def eq_spectrum_gpu(self, Tgas, mole_fraction=None, diluent=None, path_length=None, pressure=None,
                        name=None, backend="hip", device_id=0, exit_gpu=False, verbose=None):
        """Equilibrium spectrum on the GPU; radis/lbl/factory.py:1027-1391."""
        if backend not in ("hip", "vulkan"):
            raise NotImplementedError(f"Unknown GPU backend {backend!r}")
        if diluent not in (None, "air") and diluent != {"air": 1 - (mole_fraction or self.mole_fraction)}:
            raise NotImplementedError("only air as diluent")
        if self.lines is None:
            raise ValueError("no line database: call load_lines() first")
        if not np.isscalar(Tgas):
            raise ValueError("Tgas should be float.")
        if mole_fraction is not None:
            self.mole_fraction = mole_fraction
        if pressure is not None:
            self.pressure = pressure
        if path_length is not None:
            self.path_length = path_length
        t0 = time.perf_counter()
        cols = self._line_columns()
        w = self.wavenumber
        app = _LineApp(self, cols, w.size, device_id)
        abscoeff, _ = app.iterate(self.pressure, float(Tgas), self.mole_fraction, path_length=self.path_length)
        NwL, NwG = app.get_griddims()
        conditions = dict(
            Tgas=float(Tgas), Tvib=float(Tgas), Trot=float(Tgas), pressure=self.pressure,
            mole_fraction=self.mole_fraction, path_length=self.path_length, wstep=self.wstep,
            wavenum_min=self.wavenum_min, wavenum_max=self.wavenum_max, dxG=self.dxG, dxL=self.dxL,
            NwL=NwL, NwG=NwG, lines_calculated=int(cols["nu0"].size), thermal_equilibrium=True,
            optimization=self.optimization, broadening_method="fft",
            diluents={"air": 1 - self.mole_fraction}, gpu_backend="hip", spectral_points=int(w.size),
            add_at_used="gpu-backend", woutrange=(0, int(w.size)), calculation_time=time.perf_counter() - t0,
        )
        s = Spectrum(w, abscoeff, app.derived, conditions, gpu_app=None if exit_gpu else app, name=name)
        if exit_gpu:
            app.__del__()
        return s

    def eq_spectrum_gpu_interactive(self, var="transmittance", slit_function=0.0, mpl_backend="", plotkwargs={},
                                    *vargs, **kwargs):
        """``eq_spectrum_gpu`` with sliders for every condition given as a
        ``plot_tools.ParamRange`` (and for ``slit_function``): each slider
        move writes the condition and recomputes the plotted ``var`` on the
        resident GPU application (``Spectrum.recalc_gpu``); closing the
        figure releases it.  radis/lbl/factory.py:1393-1523."""
        from matplotlib import use

        from .plot_tools import is_range

        if mpl_backend:
            use(mpl_backend)
        import matplotlib.pyplot as plt
        from matplotlib.widgets import Slider

        self.interactive_params = {}
        for key in kwargs:
            if is_range(kwargs[key]):
                param_range = kwargs[key]
                self.interactive_params[key] = param_range
                param_range.name = key
                kwargs[key] = param_range.valinit
        kwargs["exit_gpu"] = False
        s = self.eq_spectrum_gpu(*vargs, **kwargs)
        if is_range(slit_function):
            self.interactive_params["slit_function"] = slit_function
            slit_function.name = "slit_function"
            slit_function = slit_function.valinit
        s.apply_slit(slit_function, unit="cm-1")  # the slit quantities: 'radiance', 'transmittance'
        s.conditions["slit_function"] = slit_function

        was_interactive = plt.isinteractive()
        plt.ion()
        line = s.plot(var, show=True, **plotkwargs)
        fig = line.figure

        def update_plot(val):
            # the sliders write s.conditions; recalc_gpu reads them
            new_y = s.recalc_gpu(var, wunit=plotkwargs.get("wunit", "default"),
                                 Iunit=plotkwargs.get("Iunit", "default"))
            line.set_ydata(new_y)
            fig.canvas.draw_idle()

        n_sliders = 0
        for key, param_range in self.interactive_params.items():
            slider_axis = fig.add_axes([0.25, 0.05 * n_sliders + 0.05, 0.65, 0.03])
            slider = Slider(ax=slider_axis, label=key, valmin=param_range.valmin, valmax=param_range.valmax,
                            valinit=param_range.valinit)
            param_range.set_widget(slider, s, update_plot)
            n_sliders += 1
        fig.subplots_adjust(bottom=0.05 * n_sliders + 0.15)
        fig.canvas.mpl_connect("close_event", s.exit_gpu)
        if not was_interactive:
            plt.ioff()
        return s


class _LineApp:
    """gpuApp-like holder built directly from S(Tref) line columns (the
    factory path does not go through the Einstein-A conversion)."""

    def __init__(self, sf, cols, Nv, device_id):
        self._dev = engine.Device(int(device_id))
        self._db = engine.LineDatabase(self._dev, cols)
        if sf.parfunc is not None:
            isos = [s for s in range(1, self._db.n_iso_slots) if self._db.molar_mass[s] > 0]
            partfunc.set_cdsd_co2_tables(self._db, sf.parfunc, isotopes=isos)
        self.sf = sf
        self.Nv = Nv
        self._batch = None
        self._outs = None
        self._griddims = (0, 0)
        self._path_length = None
        self.derived = {}

    def _req(self, p, T, x):
        sf = self.sf
        return engine.make_request(self._db, sf.wavenum_min, sf.wavenum_max, sf.wstep, T, p, x, sf.dxG, sf.dxL,
                                   sf.Tref, sf.folding_thresh, optimization=sf.optimization)

    def iterate(self, p, T, x, path_length=None):
        """abscoeff (cm-1) at (p bar, T K, x); with ``path_length`` (cm) the
        last pass also writes absorbance, emissivity, transmittance and
        radiance, left in ``self.derived``."""
        req = self._req(p, T, x)
        if self._batch is None:
            self._batch = engine.SpectrumBatch(self._db, [req])
            self._outs = self._batch.alloc_outputs()
        else:
            self._batch.set_request(0, req)
        if path_length != self._path_length:
            self._batch.set_derived(0, path_length, which=engine.SpectrumBatch.DERIVED if path_length is not None
                                    else ())
            self._path_length = path_length
        self._batch.run(self._outs)
        out = self._batch.results(self._outs)[0]
        self.derived = self._batch.derived_results(0)
        inf = self._batch.info(0)
        self._griddims = (int(inf.n_l), int(inf.n_g))
        return out, {}

    def get_griddims(self):
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


def calc_spectrum(wavenum_min, wavenum_max, Tgas, pressure=1.01325, mole_fraction=1, path_length=1,
                  wstep=0.01, lines=None, mode="gpu", name=None, **kwargs):
    """``calc_spectrum(..., mode='gpu')`` for one molecule whose lines are
    given; radis/lbl/calc.py:37, 1014-1028."""
    if mode != "gpu":
        raise NotImplementedError("this package implements mode='gpu'")
    sf = SpectrumFactory(wavenum_min, wavenum_max, wstep=wstep, pressure=pressure, mole_fraction=mole_fraction,
                         path_length=path_length, lines=lines, **kwargs)
    return sf.eq_spectrum_gpu(Tgas, name=name, exit_gpu=True)
