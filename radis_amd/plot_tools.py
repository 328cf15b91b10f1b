"""Slider ranges of ``SpectrumFactory.eq_spectrum_gpu_interactive``: the
``ParamRange`` of radis/tools/plot_tools.py:31-56 (same constructor, the same
widget callback that writes the slider value into the spectrum's conditions
and calls the update)."""
from __future__ import annotations


class ParamRange:
    """A condition given as a range: ``valmin`` .. ``valmax``, starting at
    ``valinit`` (``valmin`` when None)."""

    def __init__(self, valmin=0, valmax=1, valinit=None):
        self.valmin = valmin
        self.valmax = valmax
        if valinit is None:
            valinit = valmin
        self.valinit = valinit
        self.val = self.valinit
        self.name = None
        self.widget = None
        self.spec = None
        self.update_callback = None

    def set_widget(self, w, spec, func=lambda x: 0):
        self.widget = w
        self.spec = spec
        self.update_callback = func
        w.on_changed(self.widget_callback)

    def widget_callback(self, val):
        self.val = val
        self.spec.conditions[self.name] = val
        self.update_callback(val)

    def __repr__(self):
        return f"ParamRange({self.valmin!r} .. {self.valmax!r} [{self.valinit!r}] @ {self.val!r})"


def is_range(x):
    return isinstance(x, ParamRange)
