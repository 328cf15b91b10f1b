"""ParamRange (radis/tools/plot_tools.py:31-56) on the CPU: a widget move
writes the value into the spectrum's conditions and calls the update."""
from radis_amd.plot_tools import ParamRange, is_range


class _Widget:
    def __init__(self):
        self.cb = None

    def on_changed(self, cb):
        self.cb = cb


class _Spec:
    def __init__(self):
        self.conditions = {}


def test_param_range_defaults_and_callback():
    r = ParamRange(300.0, 2000.0)
    assert (r.valinit, r.val) == (300.0, 300.0) and is_range(r) and not is_range(300.0)
    r.name = "Tgas"
    seen = []
    w, s = _Widget(), _Spec()
    r.set_widget(w, s, seen.append)
    w.cb(1500.0)
    assert s.conditions == {"Tgas": 1500.0} and seen == [1500.0] and r.val == 1500.0
    assert "1500.0" in repr(r)
