"""Host engine over the C ABI: device context, device-resident line lists and
spectrum batches.

A :class:`SpectrumBatch` is one launch set: every request in it runs through
the same seven kernels (axis reduction, axis finalisation, tile bounds, LDS
fill, FFT pass A, fused FFT pass B with the lineshape reduction, FFT pass C)
in one stream submission, optionally captured once into a hipGraph and
replayed.  Outputs are torch float32 tensors on the device.
"""
from __future__ import annotations

import ctypes
import math
from ctypes import POINTER, byref, c_float, c_int64, c_uint8, c_void_p

import numpy as np

from . import _native
from ._native import MAX_ISO, STAGE_NAMES, check

DXL_DEFAULT = float(np.log((0.01 / 0.20) ** 0.5 + 1))  # radis/lbl/loader.py:452-454, 538-540
DXG_DEFAULT = float(np.log((0.01 / 0.46) ** 0.5 + 1))  # radis/lbl/loader.py:456-458, 541-543
FOLDING_THRESH_DEFAULT = 1e-6  # radis/lbl/factory.py:437


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise _native.NativeLibraryError("no HIP device visible to torch")
    return torch


def _stream_handle(stream):
    if stream is None:
        torch = _torch()
        return c_void_p(torch.cuda.current_stream().cuda_stream)
    if hasattr(stream, "cuda_stream"):
        return c_void_p(stream.cuda_stream)
    return c_void_p(int(stream))


class Device:
    """One HIP device (``radis_ldm_ctx``)."""

    def __init__(self, device=0, budget_bytes=0):
        torch = _torch()
        torch.cuda.set_device(device)
        self._lib = _native.load()
        h = c_void_p()
        check(self._lib.radis_ldm_ctx_create(int(device), byref(h)))
        self._h = h
        self.index = int(device)
        if budget_bytes:
            self.set_budget(budget_bytes)

    def set_budget(self, nbytes):
        check(self._lib.radis_ldm_ctx_set_budget(self._h, int(nbytes)))

    def bytes_in_use(self):
        v = c_int64()
        check(self._lib.radis_ldm_ctx_bytes_in_use(self._h, byref(v)))
        return int(v.value)

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("device context closed")
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.radis_ldm_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_F32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)


class LineDatabase:
    """Device-resident, nu0-sorted line list (``radis_ldm_db``).

    ``lines`` carries ``nu0, S_ref, El, g_air, g_self, n_air`` and optionally
    ``n_self, delta, iso``, plus ``molar_mass`` (g/mol per iso slot, slot 0
    unused) and a partition function ``Q(iso, T)``.  Two-temperature requests
    also need the vibrational energies ``Evib_l, Evib_u`` of the lower and
    upper states and ``Qneq(iso, Tvib, Trot)``.  Lines are sorted by
    ``nu0`` here if needed; ``order`` maps device rows to input rows."""

    def __init__(self, device: Device, lines):
        self.device = device
        self._lib = device._lib
        get = (lambda k: getattr(lines, k, None)) if not isinstance(lines, dict) else lines.get
        nu0 = np.ascontiguousarray(get("nu0"), dtype=np.float64)
        order = None
        if nu0.size > 1 and not np.all(nu0[1:] >= nu0[:-1]):
            order = np.argsort(nu0, kind="stable")
            nu0 = nu0[order]
        pick = (lambda a: a) if order is None else (lambda a: a[order])
        cols = {}
        for k in ("S_ref", "El", "g_air", "g_self", "n_air"):
            v = get(k)
            if v is None:
                raise ValueError(f"line list lacks column {k!r}")
            cols[k] = _F32(pick(np.asarray(v)))
        for k in ("n_self", "delta"):
            v = get(k)
            cols[k] = None if v is None else _F32(pick(np.asarray(v)))
        iso = get("iso")
        cols["iso"] = None if iso is None else np.ascontiguousarray(pick(np.asarray(iso)), dtype=np.uint8)
        for k in ("Evib_l", "Evib_u"):
            v = get(k)
            cols[k] = None if v is None else _F32(pick(np.asarray(v)))
        mm = np.zeros(MAX_ISO)
        mma = np.asarray(get("molar_mass"), dtype=np.float64).ravel()
        if mma.size > MAX_ISO:
            raise ValueError(f"at most {MAX_ISO} isotopologue slots")
        mm[: mma.size] = mma
        self.n_iso_slots = max(2, int(mma.size))
        self.molar_mass = mm
        self.Q = get("Q")
        self.Qneq = get("Qneq")
        self.q_tables = {}  # iso slot -> (knots, coefficients, degree, T_lo, T_hi) on the device
        self.order = order
        self.nu0 = nu0
        self.n = int(nu0.size)
        self.has_iso = cols["iso"] is not None
        self.has_delta = cols["delta"] is not None
        self.has_nself = cols["n_self"] is not None
        self.has_evib = cols["Evib_l"] is not None
        self._keep = (nu0, cols, mm)
        st = _native.Lines()
        fp = lambda a: None if a is None else a.ctypes.data_as(POINTER(c_float))
        st.nu0 = nu0.ctypes.data_as(POINTER(ctypes.c_double))
        st.S_ref, st.El, st.g_air = fp(cols["S_ref"]), fp(cols["El"]), fp(cols["g_air"])
        st.g_self, st.n_air = fp(cols["g_self"]), fp(cols["n_air"])
        st.n_self, st.delta = fp(cols["n_self"]), fp(cols["delta"])
        st.iso = None if cols["iso"] is None else cols["iso"].ctypes.data_as(POINTER(c_uint8))
        st.n = self.n
        st.molar_mass = mm.ctypes.data_as(POINTER(ctypes.c_double))
        st.n_iso_slots = self.n_iso_slots
        st.Evib_l, st.Evib_u = fp(cols["Evib_l"]), fp(cols["Evib_u"])
        h = c_void_p()
        check(self._lib.radis_ldm_db_upload(device.handle, byref(st), byref(h)))
        self._h = h
        self._keep = None  # the library copied everything

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("line database closed")
        return self._h

    def window(self, wmin, wmax):
        """(i0, i1): lines with wmin <= nu0 <= wmax (radis/lbl/loader.py:2707)."""
        return int(np.searchsorted(self.nu0, wmin, "left")), int(np.searchsorted(self.nu0, wmax, "right"))

    def request_bytes(self, req):
        """(workspace bytes, planned RequestInfo) of one request, nothing allocated."""
        b = c_int64()
        inf = _native.RequestInfo()
        check(self._lib.radis_ldm_request_bytes(self.handle, byref(req), byref(b), byref(inf)))
        return int(b.value), inf

    def set_partition_table(self, iso, T, Q, k=3):
        """Tabulated Q(T) of isotopologue slot ``iso``, interpolated as
        PartFuncCO2_CDSDtab does (radis/levels/partfunc_cdsd.py:84-117):
        ``scipy.interpolate.splrep(T, Q)`` (cubic), valid from 0.95 min(T) to
        1.05 max(T).  The spline goes to the device, which evaluates
        Q(Tref)/Q(T) itself for requests made with ``device_q``."""
        from scipy.interpolate import splrep

        T = np.asarray(T, dtype=np.float64).ravel()
        Q = np.asarray(Q, dtype=np.float64).ravel()
        if T.size != Q.size or T.size < k + 1:
            raise ValueError("T and Q must have the same length, above the spline degree")
        t, c, kk = splrep(T, Q, k=k)
        t = np.ascontiguousarray(t, dtype=np.float64)
        c = np.ascontiguousarray(np.asarray(c, dtype=np.float64)[: t.size - kk - 1])
        lo, hi = 0.95 * float(T.min()), 1.05 * float(T.max())  # partfunc_cdsd.py:101-103
        check(self._lib.radis_ldm_db_set_qspline(self.handle, int(iso), t.ctypes.data_as(POINTER(ctypes.c_double)),
                                                 t.size, c.ctypes.data_as(POINTER(ctypes.c_double)), int(kk), lo, hi))
        self.q_tables[int(iso)] = (t, c, int(kk), lo, hi)

    def partition_function(self, iso, T, stream=None):
        """Q(T) of slot ``iso`` from its table, evaluated on the device."""
        torch = _torch()
        Tn = np.asarray(T, dtype=np.float64)
        Td = torch.as_tensor(np.ascontiguousarray(Tn.ravel()), device="cuda")
        Qd = torch.empty_like(Td)
        check(self._lib.radis_ldm_qspline_eval(self.handle, int(iso), c_void_p(Td.data_ptr()), Td.numel(),
                                               c_void_p(Qd.data_ptr()), _stream_handle(stream)))
        return Qd.cpu().numpy().reshape(Tn.shape)

    def tables_cover_isotopologues(self):
        """True when every isotopologue slot with a molar mass has a Q table."""
        slots = [s for s in range(1, self.n_iso_slots) if self.molar_mass[s] > 0]
        return bool(slots) and all(s in self.q_tables for s in slots)

    def q_ratio(self, T, Tref):
        """Q(Tref)/Q(T) per iso slot; slots without lines get 1."""
        out = np.ones(MAX_ISO)
        if self.Q is None:
            return out
        for s in range(1, self.n_iso_slots):
            if self.molar_mass[s] > 0:
                out[s] = float(self.Q(s, Tref)) / float(self.Q(s, T))
        return out

    def q_ratio_noneq(self, Tvib, Trot, Tref):
        """Q(Tref)/Q_neq(Tvib, Trot) per iso slot; slots without lines get 1."""
        if self.Qneq is None or self.Q is None:
            raise ValueError("two-temperature requests need Q(iso, T) and Qneq(iso, Tvib, Trot) with the lines")
        out = np.ones(MAX_ISO)
        for s in range(1, self.n_iso_slots):
            if self.molar_mass[s] > 0:
                out[s] = float(self.Q(s, Tref)) / float(self.Qneq(s, Tvib, Trot))
        return out

    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.radis_ldm_db_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_request(db: LineDatabase, wmin, wmax, wstep, Tgas, pressure_bar, mole_fraction,
                 dxG=DXG_DEFAULT, dxL=DXL_DEFAULT, Tref=296.0, folding_thresh=FOLDING_THRESH_DEFAULT,
                 fft_len=0, q_ratio=None, n_v=0, line_window=None, fp64_widths=False, fft_n1=0, tile_w=0,
                 generic_fft=False, optimization="simple", Tvib=None, Trot=None, device_q=None,
                 period_fft_len=False):
    """Fill a ``radis_ldm_request``; Q(Tref)/Q(T) comes from the database's
    partition function unless ``q_ratio`` is given.  ``n_v`` fixes the grid
    length (default: len(arange(wmin, wmax+wstep, wstep))); ``line_window``
    (lo, hi) overrides the line selection [wmin, wmax]; ``fft_n1``/``tile_w``
    fix the FFT split and the fill tile width (default: the autotuner's cache,
    else the planner); ``generic_fft`` selects the run-time-sized FFT passes;
    ``optimization`` is the LDM weighting, ``"simple"`` or ``"min-RMS"``
    (radis/lbl/broadening.py:2265-2318).  ``Tvib`` and ``Trot`` select
    two-temperature line strengths (``Tgas`` is then the translational
    temperature of the widths; RADIS's ``non_eq_spectrum`` takes
    ``Ttrans = Trot`` by default, radis/lbl/factory.py:1685-1687).
    ``device_q`` takes Q(Tref)/Q(T) from the database's partition tables,
    evaluated on the device (default: when no ``q_ratio`` is given and the
    tables cover every isotopologue slot of an equilibrium request).
    The spectrum is RADIS's at FFT length 2 N_v whatever transform length
    runs (the library corrects the circular images of a longer one);
    ``period_fft_len`` instead returns the circular result at the transform
    length itself (the reference GPU's semantics)."""
    if optimization not in ("simple", "min-RMS"):
        raise ValueError(f"optimization must be 'simple' or 'min-RMS', not {optimization!r}")
    r = _native.Request()
    r.wmin, r.wmax, r.wstep = float(wmin), float(wmax), float(wstep)
    r.Tgas, r.pressure_bar, r.mole_fraction = float(Tgas), float(pressure_bar), float(mole_fraction)
    r.Tref, r.dxG, r.dxL, r.folding_thresh = float(Tref), float(dxG), float(dxL), float(folding_thresh)
    noneq = Tvib is not None or Trot is not None
    if noneq and (Tvib is None or Trot is None):
        raise ValueError("give both Tvib and Trot")
    if device_q is None:  # the database's Q tables when they cover its isotopologues
        device_q = q_ratio is None and not noneq and db.tables_cover_isotopologues()
    if device_q:
        if noneq:
            raise ValueError("device_q takes equilibrium partition functions only")
        q = np.ones(MAX_ISO)  # ignored for the slots with a table
    elif q_ratio is not None:
        q = np.asarray(q_ratio, dtype=np.float64)
    elif noneq:
        q = db.q_ratio_noneq(Tvib, Trot, Tref)
    else:
        q = db.q_ratio(Tgas, Tref)
    for s in range(MAX_ISO):
        r.q_ratio[s] = float(q[s]) if s < q.size else 1.0
    r.fft_len = int(fft_len or 0)
    r.n_v = int(n_v or 0)
    r.flags = ((_native.FP64_WIDTHS if fp64_widths else 0) | (_native.GENERIC_FFT if generic_fft else 0)
               | (_native.MIN_RMS if optimization == "min-RMS" else 0) | (_native.NONEQ if noneq else 0)
               | (_native.DEVICE_Q if device_q else 0) | (_native.PERIOD_L if period_fft_len else 0))
    if noneq:
        r.Tvib, r.Trot = float(Tvib), float(Trot)
    r.fft_n1, r.tile_w = int(fft_n1 or 0), int(tile_w or 0)
    if line_window is not None:
        r.line_lo, r.line_hi = float(line_window[0]), float(line_window[1])
    if not (fft_len or fft_n1 or tile_w or generic_fft):
        from . import autotune

        autotune.apply(db, r)
    return r


def request_from_conditions(db, cond, fft_len=0, fp64_widths=False, period_fft_len=False):
    return make_request(db, cond.wmin, cond.wmax, cond.wstep, cond.Tgas, cond.pressure_bar,
                        cond.mole_fraction, cond.dxG, cond.dxL, cond.Tref, fft_len=fft_len,
                        fp64_widths=fp64_widths, period_fft_len=period_fft_len)


def plan_grid(wmin, wmax, wstep):
    """(N_v, L) for a grid: length of arange(wmin, wmax+wstep, wstep) and the
    planner's FFT length."""
    lib = _native.load()
    nv, L = c_int64(), c_int64()
    check(lib.radis_ldm_plan_grid(float(wmin), float(wmax), float(wstep), byref(nv), byref(L)))
    return int(nv.value), int(L.value)


class SpectrumBatch:
    """Requests planned together over one line database (``radis_ldm_batch``)."""

    def __init__(self, db: LineDatabase, requests):
        self.db = db
        self._lib = db._lib
        self.n = len(requests)
        if self.n == 0:
            raise ValueError("empty batch")
        arr = (_native.Request * self.n)(*requests)
        h = c_void_p()
        check(self._lib.radis_ldm_batch_create(db.device.handle, db.handle, self.n, arr, byref(h)))
        self._h = h
        self._reqs = list(requests)
        self.infos = [self._info_nosync(i) for i in range(self.n)]
        self._outs = None
        self._graph_outs = None

    # -- sizes ----------------------------------------------------------------
    def _info_nosync(self, i):
        inf = _native.RequestInfo()
        check(self._lib.radis_ldm_batch_info(self._h, i, byref(inf), None))
        return inf

    def info(self, i):
        """Planned sizes and device results (N_G, N_L, status) of request i."""
        return self._info_nosync(i)

    def n_v(self, i):
        return int(self.infos[i].n_v)

    def workspace_bytes(self):
        v = c_int64()
        check(self._lib.radis_ldm_batch_workspace_bytes(self._h, byref(v)))
        return int(v.value)

    # -- conditions -------------------------------------------------------------
    # request fields fixed by the plan: the schedule and the flags that pick
    # the FFT variant, the axis work items and the period
    _PLAN_FLAGS = _native.GENERIC_FFT | _native.SCAN_AXIS | _native.PERIOD_L

    def set_request(self, i, req):
        """New conditions for request i (same grid and line window).  The
        schedule of the batch's request stays: fft_len, fft_n1, tile_w and
        the FFT-variant, axis-scan and period flags are taken from it, so a
        request built by make_request (which may apply a different tuned
        schedule for the new conditions' size class) keeps the plan."""
        old = self._reqs[i]
        r = _native.Request()
        ctypes.pointer(r)[0] = req
        r.fft_len, r.fft_n1, r.tile_w = old.fft_len, old.fft_n1, old.tile_w
        r.flags = (int(req.flags) & ~self._PLAN_FLAGS) | (int(old.flags) & self._PLAN_FLAGS)
        check(self._lib.radis_ldm_batch_set_request(self._h, int(i), byref(r)))
        self._reqs[i] = r
        self.infos[i] = self._info_nosync(i)

    # -- execution --------------------------------------------------------------
    def alloc_outputs(self):
        torch = _torch()
        return [torch.empty(self.n_v(i), dtype=torch.float32, device="cuda") for i in range(self.n)]

    def run(self, outs=None, stream=None):
        """Enqueue all requests on ``stream``; returns the output tensors
        (stored value * 2**scale_log2 = abscoeff in cm-1)."""
        if outs is None:
            if self._outs is None:
                self._outs = self.alloc_outputs()
            outs = self._outs
        self._check_outs(outs)
        ptrs = (c_void_p * self.n)(*[o.data_ptr() for o in outs])
        check(self._lib.radis_ldm_batch_run(self._h, ptrs, _stream_handle(stream)))
        return outs

    def capture(self, outs=None, stream=None):
        """Capture the launch set into a hipGraph bound to ``outs``."""
        if outs is None:
            outs = self.alloc_outputs()
        self._check_outs(outs)
        ptrs = (c_void_p * self.n)(*[o.data_ptr() for o in outs])
        torch = _torch()
        s = stream if stream is not None else torch.cuda.Stream()
        check(self._lib.radis_ldm_batch_graph_capture(self._h, ptrs, _stream_handle(s)))
        self._graph_outs = outs
        return outs

    def replay(self, stream=None):
        """Replay the captured graph; recaptures first if a re-plan (LDM growth
        after set_request) invalidated it."""
        rc = self._lib.radis_ldm_batch_graph_replay(self._h, _stream_handle(stream))
        if rc == _native.EINVAL and self._graph_outs is not None:
            self.capture(self._graph_outs)
            rc = self._lib.radis_ldm_batch_graph_replay(self._h, _stream_handle(stream))
        check(rc)
        return self._graph_outs

    DERIVED = ("absorbance", "transmittance", "emissivity", "radiance")

    def set_derived(self, i, path_length, which=DERIVED):
        """Have the last pass also write request i's derived quantities
        (``which`` of absorbance, transmittance, emissivity, radiance) for a
        path of ``path_length`` cm, on every following run and replay.
        Returns the float32 device tensors (absorbance stored times
        2**-scale_log2, like abscoeff).  ``which=()`` turns them off."""
        torch = _torch()
        bad = set(which) - set(self.DERIVED)
        if bad:
            raise ValueError(f"unknown derived quantities {sorted(bad)}")
        nv = self.n_v(i) if 0 <= i < self.n else 0  # the library reports a bad index
        tens = {k: torch.empty(nv, dtype=torch.float32, device="cuda") for k in which}
        d = _native.Derived(*[c_void_p(tens[k].data_ptr()) if k in tens else None for k in self.DERIVED],
                            float(path_length))
        check(self._lib.radis_ldm_batch_set_derived(self._h, int(i), byref(d) if which else None))
        if not hasattr(self, "_derived"):
            self._derived = {}
        self._derived[i] = tens  # the library holds raw pointers: keep the tensors alive
        return tens

    def derived_results(self, i):
        """Request i's derived quantities from the last run as float64 numpy
        arrays (absorbance rescaled by 2**scale_log2); synchronises."""
        tens = getattr(self, "_derived", {}).get(i, {})
        inf = self.info(i)
        out = {}
        for k, t in tens.items():
            a = t[: inf.n_v].double().cpu().numpy()
            if k == "absorbance" and inf.scale_log2:
                a = np.ldexp(a, int(inf.scale_log2))
            out[k] = a
        return out

    def _check_outs(self, outs):
        torch = _torch()
        if len(outs) != self.n:
            raise ValueError("one output tensor per request")
        for i, o in enumerate(outs):
            if o.dtype != torch.float32 or not o.is_cuda or o.numel() < self.n_v(i) or not o.is_contiguous():
                raise ValueError(f"output {i} must be a contiguous float32 device tensor of {self.n_v(i)} values")

    def enable_timing(self, on=True):
        check(self._lib.radis_ldm_batch_enable_timing(self._h, 1 if on else 0))

    def stage_ms(self):
        """Per-stage device time (ms) of the last timed run."""
        buf = (c_float * _native.N_STAGES)()
        check(self._lib.radis_ldm_batch_stage_ms(self._h, buf))
        return dict(zip(STAGE_NAMES, [float(x) for x in buf]))

    def line_terms(self, i, stream=None):
        """Per-line (k0, tv, l0, tg, m0, tl, I) the fill used for request i
        in the last run, as numpy arrays (synchronises)."""
        torch = _torch()
        inf = self.info(i)
        n = int(inf.line_end - inf.line_begin)
        k0 = torch.zeros(max(n, 1), dtype=torch.int64, device="cuda")
        l0 = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
        m0 = torch.zeros_like(l0)
        tv, tg, tl, I = (torch.zeros(max(n, 1), dtype=torch.float32, device="cuda") for _ in range(4))
        check(self._lib.radis_ldm_batch_line_terms(self._h, int(i), *[c_void_p(t.data_ptr()) for t in
                                                                     (k0, l0, m0, tv, tg, tl, I)],
                                                   _stream_handle(stream)))
        torch.cuda.synchronize()
        cv = lambda t: t[:n].cpu().numpy()
        return dict(k0=cv(k0), tv=cv(tv), l0=cv(l0), tg=cv(tg), m0=cv(m0), tl=cv(tl), I=cv(I))

    def ldm(self, i, stream=None):
        """Request i's LDM from the last run as a float32 numpy array of shape
        (N_G*N_L, N_v+1), rows r = l*N_L + m (synchronises)."""
        torch = _torch()
        inf = self.info(i)
        rows, cols = int(inf.n_g * inf.n_l), int(inf.n_v + 1)
        if rows == 0:
            return np.zeros((0, cols), dtype=np.float32)
        dst = torch.empty((rows, cols), dtype=torch.float32, device="cuda")
        check(self._lib.radis_ldm_batch_copy_ldm(self._h, int(i), c_void_p(dst.data_ptr()),
                                                 _stream_handle(stream)))
        torch.cuda.synchronize()
        return dst.cpu().numpy()

    def results(self, outs):
        """Synchronise and return float64 numpy abscoeff per request, after
        checking the device status of every request."""
        res = []
        for i in range(self.n):
            inf = self.info(i)
            if inf.status:
                raise _native.LdmError(_native.EDEVICE, f"request {i}: device status {inf.status:#x}"
                                       f" ({describe_status(inf.status)})")
            a = outs[i][: inf.n_v].double().cpu().numpy()
            if inf.scale_log2:
                a = np.ldexp(a, int(inf.scale_log2))
            res.append(a)
        return res

    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.radis_ldm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_noalloc(db: LineDatabase, req, out, result=None, stream=None):
    """One request through ``radis_ldm_run_noalloc``: the library allocates
    nothing, its workspace is the device's fixed scratch array.  ``out`` is a
    float32 device tensor of at least N_v values (abscoeff times
    2**-scale_log2); ``result``, if given, a device tensor of at least 48
    bytes that receives the ``radis_ldm_result`` (N_G, N_L, status, width
    extremes).  Returns the planned RequestInfo (scale_log2, n_v, fft_len).
    Nothing is synchronised."""
    torch = _torch()
    if out.dtype != torch.float32 or not out.is_cuda or not out.is_contiguous():
        raise ValueError("out must be a contiguous float32 device tensor")
    if result is not None and (not result.is_cuda or result.numel() * result.element_size() < 48):
        raise ValueError("result must be a device tensor of at least 48 bytes")
    inf = _native.RequestInfo()
    check(db._lib.radis_ldm_run_noalloc(db.device.handle, db.handle, byref(req), c_void_p(out.data_ptr()),
                                        c_void_p(result.data_ptr()) if result is not None else None, byref(inf),
                                        _stream_handle(stream)))
    return inf


def read_result(result):
    """The ``radis_ldm_result`` in a device tensor (synchronises)."""
    raw = result.view(-1).view(dtype=__import__("torch").uint8)[:48].cpu().numpy().tobytes()
    return _native.Result.from_buffer_copy(raw)


def describe_status(st):
    parts = []
    if st & _native.ST_AXIS:
        parts.append("LDM axis longer than 128 nodes")
    if st & _native.ST_CAP:
        parts.append("LDM rows above the planned capacity")
    if st & _native.ST_FOLD:
        parts.append("alias folding did not converge")
    if st & _native.ST_WIDTH:
        parts.append("non-positive broadening width")
    return ", ".join(parts) or "ok"


def synthesize(db, conditions, fft_len=0):
    """Convenience: one spectrum per condition, float64 numpy abscoeff (cm-1)."""
    reqs = [request_from_conditions(db, c, fft_len=fft_len) for c in conditions]
    b = SpectrumBatch(db, reqs)
    try:
        outs = b.run()
        return b.results(outs), [b.info(i) for i in range(b.n)]
    finally:
        b.close()


def lines_per_second(n_lines, seconds):
    return float(n_lines) / seconds if seconds > 0 else math.inf
