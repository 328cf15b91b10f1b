"""Seeded synthetic line lists and request streams for the five benchmark
workloads of the DIT/LDM path.

Line lists are generated in float64 and then rounded to the device storage
dtypes (``nu0`` float64; ``S_ref``, ``El``, ``g_air``, ``g_self``, ``n_air``
float32).  Every consumer (GPU path and CPU checker) reads the rounded
values, so both see identical inputs.  Lines are returned sorted by ``nu0``
(RADIS sorts by shifted wavenumber, lbl/base.py:2370; with no pressure
shift the order is the same).

Workloads
---------
* ``co_5k``      5,000 CO-like lines, 1900-2300 cm-1, wstep 0.002, T=700 K,
                 p=1 atm, x=0.1, dxG=dxL=0.2.
* ``co2_5m``     5 M CO2-like lines, 2000-2500 cm-1, wstep 0.001, T=1500 K,
                 p=1 atm, x=0.1, RADIS default dxG/dxL.
* ``h2o_100m``   100 M H2O/CH4-like lines, 1000-4000 cm-1, wstep 0.005,
                 T=2500 K, p=5 atm, x=0.1, RADIS default dxG/dxL.
* ``batch_256``  256 requests over the ``co2_5m`` list with seeded
                 (T, p, x, window), wstep 0.001.
* ``serve``      Poisson request stream over two resident lists.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

ATM = 1.01325  # bar per atm
DXL_DEFAULT = float(np.log((0.01 / 0.20) ** 0.5 + 1))  # lbl/loader.py:452-454, 538-540
DXG_DEFAULT = float(np.log((0.01 / 0.46) ** 0.5 + 1))  # lbl/loader.py:456-458, 541-543


@dataclasses.dataclass
class LineList:
    """Structure-of-arrays line list, sorted by ``nu0``."""

    nu0: np.ndarray  # float64 cm-1
    S_ref: np.ndarray  # float32 cm/molecule at Tref
    El: np.ndarray  # float32 cm-1
    g_air: np.ndarray  # float32 cm-1/atm (HWHM)
    g_self: np.ndarray  # float32 cm-1/atm (HWHM)
    n_air: np.ndarray  # float32
    molar_mass: np.ndarray  # float64 g/mol, indexed by iso (slot 0 unused)
    q_coef: tuple  # Q(T) = a * T**b for every iso
    n_self: np.ndarray | None = None  # float32 or None (use n_air)
    delta: np.ndarray | None = None  # float32 cm-1/atm or None (no shift)
    iso: np.ndarray | None = None  # uint8 or None (single isotopologue 1)
    name: str = ""

    def __len__(self):
        return int(self.nu0.size)

    def Q(self, iso, T):
        a, b = self.q_coef
        return a * T**b

    def as_dict(self):
        """Mapping consumed by the CPU checker."""
        return dict(
            nu0=self.nu0, S_ref=self.S_ref, El=self.El, g_air=self.g_air,
            g_self=self.g_self, n_air=self.n_air, n_self=self.n_self,
            delta=self.delta, iso=self.iso, molar_mass=self.molar_mass, Q=self.Q,
        )

    def subset(self, idx):
        """Sorted sub-list (``idx`` is an index array or slice)."""
        pick = lambda a: None if a is None else a[idx]
        return LineList(
            self.nu0[idx], self.S_ref[idx], self.El[idx], self.g_air[idx],
            self.g_self[idx], self.n_air[idx], self.molar_mass, self.q_coef,
            pick(self.n_self), pick(self.delta), pick(self.iso), self.name,
        )


@dataclasses.dataclass
class Conditions:
    """One spectrum request (units as RADIS: K, bar, cm-1)."""

    wmin: float
    wmax: float
    wstep: float
    Tgas: float
    pressure_bar: float
    mole_fraction: float
    dxG: float = DXG_DEFAULT
    dxL: float = DXL_DEFAULT
    Tref: float = 296.0
    path_length: float = 1.0


def _finish(rng_nu, S, El, ga, gs, na, molar_mass, q_coef, name):
    order = np.argsort(rng_nu, kind="stable")
    f32 = lambda a: np.ascontiguousarray(a[order], dtype=np.float32)
    return LineList(
        nu0=np.ascontiguousarray(rng_nu[order], dtype=np.float64),
        S_ref=f32(S), El=f32(El), g_air=f32(ga), g_self=f32(gs), n_air=f32(na),
        molar_mass=np.asarray(molar_mass, dtype=np.float64), q_coef=q_coef, name=name,
    )


def co_like(n=5000, seed=1000):
    """CO-like diatomic list (workload ``co_5k``)."""
    rng = np.random.default_rng(seed)
    nu = rng.uniform(1900.0, 2300.0, n)
    S = 10.0 ** rng.uniform(-30.0, -19.0, n)
    El = rng.uniform(0.0, 8000.0, n)
    ga = rng.uniform(0.03, 0.08, n)
    gs = rng.uniform(0.05, 0.10, n)
    na = rng.uniform(0.5, 0.8, n)
    return _finish(nu, S, El, ga, gs, na, [0.0, 28.0], (0.36, 1.0), "co_like")


def co2_like(n=5_000_000, seed=1001):
    """CO2-like triatomic list (workload ``co2_5m``)."""
    rng = np.random.default_rng(seed)
    nu = rng.uniform(2000.0, 2500.0, n)
    S = 10.0 ** rng.uniform(-32.0, -18.0, n)
    El = np.minimum(rng.exponential(3000.0, n), 15000.0)
    ga = rng.uniform(0.03, 0.08, n)
    gs = rng.uniform(0.05, 0.10, n)
    na = rng.uniform(0.5, 0.8, n)
    return _finish(nu, S, El, ga, gs, na, [0.0, 44.0], (0.94, 1.1), "co2_like")


def h2o_like(n=100_000_000, seed=1002, chunk=10_000_000):
    """H2O/CH4-like list (workload ``h2o_100m``), generated in chunks to bound
    host memory."""
    rng = np.random.default_rng(seed)
    parts = {k: [] for k in ("nu", "S", "El", "ga", "gs", "na")}
    left = n
    while left > 0:
        m = min(chunk, left)
        parts["nu"].append(rng.uniform(1000.0, 4000.0, m))
        parts["S"].append((10.0 ** rng.uniform(-35.0, -18.0, m)).astype(np.float32))
        parts["El"].append(rng.exponential(4000.0, m).astype(np.float32))
        parts["ga"].append(rng.uniform(0.02, 0.10, m).astype(np.float32))
        parts["gs"].append(rng.uniform(0.05, 0.10, m).astype(np.float32))
        parts["na"].append(rng.uniform(0.5, 0.8, m).astype(np.float32))
        left -= m
    cat = {k: np.concatenate(v) for k, v in parts.items()}
    return _finish(cat["nu"], cat["S"], cat["El"], cat["ga"], cat["gs"], cat["na"],
                   [0.0, 18.0], (0.05, 1.5), "h2o_like")


CONDITIONS = {
    "co_5k": Conditions(1900.0, 2300.0, 0.002, 700.0, 1.0 * ATM, 0.1, dxG=0.2, dxL=0.2),
    "co2_5m": Conditions(2000.0, 2500.0, 0.001, 1500.0, 1.0 * ATM, 0.1),
    "h2o_100m": Conditions(1000.0, 4000.0, 0.005, 2500.0, 5.0 * ATM, 0.1),
}


def workload(name, n=None):
    """(LineList, Conditions) for a single-spectrum workload; ``n`` overrides
    the line count (same distributions, same seed)."""
    if name == "co_5k":
        return co_like(n or 5000), CONDITIONS[name]
    if name == "co2_5m":
        return co2_like(n or 5_000_000), CONDITIONS[name]
    if name == "h2o_100m":
        return h2o_like(n or 100_000_000), CONDITIONS[name]
    raise KeyError(name)


def random_requests(n, lo, hi, seed=1003, wsteps=(0.001,), median_width=50.0,
                    sigma=0.8, wmin_width=10.0, wmax_width=400.0):
    """Seeded requests: T~U(300,3000) K, p~logU(0.1,20) atm, x~U(0.01,1),
    window width lognormal(median 50 cm-1, sigma 0.8) clipped to [10, 400],
    start ~ U(lo, hi-W), wstep drawn from ``wsteps``."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        T = rng.uniform(300.0, 3000.0)
        p_atm = math.exp(rng.uniform(math.log(0.1), math.log(20.0)))
        x = rng.uniform(0.01, 1.0)
        W = float(np.clip(median_width * math.exp(sigma * rng.standard_normal()), wmin_width,
                          min(wmax_width, hi - lo)))
        start = rng.uniform(lo, hi - W)
        ws = float(wsteps[rng.integers(len(wsteps))])
        out.append(Conditions(round(start, 6), round(start + W, 6), ws, T, p_atm * ATM, x))
    return out


def grid_size(wmin, wmax, wstep):
    """Number of points of ``arange(wmin, wmax+wstep, wstep)``."""
    return int(np.arange(wmin, wmax + wstep, wstep).size)
