"""FP64 NumPy restatement of RADIS's CPU LDM-fft equilibrium synthesis path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``radis_amd``) never calls it and fails loudly
when its HIP library is missing.

Every function restates one step of RADIS 0.17.1 (``radis/...`` paths are
relative to the reference checkout) in float64, with the same NumPy
operations in the same order, so that integer grid indices are bit-identical
and floating-point values agree to rounding:

=====================  ==========================================================
step                   reference
=====================  ==========================================================
grid                   lbl/factory.py:2590-2617 (``arange(wmin, wmax+wstep, wstep)``)
line selection         lbl/loader.py:2707 (``wav >= wmin & wav <= wmax``)
S(T)                   lbl/base.py:2664-2674 (``calc_linestrength_eq``)
S(Tvib, Trot)          lbl/base.py:3479-3530 (Boltzmann populations),
                       3937-3947 (``calc_linestrength_noneq``), S0 2437-2441
pressure shift         lbl/base.py:2351-2360 (``calc_lineshift``)
Doppler HWHM           lbl/broadening.py:170-172 (``doppler_broadening_HWHM``)
Lorentzian HWHM        lbl/broadening.py:341-383 (``pressure_broadening_HWHM``),
                       p_atm = p_bar/1.01325 at lbl/broadening.py:1063
LDM axes               lbl/broadening.py:1823-1852 (``_init_w_axis``)
FT templates+folding   lbl/broadening.py:1895-1926, ``voigt_FT`` 828-853,
                       ``gaussian_FT`` 252, ``lorentzian_FT`` 604
indices                lbl/broadening.py:2151-2154 (``_get_indices``), 2253-2263
weights (simple)       lbl/broadening.py:2305-2318
weights (min-RMS)      lbl/broadening.py:2265-2303
scatter                lbl/broadening.py:2340-2346, 2478-2485;
                       misc/arrays.py:472-502 (``numpy_add_at``)
FFT apply              lbl/broadening.py:2523-2530
density, abscoeff      lbl/factory.py:908-918
derived quantities     lbl/factory.py:1289-1297, spectrum/equations.py:21-44,
                       phys/blackbody.py:35-74, phys/convert.py:120-122
slit convolution       tools/slit.py:735-737 (oaconvolve 'same' x wstep),
                       1064-1071 (remove_boundary, 'valid')
tabulated Q(T)         levels/partfunc_cdsd.py:84-117 (splrep/splev of a Q(T)
                       table; FITPACK's splev/fpbspl, scipy >= 1.4.0 per the
                       reference's pyproject.toml)
=====================  ==========================================================

The one generalisation: the FFT length ``L`` is a parameter.  RADIS uses
``L = 2*N_v`` (lbl/broadening.py:1900, 2340-2346); the GPU path uses an even
``L >= 2*N_v`` with small prime factors, and parity is measured with the
oracle run at the same ``L``.
"""
from __future__ import annotations

import numpy as np

# --- physical constants, radis/phys/constants.py:22-78 ----------------------
h = 6.62607015e-34
k_b = 1.3806490e-23
c = 2.99792458e8
Na = 6.02214076e23
hc_k = h * c / k_b * 100  # constants.py:69
k_b_CGS = 1.380648813e-16  # constants.py:75
c_CGS = 2.99792458e10  # constants.py:78
ATM_PER_BAR = 1.01325  # lbl/broadening.py:1063, lbl/base.py:2351

# RADIS default LDM steps, lbl/loader.py:452-459 and 538-543
DXL_DEFAULT = float(np.log((0.01 / 0.20) ** 0.5 + 1))
DXG_DEFAULT = float(np.log((0.01 / 0.46) ** 0.5 + 1))
FOLDING_THRESH_DEFAULT = 1e-6  # lbl/factory.py:437


def wavenumber_grid(wmin, wmax, wstep):
    """Calculation grid; lbl/factory.py:2590-2592 with neighbour_lines=0."""
    return np.arange(wmin, wmax + wstep, wstep)


def select_window(nu0, wmin, wmax):
    """Index range of lines kept for a window; lbl/loader.py:2707.

    ``nu0`` must be sorted ascending.  Returns ``(i0, i1)`` such that
    ``nu0[i0:i1]`` are exactly the lines with ``wmin <= nu0 <= wmax``.
    """
    i0 = int(np.searchsorted(nu0, wmin, side="left"))
    i1 = int(np.searchsorted(nu0, wmax, side="right"))
    return i0, i1


def linestrength_eq(S_ref, El, nu0, Tgas, Tref, Qref_over_Qgas):
    """S(T); lbl/base.py:2664-2674 (operation order kept)."""
    return (
        S_ref
        * Qref_over_Qgas
        * np.exp(-hc_k * El * (1 / Tgas - 1 / Tref))
        * (1 - np.exp(-hc_k * nu0 / Tgas))
        / (1 - np.exp(-hc_k * nu0 / Tref))
    )


def s0_einstein(A, gu, nu0, Ia):
    """Intensity without the population term, ``Ia gu A / (8 pi c nu0^2)``;
    lbl/base.py:2437-2441."""
    return Ia * gu * A / (8 * np.pi * c_CGS * nu0**2)


def populations_noneq(gvibl, grotl, Evibl, Erotl, gvibu, grotu, Evibu, Erotu, Tvib, Trot, Qneq):
    """Upper and lower state populations (nu, nl) for Boltzmann vibrational
    and rotational distributions without electronic states;
    lbl/base.py:3479-3530."""
    nu_vib_x_Qvib = gvibu * np.exp(-hc_k * Evibu / Tvib)
    nl_vib_x_Qvib = gvibl * np.exp(-hc_k * Evibl / Tvib)
    nu_rot_x_Qrot = grotu * np.exp(-Erotu * hc_k / Trot)
    nl_rot_x_Qrot = grotl * np.exp(-Erotl * hc_k / Trot)
    nu = nu_vib_x_Qvib * nu_rot_x_Qrot / Qneq
    nl = nl_vib_x_Qvib * nl_rot_x_Qrot / Qneq
    return nu, nl


def linestrength_noneq(S0, gl, gu, nu, nl):
    """Non-LTE linestrength from the populations; lbl/base.py:3937-3947."""
    line_strength = S0 / gl
    line_strength = line_strength * nl
    line_strength = line_strength * (1 - gl / gu * nu / nl)
    return line_strength


def lineshift(nu0, delta_air, pressure_bar):
    """Shifted line centre; lbl/base.py:2351-2360.  ``delta_air=None``: no shift."""
    if delta_air is None:
        return np.asarray(nu0, dtype=np.float64)
    air_pressure = pressure_bar / ATM_PER_BAR
    return nu0 + (delta_air * air_pressure)


def doppler_hwhm(nu0, molar_mass, Tgas):
    """Gaussian HWHM; lbl/broadening.py:170-172 (uses unshifted ``wav``)."""
    return (nu0 / c_CGS) * np.sqrt((2 * Na * k_b_CGS * Tgas * np.log(2)) / molar_mass)


def pressure_hwhm(g_air, g_self, n_air, n_self, pressure_bar, mole_fraction, Tgas, Tref):
    """Lorentzian HWHM; lbl/broadening.py:341-383 with diluent {'air': 1-x}
    (lbl/factory.py:2062-2067).  ``n_self=None`` uses ``n_air``
    (broadening.py:376-381)."""
    pressure_atm = pressure_bar / ATM_PER_BAR
    x_air = 1 - mole_fraction
    gamma_lb = 0
    gamma_lb += (Tref / Tgas) ** n_air * g_air * pressure_atm * x_air
    n_s = n_air if n_self is None else n_self
    gamma_lb += ((Tref / Tgas) ** n_s) * ((g_self * pressure_atm * mole_fraction))
    return gamma_lb


def init_w_axis(w_dat, log_p):
    """Log-spaced LDM width axis; lbl/broadening.py:1823-1837."""
    w_min = w_dat.min()
    if w_min < 0:
        w_min = w_dat[w_dat > 0].min()
    w_max = w_dat.max() + 1e-4
    N = np.ceil((np.log(w_max) - np.log(w_min)) / log_p) + 1
    return w_min * np.exp(log_p * np.arange(N))


def voigt_FT(x, hwhmG, hwhmL):
    """Fourier transform of a Voigt profile; lbl/broadening.py:252, 604, 850-853."""
    IG = np.exp(-((2 * np.pi * x * hwhmG) ** 2) / (4 * np.log(2)))
    IL = np.exp(-2 * np.pi * x * hwhmL)
    return IG * IL


def fold_count(wstep, hwhmG, hwhmL, folding_thresh=FOLDING_THRESH_DEFAULT):
    """Number of alias terms RADIS adds; loop condition at lbl/broadening.py:1915-1918."""
    n = 1
    while voigt_FT(n / (2 * wstep), hwhmG, hwhmL) >= folding_thresh:
        n += 1
    return n - 1


def lineshape_templates(L, wstep, wG, wL, folding_thresh=FOLDING_THRESH_DEFAULT):
    """FT lineshape templates with alias folding, normalised to 1 at x=0;
    lbl/broadening.py:1895-1926 with ``2*len(w)`` replaced by ``L``.

    Returns float64 array of shape (N_G, N_L, L//2+1)."""
    x = np.fft.rfftfreq(L, wstep)
    w_fold = (x, x[::-1])
    out = np.empty((len(wG), len(wL), x.size))
    for l in range(len(wG)):
        for m in range(len(wL)):
            V = voigt_FT(x, wG[l] / 2, wL[m] / 2)
            n = 1
            while voigt_FT(n / (2 * wstep), wG[l] / 2, wL[m] / 2) >= folding_thresh:
                V += voigt_FT(n / (2 * wstep) + w_fold[n & 1], wG[l] / 2, wL[m] / 2)
                n += 1
            V /= V[0]
            out[l, m] = V
    return out


def get_indices(arr_i, axis):
    """``np.interp`` position split into (i0, i1, frac); lbl/broadening.py:2151-2154."""
    pos = np.interp(arr_i, axis, np.arange(axis.size))
    index = pos.astype(np.int32)
    return index, index + 1, pos - index


def interp_position_scalar(x, xp):
    """Pure-Python statement of NumPy's ``interp`` arithmetic for the
    position of ``x`` on the increasing grid ``xp`` with ``fp = arange``.

    This is the rule the GPU index kernel emulates (binary search for the
    largest ``j`` with ``xp[j] <= x``; clamp outside; exact hit returns ``j``;
    otherwise ``(1/(xp[j+1]-xp[j]))*(x-xp[j]) + j``).  Used by the tests to
    pin ``get_indices`` against NumPy itself on small inputs."""
    n = len(xp)
    if x < xp[0]:
        return 0.0
    if x >= xp[n - 1]:
        return float(n - 1)
    lo, hi = 0, n - 1
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if xp[mid] <= x:
            lo = mid
        else:
            hi = mid
    j = lo
    if xp[j] == x:
        return float(j)
    slope = 1.0 / (xp[j + 1] - xp[j])
    return slope * (x - xp[j]) + j


def min_rms_weights(tvi, tGi, tLi, wG_dat, wL_dat, wstep, dxG, dxL):
    """``optimization='min-RMS'`` fractions (avi, aGi, aLi) from the simple
    ones and the FWHM; lbl/broadening.py:2265-2303."""
    dxvGi = wstep / wG_dat
    C1_GG = ((6 * np.pi - 16) / (15 * np.pi - 32)) ** (1 / 1.50)
    C1_LG = ((6 * np.pi - 16) / 3 * (np.log(2) / (2 * np.pi)) ** 0.5) ** (1 / 2.25)
    C2_GG = (2 * np.log(2) / 15) ** (1 / 1.50)
    C2_LG = ((2 * np.log(2)) ** 2 / 15) ** (1 / 2.25)
    alpha_i = wL_dat / wG_dat
    R_Gv = 8 * np.log(2)
    R_GG = 2 - 1 / (C1_GG + C2_GG * alpha_i ** (2 / 1.50)) ** 1.50
    R_GL = -2 * np.log(2) * alpha_i**2
    R_LL = 1
    R_LG = 1 / (C1_LG * alpha_i ** (1 / 2.25) + C2_LG * alpha_i ** (4 / 2.25)) ** 2.25
    avi = tvi
    aGi = tGi + (R_Gv * tvi * (tvi - 1) * dxvGi**2 + R_GG * tGi * (tGi - 1) * dxG**2
                 + R_GL * tLi * (tLi - 1) * dxL**2) / (2 * dxG)
    aLi = tLi + (R_LG * tGi * (tGi - 1) * dxG**2 + R_LL * tLi * (tLi - 1) * dxL**2) / (2 * dxL)
    return avi, aGi, aLi


def ldm_scatter(L, NG, NL, ki0, ki1, avi, li0, li1, aGi, mi0, mi1, aLi, S):
    """Distribute line intensities on the (L, N_G, N_L) float64 LDM with
    the fractions (avi, aGi, aLi) along the three axes (the simple weights
    are the interpolation fractions themselves); lbl/broadening.py:2305-2318,
    2340-2346, 2478-2485."""
    awV00 = (1 - aGi) * (1 - aLi)
    awV01 = (1 - aGi) * aLi
    awV10 = aGi * (1 - aLi)
    awV11 = aGi * aLi
    Iv0 = S * (1 - avi)
    Iv1 = S * avi
    LDM = np.zeros((L, NG, NL))
    np.add.at(LDM, (ki0, li0, mi0), Iv0 * awV00)
    np.add.at(LDM, (ki0, li0, mi1), Iv0 * awV01)
    np.add.at(LDM, (ki0, li1, mi0), Iv0 * awV10)
    np.add.at(LDM, (ki0, li1, mi1), Iv0 * awV11)
    np.add.at(LDM, (ki1, li0, mi0), Iv1 * awV00)
    np.add.at(LDM, (ki1, li0, mi1), Iv1 * awV01)
    np.add.at(LDM, (ki1, li1, mi0), Iv1 * awV10)
    np.add.at(LDM, (ki1, li1, mi1), Iv1 * awV11)
    return LDM


def ldm_apply_fft(LDM, templates, Nv, wstep, fft_workers=None):
    """Sum over (l, m) of rfft(LDM[:, l, m]) * template, irfft, crop, /wstep;
    lbl/broadening.py:2523-2530.  ``fft_workers``: run the transforms with
    scipy.fft on that many threads (same arithmetic, pocketfft underneath)
    instead of numpy.fft's single thread."""
    NG, NL = LDM.shape[1], LDM.shape[2]
    if fft_workers:
        import scipy.fft as sfft

        rfft = lambda a: sfft.rfft(a, workers=fft_workers)  # noqa: E731
        irfft = lambda a: sfft.irfft(a, workers=fft_workers)  # noqa: E731
    else:
        rfft, irfft = np.fft.rfft, np.fft.irfft
    Ildm_FT = 1j * np.zeros(templates.shape[2])
    for l in range(NG):
        for m in range(NL):
            Ildm_FT += rfft(LDM[:, l, m]) * templates[l, m]
    sumoflines = irfft(Ildm_FT)[:Nv]
    sumoflines /= wstep
    return sumoflines


def number_density(mole_fraction, pressure_bar, Tgas):
    """Molecules per cm3; lbl/factory.py:908."""
    return mole_fraction * ((pressure_bar * 1e5) / (k_b * Tgas)) * 1e-6


def _noneq_strength(lines, nu0, iso, Tvib, Trot, Tref):
    """S(Tvib, Trot) of every line.  The lines give ``Evib_l``, ``Evib_u``
    and ``Qneq(iso, Tvib, Trot)``, and either ``S0`` (Einstein form) or
    ``S_ref`` with ``Q``, from which S0 follows as
    S_ref Q(Tref) / (exp(-c2 El/Tref) (1 - exp(-c2 nu0/Tref))).  Degeneracies:
    ``gl``, ``gu`` (total) with the vibrational one ``gvibl``/``gvibu``
    (default 1) and the rotational one the quotient; 1 when absent.  Level
    energies: E_l = El, E_u = El + nu0, Erot = E - Evib."""
    f64 = lambda k, d=None: d if lines.get(k) is None else np.asarray(lines[k], dtype=np.float64)  # noqa: E731
    El = f64("El")
    isoi = np.ones(nu0.shape, dtype=np.int64) if iso is None else np.asarray(iso, dtype=np.int64)
    if lines.get("S0") is not None:
        S0 = f64("S0")
    else:
        Qref = np.array([lines["Q"](int(i), Tref) if i > 0 else 1.0 for i in range(int(isoi.max()) + 1)])[isoi]
        S0 = f64("S_ref") * Qref / (np.exp(-hc_k * El / Tref) * (1 - np.exp(-hc_k * nu0 / Tref)))
    one = np.ones(nu0.shape)
    gl, gu = f64("gl", one), f64("gu", one)
    gvibl, gvibu = f64("gvibl", one), f64("gvibu", one)
    Evibl, Evibu = f64("Evib_l"), f64("Evib_u")
    Eu = El + nu0
    Qneq = np.array([lines["Qneq"](int(i), Tvib, Trot) if i > 0 else 1.0
                     for i in range(int(isoi.max()) + 1)])[isoi]
    nu, nl = populations_noneq(gvibl, gl / gvibl, Evibl, El - Evibl, gvibu, gu / gvibu, Evibu, Eu - Evibu,
                               Tvib, Trot, Qneq)
    return linestrength_noneq(S0, gl, gu, nu, nl)


def line_parameters(lines, Tgas, pressure_bar, mole_fraction, Tref=296.0, Tvib=None, Trot=None):
    """Per-line S(T), shifted centre, Gaussian and Lorentzian FWHM (float64).

    ``lines`` is a dict of arrays: ``nu0, S_ref, El, g_air, g_self, n_air``
    and optionally ``n_self``, ``delta``, ``iso``; plus ``molar_mass``
    (indexable by iso) and ``Q`` (callable ``Q(iso, T)``).  With ``Tvib``
    and ``Trot`` the line strengths are the two-temperature ones
    (``_noneq_strength``) and ``Tgas`` is the translational temperature."""
    nu0 = np.asarray(lines["nu0"], dtype=np.float64)
    iso = lines.get("iso")
    mm = np.asarray(lines["molar_mass"], dtype=np.float64)
    if iso is None:
        molar_mass = np.full(nu0.shape, mm[1] if mm.size > 1 else mm[0])
        qratio = lines["Q"](1, Tref) / lines["Q"](1, Tgas)
    else:
        iso = np.asarray(iso)
        molar_mass = mm[iso]
        qtab = np.array(
            [lines["Q"](i, Tref) / lines["Q"](i, Tgas) if i > 0 else 1.0 for i in range(mm.size)]
        )
        qratio = qtab[iso]
    f64 = lambda k: None if lines.get(k) is None else np.asarray(lines[k], dtype=np.float64)
    if Tvib is not None:
        S = _noneq_strength(lines, nu0, iso, Tvib, Trot, Tref)
    else:
        S = linestrength_eq(f64("S_ref"), f64("El"), nu0, Tgas, Tref, qratio)
    shiftwav = lineshift(nu0, f64("delta"), pressure_bar)
    wG_dat = doppler_hwhm(nu0, molar_mass, Tgas) * 2.000
    wL_dat = (
        pressure_hwhm(
            f64("g_air"), f64("g_self"), f64("n_air"), f64("n_self"),
            pressure_bar, mole_fraction, Tgas, Tref,
        )
        * 2.000
    )
    return S, shiftwav, wG_dat, wL_dat


def eq_abscoeff(
    lines,
    wmin,
    wmax,
    wstep,
    Tgas,
    pressure_bar,
    mole_fraction,
    dxG=DXG_DEFAULT,
    dxL=DXL_DEFAULT,
    Tref=296.0,
    L=None,
    folding_thresh=FOLDING_THRESH_DEFAULT,
    return_details=False,
    fft_workers=None,
    optimization="simple",
    Tvib=None,
    Trot=None,
):
    """End-to-end abscoeff (cm-1) on the calculation grid, the way
    ``SpectrumFactory.eq_spectrum`` computes it with
    ``broadening_method='fft'`` and ``optimization`` ``'simple'`` or
    ``'min-RMS'`` (lbl/factory.py:745-918); with ``Tvib`` and ``Trot`` the
    line strengths are the two-temperature ones of ``non_eq_spectrum``
    (lbl/factory.py:1525-1952, Ttrans = Trot by default at 1685-1687, calling
    lbl/base.py:3360 and 3879) and ``Tgas``
    is the translational temperature.

    ``lines['nu0']`` must be sorted; only lines with wmin <= nu0 <= wmax are
    used (lbl/loader.py:2707).  ``L`` defaults to ``2*N_v``."""
    w = wavenumber_grid(wmin, wmax, wstep)
    Nv = w.size
    if L is None:
        L = 2 * Nv
    if L < 2 * Nv or L % 2:
        raise ValueError("L must be even and >= 2*N_v")
    i0, i1 = select_window(np.asarray(lines["nu0"]), wmin, wmax)
    sub = {k: (v[i0:i1] if isinstance(v, np.ndarray) and v.shape == np.shape(lines["nu0"]) else v)
           for k, v in lines.items()}
    if i1 <= i0:
        out = np.zeros(Nv)
        if return_details:
            return out, dict(i0=i0, i1=i1, Nv=Nv, L=L, NG=0, NL=0)
        return out
    S, shiftwav, wG_dat, wL_dat = line_parameters(sub, Tgas, pressure_bar, mole_fraction, Tref, Tvib, Trot)
    wL = init_w_axis(wL_dat, dxL)
    wG = init_w_axis(wG_dat, dxG)
    ki0, ki1, tvi = get_indices(shiftwav, w)
    li0, li1, tGi = get_indices(np.log(wG_dat), np.log(wG))
    mi0, mi1, tLi = get_indices(np.log(wL_dat), np.log(wL))
    if optimization == "min-RMS":
        avi, aGi, aLi = min_rms_weights(tvi, tGi, tLi, wG_dat, wL_dat, wstep, dxG, dxL)
    elif optimization == "simple":
        avi, aGi, aLi = tvi, tGi, tLi
    else:
        raise ValueError(f"optimization {optimization!r}")
    LDM = ldm_scatter(L, wG.size, wL.size, ki0, ki1, avi, li0, li1, aGi, mi0, mi1, aLi, S)
    templates = lineshape_templates(L, wstep, wG, wL, folding_thresh)
    sumoflines = ldm_apply_fft(LDM, templates, Nv, wstep, fft_workers)
    abscoeff = sumoflines * number_density(mole_fraction, pressure_bar, Tgas)
    if return_details:
        return abscoeff, dict(
            i0=i0, i1=i1, Nv=Nv, L=L, NG=wG.size, NL=wL.size, wG=wG, wL=wL,
            ki0=ki0, tvi=tvi, li0=li0, tGi=tGi, mi0=mi0, tLi=tLi, S=S, aGi=aGi, aLi=aLi,
            wG_dat=wG_dat, wL_dat=wL_dat,
        )
    return abscoeff


def cm2nm(w_cm1):
    """Vacuum wavelength (nm) of a wavenumber (cm-1); phys/convert.py:120-122."""
    return 1 / w_cm1 * 1e9 / 100


def planck(lambda_nm, T):
    """Blackbody spectral radiance in mW/cm2/sr/nm; phys/blackbody.py:35-74."""
    lbd = np.asarray(lambda_nm, dtype=np.float64) * 1e-9
    with np.errstate(over="ignore"):
        ipl = (2 * h * c**2 / lbd**5) * (1 / (np.exp(h * c / (lbd * k_b * T)) - 1))
    return ipl * 1e-10


def derived_quantities(wavenumber, abscoeff, path_length, Tgas):
    """absorbance, emissivity, transmittance and radiance from abscoeff
    (cm-1) along a path of ``path_length`` cm at ``Tgas`` K;
    lbl/factory.py:1289-1297 with spectrum/equations.py:21-44."""
    absorbance = np.asarray(abscoeff, dtype=np.float64) * path_length
    emissivity = -np.expm1(-absorbance)
    transmittance = 1 - emissivity
    radiance = emissivity * planck(cm2nm(np.asarray(wavenumber, dtype=np.float64)), Tgas)
    return dict(absorbance=absorbance, emissivity=emissivity, transmittance=transmittance, radiance=radiance)


def slit_convolve(I, I_slit, wstep, mode="valid"):
    """Convolution with a slit sampled on the spectrum grid and the 'valid'
    boundary, as convolve_with_slit does after resampling
    (tools/slit.py:735-737 ``oaconvolve(I, I_slit, 'same') * wstep``;
    remove_boundary 1064-1071), written as a direct sum: out[i] =
    wstep * sum_j I_slit[j] * I[i + (k-1)//2 - j]."""
    I = np.asarray(I, dtype=np.float64)
    I_slit = np.asarray(I_slit, dtype=np.float64)
    n, k = I.size, I_slit.size
    full = np.convolve(I, I_slit, mode="full")
    c = (k - 1) // 2
    out = full[c:c + n] * wstep
    if mode == "valid":
        la = min(n, k)
        a, b = int((la - 1) / 2), int(la / 2)
        out[:a] = np.nan
        if b:
            out[-b:] = np.nan
    return out


def splev_fitpack(t, c, k, x):
    """B-spline value as FITPACK's splev (ext=0) computes it, the evaluation
    behind PartFuncCO2_CDSDtab.at (levels/partfunc_cdsd.py:105-117): the
    interval l is the largest l in [k, n-k-2] with t[l] <= x (the end pieces
    extrapolate), fpbspl's Cox-de Boor recursion gives the k+1 nonzero
    B-splines h, and the value is sum_j c[l-k+j] h[j].  Pure Python loops
    (small inputs)."""
    t = np.asarray(t, dtype=np.float64)
    c = np.asarray(c, dtype=np.float64)
    n = t.size
    out = []
    for xv in np.atleast_1d(np.asarray(x, dtype=np.float64)):
        l = k
        while l < n - k - 2 and t[l + 1] <= xv:
            l += 1
        h = [1.0] + [0.0] * k
        for j in range(1, k + 1):
            hh = h[:j]
            h[0] = 0.0
            for i in range(1, j + 1):
                li, lj = l + i, l + i - j
                if t[li] == t[lj]:
                    h[i] = 0.0
                    continue
                f = hh[i - 1] / (t[li] - t[lj])
                h[i - 1] = h[i - 1] + f * (t[li] - xv)
                h[i] = f * (xv - t[lj])
        out.append(sum(c[l - k + j] * h[j] for j in range(k + 1)))
    return np.asarray(out).reshape(np.shape(x))


def rel_peak_error(test, ref):
    """max |test - ref| / max |ref| — the parity metric used for every gate."""
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.max(np.abs(np.asarray(test, dtype=np.float64) - ref)) / np.max(np.abs(ref)))
