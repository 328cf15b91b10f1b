"""Two-temperature (Tvib, Trot) line strengths on the GPU
(radis/lbl/base.py:3479-3530, 3937-3947; non_eq_spectrum in
radis/lbl/factory.py:1525-1952) against the oracle, on the HITRAN CO fixture
with vibrational level energies taken from the file's own quanta.
"""
import numpy as np
import pytest

from oracle import ldm_oracle as oracle
from radis_amd import _native, engine, hitran

pytestmark = pytest.mark.gpu

FIX = "tests/golden/hitran_co_3iso_2000_2300cm.par"
REL_TOL = 1e-4
Q296 = {1: 107.12, 2: 224.08, 3: 112.47}  # radis/db/molparam.txt:51-53


def _co_noneq_lines():
    """Fixture lines with Evib of a level (iso, v) = the lowest energy of a
    line state with those quanta (its J-lowest rotational level), so
    Erot = E - Evib >= 0; Q_neq = Qvib(Tvib) Qrot(Trot) per isotopologue
    (harmonic vibration from the file's own v=1 energy, linear rotor
    scaled to Q(296))."""
    c = hitran.read_par(FIX)
    order = np.argsort(c["nu0"], kind="stable")
    c = {k: v[order] for k, v in c.items()}
    vl, vu, iso = c["gq_l"].astype(int), c["gq_u"].astype(int), c["iso"].astype(int)
    El, nu0 = c["El"].astype(np.float64), c["nu0"]
    table = {}
    for i in range(nu0.size):
        for key, E in (((iso[i], vl[i]), El[i]), ((iso[i], vu[i]), El[i] + nu0[i])):
            table[key] = min(table.get(key, np.inf), E)
    ev_l = np.array([table[(iso[i], vl[i])] for i in range(nu0.size)], dtype=np.float32)
    ev_u = np.array([table[(iso[i], vu[i])] for i in range(nu0.size)], dtype=np.float32)
    g1 = {s: table[(s, 1)] - table[(s, 0)] for s in (1, 2, 3)}

    def qvib(s, T):
        return 1.0 / (1.0 - np.exp(-oracle.hc_k * g1[s] / T))

    def qneq(s, Tvib, Trot):
        return qvib(s, Tvib) * Q296[s] / qvib(s, 296.0) * (Trot / 296.0)

    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    return dict(nu0=np.ascontiguousarray(nu0), S_ref=f32(c["S_ref"]), El=f32(El), g_air=f32(c["g_air"]),
                g_self=f32(c["g_self"]), n_air=f32(c["n_air"]), delta=f32(c["delta"]), iso=iso.astype(np.uint8),
                molar_mass=np.array(hitran.CO_MOLAR_MASS[:4]), Evib_l=ev_l, Evib_u=ev_u, gl=c["gl"], gu=c["gu"],
                Q=lambda s, T: qneq(s, T, T), Qneq=qneq)


@pytest.fixture(scope="module")
def dev():
    d = engine.Device(0)
    yield d
    d.close()


@pytest.mark.parametrize("Tvib,Trot", [(3000.0, 800.0), (600.0, 2000.0), (9000.0, 150.0)])
def test_noneq_matches_oracle(dev, Tvib, Trot):
    lines = _co_noneq_lines()
    db = engine.LineDatabase(dev, lines)
    wmin, wmax, wstep, p, x = 2000.0, 2300.0, 0.002, 1.0, 0.1
    req = engine.make_request(db, wmin, wmax, wstep, Trot, p, x, Tvib=Tvib, Trot=Trot)
    b = engine.SpectrumBatch(db, [req])
    res = b.results(b.run())[0]
    inf = b.info(0)
    ref, det = oracle.eq_abscoeff(lines, wmin, wmax, wstep, Trot, p, x, return_details=True,
                                  Tvib=Tvib, Trot=Trot)
    t = b.line_terms(0)
    np.testing.assert_array_equal(t["k0"], det["ki0"])
    np.testing.assert_array_equal(t["l0"], det["li0"])
    np.testing.assert_array_equal(t["m0"], det["mi0"])
    dens = oracle.number_density(x, p, Trot) * 2.0 ** -int(inf.scale_log2)
    I_ref = det["S"] * dens
    assert oracle.rel_peak_error(t["I"], I_ref) <= 1e-5
    err = oracle.rel_peak_error(res, ref)
    print(f"Tvib={Tvib} Trot={Trot}: err {err:.2e}, min S/max S {det['S'].min() / det['S'].max():.2e}")
    assert err <= REL_TOL
    if Tvib / Trot > 20:
        assert det["S"].min() < 0 and res.min() < 0  # inverted P-branch lines absorb negatively
    b.close()
    db.close()


def test_noneq_at_one_temperature_equals_equilibrium(dev):
    lines = _co_noneq_lines()
    db = engine.LineDatabase(dev, lines)
    args = (2000.0, 2300.0, 0.002, 1500.0, 1.0, 0.1)
    b = engine.SpectrumBatch(db, [engine.make_request(db, *args),
                                  engine.make_request(db, *args, Tvib=1500.0, Trot=1500.0)])
    eq, neq = b.results(b.run())
    assert oracle.rel_peak_error(neq, eq) <= 2e-6
    b.close()
    db.close()


def test_noneq_needs_vibrational_energies(dev):
    lines = {k: v for k, v in _co_noneq_lines().items() if k not in ("Evib_l", "Evib_u")}
    db = engine.LineDatabase(dev, lines)
    req = engine.make_request(db, 2000.0, 2300.0, 0.01, 1000.0, 1.0, 0.1, Tvib=2000.0, Trot=1000.0)
    with pytest.raises(_native.LdmError):
        engine.SpectrumBatch(db, [req])
    db.close()
    bad = dict(_co_noneq_lines())
    bad["Evib_l"] = bad["El"] + 10.0  # above the lower-state energy
    with pytest.raises(_native.LdmError):
        engine.LineDatabase(dev, bad)
