"""Particle species (change_particle, elegant.c:2933-2991): the values
eh_species_named gives, what they change on the tracked path, and GPU
parity with the checker for positrons and protons.

The CPU tests pin the species arithmetic with what physics fixes: the
proton's rest energy and classical radius, a custom species of electron
mass and charge reproducing the electron's constants, a positron's WAKE
kick equal to the electron's bit for bit (particleRelSign enters the wake
voltage and the energy change alike), a proton's RF kick equal to an
electron's at the voltage scaled by -m_e/m_p, and field-given quadrupole
strengths carrying charge/mass."""
import ctypes as C
import math

import numpy as np
import pytest

from elegant_amd import _abi as A
from elegant_amd import elements as E
from elegant_amd import lattices as Lt
from elegant_amd import native
from oracle import oracle_ctypes as O
from tests.helpers import assert_rows_identical, lossy_fodo
from tests.test_oracle import PO, wake_bunch

ME_MEV = 0.51099895000


def test_electron_and_positron():
    e = native.species("electron")
    assert (e.mass, e.charge, e.massMV, e.radius, e.relSign) == (
        9.1093837015e-31, 1.602176634e-19, ME_MEV, 2.8179403262e-15, 1.0)
    p = native.species("positron")
    assert (p.mass, p.massMV, p.radius) == (e.mass, e.massMV, e.radius)
    assert p.charge == -e.charge and p.relSign == -1.0


def test_proton_and_muon_against_textbook_values():
    p = native.species("proton")
    # elegant's table mass 1.6726485e-27 kg is 1.6e-5 above CODATA 2018's
    # 1.67262192e-27: rest energy 938.272 MeV and radius 1.53469e-18 m to that
    assert p.charge == -1.602176634e-19 and p.relSign == -1.0
    assert abs(p.massMV / 938.27208816 - 1) < 2e-5
    assert abs(p.radius / 1.53469e-18 - 1) < 2e-5
    assert abs(p.radius * p.mass / (2.8179403262e-15 * 9.1093837015e-31) - 1) < 1e-9  # r m = e^2/(4 pi eps0 c^2)
    mu = native.species("muon")
    assert mu.charge == 1.602176634e-19 and mu.relSign == 1.0
    assert abs(mu.massMV / 105.658 - 1) < 1e-4


def test_custom_species():
    c = native.species("custom", 1.0, 1.0)  # electron mass and charge, by the general formulas
    e = native.species("electron")
    assert c.relSign == 1.0 and c.charge == e.charge and c.mass == e.mass
    assert abs(c.massMV / e.massMV - 1) < 1e-9 and abs(c.radius / e.radius - 1) < 1e-9
    d = native.species("custom", 4.0, -2.0)
    assert d.relSign == -1.0 and d.charge == -2 * 1.602176634e-19
    assert abs(d.massMV / (2 * ME_MEV) - 1) < 1e-9
    for args in [("custom", 0.0, 1.0), ("custom", 1.0, 0.0), ("neutron", 1.0, 1.0)]:
        with pytest.raises(native.EHError):
            native.species(*args)


def test_positron_wake_kick_equals_the_electrons():
    rows = wake_bunch(20000, 20e-12, seed=31)
    table = Lt.resonator_wake()
    w = E.WAKE(table)
    q = 5e-9 / 20000
    a = rows.copy()
    o = O.Oracle("emath")
    o.set_macro_charge(q)
    o.wake(a, w.p.wake, PO, 0)
    b = rows.copy()
    with O.species(native.species("positron")):
        o2 = O.Oracle("emath")
        o2.set_macro_charge(q)
        o2.wake(b, w.p.wake, PO, 0)
    assert np.array_equal(a, b) and not np.array_equal(a, rows)
    # a proton's kick is m_e/m_p of the electron's
    c = rows.copy()
    sp = native.species("proton")
    with O.species(sp):
        o3 = O.Oracle("emath")
        o3.set_macro_charge(q)
        o3.wake(c, w.p.wake, PO, 0)
    d_e = a[:, 5] - rows[:, 5]
    d_p = c[:, 5] - rows[:, 5]
    big = np.abs(d_e) > 1e-3 * np.abs(d_e).max()
    assert np.allclose(d_p[big] / d_e[big], ME_MEV / sp.massMV, rtol=1e-6)


def _rf_kick(rows, volt, sp=None):
    rf = E.RFCA("RF", L=0.0, VOLT=volt, FREQ=352e6, PHASE=150.0, FIDUCIAL="light").p.rfca
    c = rows.copy()
    ptr, keep = A.rows_pointer(c)
    pc = C.c_double(PO)
    o = O.Oracle("emath")
    if sp is None:
        n = o.L.or_simple_rf_cavity(o.ctx, 0, ptr, c.shape[0], C.byref(rf), C.byref(pc), 7.0)
    else:
        with O.species(sp):
            n = o.L.or_simple_rf_cavity(o.ctx, 0, ptr, c.shape[0], C.byref(rf), C.byref(pc), 7.0)
    assert n == c.shape[0]
    return c


def test_proton_rf_kick_is_the_electrons_at_scaled_voltage():
    rows = wake_bunch(5000, 20e-12, seed=32)
    sp = native.species("proton")
    p = _rf_kick(rows, 3e6, sp)
    e = _rf_kick(rows, 3e6 * ME_MEV / sp.massMV * sp.relSign)
    assert np.allclose(p[:, 5] - rows[:, 5], e[:, 5] - rows[:, 5], rtol=1e-12, atol=0)
    assert np.abs(p[:, 5] - rows[:, 5]).max() > 0


def test_field_given_quadrupole_strength_carries_charge_over_mass():
    rows = wake_bunch(2000, 20e-12, seed=33)
    rows[:, 0] = 1e-3 * np.cos(np.arange(2000))
    rows[:, 2] = 1e-3 * np.sin(np.arange(2000))
    sp = native.species("proton")
    B, bore, L = 0.6, 0.025, 0.5
    k1 = B / bore * (sp.charge / (sp.mass * 2.99792458e8 * PO))
    out = []
    for e, s in [(E.KQUAD("Q", L=L, B=B, BORE=bore, N_KICKS=4), sp), (E.KQUAD("Q", L=L, K1=k1, N_KICKS=4), None)]:
        c = rows.copy()
        ptr, keep = A.rows_pointer(c)
        ap = A.EH_MAXAMP()
        L_ = O.lib("emath")
        if s is None:
            L_.or_multipole_tracking2(ptr, c.shape[0], C.byref(e), PO, 0.0, C.byref(ap))
        else:
            with O.species(s):
                L_.or_multipole_tracking2(ptr, c.shape[0], C.byref(e), PO, 0.0, C.byref(ap))
        out.append(c)
    assert np.array_equal(out[0], out[1]) and not np.array_equal(out[0], rows)
    assert k1 < 0  # the proton sees the electron's focusing reversed


# ------------------------------------------------------------------ GPU

@pytest.mark.gpu
@pytest.mark.parametrize("name", ["positron", "proton", "muon"])
def test_species_ring_bit_exact(tracker, name):
    """a FODO ring with field-given quadrupoles, an RF cavity and a WAKE
    after a CHARGE, 6 passes of a lossy beam: rows, losses and effort as
    the checker's under the same species"""
    table = Lt.resonator_wake()
    els = [E.CHARGE("Q", TOTAL=2e-9), E.MAXAMP("AP", X_MAX=0.02, Y_MAX=0.02)]
    for cell in range(8):
        els += [E.DRIF("D1", L=2.0), E.KQUAD("QF", L=0.5, B=0.6, BORE=0.025, N_KICKS=4),
                E.KSEXT("SF", L=0.2, K2=20.0), E.DRIF("D2", L=2.0),
                E.KQUAD("QD", L=0.5, B=-0.6, BORE=0.025, N_KICKS=4), E.KSEXT("SD", L=0.2, K2=-35.0)]
    els += [E.RFCA("RF", L=0.0, VOLT=3e6, FREQ=352e6, PHASE=150.0), E.WAKE(table, "WK")]
    bl = E.Beamline(els, keep=[table])
    _, Po, rows = lossy_fodo(1500, seed=41)
    sp = native.species(name)
    g = rows.copy()
    tracker.set_species(sp)
    try:
        ng, eg = tracker.do_tracking(g, bl, Po, 6)
    finally:
        tracker.set_species(native.species("electron"))
    c = rows.copy()
    with O.species(sp):
        o = O.Oracle("emath")
        nc = o.do_tracking(c, bl, Po, 6)
    assert ng == nc and eg == o.effort()
    assert rows.shape[0] - ng > 100
    assert_rows_identical(g, c, "%s ring" % name)
    # the species changed the result
    e = rows.copy()
    O.Oracle("emath").do_tracking(e, bl, Po, 6)
    assert not np.array_equal(e, c)
