"""MALIGN (offset_beam, do_tracking.c:3243-3306, 1818-1822) and the
beam-position monitors MONI/HMON/VMON (drifts with no aperture check after
them) and zero-strength HKICK/VKICK correctors (drifts whose loss point is
not extrapolated) on the GPU against the checker: a lossy FODO ring with beam offsets of
every kind, one applied on a single pass, and monitors in every cell,
through do_tracking and through a batch."""
import numpy as np
import pytest

from elegant_amd import elements as E
from elegant_amd import native
from oracle import oracle_ctypes as O
from tests.helpers import assert_rows_identical, lossy_fodo

pytestmark = pytest.mark.gpu


def ring_with_maligns():
    els = [E.MAXAMP("AP", X_MAX=0.02, Y_MAX=0.02)]
    for cell in range(12):
        els += [E.DRIF("D1", L=2.0), E.KQUAD("QF", L=0.5, K1=1.2, N_KICKS=4), E.KSEXT("SF", L=0.2, K2=20.0),
                E.DRIF("D2", L=2.0), E.KQUAD("QD", L=0.5, K1=-1.2, N_KICKS=4), E.KSEXT("SD", L=0.2, K2=-35.0)]
        els.append([E.MONI("BPM", L=0.1), E.HMON("HM", L=0.05, ORDER=1), E.VMON("VM")][cell % 3])
        els.append([E.HKICK("HC", L=0.1), E.VKICK("VC", L=0.1, ORDER=1)][cell % 2])
        if cell == 2:
            els.append(E.MALIGN("M1", DX=2e-4, DXP=-3e-5, DY=-1e-4, DYP=2e-5, DZ=0.05))
        if cell == 5:
            els.append(E.MALIGN("M2", DP=1e-3, DT=2e-12))
        if cell == 8:
            els.append(E.MALIGN("M3", DE=-5e-4, ON_PASS=2))
        if cell == 10:
            els.append(E.MALIGN("M4", DX=0.004, ON_PASS=3))  # pushes the halo into the aperture
    return E.Beamline(els)


def test_malign_ring_bit_exact(tracker):
    bl = ring_with_maligns()
    _, Po, rows = lossy_fodo(1500, seed=51)
    g = rows.copy()
    ng, eg = tracker.do_tracking(g, bl, Po, 6)
    c = rows.copy()
    o = O.Oracle("emath")
    nc = o.do_tracking(c, bl, Po, 6)
    assert ng == nc and eg == o.effort()
    assert rows.shape[0] - ng > 100
    assert_rows_identical(g, c, "ring with MALIGN")


def test_malign_in_a_batch_bit_exact(tracker):
    """ON_PASS counts the passes of the call (pass0 + pass in the launch)"""
    bl = ring_with_maligns()
    _, Po, rows = lossy_fodo(800, seed=52)
    b = native.Batch(tracker)
    b.add_variant(bl, Po, rows)
    b.track(0, 5)
    g, ng = b.download(0)
    b.close()
    c = rows.copy()
    nc = O.Oracle("emath").do_tracking(c, bl, Po, 5)
    assert ng == nc and rows.shape[0] - ng > 50
    assert_rows_identical(g, c, "batch with MALIGN")


def test_recirc_bit_exact(tracker):
    """an injection offset and a monitor before the recirculation point act
    on pass 0 only (do_tracking.c:459-465); the rest of the ring every pass"""
    ring = ring_with_maligns()
    els = [E.MALIGN("INJ", DX=0.003, DXP=-1e-4), E.MONI("B0", L=0.2), E.RECIRC("RC")] + list(ring.elements)
    bl = E.Beamline(els)
    _, Po, rows = lossy_fodo(1200, seed=53)
    g = rows.copy()
    ng, eg = tracker.do_tracking(g, bl, Po, 5)
    c = rows.copy()
    o = O.Oracle("emath")
    nc = o.do_tracking(c, bl, Po, 5)
    assert ng == nc and eg == o.effort()
    assert_rows_identical(g, c, "ring with RECIRC")
    # continuing from pass 5 starts at the recirculation point
    g2, c2 = g[:ng].copy(), c[:nc].copy()
    ng2, _ = tracker.do_tracking(g2, bl, Po, 2, pass_offset=5)
    nc2 = O.Oracle("emath").do_tracking(c2, bl, Po, 2, pass_offset=5)
    assert ng2 == nc2
    assert_rows_identical(g2, c2, "ring with RECIRC, later passes")

