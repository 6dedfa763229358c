"""WAKE and ZLONGIT on the GPU vs the CPU checker: integer histograms,
voltages and kicked coordinates bit-identical, then a full ring with both
collective elements over several passes."""
import numpy as np
import pytest

from elegant_amd import _abi as A
from elegant_amd import elements as E
from elegant_amd import lattices as Lt
from oracle import oracle_ctypes as O
from tests.helpers import assert_rows_bitwise, assert_rows_identical
from tests.test_oracle import PO, wake_bunch

pytestmark = pytest.mark.gpu


def test_wake_step_bit_exact(tracker):
    n = 300000
    rows = wake_bunch(n, 20e-12, seed=12)
    table = Lt.resonator_wake()
    w = E.WAKE(table)
    q = 5e-9 / n
    g = rows.copy()
    h, v, tmin = tracker.collective_probe(g, w, PO, 0, q)
    o = O.Oracle("emath")
    o.set_macro_charge(q)
    c = rows.copy()
    o.wake(c, w.p.wake, PO, 0)
    hc, vc, tminc, dt = o.last_hist()
    assert tmin == tminc
    assert np.array_equal(h, hc) and h.sum() == n
    assert np.array_equal(v, vc)
    assert_rows_identical(g, c, "wake kick")


def test_zlongit_step_bit_exact(tracker):
    n = 300000
    rows = wake_bunch(n, 20e-12, seed=13)
    zl = E.ZLONGIT(BROAD_BAND=1, RS=1e4, Q=1.0, FREQ=20e9, N_BINS=8192, BIN_SIZE=2e-12)
    q = 5e-9 / n
    g = rows.copy()
    h, v, tmin = tracker.collective_probe(g, zl, PO, 0, q)
    o = O.Oracle("emath")
    o.set_macro_charge(q)
    c = rows.copy()
    o.zlongit(c, zl.p.zlongit, PO, 0)
    hc, vc, tminc, dt = o.last_hist()
    assert h.size == 8192 and tmin == tminc
    assert np.array_equal(h, hc) and h.sum() == n
    assert np.array_equal(v, vc)
    assert_rows_identical(g, c, "zlongit kick")


def test_wake_table_from_the_example_sdds_file(tracker):
    """the wake file the reference ships with its examples
    (tests/golden/wake_Sx_200um_6mm_xband.sdds: 31 points 0.667 ps apart),
    read by the SDDS reader, as the WAKE Green function of a 1.5 ps bunch:
    histogram, voltages and kicked rows bit for bit as the checker's"""
    import os
    from elegant_amd import sdds
    t, W = sdds.wake_table(os.path.join(os.path.dirname(__file__), "golden", "wake_Sx_200um_6mm_xband.sdds"))
    w = E.WAKE(E.WakeTable(t, W))
    n = 200000
    rows = wake_bunch(n, 1.5e-12, seed=19)
    q = 1e-9 / n
    g = rows.copy()
    h, v, tmin = tracker.collective_probe(g, w, PO, 0, q)
    o = O.Oracle("emath")
    o.set_macro_charge(q)
    c = rows.copy()
    o.wake(c, w.p.wake, PO, 0)
    hc, vc, tminc, dt = o.last_hist()
    assert h.size == hc.size > 10 and tmin == tminc
    assert np.array_equal(h, hc) and h.sum() == n
    assert np.array_equal(v, vc) and np.abs(v).max() > 0
    assert_rows_identical(g, c, "example wake file")


@pytest.mark.parametrize("nbins,bin_size", [(16384, 1e-12), (65536, 2e-13), (131072, 1e-13)])
def test_zlongit_beyond_the_lds_transform_bit_exact(tracker, nbins, bin_size):
    """ZLONGIT with more bins than the one-workgroup LDS transform holds
    (8192): the global-memory transform of the packed m = nb/2 points, with
    odd (m = 2^13, 2^15) and even (m = 2^16) stage counts, gives the
    checker's histogram, voltages and kicked rows bit for bit"""
    n = 300000
    rows = wake_bunch(n, 20e-12, seed=17)
    zl = E.ZLONGIT(BROAD_BAND=1, RS=1e4, Q=1.0, FREQ=20e9, N_BINS=nbins, BIN_SIZE=bin_size)
    q = 5e-9 / n
    g = rows.copy()
    h, v, tmin = tracker.collective_probe(g, zl, PO, 0, q)
    o = O.Oracle("emath")
    o.set_macro_charge(q)
    c = rows.copy()
    o.zlongit(c, zl.p.zlongit, PO, 0)
    hc, vc, tminc, dt = o.last_hist()
    assert h.size == nbins and tmin == tminc
    assert np.array_equal(h, hc) and h.sum() == n
    assert np.array_equal(v, vc)
    assert_rows_identical(g, c, "zlongit kick, %d bins" % nbins)


def test_ring_with_wakes_bit_exact_vs_emath_checker(tracker):
    bl, Po = Lt.storage_ring_wake()
    rows = Lt.matched_beam(bl, 3000, 1e-9, 1e-11, 1e-3, 20e-12, 31, Po)
    rows[:300, 0] += 0.0149  # a halo at the edge of the 15 mm aperture: losses within a few passes
    g = rows.copy()
    ng, eg = tracker.do_tracking(g, bl, Po, 3)
    c = rows.copy()
    o = O.Oracle("emath")
    nc = o.do_tracking(c, bl, Po, 3)
    print("ring+wakes: survivors", ng, "effort", eg)
    assert ng == nc and eg == o.effort()
    assert rows.shape[0] - ng > 10
    assert_rows_identical(g, c, "ring with wakes")


def test_ring_with_wakes_and_an_elliptical_aperture(tracker):
    """the check after WAKE and ZLONGIT (inside the kick kernel) with an
    elliptical MAXAMP open on the +x side"""
    bl, Po = Lt.storage_ring_wake()
    assert bl.elements[1].type == A.EH_T_MAXAMP
    els = list(bl.elements)
    els[1] = E.MAXAMP("AP", X_MAX=0.015, Y_MAX=0.008, ELLIPTICAL=1, EXPONENT=2, OPEN_SIDE="+x")
    ring = E.Beamline(els, keep=bl.keep)
    rows = Lt.matched_beam(ring, 3000, 1e-9, 1e-11, 1e-3, 20e-12, 31, Po)
    rows[:300, 0] += 0.0149
    rows[300:400, 2] += 0.0079
    g = rows.copy()
    ng, eg = tracker.do_tracking(g, ring, Po, 3)
    c = rows.copy()
    o = O.Oracle("emath")
    nc = o.do_tracking(c, ring, Po, 3)
    assert ng == nc and eg == o.effort()
    assert rows.shape[0] - ng > 10
    assert_rows_bitwise(g, c, "ring with wakes, elliptical aperture")


def test_zlongit_from_impedance_tables_bit_exact(tracker):
    """ZLONGIT with Z(f) tables (BROAD_BAND=0): the resonator's samples as a
    table give the broad-band element's step on the GPU bit for bit, and a
    table with a DC term and a Nyquist term (both doubled in the packed
    spectrum) matches the checker"""
    from tests.test_oracle import _resonator_table
    n = 300000
    rows = wake_bunch(n, 20e-12, seed=44)
    q = 5e-9 / n
    table = _resonator_table(8192, 2e-12, 20e9, 1.0, 1e4)
    bb = E.ZLONGIT(BROAD_BAND=1, RS=1e4, Q=1.0, FREQ=20e9, N_BINS=8192, BIN_SIZE=2e-12)
    res = []
    for e in (bb, E.ZLONGIT_TABLE(table)):
        g = rows.copy()
        h, v, tmin = tracker.collective_probe(g, e, PO, 0, q)
        res.append((g, h, v, tmin))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][2], res[1][2])
    t2 = E.ImpedanceTable(table.f, table.Zreal + 30.0, table.Zimag)  # DC and Nyquist now nonzero
    t2.Zimag[-1] = 7.0  # ignored: the Nyquist term is real
    e = E.ZLONGIT_TABLE(t2)
    g = rows.copy()
    h, v, tmin = tracker.collective_probe(g, e, PO, 0, q)
    o = O.Oracle("emath")
    o.set_macro_charge(q)
    c = rows.copy()
    o.zlongit(c, e.p.zlongit, PO, 0)
    hc, vc, tminc, dt = o.last_hist()
    assert tmin == tminc and np.array_equal(h, hc) and np.array_equal(v, vc)
    assert_rows_identical(g, c, "zlongit from tables")
    assert not np.array_equal(g, res[0][0])


def test_ring_with_wake_change_p0_bit_exact(tracker):
    """the ring with WAKE CHANGE_P0 (wake.c:288-289): the host's match step
    after the wake kick, before the checks that follow the element, moves the
    central momentum every pass and the program is recompiled for it; rows,
    losses, effort and the final central momentum as the checker's"""
    bl, Po = Lt.storage_ring_wake(zlongit=False)
    els = [A.eh_element.from_buffer_copy(e) for e in bl.elements]
    for e in els:
        if e.type == A.EH_T_WAKE:
            e.p.wake.change_p0 = 1
    ring = E.Beamline(els, keep=bl.keep)
    rows = Lt.matched_beam(ring, 3000, 1e-9, 1e-11, 1e-3, 20e-12, 33, Po)
    rows[:300, 0] += 0.0149
    g = rows.copy()
    ng, eg = tracker.do_tracking(g, ring, Po, 4)
    c = rows.copy()
    o = O.Oracle("emath")
    nc = o.do_tracking(c, ring, Po, 4)
    assert ng == nc and eg == o.effort()
    assert rows.shape[0] - ng > 10
    assert tracker.P_central == o.P_central != Po
    assert_rows_identical(g, c, "ring with wake change_p0")
