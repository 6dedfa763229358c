"""Bunched beams: WAKE and ZLONGIT acting per bucket of the bunch index in
column 8 (index_bunch_assignments, lrwake.c:17-233; wake.c:84-117;
zlongit.c:142-195).

CPU tests pin the checker with what the reference's bucket loop fixes: a
bucket's result is that of its bunch tracked alone (the bucket sees only its
own rows), START_BUNCH/END_BUNCH leave buckets untouched, the ZLONGIT bin
settings follow the first bucket, and without id slots a train longer than
the wake window is refused as in the single-bunch code.  GPU tests compare
the library with the checker bit for bit at the element entry point and in
a ring with losses."""
import math

import numpy as np
import pytest

from elegant_amd import _abi as A
from elegant_amd import elements as E
from elegant_amd import lattices as Lt
from oracle import oracle_ctypes as O
from tests.helpers import assert_rows_identical
from tests.test_oracle import PO

SPACING = 2e-9  # bunch spacing (s), longer than the 1 ns wake table


def bunch_train(n_per, nbunch, sigma_t, seed, first_index=0):
    """nbunch Gaussian bunches of n_per particles, SPACING apart in time,
    rows of 10 columns (bunch index in column 8, weight 1 in column 9);
    bunches interleaved in row order so that each bucket's rows are spread
    over the whole array"""
    rng = Lt.Rng(seed)
    beta = PO / math.sqrt(PO * PO + 1)
    n = n_per * nbunch
    z = rng.normal(2 * n).reshape(2, n)
    b = np.arange(n) % nbunch
    r = np.zeros((n, 10))
    r[:, 4] = 0.3 + 2.99792458e8 * beta * (sigma_t * z[0] + SPACING * b)
    r[:, 5] = 1e-3 * z[1]
    r[:, 6] = np.arange(1, n + 1)
    r[:, 7] = -1
    r[:, 8] = b + first_index
    r[:, 9] = 1
    return r


def _oracle(id_slots, q):
    o = O.Oracle("emath")
    o.set_row_width(10)
    o.set_id_slots_per_bunch(id_slots)
    o.set_macro_charge(q)
    return o


def test_wake_buckets_equal_bunches_alone():
    rows = bunch_train(4000, 3, 15e-12, seed=5, first_index=7)
    table = Lt.resonator_wake()  # the element points into it: keep it alive
    w = E.WAKE(table)
    q = 1e-9 / 4000
    all_ = rows.copy()
    _oracle(4000, q).wake(all_, w.p.wake, PO, 0)
    for b in range(3):
        sel = rows[:, 8] == 7 + b
        alone = np.ascontiguousarray(rows[sel])
        _oracle(0, q).wake(alone, w.p.wake, PO, 0)
        assert np.array_equal(all_[sel], alone), "bucket %d" % b
        assert not np.array_equal(alone, rows[sel])


def test_wake_train_without_id_slots_is_a_long_beam():
    rows = bunch_train(2000, 2, 15e-12, seed=6)
    table = Lt.resonator_wake()
    w = E.WAKE(table)
    with pytest.raises(RuntimeError):
        _oracle(0, 1e-12).wake(rows.copy(), w.p.wake, PO, 0)


def test_wake_start_and_end_bunch():
    rows = bunch_train(3000, 4, 15e-12, seed=8, first_index=2)
    table = Lt.resonator_wake()
    w = E.WAKE(table, START_BUNCH=1, END_BUNCH=2)
    w_all = E.WAKE(table)
    q = 1e-9 / 3000
    out = rows.copy()
    _oracle(3000, q).wake(out, w.p.wake, PO, 0)
    for b in range(4):
        sel = rows[:, 8] == 2 + b
        if b in (1, 2):
            alone = np.ascontiguousarray(rows[sel])
            _oracle(0, q).wake(alone, w_all.p.wake, PO, 0)
            assert np.array_equal(out[sel], alone)
        else:
            assert np.array_equal(out[sel], rows[sel]), "bucket %d left out" % b


def test_wake_with_n_bins_takes_a_plain_mean_per_bucket():
    """N_BINS > 0 centres the bins on the bucket's mean time, a plain sum
    over its rows (wake.c:110-116): the kicked rows equal a bunch alone
    whose mean is that plain sum; the histogram is centred on it"""
    rows = bunch_train(5000, 2, 4e-12, seed=9)  # within the 51.2 ps of 512 bins
    table = Lt.resonator_wake()
    w = E.WAKE(table, N_BINS=512)
    q = 1e-9 / 5000
    out = rows.copy()
    o = _oracle(5000, q)
    o.wake(out, w.p.wake, PO, 0)
    h, v, tmin, dt = o.last_hist()  # the last bucket
    sel = rows[:, 8] == 1
    P = PO * (rows[sel, 5] + 1)
    t = rows[sel, 4] * np.sqrt(P * P + 1) / (2.99792458e8 * P)
    tmean = 0.0
    for x in t:
        tmean += x
    tmean /= t.size
    assert h.size == 512 and h.sum() == t.size
    assert tmin == tmean - dt * 512 / 2.0
    assert not np.array_equal(out[sel], rows[sel])


def test_zlongit_buckets_follow_the_first_buckets_bins():
    rows = bunch_train(4000, 2, 15e-12, seed=10)
    # the second bunch twice as long: alone it would choose other bins
    sel1 = rows[:, 8] == 1
    mid = 0.3 + 2.99792458e8 * (PO / math.sqrt(PO * PO + 1)) * SPACING
    rows[sel1, 4] = mid + 2 * (rows[sel1, 4] - mid)
    # a low Q keeps the optimiser's resolution factor below 1: the bin count
    # follows the time span (256 bins for bunch 0, 512 for bunch 1 alone);
    # bunch 1 is longer than half of bunch 0's window, so ALLOW_LONG_BEAM
    zl = E.ZLONGIT(BROAD_BAND=1, RS=1e4, Q=0.02, FREQ=20e9, N_BINS=128, BIN_SIZE=1e-12, ALLOW_LONG_BEAM=1)
    q = 1e-9 / 4000
    out = rows.copy()
    o = _oracle(4000, q)
    o.zlongit(out, zl.p.zlongit, PO, 0)
    assert o.last_hist()[0].size == 256
    ref = _oracle(0, q)  # one context: bunch 0 sets the bins, bunch 1 reuses them
    for b in range(2):
        sel = rows[:, 8] == b
        alone = np.ascontiguousarray(rows[sel])
        ref.zlongit(alone, zl.p.zlongit, PO, 0)
        assert np.array_equal(out[sel], alone), "bucket %d" % b
    fresh = np.ascontiguousarray(rows[sel1])
    o1 = _oracle(0, q)
    o1.zlongit(fresh, zl.p.zlongit, PO, 0)
    assert o1.last_hist()[0].size == 512
    assert not np.array_equal(fresh, out[sel1])


def test_wide_rows_leave_the_checkers_row_width_at_eight():
    """the checker's row width is a library global (elegant's
    totalPropertiesPerParticle): a context with 10-column rows sets it for
    its own calls only, so later direct calls on 8-column rows stay in
    bounds"""
    rows = bunch_train(500, 2, 15e-12, seed=12)
    table = Lt.resonator_wake()
    w = E.WAKE(table)
    o = _oracle(500, 1e-12)
    o.wake(rows, w.p.wake, PO, 0)
    assert o.L.or_row_width() == 8
    with pytest.raises(ValueError):
        o.wake(np.ascontiguousarray(rows[:, :8]), w.p.wake, PO, 0)
    assert o.L.or_row_width() == 8


# ------------------------------------------------------------------ GPU

def _bunched(tracker, id_slots):
    tracker.set_particle_properties(10)
    tracker.set_id_slots_per_bunch(id_slots)


def _reset(tracker):
    tracker.set_id_slots_per_bunch(0)
    tracker.set_particle_properties(8)


@pytest.mark.gpu
@pytest.mark.parametrize("n_bins,start,end", [(0, -1, -1), (1024, -1, -1), (0, 1, 2)])
def test_wake_buckets_on_the_gpu_bit_exact(tracker, n_bins, start, end):
    rows = bunch_train(50000, 4, 4e-12, seed=21, first_index=3)
    table = Lt.resonator_wake()
    w = E.WAKE(table, N_BINS=n_bins, START_BUNCH=start, END_BUNCH=end)
    q = 1e-9 / 50000
    g = rows.copy()
    _bunched(tracker, 50000)
    try:
        h, v, tmin = tracker.collective_probe(g, w, PO, 0, q)
    finally:
        _reset(tracker)
    c = rows.copy()
    o = _oracle(50000, q)
    o.wake(c, w.p.wake, PO, 0)
    hc, vc, tminc, dt = o.last_hist()
    assert tmin == tminc and np.array_equal(h, hc) and np.array_equal(v, vc)
    assert_rows_identical(g, c, "bunched wake, n_bins %d, bunches %d..%d" % (n_bins, start, end))
    assert np.array_equal(g[:, 8:], rows[:, 8:])


@pytest.mark.gpu
def test_zlongit_buckets_on_the_gpu_bit_exact(tracker):
    rows = bunch_train(50000, 3, 15e-12, seed=22)
    zl = E.ZLONGIT(BROAD_BAND=1, RS=1e4, Q=1.0, FREQ=20e9, N_BINS=1024, BIN_SIZE=1e-12)
    q = 1e-9 / 50000
    g = rows.copy()
    _bunched(tracker, 50000)
    try:
        h, v, tmin = tracker.collective_probe(g, zl, PO, 0, q)
    finally:
        _reset(tracker)
    c = rows.copy()
    o = _oracle(50000, q)
    o.zlongit(c, zl.p.zlongit, PO, 0)
    hc, vc, tminc, dt = o.last_hist()
    assert tmin == tminc and np.array_equal(h, hc) and np.array_equal(v, vc)
    assert_rows_identical(g, c, "bunched zlongit")


@pytest.mark.gpu
def test_ring_with_a_bunch_train_bit_exact(tracker):
    """three bunches one RF period apart through the ring with WAKE, ZLONGIT and the
    15 mm aperture, with a halo that is lost over the passes: survivors,
    effort and rows (bunch column included) as the checker's"""
    bl, Po = Lt.storage_ring_wake()
    freq = [e.p.rfca.freq for e in bl.elements if e.type == A.EH_T_RFCA][0]
    spacing = 1 / freq  # one RF period: every bunch in the same phase
    n_per = 2000
    beam = Lt.matched_beam(bl, 3 * n_per, 1e-9, 1e-11, 1e-3, 20e-12, 41, Po)
    beta = Po / math.sqrt(Po * Po + 1)
    rows = np.zeros((3 * n_per, 10))
    rows[:, :8] = beam
    b = np.arange(3 * n_per) % 3
    rows[:, 4] += 2.99792458e8 * beta * spacing * b
    rows[:, 8] = b
    rows[:, 9] = 1
    rows[:300, 0] += 0.0149
    g = rows.copy()
    _bunched(tracker, n_per)
    try:
        ng, eg = tracker.do_tracking(g, bl, Po, 3)
    finally:
        _reset(tracker)
    c = rows.copy()
    o = O.Oracle("emath")
    o.set_row_width(10)
    with pytest.raises(RuntimeError):  # as one bucket the train outruns the 1 ns wake
        o.do_tracking(rows.copy(), bl, Po, 3)
    o = O.Oracle("emath")
    o.set_row_width(10)
    o.set_id_slots_per_bunch(n_per)
    nc = o.do_tracking(c, bl, Po, 3)
    assert ng == nc and eg == o.effort()
    assert rows.shape[0] - ng > 10
    assert_rows_identical(g, c, "ring with a bunch train")
    assert np.array_equal(g[:, 8:], c[:, 8:])
