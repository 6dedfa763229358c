"""The reference's element API around the rows: accepted rows permuted with
the particles (multipole.c:1075-1076, limit_amplitudes.c:263-264, 308-309,
do_tracking.c:360-361), whole rows of totalPropertiesPerParticle columns
moved by the compaction (elegant.c:2663-2670), and do_tracking's stopping
conditions (stop_tracking_particle_limit and the interrupt file,
do_tracking.c:494-520), against the CPU checker."""
import ctypes as C
import os
import threading
import time

import numpy as np
import pytest

from elegant_amd import _abi as A
from elegant_amd import elements as E
from elegant_amd import lattices as Lt
from oracle import oracle_ctypes as O
from tests.helpers import assert_rows_bitwise, assert_rows_identical, lossy_fodo, lossy_ring
from tests.test_elements_gpu import beam, csbend_edge_rows, rows_ptr

pytestmark = pytest.mark.gpu

PO = Lt.momentum(6000.0)
W = 10  # x, x', y, y', s, delta, ID, pass of loss, bunch, weight (track.h:81-101)


def wide(rows):
    """rows with elegant's four basic properties: bunch and weight columns
    hold values derived from the ID, so that a row that moved without its
    particle shows"""
    w = np.zeros((rows.shape[0], W))
    w[:, :8] = rows[:, :8]
    w[:, 7] = -7.0 - rows[:, 6]           # a caller's value the element routines do not touch
    w[:, 8] = 3.0 * rows[:, 6]
    w[:, 9] = 0.5 + rows[:, 6]
    return w


def check_followers(initial, out, acc, what):
    """columns 7.. and the accepted rows belong to the particle now in the row"""
    idx = out[:, 6].astype(np.int64) - 1          # IDs are 1..n in initial row order
    assert (initial[idx, 6] == out[:, 6]).all()
    assert_rows_bitwise(out[:, 7:], initial[idx, 7:], what + " (columns 7..)")
    if acc is not None:
        assert_rows_bitwise(acc, initial[idx], what + " (accepted)")


@pytest.fixture()
def wide_tracker(tracker):
    tracker.set_particle_properties(W)
    yield tracker
    tracker.set_particle_properties(8)


def _oracle(fn, rows, *args):
    ptr, keep = rows_ptr(rows)
    return getattr(O.lib("emath"), fn)(ptr, rows.shape[0], *args)


def test_element_entry_points_permute_whole_rows_and_accepted(wide_tracker):
    tr = wide_tracker
    ap = A.EH_MAXAMP(0.005, 0.004, 0, 2, 2, 0)
    quad = E.KQUAD(L=0.4, K1=1.8, N_KICKS=8, DX=2e-4)
    bend = E.CSBEND(L=1.0, ANGLE=0.08, N_KICKS=16).p.csbend
    cases = [
        ("multipole", lambda g, a: tr.multipole_tracking2(g, quad, PO, 12.5, ap, accepted=a),
         lambda c: _oracle("or_multipole_tracking2", c, C.byref(quad), PO, 12.5, C.byref(ap))),
        ("csbend", lambda g, a: tr.track_through_csbend(g, bend, PO, 3.0, ap, accepted=a),
         lambda c: _oracle("or_track_through_csbend", c, C.byref(bend), PO, 3.0, C.byref(ap))),
        ("limit_amplitudes", lambda g, a: tr.limit_amplitudes(g, 0.004, 0.003, 2.0, PO, 1, accepted=a),
         lambda c: _oracle_limit(c, 0.004, 0.003, 2.0, PO, 1, 0)),
        ("elimit_amplitudes", lambda g, a: tr.elimit_amplitudes(g, 0.004, 0.003, 2.0, PO, 1, 0, 4, 2, accepted=a),
         lambda c: _oracle_elimit(c, 0.004, 0.003, 2.0, PO, 1, 0, 4, 2)),
    ]
    for name, gpu, cpu in cases:
        r = beam(2001, 5)
        initial = wide(r)
        g = initial.copy()
        acc = initial.copy()             # the caller's accepted copy, as do_tracking makes it
        ng = gpu(g, acc)
        c = r.copy()
        nc = cpu(c)
        assert ng == nc and 0 < nc < r.shape[0], name
        assert_rows_identical(g[:, :7], c[:, :7], name)
        check_followers(initial, g, acc, name)
        # without accepted the rows still move whole
        g2 = initial.copy()
        assert gpu(g2, None) == ng
        assert_rows_bitwise(g2, g, name + " without accepted")


def _oracle_limit(rows, xmax, ymax, z, Po, extrap, open_code):
    ptr, keep = rows_ptr(rows)
    return O.lib("emath").or_limit_amplitudes(ptr, xmax, ymax, rows.shape[0], z, Po, extrap, open_code)


def _oracle_elimit(rows, xmax, ymax, z, Po, extrap, open_code, xe, ye):
    ptr, keep = rows_ptr(rows)
    return O.lib("emath").or_elimit_amplitudes(ptr, xmax, ymax, rows.shape[0], z, Po, extrap, open_code, xe, ye)


def test_rf_cavity_permutes_accepted(wide_tracker):
    rf = E.RFCA(L=0.0, VOLT=5e6, PHASE=160.0, FREQ=352e6, T_REFERENCE=2e-8).p.rfca
    r = csbend_edge_rows(1000, 31)
    initial = wide(r)
    g = initial.copy()
    acc = initial.copy()
    ng, ph = wide_tracker.simple_rf_cavity(g, rf, PO, 7.0, accepted=acc)
    c = r.copy()
    o = O.Oracle("emath")
    pc = C.c_double(PO)
    ptr, keep = rows_ptr(c)
    nc = o.L.or_simple_rf_cavity(o.ctx, 0, ptr, c.shape[0], C.byref(rf), C.byref(pc), 7.0)
    assert ng == nc and 0 < nc < r.shape[0]
    assert_rows_bitwise(g[:, :7], c[:, :7], "rfca")
    check_followers(initial, g, acc, "rfca")


@pytest.mark.parametrize("ring", [False, True])
def test_do_tracking_accepted_and_wide_rows(wide_tracker, ring):
    """accepted ends holding each particle's initial row at the particle's
    final row; bunch and weight columns follow their particle; column 7 is
    the pass of loss"""
    bl, Po, rows = lossy_ring(900, seed=3) if ring else lossy_fodo(1500, seed=4)
    initial = wide(rows)
    g = initial.copy()
    acc = np.zeros_like(initial)
    ng, eg = wide_tracker.do_tracking(g, bl, Po, 3 if ring else 40, accepted=acc)
    c = rows.copy()
    o = O.Oracle("emath")
    nc = o.do_tracking(c, bl, Po, 3 if ring else 40)
    assert ng == nc and eg == o.effort() and 0 < nc < rows.shape[0]
    assert_rows_identical(g[:, :8], c[:, :8], "do_tracking")
    idx = g[:, 6].astype(np.int64) - 1
    assert_rows_bitwise(g[:, 8:], initial[idx, 8:], "bunch and weight")
    assert_rows_bitwise(acc, initial[idx], "accepted")


@pytest.mark.parametrize("chunk", [64, 7])
def test_particle_limit_stops_like_the_reference(tracker, chunk):
    """stop_tracking_particle_limit: the pass loop ends before the first pass
    that would start with fewer particles; a fused launch that crossed the
    limit is rolled back and rerun up to that pass"""
    bl, Po, rows = lossy_fodo(2000, seed=8)
    o = O.Oracle("glibc")
    c = rows.copy()
    nfull = o.do_tracking(c, bl, Po, 60)
    limit = (rows.shape[0] + nfull) // 2
    tracker.set_chunk_passes(chunk)
    tracker.set_particle_limit(limit)
    try:
        g = rows.copy()
        ng, eg = tracker.do_tracking(g, bl, Po, 60)
    finally:
        tracker.set_particle_limit(0)
        tracker.set_chunk_passes(64)
    o2 = O.Oracle("glibc")
    o2.set_particle_limit(limit)
    c2 = rows.copy()
    nc = o2.do_tracking(c2, bl, Po, 60)
    assert ng == nc and eg == o2.effort()
    assert nfull < nc < rows.shape[0] and eg < o.effort()
    assert_rows_identical(g, c2, "particle limit %d" % limit)


def test_particle_limit_on_passes_that_need_the_host(tracker):
    """the storage ring's first pass stops at the RF cavity for its TMEAN
    fiducial: the limit is tested before every such pass too"""
    bl, Po, rows = lossy_ring(900, seed=11)
    o = O.Oracle("emath")
    c = rows.copy()
    n1 = o.do_tracking(c, bl, Po, 1)
    tracker.set_particle_limit(n1 + 1)      # reached after the first pass
    try:
        g = rows.copy()
        ng, eg = tracker.do_tracking(g, bl, Po, 4)
    finally:
        tracker.set_particle_limit(0)
    assert ng == n1 and eg == o.effort()
    assert_rows_identical(g, c, "limit after pass 0")


def test_interrupt_file_and_flag(tracker, tmp_path):
    bl, Po, rows = lossy_fodo(200, seed=9)
    path = str(tmp_path / "stop")
    tracker.load(bl, Po)
    tracker.upload(rows)
    tracker.set_interrupt_file(path)
    try:
        st = tracker.track(0, 5)
        assert st.passes_done == 5
        open(path, "w").close()
        st = tracker.track(5, 5)
        assert st.passes_done == 0 and st.effort == 0
    finally:
        tracker.set_interrupt_file(None)
    # a file that already exists only counts once it is modified again
    tracker.set_interrupt_file(path)
    try:
        st = tracker.track(5, 5)
        assert st.passes_done == 5
        later = time.time() + 5
        os.utime(path, (later, later))
        assert tracker.track(10, 5).passes_done == 0
    finally:
        tracker.set_interrupt_file(None)
    g, ng = tracker.download()
    c = rows.copy()
    o = O.Oracle("glibc")
    nc = o.do_tracking(c, bl, Po, 10)
    assert ng == nc
    assert_rows_identical(g, c, "10 passes around the interrupts")

    flag = C.c_int32(1)
    tracker.set_interrupt_flag(flag)
    try:
        assert tracker.track(10, 5).passes_done == 0
        # set from another thread while passes run one launch at a time: the
        # run ends at a pass boundary, with every pass done complete
        flag.value = 0
        tracker.set_chunk_passes(1)
        t = threading.Timer(0.02, lambda: setattr(flag, "value", 1))
        t.start()
        st = tracker.track(10, 100000)
        t.join()
    finally:
        tracker.set_interrupt_flag(None)
        tracker.set_chunk_passes(64)
    done = int(st.passes_done)
    assert 0 < done < 100000
    g, ng = tracker.download()
    c = rows.copy()
    o = O.Oracle("glibc")
    nc = o.do_tracking(c, bl, Po, 10 + done)
    assert ng == nc
    assert_rows_identical(g, c, "%d passes before the flag" % (10 + done))
