"""GPU path vs the CPU checker on the same seeded inputs.

Rows come back from the GPU in elegant's row order (survivors, then lost
rows in swap-to-tail order), so the comparison is row by row: coordinates,
particle IDs and pass of loss must all be identical."""
import numpy as np
import pytest

from elegant_amd import lattices as Lt
from oracle import oracle_ctypes as O
from tests.helpers import assert_rows_identical, lossy_fodo, lossy_ring

pytestmark = pytest.mark.gpu


def run_both(tracker, bl, Po, rows, passes, mode):
    g = rows.copy()
    ng, eff_g = tracker.do_tracking(g, bl, Po, passes)
    c = rows.copy()
    o = O.Oracle(mode)
    nc = o.do_tracking(c, bl, Po, passes)
    assert tracker.P_central == o.P_central  # equal to Po unless CHANGE_P0 moved it
    return g, ng, eff_g, c, nc, o.effort()


def test_fodo_bit_exact_vs_glibc_checker(tracker):
    bl, Po, rows = lossy_fodo(1500, seed=21)
    g, ng, eg, c, nc, ec = run_both(tracker, bl, Po, rows, 60, "glibc")
    print("fodo: survivors", ng, "lost", rows.shape[0] - ng, "effort", eg)
    assert ng == nc and eg == ec
    assert rows.shape[0] - ng > 100  # losses were exercised
    assert_rows_identical(g, c, "fodo")


def test_ring_bit_exact_vs_emath_checker(tracker):
    bl, Po, rows = lossy_ring(600, seed=5)
    g, ng, eg, c, nc, ec = run_both(tracker, bl, Po, rows, 3, "emath")
    print("ring: survivors", ng, "lost", rows.shape[0] - ng)
    assert ng == nc and eg == ec
    assert rows.shape[0] - ng > 20
    assert_rows_identical(g, c, "ring")


def test_ring_vs_glibc_checker_one_turn(tracker):
    """glibc and the deterministic math differ in the last bit on ~2% of
    calls, and the bend's X = rho0 (f - 1) + f X amplifies such a difference
    in cos by rho0/x ~ 1e4: after one turn coordinates agree to 5e-10
    relative (the two CPU builds differ by the same amount, test_oracle.py)."""
    bl, Po, rows = Lt.storage_ring_workload(n=800, seed=9)
    g, ng, eg, c, nc, ec = run_both(tracker, bl, Po, rows, 1, "glibc")
    assert ng == nc
    gi = g[np.argsort(g[:ng, 6])]
    ci = c[np.argsort(c[:nc, 6])]
    for col, scale in [(0, 1e-3), (1, 1e-4), (2, 1e-4), (3, 1e-5), (4, 1.0), (5, 1e-3)]:
        err = np.abs(gi[:, col] - ci[:, col]) / np.maximum(np.abs(ci[:, col]), scale)
        assert err.max() <= 5e-10, (col, err.max())


def test_ring_edge_rows_bitwise(tracker):
    """do_tracking over the storage ring with rows of every kind mixed into
    an ordinary beam (zeros of both signs, subnormals, large slopes, 1 + delta
    at 0, huge offsets, NaN and infinities): the initial and per-element
    culls, the branch-free bodies and the general code together reproduce
    the checker bit for bit, signs of zeros included"""
    from tests.helpers import assert_rows_bitwise
    from tests.test_elements_gpu import csbend_edge_rows
    bl, Po, _ = Lt.storage_ring_workload(n=10, seed=9)
    rows = csbend_edge_rows(700, 41)
    rows[:, :4] *= 0.5
    g, ng, eg, c, nc, ec = run_both(tracker, bl, Po, rows, 2, "emath")
    assert ng == nc and eg == ec and 0 < nc < rows.shape[0]
    assert_rows_bitwise(g, c, "ring edges")


def test_example_deck_bit_exact_vs_emath_checker(tracker):
    """the reference's example deck (parTrack.lte: bends with K2 and K3 and
    HGAP/FINT edges, quadrupoles and sextupoles transmuted to KQUAD/KSEXT)
    over 50 turns, behind a MAXAMP of 10 x 5 mm so that a large-amplitude
    halo is lost along the ring: rows, order, losses and effort identical to
    the checker"""
    from elegant_amd import elements as E
    from tests.test_lte import par_ring
    from tests.test_elements_gpu import beam
    deck, Po = par_ring()
    bl = E.Beamline([E.MAXAMP("AP", X_MAX=0.01, Y_MAX=0.005)] + list(deck.elements), keep=deck.keep)
    rows = beam(2000, 21, halo=0.3, xshift=0.006)
    g = rows.copy()
    ng, eg = tracker.do_tracking(g, bl, Po, 50)
    o = O.Oracle("emath")
    c = rows.copy()
    nc = o.do_tracking(c, bl, Po, 50)
    print("parTrack deck: survivors", ng, "effort", eg)
    assert ng == nc and eg == o.effort()
    assert rows.shape[0] - ng > 10
    assert_rows_identical(g, c, "parTrack deck")


def test_repeated_cells_with_cavities_refold(tracker):
    """twelve identical cells, each with its own RF cavity whose TMEAN
    fiducial is set on the first pass from the beam's mean arrival time
    there: before the fiducials are known the cells fold into one body of
    records; each fiducial then makes its cavity's record differ, and the
    program refolds to the flat beamline.  Rows, order, losses and effort
    identical to the checker, with losses at every kind of element"""
    from elegant_amd import elements as E
    from elegant_amd import native
    els = [E.MAXAMP("AP", X_MAX=0.02, Y_MAX=0.02)]
    for _ in range(12):
        els += [E.DRIF("D1", L=2.0), E.KQUAD("QF", L=0.5, K1=1.2, N_KICKS=4), E.KSEXT("SF", L=0.2, K2=20.0),
                E.DRIF("D2", L=2.0), E.KQUAD("QD", L=0.5, K1=-1.2, N_KICKS=4), E.KSEXT("SD", L=0.2, K2=-35.0),
                E.RFCA("RF", L=0.0, VOLT=2e5, FREQ=352e6, PHASE=180.0)]
    bl = E.Beamline(els)
    _, Po, rows = lossy_fodo(1200, seed=33)
    assert native.program_runs(bl, Po) == [(0, 1, 1, 0), (1, 7, 12, 1)]
    g, ng, eg, c, nc, ec = run_both(tracker, bl, Po, rows, 8, "emath")
    print("cells with cavities: survivors", ng, "lost", rows.shape[0] - ng, "runs", tracker.program_runs())
    assert tracker.program_runs() == 1
    assert ng == nc and eg == ec
    assert rows.shape[0] - ng > 100
    assert_rows_identical(g, c, "cells with cavities")


def test_ring_with_change_p0_bit_exact(tracker):
    """a FODO ring whose quadrupoles are given by field and bore (B/BORE:
    their strengths scale with the central momentum, multipole.c) and two
    RF cavities with CHANGE_P0, each moving the central momentum to the
    beam's mean every pass: rows, losses, effort and the final central
    momentum identical to the checker over 12 passes of a lossy beam"""
    from elegant_amd import elements as E
    els = [E.MAXAMP("AP", X_MAX=0.02, Y_MAX=0.02)]
    for cell in range(12):
        els += [E.DRIF("D1", L=2.0), E.KQUAD("QF", L=0.5, B=0.6, BORE=0.025, N_KICKS=4),
                E.KSEXT("SF", L=0.2, K2=20.0), E.DRIF("D2", L=2.0),
                E.KQUAD("QD", L=0.5, B=-0.6, BORE=0.025, N_KICKS=4), E.KSEXT("SD", L=0.2, K2=-35.0)]
        if cell in (0, 6):
            els.append(E.RFCA("RF", L=0.0, VOLT=3e6, FREQ=352e6, PHASE=150.0, CHANGE_P0=1))
    bl = E.Beamline(els)
    _, Po, rows = lossy_fodo(1500, seed=37)
    g, ng, eg, c, nc, ec = run_both(tracker, bl, Po, rows, 12, "emath")
    print("change_p0 ring: survivors", ng, "lost", rows.shape[0] - ng, "P_central", Po, "->", tracker.P_central)
    assert ng == nc and eg == ec
    assert rows.shape[0] - ng > 100
    assert tracker.P_central != Po
    assert_rows_identical(g, c, "change_p0 ring")
