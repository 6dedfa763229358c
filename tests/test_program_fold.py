"""Folding of repeated cells into runs (elegant_amd/host/eh_runs.c): the
tracking kernel walks a body of element records reps times instead of the
flat beamline, so the runs must cover the flat program exactly, in order, and
fold only records that are identical apart from their positions."""
import numpy as np

from elegant_amd import elements as E
from elegant_amd import lattices as Lt
from elegant_amd import native


def covers(runs, n):
    """runs tile [0, n) in order and their bodies are laid out back to back"""
    flat, first = 0, 0
    for f, L, R, fl in runs:
        assert fl == flat and f == first and L >= 1 and R >= 1
        flat += L * R
        first += L
    assert flat == n


def test_fodo_ring_folds_to_one_cell():
    bl, Po, _ = Lt.fodo_workload(n=4)
    runs = native.program_runs(bl, Po)
    covers(runs, 601)
    assert runs == [(0, 1, 1, 0), (1, 6, 100, 1)]


def test_storage_ring_folds_its_cells():
    bl, Po = Lt.storage_ring(seed=1)
    runs = native.program_runs(bl, Po)
    covers(runs, len(bl))
    assert (1, 27, 40, 1) in runs and len(runs) == 3


def test_distinct_records_stay_flat():
    """a strength, an aperture, a misalignment or a slice count that differs
    between otherwise equal elements keeps them apart"""
    els = [E.DRIF("D", L=1.0)]
    els += [E.KQUAD("Q", L=0.5, K1=1.0 + 1e-16 * 8), E.KQUAD("Q", L=0.5, K1=1.0)]
    els += [E.KQUAD("Q", L=0.5, K1=1.0, DX=1e-6), E.KQUAD("Q", L=0.5, K1=1.0, N_KICKS=8)]
    els += [E.KSEXT("S", L=0.2, K2=3.0), E.KSEXT("S", L=0.2, K2=3.0, TILT=1e-3)]
    bl = E.Beamline(els)
    runs = native.program_runs(bl, Lt.momentum(1000.0))
    covers(runs, len(bl))
    assert runs == [(0, len(bl), 1, 0)]


def test_greedy_choice_and_literal_runs():
    """A (B C)x3 D D D (B C)x2: the longest cover at each position, the
    shortest block among equal covers, single elements between repeats
    gathered into one single-pass run"""
    A = E.DRIF("A", L=0.3)
    B = E.KQUAD("B", L=0.5, K1=0.7)
    Cq = E.DRIF("C", L=1.1)
    D = E.KSEXT("D", L=0.2, K2=5.0)
    X = E.DRIF("X", L=0.9)
    els = [A, B, Cq, B, Cq, B, Cq, D, D, D, X, A, B, Cq, B, Cq]
    bl = E.Beamline(els)
    runs = native.program_runs(bl, Lt.momentum(1000.0))
    covers(runs, len(bl))
    assert runs == [(0, 1, 1, 0), (1, 2, 3, 1), (3, 1, 3, 7), (4, 2, 1, 10), (6, 2, 2, 12)]


def test_positions_do_not_split_cells():
    """the same cell placed at different positions folds: only (z_start,
    z_end, element index) differ, and they come from the position table"""
    cell = [E.DRIF("D", L=0.7), E.KQUAD("Q", L=0.4, K1=0.9)]
    bl = E.Beamline(cell * 50)
    runs = native.program_runs(bl, Lt.momentum(500.0))
    assert runs == [(0, 2, 50, 0)]
    assert np.sum([L * R for _, L, R, _ in runs]) == 100
