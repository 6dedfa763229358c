"""The measured workloads at their full sizes on the GPU, checked through
properties that hold at any size:

* the storage ring with 10^6 particles: particles are independent apart
  from the RF cavity's fiducial time (the mean arrival time of all
  particles on the first pass), so every particle of a seeded sample,
  tracked alone on the CPU checker with the cavity given the phase the GPU
  run fixed, must come out bit-identical (coordinates, ID, pass of loss)
  to its row in the full-size GPU run; row IDs stay a permutation;
  the effort equals the survivors summed over passes as the loss passes
  record them; a second run gives the same bits in the same order;
* the ring with WAKE and ZLONGIT with 2*10^6 particles: particles interact
  through the histograms, so the run is checked for determinism (the
  fixed-point sums and integer histograms make it independent of thread
  order), row bookkeeping and effort."""
import numpy as np
import pytest

from elegant_amd import _abi as A
from elegant_amd import lattices as Lt
from oracle import oracle_ctypes as O
from tests.helpers import assert_rows_bitwise

pytestmark = pytest.mark.gpu


def check_bookkeeping(rows0, g, ng, effort, passes):
    n = rows0.shape[0]
    assert np.array_equal(np.sort(g[:, 6]), np.arange(1, n + 1))
    assert (g[:ng, 7] == -1).all() and (g[ng:, 7] >= 0).all() and (g[ng:, 7] < passes).all()
    # do_tracking's effort: the particles left after each pass (do_tracking.c:2758-2765)
    alive = sum(int(np.sum((g[:, 7] == -1) | (g[:, 7] > p))) for p in range(passes))
    assert effort == alive


def test_storage_ring_full_size_sampled_parity(tracker):
    bl, Po, rows = Lt.storage_ring_workload()
    n = rows.shape[0]
    assert n == 1_000_000 and np.array_equal(rows[:, 6], np.arange(1, n + 1))
    # a halo pushed 10 mm out: about 70 % of it is lost in the two passes,
    # on both of them, so the sample holds lost and surviving halo particles
    h = n // 40
    rows[:h, 0] += np.sign(rows[:h, 0]) * 0.010
    passes = 2
    print("storage ring, %d particles, %d passes" % (n, passes), flush=True)
    g = rows.copy()
    ng, effort = tracker.do_tracking(g, bl, Po, passes)
    assert 0 < n - ng < h
    check_bookkeeping(rows, g, ng, effort, passes)
    # a seeded sample, a third of it from the halo, tracked alone on the checker
    rng = np.random.default_rng(5)
    ids = np.unique(np.concatenate([rng.choice(np.arange(1, h + 1), 500, replace=False),
                                    rng.choice(np.arange(h + 1, n + 1), 1000, replace=False)]))
    sub = rows[ids - 1].copy()
    o = O.Oracle("emath")
    rf = [i for i, e in enumerate(bl.elements) if e.type == A.EH_T_RFCA]
    assert len(rf) == 1
    phase = tracker.rfca_fiducial(rf[0])
    assert np.isfinite(phase)
    o.set_fiducial(rf[0], phase)
    o.do_tracking(sub, bl, Po, passes)
    where = np.empty(n + 1, dtype=np.int64)
    where[g[:, 6].astype(np.int64)] = np.arange(n)
    got = g[where[sub[:, 6].astype(np.int64)]]
    assert (sub[:, 7] >= 0).sum() > 50
    assert_rows_bitwise(got, sub, "sampled particles of the 10^6 run")
    # the same input gives the same bits in the same order
    g2 = rows.copy()
    ng2, effort2 = tracker.do_tracking(g2, bl, Po, passes)
    assert ng2 == ng and effort2 == effort
    assert_rows_bitwise(g2, g, "second run")


def test_ring_with_wakes_full_size_is_deterministic(tracker):
    bl, Po, rows = Lt.storage_ring_wake_workload()
    n = rows.shape[0]
    assert n == 2_000_000
    h = n // 40
    rows[:h, 0] += np.sign(rows[:h, 0]) * 0.010
    passes = 2
    print("ring with wakes, %d particles, %d passes" % (n, passes), flush=True)
    g = rows.copy()
    ng, effort = tracker.do_tracking(g, bl, Po, passes)
    assert 0 < n - ng < h
    check_bookkeeping(rows, g, ng, effort, passes)
    g2 = rows.copy()
    ng2, effort2 = tracker.do_tracking(g2, bl, Po, passes)
    assert ng2 == ng and effort2 == effort
    assert_rows_bitwise(g2, g, "second run with wakes")
