"""The no-allocation WAKE step (eh_track_through_wake_noalloc): device-resident
particles, the Green function on the device and the library's static scratch,
against eh_track_through_wake (host rows) on the same bunch, bit for bit, on
two calls in a row, the second with fewer particles and a fixed-N_BINS wake."""
import numpy as np
import pytest

from elegant_amd import _abi as A
from elegant_amd import elements as E
from elegant_amd import lattices as Lt
from elegant_amd import native
from tests.helpers import assert_rows_bitwise
from tests.test_oracle import PO, wake_bunch

pytestmark = pytest.mark.gpu


def bunch(n, seed):
    rows = wake_bunch(n, 20e-12, seed=seed)
    rng = np.random.default_rng(seed)
    rows[:, 0] = 1e-4 * rng.standard_normal(n)
    rows[:, 1] = 2e-5 * rng.standard_normal(n)
    rows[:, 2] = 5e-5 * rng.standard_normal(n)
    rows[:, 3] = 1e-5 * rng.standard_normal(n)
    return rows


def run_noalloc(tr, rows, wake_elem, table, i_pass, q):
    n = rows.shape[0]
    cols = [tr.device_array(rows[:, k]) for k in range(6)]
    wdev = tr.device_array(table.W)
    try:
        ret = tr.track_through_wake_noalloc(cols, n, wake_elem.p.wake, wdev, PO, i_pass, q)
        out = rows.copy()
        for k in range(6):
            out[:, k] = tr.read_device_array(cols[k], n)
    finally:
        for a in cols + [wdev]:
            tr.free_device_array(a)
    return ret, out


def test_noalloc_wake_matches_host_row_entry_point(tracker):
    table = Lt.resonator_wake()
    cases = [(bunch(300000, 21), E.WAKE(table, INTERPOLATE=1), 0),
             (bunch(40000, 22), E.WAKE(table, N_BINS=4000, RAMP_PASSES=5), 2)]
    for rows, w, i_pass in cases:
        n = rows.shape[0]
        q = 5e-9 / n
        ref = rows.copy()
        r_ref = tracker.track_through_wake(ref, w.p.wake, PO, i_pass, q)
        r_na, got = run_noalloc(tracker, rows, w, table, i_pass, q)
        assert r_ref == n and r_na == n
        assert not np.array_equal(got[:, 5], rows[:, 5]), "the step kicked nothing"
        assert_rows_bitwise(got[:, :6], ref[:, :6], "no-allocation WAKE step, n=%d" % n)


def test_noalloc_wake_rejects_tables_beyond_its_scratch(tracker):
    table = Lt.resonator_wake(points=A.EH_NOALLOC_WAKE_POINTS + 1)
    w = E.WAKE(table)
    with pytest.raises(native.EHError) as ei:
        tracker.track_through_wake_noalloc([1] * 6, 10, w.p.wake, 1, PO, 0, 1e-12)
    assert ei.value.code == A.EH_ERR_INVALID
