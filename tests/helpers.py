"""Shared workload builders for the parity tests."""
import numpy as np

from elegant_amd import lattices as Lt


def lossy_fodo(n, seed):
    """FODO ring beam with a halo reaching the 20 mm aperture, so that losses
    happen in drifts (extrapolated), inside multipoles and at the apertures."""
    bl, Po, rows = Lt.fodo_workload(n=n, seed=seed)
    rng = Lt.Rng(seed + 99)
    k = n // 3
    rows[:k, 0] = (rng.uniform(k) * 2 - 1) * 0.03
    rows[:k, 2] = (rng.uniform(k) * 2 - 1) * 0.03
    rows[:k, 1] = (rng.uniform(k) * 2 - 1) * 3e-3
    rows[:k, 5] = (rng.uniform(k) * 2 - 1) * 0.05
    return bl, Po, rows


def lossy_ring(n, seed, lattice_seed=1):
    bl, Po = Lt.storage_ring(seed=lattice_seed)
    rows = Lt.matched_beam(bl, n, 1e-9, 1e-11, 1e-3, 10e-12, seed, Po)
    rng = Lt.Rng(seed + 7)
    k = n // 3
    rows[:k, 0] = (rng.uniform(k) * 2 - 1) * 0.02
    rows[:k, 2] = (rng.uniform(k) * 2 - 1) * 0.012
    rows[:k, 5] = (rng.uniform(k) * 2 - 1) * 0.04
    return bl, Po, rows


def assert_rows_bitwise(a, b, what=""):
    """Bit-identical rows including the sign of zero (any NaN equals any NaN)."""
    assert a.shape == b.shape
    same = (a.view(np.int64) == b.view(np.int64)) | (np.isnan(a) & np.isnan(b))
    if not same.all():
        bad = np.argwhere(~same)
        i, j = bad[0]
        raise AssertionError("%s: %d values differ; first row %d col %d: %r vs %r (ids %r %r)" % (
            what, len(bad), i, j, a[i, j], b[i, j], a[i, 6], b[i, 6]))


def assert_rows_identical(a, b, what=""):
    """Bit-identical rows (signed zeros compare equal, NaN equals NaN)."""
    assert a.shape == b.shape
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    if not same.all():
        bad = np.argwhere(~same)
        i, j = bad[0]
        raise AssertionError("%s: %d values differ; first row %d col %d: %r vs %r (ids %r %r)" % (
            what, len(bad), i, j, a[i, j], b[i, j], a[i, 6], b[i, 6]))
