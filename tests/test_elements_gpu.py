"""The element-by-element entry points of the C ABI (the routines elegant's
element loop calls) against the same routines of the CPU checker, on rows
with losses, and do_tracking on the edge cases: no particles, one particle,
a ragged count, everything lost at once, invalid coordinates."""
import ctypes as C
import math

import numpy as np
import pytest

from elegant_amd import _abi as A
from elegant_amd import elements as E
from elegant_amd import lattices as Lt
from oracle import oracle_ctypes as O
from tests.helpers import assert_rows_bitwise, assert_rows_identical

pytestmark = pytest.mark.gpu

PO = Lt.momentum(6000.0)


def beam(n, seed, halo=0.3, xshift=0.004):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8))
    r[:, 0] = rng.normal(0, 1e-3, n)
    r[:, 1] = rng.normal(0, 2e-4, n)
    r[:, 2] = rng.normal(0, 5e-4, n)
    r[:, 3] = rng.normal(0, 1e-4, n)
    r[:, 4] = rng.normal(0, 3e-3, n)
    r[:, 5] = rng.normal(0, 1e-3, n)
    k = int(halo * n)
    r[:k, 0] += xshift * np.sign(r[:k, 0])
    r[:, 6] = np.arange(1, n + 1)
    r[:, 7] = -1
    return r


def rows_ptr(r):
    return A.rows_pointer(r)


@pytest.mark.parametrize("make", [
    lambda: E.KQUAD(L=0.4, K1=1.8, N_KICKS=8, DX=2e-4, DY=-1e-4, TILT=3e-3),
    lambda: E.KSEXT(L=0.2, K2=300.0, N_KICKS=4, DX=1e-4),
    lambda: E.KOCT(L=0.1, K3=4e4, TILT=1e-3),
    lambda: E.KQUAD(L=0.5, LEFFECTIVE=0.4, K1=1.8, N_KICKS=8, DX=2e-4, DY=-1e-4, TILT=3e-3),
])
def test_multipole_tracking2(tracker, make):
    e = make()
    ap = A.EH_MAXAMP(0.005, 0.004, 0, 2, 2, 0)
    r = beam(3001, 1)
    g = r.copy()
    c = r.copy()
    ng = tracker.multipole_tracking2(g, e, PO, 12.5, ap)
    ptr, keep = rows_ptr(c)
    nc = O.lib("emath").or_multipole_tracking2(ptr, c.shape[0], C.byref(e), PO, 12.5, C.byref(ap))
    assert ng == nc and 0 < nc < r.shape[0]
    assert_rows_identical(g[:, :7], c[:, :7], "multipole")


@pytest.mark.parametrize("make,aperture", [
    (lambda: E.KQUAD(L=0.4, K1=1.8, N_KICKS=8, DX=2e-4, DY=-1e-4, TILT=3e-3), True),
    (lambda: E.KQUAD(L=0.4, K1=-1.8, N_KICKS=8), False),
    (lambda: E.KSEXT(L=0.2, K2=300.0, N_KICKS=4, DX=1e-4), False),
    (lambda: E.KOCT(L=0.0, K3=4e4), False),
    (lambda: E.KQUAD(L=0.4, K1=1.8, N_KICKS=8, EXPAND_HAMILTONIAN=1), False),
    (lambda: E.KQUAD(L=0.5, LEFFECTIVE=0.4, K1=1.8, N_KICKS=8, DX=2e-4, TILT=3e-3), True),
])
def test_multipole_fast_path_edges(tracker, make, aperture):
    """multipole_tracking2 with rows of every kind in the same waves as
    ordinary ones (the bend's edge set: zeros of both signs, tiny, huge,
    NaN and infinite values, slopes that end the particle): the branch-free
    common path and the general code together match the checker bit for
    bit, signs of zeros included; a thin element and an expanded
    Hamiltonian take the general code"""
    e = make()
    ap = A.EH_MAXAMP(0.05, 0.04, 0, 2, 2, 0) if aperture else None
    r = csbend_edge_rows(1000, 21)
    g = r.copy()
    c = r.copy()
    ng = tracker.multipole_tracking2(g, e, PO, 12.5, ap)
    ptr, keep = rows_ptr(c)
    nc = O.lib("emath").or_multipole_tracking2(ptr, c.shape[0], C.byref(e), PO, 12.5,
                                                C.byref(ap) if ap is not None else None)
    assert_same_survivors(r, g, ng, c, nc)
    assert_rows_bitwise(g[:, :7], c[:, :7], "multipole edges")


def test_track_through_csbend(tracker):
    b = E.CSBEND(L=1.0, ANGLE=0.08, N_KICKS=16, E1=0.02, E2=0.04, DX=1e-4, TILT=1e-3).p.csbend
    ap = A.EH_MAXAMP(0.005, 0.004, 0, 2, 2, 0)
    r = beam(2500, 2)
    g = r.copy()
    c = r.copy()
    ng = tracker.track_through_csbend(g, b, PO, 3.0, ap)
    ptr, keep = rows_ptr(c)
    nc = O.lib("emath").or_track_through_csbend(ptr, c.shape[0], C.byref(b), PO, 3.0, C.byref(ap))
    assert ng == nc and 0 < nc < r.shape[0]
    assert_rows_identical(g[:, :7], c[:, :7], "csbend")


def assert_same_survivors(r, g, ng, c, nc):
    """the same particles survive on both sides (names the rows that differ)"""
    kept_g = set(g[:ng, 6].astype(int).tolist())
    kept_c = set(c[:nc, 6].astype(int).tolist())
    diff = sorted(kept_g ^ kept_c)
    assert not diff, "survivors differ (ids, initial rows, kept on the GPU): %r" % [
        (i, r[i - 1, :6].tolist(), i in kept_g) for i in diff[:6]]
    assert ng == nc and 0 < nc < r.shape[0]


def csbend_edge_rows(n, seed):
    """an ordinary beam with, every 7th row, a case the bend's branch-free
    path must either reproduce bit for bit or hand to the general code:
    exact and negative zeros, tiny and subnormal values, slopes outside the
    narrow polynomial bands, momenta that end the particle inside the bend,
    1 + delta at or near 0, huge offsets, NaN and infinities"""
    r = beam(n, seed, halo=0.0)
    nan, inf = float("nan"), float("inf")
    cases = [
        (0, 0, 0, 0, 0), (-0.0, -0.0, -0.0, -0.0, -0.0), (0, 1e-4, 0, 0, 0), (1e-3, 0, 0, -0.0, 0),
        (1e-300, 1e-310, 0, -1e-300, 0), (5e-324, -5e-324, 1e-320, 0, 0), (0, 0.2, 0, 0, 0),
        (0, -0.9, 0, 0, 0), (0, 0, 0, 2.0, 0), (0, 1.5, 0, 0, 0), (1e-3, 1e-4, 0, 0, -1.0),
        (1e-3, 1e-4, 0, 0, -0.9999999), (1e200, 0, 0, 0, 0), (1e300, 0, 0, 0, 0), (-12.7, 0, 0, 0, 0),
        (nan, 0, 0, 0, 0), (0, inf, 0, 0, 0), (0, 0, -inf, 0, 0), (1e-3, 0, 0, 0, nan), (0, 0, 0, 0, 1e6),
        (0.03, 0.0312, 0, 0, 0), (0, 0.03125, 0, 0.0, 0), (2e-3, -2e-3, 1e-3, 1e-3, 0.05),
    ]
    for k, (x, xp, y, yp, dp) in enumerate(cases):
        i = 7 * k + 3
        r[i, [0, 1, 2, 3, 5]] = (x, xp, y, yp, dp)
    return r


@pytest.mark.parametrize("angle,aperture", [(0.08, True), (0.08, False), (-0.08, False), (0.3, False)])
def test_csbend_fast_path_edges(tracker, angle, aperture):
    """CSBEND rows of every kind in the same waves as ordinary ones: the
    branch-free common path and the general code together must match the
    checker bit for bit, signs of zeros included (a negative bend and a
    strong one exercise the general code for whole elements and for most
    substeps)"""
    b = E.CSBEND(L=1.0, ANGLE=angle, N_KICKS=16, E1=0.02, E2=0.04, DX=1e-4, TILT=1e-3).p.csbend
    ap = A.EH_MAXAMP(0.05, 0.04, 0, 2, 2, 0) if aperture else None
    r = csbend_edge_rows(1000, 12)
    g = r.copy()
    c = r.copy()
    ng = tracker.track_through_csbend(g, b, PO, 3.0, ap)
    ptr, keep = rows_ptr(c)
    nc = O.lib("emath").or_track_through_csbend(ptr, c.shape[0], C.byref(b), PO, 3.0,
                                                 C.byref(ap) if ap is not None else None)
    assert_same_survivors(r, g, ng, c, nc)
    assert_rows_bitwise(g[:, :7], c[:, :7], "csbend edges angle %g" % angle)


def test_simple_rf_cavity_two_calls(tracker):
    """the first call sets the fiducial phase (TMEAN), the second reuses it"""
    rf = E.RFCA(L=0.0, VOLT=5e6, PHASE=160.0, FREQ=352e6).p.rfca
    r = beam(2000, 3, halo=0.0)
    r[:5, 5] = -1.5  # invalid momenta are culled after the cavity
    g = r.copy()
    c = r.copy()
    ng1, ph = tracker.simple_rf_cavity(g, rf, PO, 7.0)
    o = O.Oracle("emath")
    ptr, keep = rows_ptr(c)
    pc = C.c_double(PO)
    nc1 = o.L.or_simple_rf_cavity(o.ctx, 0, ptr, c.shape[0], C.byref(rf), C.byref(pc), 7.0)
    assert ng1 == nc1
    assert_rows_identical(g[:, :7], c[:, :7], "rfca pass 0")
    g2 = g[:ng1].copy()
    c2 = c[:nc1].copy()
    ng2, ph2 = tracker.simple_rf_cavity(g2, rf, PO, 7.0, ph)
    assert ph2 == ph
    ptr, keep = rows_ptr(c2)
    nc2 = o.L.or_simple_rf_cavity(o.ctx, 0, ptr, c2.shape[0], C.byref(rf), C.byref(pc), 7.0)
    assert ng2 == nc2
    assert_rows_identical(g2[:, :7], c2[:, :7], "rfca pass 1")


def test_simple_rf_cavity_change_p0(tracker):
    """CHANGE_P0 (simple_rfca.c:977-982, do_tracking.c:3371-3417): after the
    cavity's own culls the central momentum becomes the KahanPlus mean
    momentum of the rows left, and every delta is re-expressed against it;
    rows and the new central momentum bit for bit against the checker"""
    rf = E.RFCA(L=0.0, VOLT=5e6, PHASE=160.0, FREQ=352e6, CHANGE_P0=1).p.rfca
    r = beam(2000, 3, halo=0.0)
    r[:5, 0] = np.nan  # removed by the cavity's removeInvalidParticles before the mean
    g = r.copy()
    c = r.copy()
    ng, ph = tracker.simple_rf_cavity(g, rf, PO, 7.0)
    o = O.Oracle("emath")
    ptr, keep = rows_ptr(c)
    pc = C.c_double(PO)
    nc = o.L.or_simple_rf_cavity(o.ctx, 0, ptr, c.shape[0], C.byref(rf), C.byref(pc), 7.0)
    assert ng == nc == r.shape[0] - 5
    assert tracker.P_central == pc.value and pc.value != PO
    assert_rows_identical(g[:, :7], c[:, :7], "rfca change_p0")


@pytest.mark.parametrize("volt", [5e6, 0.0])
def test_simple_rf_cavity_edges(tracker, volt):
    """the cavity with rows of every kind: its own removeInvalidParticles
    (any NaN coordinate, delta <= -1) is the only cull; infinite offsets pass
    through as the reference lets them (bit for bit, signs of zeros
    included).  T_REFERENCE fixes the phase without averaging the rows."""
    rf = E.RFCA(L=0.0, VOLT=volt, PHASE=160.0, FREQ=352e6, T_REFERENCE=2e-8).p.rfca
    r = csbend_edge_rows(1000, 31)
    g = r.copy()
    c = r.copy()
    ng, ph = tracker.simple_rf_cavity(g, rf, PO, 7.0)
    o = O.Oracle("emath")
    ptr, keep = rows_ptr(c)
    pc = C.c_double(PO)
    nc = o.L.or_simple_rf_cavity(o.ctx, 0, ptr, c.shape[0], C.byref(rf), C.byref(pc), 7.0)
    if volt == 0:
        assert ng == nc == r.shape[0]
    else:
        assert_same_survivors(r, g, ng, c, nc)
    assert_rows_bitwise(g[:, :7], c[:, :7], "rfca edges volt %g" % volt)


def test_drift_and_limit_amplitudes(tracker):
    d = E.DRIF(L=1.3).p.drift
    r = beam(4097, 4)
    g = r.copy()
    c = r.copy()
    assert tracker.track_particles_drift(g, d) == r.shape[0]
    ptr, keep = rows_ptr(c)
    O.lib("glibc").or_track_particles_drift(ptr, c.shape[0], 1.3, 2)
    assert_rows_identical(g[:, :7], c[:, :7], "drift")
    ng = tracker.limit_amplitudes(g, 0.004, 0.002, 9.0, PO, 1)
    nc = O.lib("glibc").or_limit_amplitudes(ptr, 0.004, 0.002, c.shape[0], 9.0, PO, 1, 0)
    assert ng == nc and 0 < nc < r.shape[0]
    assert_rows_identical(g[:, :7], c[:, :7], "limit_amplitudes")


def test_drift_and_limit_amplitudes_edges(tracker):
    """the element entry points leave NaN and infinite rows to limit_amplitudes,
    which removes them at its own z (bit for bit, signs of zeros included)"""
    d = E.DRIF(L=1.3).p.drift
    r = csbend_edge_rows(1000, 8)
    g = r.copy()
    c = r.copy()
    assert tracker.track_particles_drift(g, d) == r.shape[0]
    ptr, keep = rows_ptr(c)
    O.lib("glibc").or_track_particles_drift(ptr, c.shape[0], 1.3, 2)
    assert_rows_bitwise(g[:, :7], c[:, :7], "drift edges")
    ng = tracker.limit_amplitudes(g, 0.004, 0.002, 9.0, PO, 1)
    nc = O.lib("glibc").or_limit_amplitudes(ptr, 0.004, 0.002, c.shape[0], 9.0, PO, 1, 0)
    assert_same_survivors(r, g, ng, c, nc)
    assert_rows_bitwise(g[:, :7], c[:, :7], "limit_amplitudes edges")


# (elliptical, exponent, yExponent, open side, extrapolate): a true ellipse
# with the drift loss point, super-ellipses, each open side, a rectangle
# with an open side, one half-axis non-positive
SHAPES = [(1, 2, 0, 0, 1), (1, 2, 0, 0, 0), (1, 4, 0, 0, 1), (1, 6, 2, 0, 0), (1, 10, 40, 0, 0),
          (1, 2, 0, A.EH_OPEN_PLUS_X, 1), (1, 4, 0, A.EH_OPEN_MINUS_Y, 0), (0, 2, 0, A.EH_OPEN_MINUS_X, 1),
          (0, 2, 0, A.EH_OPEN_PLUS_Y, 0)]


def shaped_rows(n, seed):
    """edge rows plus a halo spread over the aperture's boundary in both
    planes, with slopes of both signs (the drift loss point)"""
    r = csbend_edge_rows(n, seed)
    rng = np.random.default_rng(seed + 1)
    k = n // 3
    r[:k, 0] = rng.uniform(-0.006, 0.006, k)
    r[:k, 2] = rng.uniform(-0.003, 0.003, k)
    r[:k, 1] = rng.normal(0, 2e-3, k)
    r[:k, 3] = rng.normal(0, 1e-3, k)
    return r


@pytest.mark.parametrize("shape", SHAPES)
def test_shaped_limit_amplitudes(tracker, shape):
    """limit_amplitudes with an open side and elimit_amplitudes for the
    ellipse and super-ellipses (limit_amplitudes.c:193-272, 472-598)
    against the checker: survivors, order, loss point and the ellipse's
    sqrt((Po(1+delta))^2 + 1) in the lost rows, bit for bit"""
    ell, xe, ye, op, ex = shape
    r = shaped_rows(1200, 31)
    g = r.copy()
    c = r.copy()
    ptr, keep = rows_ptr(c)
    if ell:
        ng = tracker.elimit_amplitudes(g, 0.004, 0.002, 9.0, PO, ex, op, xe, ye)
        nc = O.lib("glibc").or_elimit_amplitudes(ptr, 0.004, 0.002, c.shape[0], 9.0, PO, ex, op, xe, ye)
    else:
        ng = tracker.limit_amplitudes(g, 0.004, 0.002, 9.0, PO, ex, op)
        nc = O.lib("glibc").or_limit_amplitudes(ptr, 0.004, 0.002, c.shape[0], 9.0, PO, ex, op)
    assert_same_survivors(r, g, ng, c, nc)
    assert_rows_bitwise(g[:, :7], c[:, :7], "shape %r" % (shape,))


def test_shaped_limit_amplitudes_early_returns(tracker):
    """a non-positive half-axis: the ellipse falls back to the rectangle in
    the other plane, none positive checks nothing (NaN rows stay); odd
    exponents are refused"""
    r = shaped_rows(600, 5)
    for xmax, ymax in ((0.004, 0.0), (0.0, -1.0), (-1.0, -1.0)):
        g = r.copy()
        c = r.copy()
        ptr, keep = rows_ptr(c)
        ng = tracker.elimit_amplitudes(g, xmax, ymax, 2.0, PO, 1, 0, 4, 0)
        nc = O.lib("glibc").or_elimit_amplitudes(ptr, xmax, ymax, c.shape[0], 2.0, PO, 1, 0, 4, 0)
        assert ng == nc
        assert_rows_bitwise(g[:, :7], c[:, :7], "early return %g %g" % (xmax, ymax))
    with pytest.raises(Exception):
        tracker.elimit_amplitudes(r.copy(), 0.004, 0.002, 2.0, PO, 1, 0, 3, 0)


@pytest.mark.parametrize("shape", [(1, 2, 0, 0), (1, 4, 2, 0), (1, 2, 0, A.EH_OPEN_PLUS_X),
                                   (0, 2, 0, A.EH_OPEN_MINUS_Y)])
def test_shaped_apertures_inside_elements(tracker, shape):
    """checkMultAperture's ellipse (>= 1) and open sides inside the
    multipole and bend integrators (multipole.c:1786-1833), both elements
    misaligned and tilted so the aperture frame is rotated"""
    ell, xe, ye, op = shape
    ap = A.EH_MAXAMP(0.005, 0.0025, ell, xe, ye, op)
    q = E.KQUAD(L=0.4, K1=1.8, N_KICKS=8, DX=2e-4, DY=-1e-4, TILT=3e-3)
    b = E.CSBEND(L=1.0, ANGLE=0.08, N_KICKS=16, E1=0.02, E2=0.04, DX=1e-4, TILT=1e-3).p.csbend
    r = shaped_rows(1500, 17)
    g = r.copy()
    c = r.copy()
    ng = tracker.multipole_tracking2(g, q, PO, 12.5, ap)
    ptr, keep = rows_ptr(c)
    nc = O.lib("emath").or_multipole_tracking2(ptr, c.shape[0], C.byref(q), PO, 12.5, C.byref(ap))
    assert_same_survivors(r, g, ng, c, nc)
    assert_rows_bitwise(g[:, :7], c[:, :7], "multipole shape %r" % (shape,))
    g = r.copy()
    c = r.copy()
    ng = tracker.track_through_csbend(g, b, PO, 3.0, ap)
    ptr, keep = rows_ptr(c)
    nc = O.lib("emath").or_track_through_csbend(ptr, c.shape[0], C.byref(b), PO, 3.0, C.byref(ap))
    assert_same_survivors(r, g, ng, c, nc)
    assert_rows_bitwise(g[:, :7], c[:, :7], "csbend shape %r" % (shape,))


@pytest.mark.parametrize("shape", [(1, 2, 0, 0), (1, 6, 4, A.EH_OPEN_PLUS_X), (0, 2, 0, A.EH_OPEN_MINUS_Y)])
def test_ring_with_shaped_aperture(tracker, shape):
    """do_tracking over the storage ring with its MAXAMP made elliptical or
    given an open side: every element's in-integrator check and the check
    after each element (drifts with the loss point), three passes"""
    ell, xe, ye, op = shape
    bl, Po = Lt.storage_ring(seed=4)
    els = [E.MAXAMP("AP", X_MAX=0.012, Y_MAX=0.006, ELLIPTICAL=ell, EXPONENT=xe, YEXPONENT=ye,
                    OPEN_SIDE=op)] + bl.elements[1:]
    assert bl.elements[0].type == A.EH_T_MAXAMP
    bl = E.Beamline(els, keep=bl.keep)
    rows = Lt.matched_beam(bl, 1000, 1e-9, 1e-11, 1e-3, 10e-12, 9, Po)
    rows[:300, 0] += np.sign(rows[:300, 0]) * 0.0105
    rows[300:450, 2] += np.sign(rows[300:450, 2]) * 0.0052
    g = rows.copy()
    c = rows.copy()
    ng, eg = tracker.do_tracking(g, bl, Po, 3)
    o = O.Oracle("emath")
    nc = o.do_tracking(c, bl, Po, 3)
    assert ng == nc and eg == o.effort() and 0 < nc < 1000
    assert_rows_bitwise(g, c, "ring shape %r" % (shape,))


def test_wake_and_zlongit_entry_points(tracker):
    table = Lt.resonator_wake(points=4000)
    w = E.WAKE(table)
    # 8,192 bins of 2 ps: the bin optimiser keeps this layout for the resonator
    zl = E.ZLONGIT(BROAD_BAND=1, RS=1e4, Q=1.0, FREQ=20e9, N_BINS=8192, BIN_SIZE=2e-12)
    r = beam(20000, 5, halo=0.0)
    r[:, 4] = 0.2 + r[:, 4] * 2
    q = 2e-9 / r.shape[0]
    for el, kind in ((w, "wake"), (zl, "zlongit")):
        g = r.copy()
        c = r.copy()
        o = O.Oracle("emath")
        o.set_macro_charge(q)
        if kind == "wake":
            ng = tracker.track_through_wake(g, el.p.wake, PO, 0, q)
            nc = o.wake(c, el.p.wake, PO, 0)
        else:
            ng = tracker.track_through_zlongit(g, el.p.zlongit, PO, 0, q)
            nc = o.zlongit(c, el.p.zlongit, PO, 0)
        assert ng == nc == r.shape[0]
        assert_rows_identical(g[:, :7], c[:, :7], kind)


@pytest.mark.parametrize("n", [0, 1, 257, 1000])
def test_do_tracking_counts(tracker, n):
    bl, Po = Lt.storage_ring(seed=4)
    rows = Lt.matched_beam(bl, n, 1e-9, 1e-11, 1e-3, 10e-12, 9, Po) if n else np.zeros((0, 8))
    if n:
        rows[: n // 3, 0] += 0.0149
    g = rows.copy()
    c = rows.copy()
    ng, eg = tracker.do_tracking(g, bl, Po, 2)
    o = O.Oracle("emath")
    nc = o.do_tracking(c, bl, Po, 2)
    assert ng == nc and eg == o.effort()
    assert_rows_identical(g, c, "n=%d" % n)


def test_everything_lost_and_invalid_inputs(tracker):
    bl, Po = Lt.storage_ring(seed=4)
    rows = Lt.matched_beam(bl, 300, 1e-9, 1e-11, 1e-3, 10e-12, 10, Po)
    rows[:, 0] += 0.02          # every particle outside the 15 mm aperture
    rows[7, 1] = float("nan")   # invalid slope
    rows[8, 5] = -1.0           # delta = -1
    g = rows.copy()
    c = rows.copy()
    ng, eg = tracker.do_tracking(g, bl, Po, 3)
    o = O.Oracle("emath")
    nc = o.do_tracking(c, bl, Po, 3)
    assert ng == nc == 0 and eg == o.effort()
    assert_rows_identical(g, c, "all lost")


@pytest.mark.parametrize("end1,end2,fid", [(0, 0, "tref"), (1, 1, "tref"), (1, 0, "tmean")])
def test_simple_rf_cavity_matrix_method(tracker, end1, end2, fid):
    """RFCA with L > 0 and N_KICKS = 0: the matrix method (simple_rfca.c:
    403-407, 824-941), with and without the end-focus kicks, on rows of
    every kind (its own removeInvalidParticles is the only cull), bit for
    bit against the checker, signs of zeros included"""
    kw = dict(L=0.6, VOLT=4e6, PHASE=150.0, FREQ=352e6, N_KICKS=0, END1_FOCUS=end1, END2_FOCUS=end2)
    if fid == "tref":
        kw["T_REFERENCE"] = 2e-8
    rf = E.RFCA(**kw).p.rfca
    r = csbend_edge_rows(1000, 37) if fid == "tref" else beam(1500, 38, halo=0.0)
    if fid != "tref":
        r[:4, 5] = -1.5
    g = r.copy()
    c = r.copy()
    ng, ph = tracker.simple_rf_cavity(g, rf, PO, 9.0)
    o = O.Oracle("emath")
    ptr, keep = rows_ptr(c)
    pc = C.c_double(PO)
    nc = o.L.or_simple_rf_cavity(o.ctx, 0, ptr, c.shape[0], C.byref(rf), C.byref(pc), 9.0)
    assert_same_survivors(r, g, ng, c, nc)
    assert_rows_bitwise(g[:, :7], c[:, :7], "rfca matrix end %d%d %s" % (end1, end2, fid))
    ids = g[:ng, 6].astype(int) - 1  # rows were given with IDs 1..n in order
    assert np.nanmax(np.abs(g[:ng, 5] - r[ids, 5])) > 1e-5  # the cavity accelerated


@pytest.mark.parametrize("aperture", [False, True])
def test_csbend_zero_angle_is_the_exact_drift(tracker, aperture):
    """CSBEND with ANGLE = 0 is exactDrift (csbend.c:801-803, 3767-3790),
    through the element entry point, on rows of every kind"""
    b = E.CSBEND(L=0.7, ANGLE=0.0, N_KICKS=16).p.csbend
    ap = A.EH_MAXAMP(0.05, 0.04, 0, 2, 2, 0) if aperture else None
    r = csbend_edge_rows(1000, 13)
    g = r.copy()
    c = r.copy()
    ng = tracker.track_through_csbend(g, b, PO, 3.0, ap)
    ptr, keep = rows_ptr(c)
    nc = O.lib("emath").or_track_through_csbend(ptr, c.shape[0], C.byref(b), PO, 3.0,
                                                 C.byref(ap) if ap is not None else None)
    assert ng == nc == r.shape[0]
    assert_rows_bitwise(g[:, :7], c[:, :7], "csbend angle 0")


def test_ring_with_exact_drifts_and_matrix_cavity(tracker):
    """do_tracking over the storage ring with every drift an EDRIFT, one
    zero-angle CSBEND per cell and the RF cavity on the matrix method
    (L = 0.5, N_KICKS = 0, TMEAN fiducial): rows, order, losses and effort
    identical to the checker over three passes, with a lossy halo"""
    bl, Po = Lt.storage_ring(seed=4)
    els = []
    for e in bl.elements:
        if e.type == A.EH_T_DRIF:
            els.append(E.EDRIFT(L=e.p.drift.length))
        elif e.type == A.EH_T_RFCA:
            els.append(E.RFCA("RF", L=0.5, VOLT=6e6, FREQ=352e6, PHASE=180.0, N_KICKS=0))
        else:
            els.append(e)
            if e.type == A.EH_T_KOCT:
                els.append(E.CSBEND(L=0.05, ANGLE=0.0))
    ring = E.Beamline(els)
    rows = Lt.matched_beam(bl, 1200, 1e-9, 1e-11, 1e-3, 10e-12, 21, Po)
    rows[:400, 0] += 0.0148
    g = rows.copy()
    c = rows.copy()
    ng, eg = tracker.do_tracking(g, ring, Po, 3)
    o = O.Oracle("emath")
    nc = o.do_tracking(c, ring, Po, 3)
    assert ng == nc and eg == o.effort() and 0 < nc < rows.shape[0]
    assert_rows_identical(g, c, "ring with EDRIFT and matrix RFCA")


@pytest.mark.parametrize("kw", [
    dict(L=1.0, ANGLE=0.1, N_KICKS=8, ETILT=1e-3),
    dict(L=1.0, ANGLE=-0.1, N_KICKS=8, ETILT=-2e-3, TILT=0.3, E1=0.05, E2=0.02),
    dict(L=0.8, ANGLE=0.2, N_KICKS=4, ETILT=5e-4, TILT=1.5707963267948966, K1=0.4, HGAP=0.01, FINT=0.5),
    dict(L=1.0, ANGLE=0.15, N_KICKS=8, H1=0.3, H2=-0.2, E1=0.04, E2=0.04),
])
def test_csbend_error_tilt_and_pole_face_curvature(tracker, kw):
    """ETILT: the tilt + error tilt rotation and the exit offsets of
    computeEtiltCentroidOffset (csbend.c:1054-1075, 1359-1364), for both
    signs of the angle, with TILT (a vertical bend among them), edges and a
    gradient; H1/H2, which the first-order edges do not use: bit for bit
    against the checker on rows of every kind, signs of zeros included"""
    b = E.CSBEND(**kw).p.csbend
    r = csbend_edge_rows(1000, 15)
    g = r.copy()
    c = r.copy()
    ng = tracker.track_through_csbend(g, b, PO, 3.0, None)
    ptr, keep = rows_ptr(c)
    nc = O.lib("emath").or_track_through_csbend(ptr, c.shape[0], C.byref(b), PO, 3.0, None)
    assert_same_survivors(r, g, ng, c, nc)
    assert_rows_bitwise(g[:, :7], c[:, :7], "csbend %r" % kw)
    if "H1" in kw:
        plain = dict(kw)
        plain.pop("H1")
        plain.pop("H2")
        h = r.copy()
        tracker.track_through_csbend(h, E.CSBEND(**plain).p.csbend, PO, 3.0, None)
        assert_rows_bitwise(g[:, :7], h[:, :7], "H1/H2 change nothing at EDGE_ORDER=1")


@pytest.mark.parametrize("kw", [
    dict(L=1.0, ANGLE=0.08, K1=0.7, N_KICKS=16, E1=0.02, E2=0.04),
    dict(L=0.8, ANGLE=-0.785398163397, E1=-0.445, E2=-0.445, K2=0.1375, K3=100.0, HGAP=0.0225, FINT=0.41,
         N_KICKS=10),
    dict(L=0.5, ANGLE=0.05, K1=-0.3, K2=20.0, K3=-300.0, K4=5000.0, N_KICKS=8, FSE=1e-3, FSE_QUADRUPOLE=2e-3,
         FSE_DIPOLE=-1e-3, XKICK=1e-4, DX=1e-4, TILT=2e-3),
    dict(L=0.5, ANGLE=0.05, K2=20.0, NONLINEAR=0, N_KICKS=8),
    dict(L=0.5, ANGLE=0.05, XKICK=2e-4, N_KICKS=8),
])
def test_csbend_with_multipoles(tracker, kw):
    """CSBEND with normal multipoles K1..K4 and XKICK (the field expansion
    of csbend.c:100-141, 186-660 in the kick of csbend.c:1586-1602): the
    general bend code on rows of every kind, bit for bit against the
    checker, signs of zeros included"""
    b = E.CSBEND(**kw).p.csbend
    r = csbend_edge_rows(1000, 14)
    g = r.copy()
    c = r.copy()
    ng = tracker.track_through_csbend(g, b, PO, 3.0, None)
    ptr, keep = rows_ptr(c)
    nc = O.lib("emath").or_track_through_csbend(ptr, c.shape[0], C.byref(b), PO, 3.0, None)
    assert_same_survivors(r, g, ng, c, nc)
    assert_rows_bitwise(g[:, :7], c[:, :7], "csbend with multipoles %r" % kw)
