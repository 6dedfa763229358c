"""The CPU checker against closed forms, invariants and brute force.

Each test pins one restated routine with something a plausible restatement
mistake would violate: textbook maps, exact arithmetic, numpy brute force,
or the physics a convention fixes."""
import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest

from elegant_amd import _abi as A
from elegant_amd import elements as E
from elegant_amd import lattices as Lt
from oracle import oracle_ctypes as O

PO = Lt.momentum(6000.0)


def rows_of(pts):
    r = np.zeros((len(pts), 8))
    r[:, :6] = np.asarray(pts, dtype=np.float64)
    r[:, 6] = np.arange(1, len(pts) + 1)
    r[:, 7] = -1
    return r


def track(elems, rows, passes=1, mode="glibc", Po=PO):
    bl = E.Beamline(elems)
    o = O.Oracle(mode)
    r = rows.copy()
    n = o.do_tracking(r, bl, Po, passes)
    return r, n, o


def test_drift_matrix_second_order():
    """DRIF is the order-2 matrix: x += L x', s += L + L/2 (x'^2 + y'^2)."""
    L = 1.7
    pts = [[1e-3, 2e-3, -3e-4, 5e-4, 0.1, 1e-3], [0, 0, 0, 0, 0, 0], [-2e-3, -1e-3, 4e-4, 0, 0, -2e-3]]
    r, n, _ = track([E.DRIF(L=L)], rows_of(pts))
    for p, q in zip(pts, r):
        assert q[0] == L * p[1] + p[0] if p[1] else q[0] == p[0]
        exact_s = Fraction(p[4]) + Fraction(L) + Fraction(L / 2) * (Fraction(p[1]) ** 2 + Fraction(p[3]) ** 2)
        assert abs(float(exact_s) - q[4]) <= 2e-16 * max(1, abs(q[4]))
        assert q[1] == p[1] and q[3] == p[3] and q[5] == p[5]


def test_thick_quadrupole_linear_limit():
    """KQUAD with 4th-order integration converges to the cos/sin/cosh/sinh map
    with the Yoshida O(h^4) error; halving the slice length cuts the error ~16x."""
    k, L = 1.2, 0.5
    x0 = 1e-6
    errs = []
    for nk in (4, 8, 16):
        r, n, _ = track([E.KQUAD(L=L, K1=k, N_KICKS=nk)], rows_of([[x0, 0, x0, 0, 0, 0]]))
        s = math.sqrt(k)
        ex = x0 * math.cos(s * L)
        ey = x0 * math.cosh(s * L)
        errs.append((abs(r[0, 0] - ex), abs(r[0, 2] - ey)))
    assert errs[0][0] < 3e-3 * x0 and errs[0][1] < 3e-3 * x0
    assert 10 < errs[0][0] / errs[1][0] < 22 and 10 < errs[1][1] / errs[2][1] < 22


def test_multipole_kick_coefficients():
    """Thin-lens limit: sextupole kick dx' = -K2 L (x^2 - y^2)/2, dy' = K2 L x y;
    octupole dx' = -K3 L (x^3 - 3 x y^2)/6 (coefficients of multipole.h:137-165)."""
    L, K2, K3 = 1e-4, 40.0, 900.0
    x, y = 2e-3, -1e-3
    r, n, _ = track([E.KSEXT(L=L, K2=K2)], rows_of([[x, 0, y, 0, 0, 0]]))
    assert abs(r[0, 1] - (-K2 * L * (x * x - y * y) / 2)) < 1e-6 * abs(K2 * L * x * x)
    assert abs(r[0, 3] - (K2 * L * x * y)) < 1e-6 * abs(K2 * L * x * y)
    r, n, _ = track([E.KOCT(L=L, K3=K3)], rows_of([[x, 0, y, 0, 0, 0]]))
    assert abs(r[0, 1] - (-K3 * L * (x ** 3 - 3 * x * y * y) / 6)) < 1e-6 * abs(K3 * L * x ** 3)
    assert abs(r[0, 3] - (K3 * L * (3 * x * x * y - y ** 3) / 6)) < 1e-6 * abs(K3 * L * x ** 3)


def jacobian(elems, p0, h=1e-9, mode="glibc"):
    J = np.zeros((6, 6))
    for j in range(6):
        a = np.array(p0, dtype=float)
        b = np.array(p0, dtype=float)
        a[j] += h
        b[j] -= h
        ra, _, _ = track(elems, rows_of([a]), mode=mode)
        rb, _, _ = track(elems, rows_of([b]), mode=mode)
        J[:, j] = (ra[0, :6] - rb[0, :6]) / (2 * h)
    return J


def canonical(p):
    # (x, px, y, py, -s?, delta): use slopes -> momenta qx = (1+d) x'/sqrt(1+x'^2+y'^2)
    x, xp, y, yp, s, d = p
    den = math.sqrt(1 + xp * xp + yp * yp)
    return np.array([x, (1 + d) * xp / den, y, (1 + d) * yp / den])


@pytest.mark.parametrize("elem", [
    lambda: E.KQUAD(L=0.5, K1=1.3, N_KICKS=8),
    lambda: E.KSEXT(L=0.2, K2=40.0, N_KICKS=4),
    lambda: E.KOCT(L=0.2, K3=500.0),
    lambda: E.CSBEND(L=1.0, ANGLE=2 * math.pi / 80, N_KICKS=16, E1=0.02, E2=0.03),
])
def test_transverse_map_is_symplectic(elem):
    """The 4x4 transverse Jacobian in canonical variables (fixed delta)
    satisfies J^T S J = S (multipoles and the exact-sector bend body; the
    edge kicks of edge order 1 are linear and symplectic too)."""
    p0 = [1.5e-3, 2e-4, -8e-4, 1e-4, 0.0, 1e-3]
    h = 1e-8
    J = np.zeros((4, 4))
    base = canonical(p0)
    for j in range(4):
        pa = list(p0)
        pb = list(p0)
        # perturb canonical coordinate j, convert back to slopes
        def to_slopes(q):
            x, px, y, py = q
            d = p0[5]
            pz = math.sqrt((1 + d) ** 2 - px * px - py * py)
            return [x, px / pz, y, py / pz, 0.0, d]
        qa = base.copy()
        qb = base.copy()
        qa[j] += h
        qb[j] -= h
        ra, _, _ = track([elem()], rows_of([to_slopes(qa)]))
        rb, _, _ = track([elem()], rows_of([to_slopes(qb)]))
        J[:, j] = (canonical(ra[0, :6]) - canonical(rb[0, :6])) / (2 * h)
    S = np.array([[0, 1, 0, 0], [-1, 0, 0, 0], [0, 0, 0, 1], [0, 0, -1, 0]], dtype=float)
    assert np.abs(J.T @ S @ J - S).max() < 5e-6


def test_bend_reference_orbit_closes():
    """An on-momentum particle on axis stays on axis through a sector bend
    (no edge angles), to integrator accuracy."""
    r, n, _ = track([E.CSBEND(L=1.0, ANGLE=0.1, N_KICKS=16)], rows_of([[0, 0, 0, 0, 0, 0]]))
    assert abs(r[0, 0]) < 1e-10 and abs(r[0, 1]) < 1e-10
    assert abs(r[0, 4] - 1.0) < 1e-9


def test_add_to_particle_energy_conserves_transverse_momentum():
    L = O.lib("glibc")
    p = (C.c_double * 8)(1e-3, 2e-4, -1e-3, 3e-4, 0.5, 1e-3, 1, -1)
    P0 = PO * (1 + p[5])
    px0 = P0 * p[1] / math.sqrt(1 + p[1] ** 2 + p[3] ** 2)
    g0 = math.sqrt(P0 * P0 + 1)
    dg = 37.5
    L.or_add_to_particle_energy(p, 1.2e-9, PO, dg)
    P1 = PO * (1 + p[5])
    px1 = P1 * p[1] / math.sqrt(1 + p[1] ** 2 + p[3] ** 2)
    assert abs(math.sqrt(P1 * P1 + 1) - (g0 + dg)) < 1e-9
    assert abs(px1 - px0) < 1e-12 * abs(px0)


def test_rfca_zero_length_energy_gain():
    """RFCA with L=0: delta' = (sqrt((gamma + dgamma)^2 - 1) - Po)/Po with
    dgamma = V sin(phase) / (me c^2) for a particle at the fiducial time."""
    V, phase = 5e6, 30.0
    rf = E.RFCA(L=0.0, VOLT=V, PHASE=phase, FREQ=352e6)
    r, n, _ = track([rf], rows_of([[0, 0, 0, 0, 0, 0]]))
    g = math.sqrt(PO * PO + 1) + V / (1e6 * Lt.ME_MEV) * math.sin(math.radians(phase))
    assert abs(r[0, 5] - (math.sqrt(g * g - 1) - PO) / PO) < 1e-14


def _rf_point_time(pt, L, tref, Po=PO):
    """arrival time at the cavity centre, s = beta c t (simple_rfca.c:848-850)"""
    P = Po * (1 + pt[5])
    beta = P / math.sqrt(P * P + 1)
    return (pt[4] + L / 2 * math.sqrt(1 + pt[1] ** 2 + pt[3] ** 2)) / (Lt.C_MKS * beta) - tref


@pytest.mark.parametrize("mode", ["glibc", "emath"])
def test_rfca_matrix_method_is_an_accelerating_drift(mode):
    """RFCA with L > 0 and N_KICKS = 0 takes the matrix method
    (simple_rfca.c:403-407, 824-941): the gain dgamma = V sin(phi) is spread
    evenly along the cavity and there is no transverse force.  So P x' is
    conserved while P rises linearly from P to P1 = sqrt((gamma+dgamma)^2 - 1):
    x' leaves as x' P/P1, x advances by x' * integral_0^L P/P(z) dz (done here
    by numerical quadrature at 40 digits), delta = (P1 - Po)/Po, and s = beta c t keeps the
    arrival time while the path (half the cavity at each slope) is covered."""
    import mpmath
    mpmath.mp.dps = 40
    V, L, freq, tref, ph = 4e6, 0.6, 352e6, 1.7e-9, 40.0
    rf = E.RFCA(L=L, VOLT=V, PHASE=ph, FREQ=freq, T_REFERENCE=tref, N_KICKS=0)
    pts = [[1e-3, 2e-4, -5e-4, -1e-4, 0.2, 1e-3], [-2e-3, -3e-4, 1e-3, 2e-4, 0.5, -2e-3],
           [0.0, 1e-3, 0.0, -7e-4, 0.01, 0.0]]
    r, n, _ = track([rf], rows_of(pts), mode=mode)
    assert n == len(pts)
    for pt, out in zip(pts, r):
        P = PO * (1 + pt[5])
        g = math.sqrt(P * P + 1)
        phi = 2 * math.pi * freq * _rf_point_time(pt, L, tref) + math.radians(ph)
        dg = V / (1e6 * Lt.ME_MEV) * math.sin(phi)
        P1 = math.sqrt((g + dg) ** 2 - 1)
        assert abs(P1 - P) > 1e-6 * P  # a real acceleration, not the L*x' limit
        reach = float(mpmath.quad(lambda z: P / (P + (P1 - P) * z / L), [0, L]))
        # the reference forms log(1 + u), u = dP/P, from the rounded 1 + u: a
        # relative error up to 2^-53/|u| in the reach, against O(u) for a
        # wrong formula (P/P1 for P/dP, a missing log, ...)
        u = (P1 - P) / P
        tol = 2.0 ** -52 / abs(u) + 1e-15
        for c in (0, 2):
            if pt[c + 1]:
                assert abs((out[c] - pt[c]) / pt[c + 1] - reach) <= tol * reach
            else:
                assert out[c] == pt[c]
            assert abs(out[c + 1] - pt[c + 1] * P / P1) <= 1e-15 * abs(pt[c + 1])
        assert abs(out[5] - (P1 - PO) / PO) <= 2e-16
        path = L / 2 * (math.sqrt(1 + pt[1] ** 2 + pt[3] ** 2) +
                        math.sqrt(1 + (pt[1] * P / P1) ** 2 + (pt[3] * P / P1) ** 2))
        beta0, beta1 = P / g, P1 / math.sqrt(P1 * P1 + 1)
        assert abs(out[4] - (pt[4] + path) * beta1 / beta0) <= 1e-14 * abs(out[4])


def test_rfca_matrix_method_at_vanishing_voltage_is_a_drift():
    """With |dP/P| <= 1e-14 the matrix method takes R12 = L (simple_rfca.c:
    921-924): a cavity with a vanishing voltage moves x and y exactly as the
    exact drift does, and leaves slopes, delta and s unchanged to rounding
    (delta to the rounding of P = Po(1 + delta) it is re-formed from)."""
    pts = [[1e-3, 2e-4, -5e-4, -1e-4, 0.2, 1e-3], [-2e-3, -3e-4, 1e-3, 2e-4, 0.5, -2e-3]]
    L = 0.8
    rf = E.RFCA(L=L, VOLT=1e-9, PHASE=90.0, FREQ=352e6, T_REFERENCE=0.0, N_KICKS=0)
    r, n, _ = track([rf], rows_of(pts))
    d, nd, _ = track([E.EDRIFT(L=L)], rows_of(pts))
    assert n == nd == 2
    assert (r[:, [0, 2]] == d[:, [0, 2]]).all()
    assert np.allclose(r[:, [1, 3]], d[:, [1, 3]], rtol=1e-15, atol=0)
    assert np.allclose(r[:, 4], d[:, 4], rtol=1e-15, atol=0)
    # delta is re-formed as (P + dP - Po)/Po: a rounding of P, not of delta
    assert np.abs(r[:, 5] - d[:, 5]).max() <= 2 * np.spacing(PO) / PO


def test_exact_drift_path_length():
    """exactDrift (csbend.c:3767-3790), for EDRIFT and for a CSBEND with
    ANGLE = 0 (csbend.c:801-803): x and y advance by the slopes times L and s
    by the length of the straight path, sqrt(L^2 + (x'L)^2 + (y'L)^2),
    checked here in exact rational arithmetic to within an ulp."""
    from decimal import Decimal, getcontext
    getcontext().prec = 60
    pts = [[1e-3, 2e-4, -5e-4, -1e-4, 0.2, 1e-3], [0.0, 0.3, 0.0, -0.2, 1.5, 0.0],
           [-1e-2, 1e-7, 2e-3, 0.0, 0.0, -2e-2]]
    L = 1.3
    for elems in ([E.EDRIFT(L=L)], [E.CSBEND(L=L, ANGLE=0.0)]):
        r, n, _ = track(elems, rows_of(pts))
        assert n == len(pts)
        for pt, out in zip(pts, r):
            assert out[0] == pt[0] + pt[1] * L and out[2] == pt[2] + pt[3] * L
            dL = Decimal(L)
            exact = (dL * dL + (Decimal(pt[1]) * dL) ** 2 + (Decimal(pt[3]) * dL) ** 2).sqrt()
            assert abs(Decimal(out[4]) - (Decimal(pt[4]) + exact)) <= Decimal(np.spacing(out[4])) * 2
            assert out[1] == pt[1] and out[3] == pt[3] and out[5] == pt[5]
    e, _, _ = track([E.EDRIFT(L=L)], rows_of(pts))
    b, _, _ = track([E.CSBEND(L=L, ANGLE=0.0)], rows_of(pts))
    assert (e.view(np.int64) == b.view(np.int64)).all()


def test_kahan_plus_against_exact_sum():
    L = O.lib("glibc")
    rng = np.random.default_rng(4)
    v = (rng.uniform(1, 2, 20000) * 1e-6).tolist()
    s, err = 0.0, C.c_double(0.0)
    for x in v:
        s = L.or_kahan_plus(s, x, C.byref(err))
    exact = sum(Fraction(x) for x in v)
    naive = 0.0
    for x in v:
        naive += x
    assert abs(Fraction(s) - exact) <= abs(Fraction(naive) - exact)
    assert abs(float(Fraction(s) - exact)) <= np.spacing(s)


def test_kahan_plus_carries_the_error_into_the_next_addend():
    """KahanPlus (kahan.c:27-36) adds the running error to the next addend
    *before* that addend meets the sum: Y = b + error, sum = oldSum + Y,
    error = Y - (sum - oldSum).  On [-5/512, -5/32, 1.25*2^-56, -1/16] this
    rounds the carried -0.75*2^-56 into Y = -1/16 - 2^-56, and the last sum
    -117/512 + 2^-56 is a tie that goes to the even -117/512: the result is
    -117/512, one ulp from the correctly rounded sum -117/512 + 2^-55, with
    the dropped 2^-56 left in the error.
    Plausible restatement mistakes give other results on the same sequence:
    plain summation, the error applied once at the end (Neumaier's or a
    separate running correction) and the carry with its sign flipped."""
    L = O.lib("glibc")
    v = [-5 / 512, -5 / 32, 1.25 * 2.0 ** -56, -1 / 16]
    s, err = 0.0, C.c_double(0.0)
    for x in v:
        s = L.or_kahan_plus(s, x, C.byref(err))
    assert s == -117 / 512
    # error = Y - (sum - oldSum) = (-1/16 - 2^-56) - (-1/16 - 2^-55): the 2^-56 the tie dropped
    assert err.value == 2.0 ** -56
    exact = float(sum(Fraction(x) for x in v))
    assert exact == -117 / 512 + 2.0 ** -55

    def naive(v):
        t = 0.0
        for x in v:
            t += x
        return t

    def end_corrected(v):  # errors summed apart and added once at the end
        t, c = 0.0, 0.0
        for x in v:
            u = t + x
            c += x - (u - t)
            t = u
        return t + c

    def neumaier(v):
        t, c = 0.0, 0.0
        for x in v:
            u = t + x
            c += ((t - u) + x) if abs(t) >= abs(x) else ((x - u) + t)
            t = u
        return t + c

    def sign_flipped(v):
        t, e = 0.0, 0.0
        for x in v:
            y = x - e
            u = t + y
            e = y - (u - t)
            t = u
        return t

    for alt in (naive, end_corrected, neumaier, sign_flipped):
        assert alt(v) != s, alt.__name__


def test_fixed_point_mean_is_exact_and_order_free():
    rng = np.random.default_rng(8)
    v = 1e-6 + rng.normal(0, 1e-11, 10001)
    m = O.compute_average(v)
    exact = sum(Fraction(x) for x in v) / len(v)
    assert abs(Fraction(m) - exact) <= Fraction(np.spacing(m)) / 2 + Fraction(np.spacing(m)) / 2
    assert O.compute_average(v[::-1]) == m
    assert O.compute_average(rng.permutation(v)) == m
    assert O.compute_average(np.array([-1.5, 2.5, 3.0])) == float(Fraction(4) / 3)


def test_real_fft_matches_dft():
    rng = np.random.default_rng(1)
    for n in (16, 32, 1024):
        x = rng.normal(size=n)
        X = O.real_fft(x)
        ref = np.fft.rfft(x) / n
        packed = np.empty(n)
        packed[0] = ref[0].real
        packed[1:n - 1:2] = ref[1:n // 2].real
        packed[2:n - 1:2] = ref[1:n // 2].imag
        packed[n - 1] = ref[n // 2].real
        assert np.abs(X - packed).max() < 1e-14 * np.abs(packed).max()
        back = O.real_fft(X, inverse=True)
        assert np.abs(back - x).max() < 1e-12
    # direct O(n^2) DFT for n = 16
    n = 16
    x = rng.normal(size=n)
    j = np.arange(n)
    X = O.real_fft(x)
    for k in range(1, n // 2):
        ref = sum(x[jj] * np.exp(-2j * np.pi * jj * k / n) for jj in range(n)) / n
        assert abs(X[2 * k - 1] - ref.real) < 1e-15 and abs(X[2 * k] - ref.imag) < 1e-15


def wake_bunch(n, sigma_t, seed=3):
    rng = Lt.Rng(seed)
    beta = PO / math.sqrt(PO * PO + 1)
    z = rng.normal(2 * n).reshape(2, n)
    r = np.zeros((n, 8))
    r[:, 4] = 0.3 + 2.99792458e8 * beta * sigma_t * z[0]
    r[:, 5] = 1e-3 * z[1]
    r[:, 6] = np.arange(1, n + 1)
    r[:, 7] = -1
    return r


def test_wake_histogram_and_convolution_brute_force():
    table = Lt.resonator_wake(points=4000)
    w = E.WAKE(table)
    rows = wake_bunch(20000, 20e-12)
    o = O.Oracle()
    o.set_macro_charge(5e-9 / 20000)
    before = rows.copy()
    o.wake(rows, w.p.wake, PO, 0)
    hist, volt, tmin, dt = o.last_hist()
    P = PO * (before[:, 5] + 1)
    t = before[:, 4] * np.sqrt(P * P + 1) / (2.99792458e8 * P)
    ib = ((t - tmin) / dt + 0.5).astype(np.int64)
    assert hist.sum() == rows.shape[0]
    assert np.array_equal(np.bincount(ib, minlength=hist.size)[:hist.size], hist)
    conv = np.convolve(hist, table.W[:hist.size])[:hist.size] * (5e-9 / 20000)
    assert np.abs(conv - volt).max() < 1e-10 * np.abs(volt).max()


def test_wake_and_zlongit_agree_for_the_same_resonator():
    """The time-domain WAKE with the resonator's Green function and the
    frequency-domain ZLONGIT with its impedance give the same energy kicks
    (pins the FFT normalization and sign conventions)."""
    n = 200000
    rows = wake_bunch(n, 20e-12, seed=9)
    q = 5e-9 / n
    ow = O.Oracle()
    ow.set_macro_charge(q)
    rw = rows.copy()
    table = Lt.resonator_wake(dt=0.2e-12, points=5000)
    ow.wake(rw, E.WAKE(table).p.wake, PO, 0)
    oz = O.Oracle()
    oz.set_macro_charge(q)
    rz = rows.copy()
    zl = E.ZLONGIT(BROAD_BAND=1, RS=1e4, Q=1.0, FREQ=20e9, N_BINS=8192, BIN_SIZE=0.2e-12)
    oz.zlongit(rz, zl.p.zlongit, PO, 0)
    dw = rw[:, 5] - rows[:, 5]
    dz = rz[:, 5] - rows[:, 5]
    assert np.corrcoef(dw, dz)[0, 1] > 0.99
    assert abs(dw.mean() - dz.mean()) < 0.02 * abs(dw.mean())
    assert dw.mean() < 0  # the wake decelerates the bunch


def test_glibc_and_emath_checkers_diverge_like_the_gpu():
    """The CPU checker built with glibc and with the shared deterministic math
    differ after one turn of the ring by the same ~1e-10 the GPU shows: the
    bend's X = rho0 (f - 1) + f X amplifies last-bit cos differences by rho0/x."""
    bl, Po, rows = Lt.storage_ring_workload(n=200, seed=9)
    a = rows.copy()
    b = rows.copy()
    O.Oracle("glibc").do_tracking(a, bl, Po, 1)
    O.Oracle("emath").do_tracking(b, bl, Po, 1)
    rel = np.abs(a[:, 0] - b[:, 0]) / np.maximum(np.abs(b[:, 0]), 1e-3)
    assert rel.max() < 5e-10
    assert rel.max() > 0


def jacobian6(elems, h=1e-7, mode="glibc"):
    """central-difference 6x6 Jacobian at the origin in (x, x', y, y', s, delta)"""
    J = np.zeros((6, 6))
    for j in range(6):
        a = np.zeros(6)
        b = np.zeros(6)
        a[j] = h
        b[j] = -h
        ra, na, _ = track(elems, rows_of([a]), mode=mode)
        rb, nb, _ = track(elems, rows_of([b]), mode=mode)
        assert na == nb == 1
        J[:, j] = (ra[0, :6] - rb[0, :6]) / (2 * h)
    return J


def test_sector_bend_linear_map():
    """Pure-dipole CSBEND without edges: the textbook sector-bend matrix
    (x block cos/rho sin, dispersion rho(1-cos), sin; path length
    ds = sin(t) x + rho (1 - cos t) x'; y block a drift)."""
    L, th = 1.0, 0.1
    rho = L / th
    J = jacobian6([E.CSBEND(L=L, ANGLE=th, N_KICKS=16, EDGE1_EFFECTS=0, EDGE2_EFFECTS=0)])
    c, s = math.cos(th), math.sin(th)
    want = {(0, 0): c, (0, 1): rho * s, (1, 0): -s / rho, (1, 1): c, (0, 5): rho * (1 - c), (1, 5): s,
            (2, 2): 1.0, (2, 3): L, (3, 3): 1.0, (4, 0): s, (4, 1): rho * (1 - c)}
    for (i, j), v in want.items():
        assert abs(J[i, j] - v) < 1e-6 * max(1.0, abs(v)), (i, j, J[i, j], v)
    assert abs(J[0, 2]) < 1e-9 and abs(J[2, 0]) < 1e-9


def test_rectangular_bend_edge_focusing():
    """Edge angles E1 = E2 = angle/2 with first-order edges (no gap): the map
    is Edge * Sector * Edge with horizontal edge kick +tan(e)/rho x and
    vertical -tan(e)/rho y."""
    L, th = 1.0, 0.1
    rho = L / th
    e = th / 2
    J = jacobian6([E.CSBEND(L=L, ANGLE=th, N_KICKS=16, E1=e, E2=e)])
    c, s = math.cos(th), math.sin(th)
    Ex = np.array([[1, 0], [math.tan(e) / rho, 1]])
    Ey = np.array([[1, 0], [-math.tan(e) / rho, 1]])
    Mx = Ex @ np.array([[c, rho * s], [-s / rho, c]]) @ Ex
    My = Ey @ np.array([[1, L], [0, 1]]) @ Ey
    assert np.abs(J[0:2, 0:2] - Mx).max() < 1e-6 * rho
    assert np.abs(J[2:4, 2:4] - My).max() < 1e-6 * rho
    # the rectangular bend focuses vertically: |trace| < 2 would not hold for a drift with these kicks
    assert J[3, 2] < 0


@pytest.mark.parametrize("ctor,key,k,tilt", [
    (E.KQUAD, "K1", 1.2, math.pi / 2),
    (E.KSEXT, "K2", 40.0, math.pi / 3),
    (E.KOCT, "K3", 800.0, math.pi / 4),
])
def test_roll_by_symmetry_angle_flips_the_multipole(ctor, key, k, tilt):
    """A 2n-pole rolled by pi/n is the same magnet with the opposite sign
    (pins the roll direction and the rotation in/out of the element frame)."""
    pts = [[1.1e-3, 2e-4, -7e-4, -1e-4, 0, 5e-4], [-2e-3, 0, 1e-3, 3e-4, 0, -1e-3]]
    ra, _, _ = track([ctor(L=0.3, TILT=tilt, N_KICKS=4, **{key: k})], rows_of(pts))
    rb, _, _ = track([ctor(L=0.3, N_KICKS=4, **{key: -k})], rows_of(pts))
    assert np.abs(ra[:, :4] - rb[:, :4]).max() < 1e-15
    assert np.abs(ra[:, 4] - rb[:, 4]).max() < 1e-14


def test_offset_quadrupole_is_a_translation():
    """DX/DY move the magnet: a particle on the displaced axis feels no
    field, and any orbit is the centred magnet's orbit shifted by (DX, DY)."""
    dx, dy = 1e-3, -5e-4
    q = dict(L=0.5, K1=1.4, N_KICKS=8)
    r, _, _ = track([E.KQUAD(DX=dx, DY=dy, **q)], rows_of([[dx, 0, dy, 0, 0, 0]]))
    assert r[0, 1] == 0 and r[0, 3] == 0
    assert abs(r[0, 0] - dx) < 1e-18 and abs(r[0, 2] - dy) < 1e-18
    pts = [[2e-4, 1e-4, 3e-4, -2e-4, 0, 1e-3]]
    shifted = [[p[0] + dx, p[1], p[2] + dy, p[3], p[4], p[5]] for p in pts]
    ra, _, _ = track([E.KQUAD(DX=dx, DY=dy, **q)], rows_of(shifted))
    rb, _, _ = track([E.KQUAD(**q)], rows_of(pts))
    assert abs((ra[0, 0] - dx) - rb[0, 0]) < 1e-17 and abs((ra[0, 2] - dy) - rb[0, 2]) < 1e-17
    assert abs(ra[0, 1] - rb[0, 1]) < 1e-16 and abs(ra[0, 3] - rb[0, 3]) < 1e-16


def test_drift_aperture_loss_point_is_extrapolated():
    """A particle leaving a drift beyond X_MAX is lost where its straight
    line crossed the aperture (limit_amplitudes.c:232-268); a particle
    exactly on the aperture survives (the test is strict)."""
    els = [E.MAXAMP(X_MAX=0.02, Y_MAX=0.01), E.DRIF(L=2.0)]
    r, n, _ = track(els, rows_of([[0.02, 0, 0, 0, 0, 0], [0.019, 0.001, 0, 0, 0, 2e-3],
                                  [0, 0, 0.009, 0.002, 0, 0]]))
    assert n == 1 and r[0, 0] == 0.02
    lost = {int(row[6]): row for row in r[1:]}
    a = lost[2]
    assert abs(a[4] - 1.0) < 1e-12 and abs(a[0] - 0.02) < 1e-15
    assert a[5] == PO * (1 + 2e-3)
    b = lost[3]
    assert abs(b[4] - 0.5) < 1e-12 and abs(b[2] - 0.01) < 1e-15


def test_initial_nan_cull_precedes_the_first_element():
    """do_tracking culls NaN/infinite x or y once before the pass loop, at
    z = 0, recording pass 0 (do_tracking.c:398, 468-473): such rows are
    lost before, and swapped to the tail ahead of, the losses of the first
    element's aperture; delta becomes Po(1 + delta), s = 0"""
    els = [E.MAXAMP(X_MAX=0.02, Y_MAX=0.01), E.DRIF(L=1.0)]
    r, n, _ = track(els, rows_of([[0.03, 0, 0, 0, 0, 0], [float("nan"), 0, 0, 0, 0.7, 1e-3],
                                  [0, 0, float("-inf"), 0, 0, 0], [0.001, 0, 0, 0, 0, 0]]))
    assert n == 1 and r[0, 6] == 4
    # swap-to-tail: the first cull loses id 2 (to the last row), then id 3;
    # the aperture after MAXAMP then loses id 1
    assert [int(x) for x in r[1:, 6]] == [1, 3, 2]
    b = r[3]
    assert b[7] == 0 and b[4] == 0.0 and b[5] == PO * (1 + 1e-3)


# ---------------------------------------------------------------- MAXAMP shapes

def ipow_listed(x, p):
    """x^p with the multiplications the reference's device copy of the
    toolkit's ipow lists (gpuElegant/gpu_bin_transverse_distribution.cu:8-56)"""
    x = np.float64(x)
    if x == 0:
        return np.float64(0.0)
    if p <= 8:
        h = x * x
        return {1: x, 2: h, 3: h * x, 4: h * h, 5: h * h * x, 6: h * h * h, 7: (h * x) * (h * x) * x,
                8: (h * h) * (h * h)}[p]
    hp = ipow_listed(x, p // 2)
    return hp * hp * x if p % 2 else hp * hp


def test_ipow_follows_the_listed_multiplication_order():
    """The aperture power keeps the toolkit's multiplication order, which
    differs in the last bits from plain repeated multiplication and from
    numpy's power, and stays within (p - 1) half-ulps of the exact power."""
    rng = np.random.default_rng(8)
    x = rng.uniform(-3, 3, 4000)
    exps = rng.integers(1, 80, x.size)
    got = O.math_eval(7, x, exps.astype(np.float64), mode="glibc")
    want = np.array([ipow_listed(a, int(p)) for a, p in zip(x, exps)])
    assert (got.view(np.int64) == want.view(np.int64)).all()
    naive = np.array([np.prod(np.full(int(p), a)) for a, p in zip(x, exps)])
    assert (got != naive).any()
    for a, p, g in list(zip(x, exps, got))[:300]:
        exact = Fraction(float(a)) ** int(p)
        assert abs(Fraction(float(g)) - exact) <= abs(exact) * Fraction(int(p), 2 ** 53)
    # integers whose powers are exact in double precision come out exact
    assert O.math_eval(7, np.array([3.0, -2.0, 0.0, -0.0]), np.array([20.0, 31.0, 4.0, 2.0]), mode="glibc").tolist() \
        == [3.0 ** 20, -(2.0 ** 31), 0.0, 0.0]


def elimit(rows, xmax, ymax, z, extrap, open_code=0, xe=2, ye=0):
    r = rows.copy()
    ptr, keep = A.rows_pointer(r)
    n = O.lib("glibc").or_elimit_amplitudes(ptr, xmax, ymax, r.shape[0], z, PO, extrap, open_code, xe, ye)
    return r, n


def test_elliptical_aperture_against_exact_arithmetic():
    """elimit_amplitudes (limit_amplitudes.c:472-598) loses exactly the
    particles with (x/a)^p + (y/b)^q > 1 in exact rational arithmetic (points
    within 1e-9 of the boundary left out), for true and super-ellipses; lost
    rows are swapped to the tail with s = z and delta replaced by
    sqrt((Po(1 + delta))^2 + 1)"""
    rng = np.random.default_rng(12)
    a, b = 0.013, 0.007
    for xe, ye in ((2, 0), (4, 0), (6, 2), (10, 40), (40, 40)):
        q = ye or xe
        pts = np.zeros((3000, 6))
        pts[:, 0] = rng.uniform(-1.6 * a, 1.6 * a, 3000)
        pts[:, 2] = rng.uniform(-1.6 * b, 1.6 * b, 3000)
        pts[:, 5] = rng.normal(0, 1e-3, 3000)
        lhs = [(Fraction(float(x)) / Fraction(a)) ** xe + (Fraction(float(y)) / Fraction(b)) ** q
               for x, y in zip(pts[:, 0], pts[:, 2])]
        keep = np.array([abs(v - 1) > Fraction(1, 10 ** 9) for v in lhs])
        pts, lhs = pts[keep], [v for v, k in zip(lhs, keep) if k]
        rows = rows_of(pts)
        r, n = elimit(rows, a, b, 4.5, 0, 0, xe, ye)
        lost_ids = {i + 1 for i, v in enumerate(lhs) if v > 1}
        assert set(r[n:, 6].astype(int)) == lost_ids and n == len(lhs) - len(lost_ids)
        for row in r[n:]:
            P = float(PO) * (1 + float(rows[int(row[6]) - 1, 5]))
            assert row[4] == 4.5 and row[5] == math.sqrt(P * P + 1)
    # a point exactly on the ellipse survives the check after an element (> 1)
    r, n = elimit(rows_of([[0.013, 0, 0, 0, 0, 0], [0, 0, -0.007, 0, 0, 0]]), a, b, 1.0, 0)
    assert n == 2
    # the powers come before the quotients: with b^254 below the smallest
    # double, y^254 / b^254 is 0/0 = NaN and the comparison keeps the particle
    r, n = elimit(rows_of([[0.02, 0, 0, 0, 0, 0]]), a, b, 1.0, 0, 0, 10, 254)
    assert n == 1


def test_elliptical_loss_point_on_a_circle():
    """With a true ellipse after a drift the loss point is moved back along
    the straight line to the last crossing of the boundary
    (limit_amplitudes.c:564-575): on a circle of radius R, x0 = 1.5 R moving
    outward with slope 1 was last at the boundary 0.5 R earlier"""
    R = 0.01
    pts = [[1.5 * R, 1.0, 0, 0, 0, 0], [0, 0, -1.2 * R, -0.5, 0, 0], [0.6 * R, 0.0, 0.9 * R, 0.0, 0, 0]]
    r, n = elimit(rows_of(pts), R, R, 2.0, 1)
    assert n == 0
    lost = {int(row[6]): row for row in r}
    assert abs(lost[1][4] - (2.0 - 0.5 * R)) < 1e-15 and abs(lost[1][0] - R) < 1e-17
    assert abs(lost[2][4] - (2.0 - 0.4 * R)) < 1e-15 and abs(lost[2][2] + R) < 1e-17
    # no slope: no crossing behind it, the particle is lost where it stands
    assert lost[3][4] == 2.0 and lost[3][0] == 0.6 * R
    # no extrapolation for super-ellipses or without the drift flag
    r, n = elimit(rows_of(pts), R, R, 2.0, 1, 0, 4)
    assert n == 1 and (r[1:, 4] == 2.0).all()
    r, n = elimit(rows_of(pts), R, R, 2.0, 0)
    assert (r[:, 4] == 2.0).all()


def test_open_side_spares_particles_leaving_through_it():
    """evaluateLostWithOpenSides (limit_amplitudes.c:1105-1128): with +x open
    a particle beyond X_MAX on the + side within Y_MAX is kept, one on the
    - side or beyond Y_MAX is lost; the same for the ellipse"""
    lib = O.lib("glibc")
    pts = [[0.03, 0, 0.001, 0, 0, 0], [-0.03, 0, 0.001, 0, 0, 0], [0.03, 0, 0.02, 0, 0, 0],
           [0.001, 0, 0.02, 0, 0, 0], [0.001, 0, -0.002, 0, 0, 0]]
    for code, kept in ((A.EH_OPEN_PLUS_X, {1, 5}), (A.EH_OPEN_MINUS_X, {2, 5}), (A.EH_OPEN_PLUS_Y, {4, 5}),
                       (A.EH_OPEN_MINUS_Y, {5}), (0, {5})):
        r = rows_of(pts)
        ptr, keep = A.rows_pointer(r)
        n = lib.or_limit_amplitudes(ptr, 0.02, 0.01, r.shape[0], 1.0, PO, 1, code)
        assert set(r[:n, 6].astype(int)) == kept, code
        if code:
            # no loss-point extrapolation with an open side
            assert (r[n:, 4] == 1.0).all()
        else:
            # closed: with zero slopes both candidate distances stay -DBL_MAX,
            # dz = +DBL_MAX, and the sentinel test (dz == -DBL_MAX,
            # limit_amplitudes.c:246-247) lets it through: s = z + DBL_MAX
            assert (r[n:, 4] == 1.0 + np.finfo(np.float64).max).all()
        r2, n2 = elimit(rows_of(pts), 0.02, 0.01, 1.0, 1, code)
        assert set(r2[:n2, 6].astype(int)) == kept, code


def test_aperture_early_returns():
    """limit_amplitudes with both half-widths negative checks nothing, not
    even NaN (limit_amplitudes.c:215-217); elimit_amplitudes with one
    non-positive half-axis falls back to the rectangle in the other plane,
    with none positive checks nothing (:523-528); odd or small exponents
    are an error (:507-521)"""
    lib = O.lib("glibc")
    pts = [[float("nan"), 0, 0, 0, 0, 0], [0.05, 0, 0.05, 0, 0, 0]]
    r = rows_of(pts)
    ptr, keep = A.rows_pointer(r)
    assert lib.or_limit_amplitudes(ptr, -1.0, -1.0, 2, 1.0, PO, 0, 0) == 2
    r, n = elimit(rows_of(pts), 0.0, 0.0, 1.0, 0)
    assert n == 2
    r, n = elimit(rows_of(pts), 0.02, 0.0, 1.0, 0)
    assert n == 0 and r[0, 5] == PO * (1 + 0.0)  # rectangle's loss rows: Po(1 + delta)
    for xe, ye in ((3, 0), (0, 0), (2, 5), (2, 1)):
        r, n = elimit(rows_of(pts), 0.02, 0.02, 1.0, 0, 0, xe, ye)
        assert n == -1  # EH_ERR_INVALID


def test_elliptical_aperture_inside_an_element_is_not_strict():
    """Inside an integrator the ellipse test is >= 1 (multipole.c:1817-1822)
    while the check after the element is > 1: a particle travelling exactly
    on the boundary through a field-free KQUAD is lost at the first slice's
    check (multipole.c:1240-1248), at the element's entrance"""
    a, b = 0.01, 0.005
    els = [E.MAXAMP(X_MAX=a, Y_MAX=b, ELLIPTICAL=1), E.KQUAD(L=0.5, K1=0.0, N_KICKS=4)]
    r, n, _ = track(els, rows_of([[a, 0, 0, 0, 0, 0], [0.99 * a, 0, 0, 0, 0, 0]]))
    assert n == 1 and r[0, 6] == 2
    assert r[1, 6] == 1 and r[1, 4] == 0.0 and r[1, 0] == a
    # the same particle survives a drift (only the strict check after it)
    r, n, _ = track([E.MAXAMP(X_MAX=a, Y_MAX=b, ELLIPTICAL=1), E.DRIF(L=0.5)],
                    rows_of([[a, 0, 0, 0, 0, 0]]))
    assert n == 1


def test_open_side_inside_an_element():
    """checkMultAperture spares a particle beyond X_MAX on the open side
    within Y_MAX (multipole.c:1811-1815); on the closed side it is lost at
    the element's entrance check; the same holds for the ellipse"""
    for ell in (0, 1):
        els = [E.MAXAMP(X_MAX=0.01, Y_MAX=0.005, ELLIPTICAL=ell, OPEN_SIDE="+x"),
               E.KQUAD(L=0.5, K1=0.0, N_KICKS=4)]
        r, n, _ = track(els, rows_of([[0.012, 0, 0.001, 0, 0, 0], [-0.012, 0, 0.001, 0, 0, 0],
                                      [0.012, 0, 0.006, 0, 0, 0]]))
        assert n == 1 and r[0, 6] == 1 and r[0, 0] == 0.012, ell
        lost = {int(row[6]): row for row in r[1:]}
        assert lost[2][4] == 0.0 and lost[3][4] == 0.0, ell


# ------------------------------------------------ CSBEND multipole fields

def _field_table():
    """the committed eh_csbend_fields.h as exact polynomials:
    {(field, i, j): {(p, m): Fraction}}"""
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = open(os.path.join(root, "elegant_amd", "csrc", "eh_csbend_fields.h")).read()
    out = {}
    for f, i, j, p, m, n, d in re.findall(r"\{(\d), (\d+), (\d+), (\d+), (\d+), \{0\}, (-?\d+)\.0, (\d+)\.0\}", text):
        out.setdefault((int(f), int(i), int(j)), {})[(int(p), int(m))] = Fraction(int(n), int(d))
    return out


def test_bend_field_table_satisfies_maxwell():
    """the tabulated field of a curved sector magnet (By = sum a_ij x^i y^j,
    Bx = sum c_ij x^i y^j, midplane By = 1 + sum b_n x^n/n!) is curl-free,
    dBx/dy = dBy/dx, and divergence-free in the bend's frame,
    (1 + hx) dBx/dx + h Bx + (1 + hx) dBy/dy = 0: every coefficient of both
    residuals through total degree 9, in exact arithmetic, for arbitrary
    rational h and b_1..b_8 (a dropped term or a wrong sign fails)"""
    tab = _field_table()
    assert len(tab) > 50
    h = Fraction(3, 7)
    b = [Fraction(0)] + [Fraction((-1) ** k * (k + 2), 5 + k) for k in range(1, 9)]

    def val(field, i, j):
        if i < 0 or j < 0:
            return Fraction(0)
        if field == 0 and (i, j) == (1, 0):
            return b[1]
        if field == 1 and (i, j) == (0, 1):
            return b[1]
        poly = tab.get((field, i, j), {})
        return sum((v * h ** p * b[m] for (p, m), v in poly.items()), Fraction(0))

    a = lambda i, j: val(0, i, j)
    c = lambda i, j: val(1, i, j)
    for i in range(0, 9):
        assert a(i, 0) == (b[i] / math.factorial(i) if i >= 1 else 0)
    for deg in range(0, 10):
        for i in range(deg + 1):
            j = deg - i
            curl = (j + 1) * c(i, j + 1) - (i + 1) * a(i + 1, j)
            div = ((i + 1) * c(i + 1, j) + h * i * c(i, j) + h * c(i, j) + (j + 1) * a(i, j + 1)
                   + h * (j + 1) * a(i - 1, j + 1))
            assert curl == 0 and div == 0, (i, j, curl, div)


def test_combined_function_bend_linear_map():
    """CSBEND with K1: the combined-function sector bend, x'' = -(h^2 + K1) x
    + h delta and y'' = K1 y (K1 > 0 focuses horizontally): cos/sin and
    cosh/sinh blocks and the dispersion, from the tracked Jacobian"""
    L, th, K1 = 1.0, 0.1, 0.5
    h = th / L
    J = jacobian6([E.CSBEND(L=L, ANGLE=th, K1=K1, N_KICKS=64, EDGE1_EFFECTS=0, EDGE2_EFFECTS=0)])
    kx = math.sqrt(h * h + K1)
    ky = math.sqrt(K1)
    want = {(0, 0): math.cos(kx * L), (0, 1): math.sin(kx * L) / kx, (1, 0): -kx * math.sin(kx * L),
            (1, 1): math.cos(kx * L), (0, 5): h * (1 - math.cos(kx * L)) / kx ** 2,
            (1, 5): h * math.sin(kx * L) / kx,
            (2, 2): math.cosh(ky * L), (2, 3): math.sinh(ky * L) / ky, (3, 2): ky * math.sinh(ky * L),
            (3, 3): math.cosh(ky * L)}
    for (i, j), v in want.items():
        assert abs(J[i, j] - v) < 2e-6 * max(1.0, abs(v)), (i, j, J[i, j], v)


@pytest.mark.parametrize("key,k,ctor,ref_key", [("K2", 40.0, E.KSEXT, "K2"), ("K3", 600.0, E.KOCT, "K3")])
def test_bend_multipole_component_matches_the_multipole_element(key, k, ctor, ref_key):
    """a CSBEND whose curvature is negligible (angle 1e-5) with a K2 or K3
    component kicks as KSEXT or KOCT of the same length and strength: the
    change of x' and y' caused by the component, at points off both axes"""
    L = 0.2
    pts = [[2e-3, 0, 1e-3, 0, 0, 0], [-1.5e-3, 1e-4, 0.5e-3, -2e-4, 0, 1e-3], [1e-3, 0, -2e-3, 0, 0, 0]]
    bend = lambda kk: E.CSBEND(L=L, ANGLE=1e-5, N_KICKS=64, EDGE1_EFFECTS=0, EDGE2_EFFECTS=0, **{key: kk})
    rb, _, _ = track([bend(k)], rows_of(pts))
    r0, _, _ = track([bend(0.0)], rows_of(pts))
    rm, _, _ = track([ctor(L=L, N_KICKS=64, **{ref_key: k})], rows_of(pts))
    rd, _, _ = track([E.EDRIFT(L=L)], rows_of(pts))
    for col in (1, 3):
        d_bend = rb[:, col] - r0[:, col]
        d_mult = rm[:, col] - rd[:, col]
        assert np.all(np.abs(d_bend - d_mult) <= 1e-4 * np.abs(d_mult) + 1e-15), (col, d_bend, d_mult)
        assert np.all(np.abs(d_mult) > 1e-9)


def test_bend_decapole_component_thin_limit():
    """K4 in a short weak bend: the kick of the thin decapole, dx' = -K4 L
    (x^4 - 6 x^2 y^2 + y^4)/24 and dy' = K4 L (4 x^3 y - 4 x y^3)/24, to a few
    per cent (thick-lens terms)"""
    L, K4 = 0.05, 2e5
    pts = [[3e-3, 0, 0, 0, 0, 0], [2e-3, 0, 1.5e-3, 0, 0, 0], [0, 0, 3e-3, 0, 0, 0]]
    bend = lambda kk: E.CSBEND(L=L, ANGLE=1e-5, N_KICKS=32, EDGE1_EFFECTS=0, EDGE2_EFFECTS=0, K4=kk)
    rb, _, _ = track([bend(K4)], rows_of(pts))
    r0, _, _ = track([bend(0.0)], rows_of(pts))
    scale = K4 * L * 3e-3 ** 4 / 24
    for p, gb, g0 in zip(pts, rb, r0):
        x, y = p[0], p[2]
        wx = -K4 * L * (x ** 4 - 6 * x * x * y * y + y ** 4) / 24
        wy = K4 * L * (4 * x ** 3 * y - 4 * x * y ** 3) / 24
        assert abs((gb[1] - g0[1]) - wx) <= 0.03 * abs(wx) + 1e-3 * scale
        assert abs((gb[3] - g0[3]) - wy) <= 0.03 * abs(wy) + 1e-3 * scale


@pytest.mark.parametrize("elem", [
    lambda: E.CSBEND(L=1.0, ANGLE=0.1, K1=0.6, K2=12.0, N_KICKS=16, E1=0.02, E2=0.03),
    lambda: E.CSBEND(L=0.8, ANGLE=-0.3, K2=0.1375, K3=100.0, N_KICKS=12, E1=-0.2, E2=-0.2),
])
def test_bend_with_multipoles_is_symplectic(elem):
    """the bend body with a multipole field, like the pure dipole, has a
    symplectic transverse map at fixed delta"""
    test_transverse_map_is_symplectic(elem)


def _match(rows, n, Po):
    o = O.Oracle("glibc")
    return o.match_energy(rows, n, Po)


def test_match_energy_exact_mean_and_momentum_conservation():
    """do_match_energy (do_tracking.c:3371-3417): with every P0 (1 + delta)
    and every partial sum exact in binary, the KahanPlus mean is the exact
    mean momentum; afterwards each particle keeps its momentum,
    P0' (1 + delta') = P0 (1 + delta) to rounding, the rows past np are not
    touched, and a second match finds nothing to change"""
    rng = np.random.default_rng(8)
    n, Po = 1024, 1024.0
    rows = np.zeros((n + 16, 8))
    k = rng.integers(-2 ** 20, 2 ** 20, n)
    rows[:n, 5] = k * 2.0 ** -30 + 3 * 2.0 ** -10
    rows[n:, 5] = 0.25
    before = rows.copy()
    P1 = _match(rows, n, Po)
    exact = sum(Fraction(Po) * (1 + Fraction(d)) for d in before[:n, 5]) / n
    assert Fraction(P1) == exact and P1 != Po
    p_old = Po * (1 + before[:n, 5])
    p_new = P1 * (1 + rows[:n, 5])
    assert np.max(np.abs(p_new - p_old) / p_old) <= 4 * 2.0 ** -52
    assert np.array_equal(rows[n:], before[n:])
    assert abs(np.mean(p_new) - P1) / P1 < 1e-14
    again = rows.copy()
    assert _match(again, n, P1) == P1 and np.array_equal(again, rows)


def test_match_energy_threshold_and_summation():
    """no change when the mean moves by at most 1e-14 relative; otherwise
    the mean of a long sum (2^19 momenta) is the compensated one: within 2
    ulps of the exact mean where plain summation is off by many"""
    Po = 3000.0
    rows = np.zeros((4, 8))
    rows[:, 5] = [2e-15, -1e-15, 3e-15, 0.0]  # mean momentum change 1e-15 relative
    before = rows.copy()
    assert _match(rows, 4, Po) == Po and np.array_equal(rows, before)
    rng = np.random.default_rng(9)
    n = 1 << 19
    rows = np.zeros((n, 8))
    rows[:, 5] = rng.uniform(-3e-3, 3e-3, n) + 1e-3
    d = rows[:, 5].copy()
    P1 = _match(rows, n, Po)
    terms = [Po * (1 + x) for x in d]
    exact = float(sum(Fraction(t) for t in terms) / n)
    naive = 0.0
    for t in terms:
        naive += t
    naive /= n
    assert abs(P1 - exact) <= 2 * np.spacing(exact)  # measured: 0
    assert abs(naive - exact) > 8 * np.spacing(exact)  # measured: 46 ulps


def test_match_energy_empty_beam():
    rows = np.zeros((3, 8))
    rows[:, 5] = 0.01
    before = rows.copy()
    assert _match(rows, 0, 100.0) == 100.0 and np.array_equal(rows, before)


def _etilt_geometry(rho0, angle, etilt, tilt):
    """where the orbit of a sector bend rotated by etilt about its entrance
    axis crosses the nominal exit plane, in the nominal exit frame (x, x',
    y, y', distance along the orbit to the plane), then turned by the bend's
    tilt into the lab frame; bend toward -x, rotation right-handed about z"""
    th = angle
    P = np.array([-(1 - math.cos(th)) * rho0, 0.0, rho0 * math.sin(th)])
    D = np.array([-math.sin(th), 0.0, math.cos(th)])
    R = np.array([[math.cos(etilt), -math.sin(etilt), 0], [math.sin(etilt), math.cos(etilt), 0], [0, 0, 1]])
    Pt, Dt = R @ P, R @ D
    t = np.dot(P - Pt, D) / np.dot(Dt, D)
    Q = Pt + t * Dt
    ze, ye = D, np.array([0.0, 1.0, 0.0])
    xe = np.cross(ye, ze)
    x, y = np.dot(Q - P, xe), np.dot(Q - P, ye)
    xp, yp = np.dot(Dt, xe) / np.dot(Dt, ze), np.dot(Dt, ye) / np.dot(Dt, ze)
    c, s = math.cos(tilt), math.sin(tilt)
    return np.array([x * c - y * s, xp * c - yp * s, x * s + y * c, xp * s + yp * c, t])


@pytest.mark.parametrize("rho0,angle,etilt,tilt", [
    (5.0, 0.3, 0.01, 0.0), (20.0, 0.05, 1e-4, 0.0), (100.0, 0.02, -0.003, 0.0), (2.0, 0.6, 2e-3, 0.3),
    (8.0, 0.25, -5e-4, math.pi / 2), (1.0, 1.2, 1e-3, -1.0)])
def test_etilt_offsets_against_geometry(rho0, angle, etilt, tilt):
    """computeEtiltCentroidOffset (csbend.c:4723-4776) with the rotation by
    -tilt: the offsets agree with the exact crossing of the rotated orbit
    and the nominal exit plane (for the small error tilts the formula is
    written for: the reference takes the unnormalised direction component as
    y', exact to first order), and vanish without an error tilt"""
    L = O.lib("glibc")
    out = (C.c_double * 5)()
    L.or_etilt_offsets(rho0, angle, etilt, tilt, out)
    got = np.array(out[:])
    want = _etilt_geometry(rho0, angle, etilt, tilt)
    scale = np.maximum(np.abs(want), 1e-6 * np.max(np.abs(want)))
    assert np.max(np.abs(got - want) / scale) < 1e-6, (got, want)
    L.or_etilt_offsets(rho0, angle, 0.0, tilt, out)
    assert list(out) == [0.0] * 5


def test_quadrupole_effective_length_is_drift_quad_drift():
    """KQUAD LEFFECTIVE (multipole.c:766-777, 1028-1030, 1098-1100): the
    body has the effective length and strength K1 LEFFECTIVE, and exact
    drifts of (L - LEFFECTIVE)/2 pad it.  Pinned on the linear map: a small
    orbit follows drift . thick quad . drift with the textbook cos/cosh
    blocks (to the integrator's O(h^4) error), and with K1 = 0 the element
    is an exact drift of L."""
    L, Le, k = 0.8, 0.5, 1.2
    lend = (L - Le) / 2
    x0, xp0 = 1e-6, 2e-7
    r, n, _ = track([E.KQUAD(L=L, LEFFECTIVE=Le, K1=k, N_KICKS=32)],
                    rows_of([[x0, xp0, x0, xp0, 0, 0]]))
    s = math.sqrt(k)
    D = np.array([[1, lend], [0, 1]])
    F = np.array([[math.cos(s * Le), math.sin(s * Le) / s], [-s * math.sin(s * Le), math.cos(s * Le)]])
    Dn = np.array([[math.cosh(s * Le), math.sinh(s * Le) / s], [s * math.sinh(s * Le), math.cosh(s * Le)]])
    ex = D @ F @ D @ [x0, xp0]
    ey = D @ Dn @ D @ [x0, xp0]
    assert abs(r[0, 0] - ex[0]) < 1e-5 * x0 and abs(r[0, 1] - ex[1]) < 1e-5 * x0
    assert abs(r[0, 2] - ey[0]) < 1e-5 * x0 and abs(r[0, 3] - ey[1]) < 1e-5 * x0
    # a mistake in the strength or the padding moves the orbit far more
    wrong = D @ np.array([[math.cos(s * L), math.sin(s * L) / s], [-s * math.sin(s * L), math.cos(s * L)]]) @ D
    assert abs((wrong @ [x0, xp0])[0] - ex[0]) > 100 * abs(r[0, 0] - ex[0])
    pts = [[1e-3, 2e-3, -3e-4, 5e-4, 0.1, 1e-3], [-2e-3, -1e-3, 4e-4, 1e-4, 0, -2e-3]]
    r, n, _ = track([E.KQUAD(L=L, LEFFECTIVE=Le, K1=0.0, N_KICKS=4)], rows_of(pts))
    for p, q in zip(pts, r):
        assert abs(q[0] - (p[0] + L * p[1])) < 1e-15 and abs(q[2] - (p[2] + L * p[3])) < 1e-15
        assert abs(q[4] - (p[4] + L * math.sqrt(1 + p[1] ** 2 + p[3] ** 2))) < 1e-14


def _resonator_table(nb, bin_size, freq, Q, Rs):
    """the broad-band resonator sampled as set_up_zlongit samples it
    (zlongit.c:610-622), as an impedance table of nb/2 + 1 points; DC and
    Nyquist zero, as the broad-band spectrum has them"""
    df_n = 1 / (nb * bin_size) / freq
    n = nb // 2 + 1
    zr = np.zeros(n)
    zi = np.zeros(n)
    Ra = 2 * Rs
    for i in range(1, n - 1):
        term = Q * (1.0 / (i * df_n) - i * df_n)
        zr[i] = Ra / 2 / (1 + term * term)
        zi[i] = zr[i] * term
    f = np.arange(n) * (1 / (nb * bin_size))
    return E.ImpedanceTable(f, zr, zi)


def test_zlongit_table_of_the_resonator_equals_the_broad_band_element():
    """ZLONGIT from Z(f) tables (zlongit.c:634-700) fed the broad-band
    resonator's own samples gives the broad-band element's result bit for
    bit: the packed-spectrum layout (real and imaginary interleaved from
    index 1, DC and Nyquist doubled) and the bin size 1/(N df) are pinned by
    an independent construction of the same spectrum"""
    nb, bs = 8192, 2e-12
    table = _resonator_table(nb, bs, 20e9, 1.0, 1e4)
    assert table.f[1] - table.f[0] == 61035156.25  # exact: 1/(N df) rounds back to BIN_SIZE
    bb = E.ZLONGIT(BROAD_BAND=1, RS=1e4, Q=1.0, FREQ=20e9, N_BINS=nb, BIN_SIZE=bs)
    tb = E.ZLONGIT_TABLE(table)
    rows = wake_bunch(20000, 20e-12, seed=41)
    q = 5e-9 / 20000
    out = []
    for e in (bb, tb):
        r = rows.copy()
        o = O.Oracle("emath")
        o.set_macro_charge(q)
        o.zlongit(r, e.p.zlongit, PO, 0)
        out.append((r, o.last_hist()))
    assert np.array_equal(out[0][0], out[1][0]) and not np.array_equal(out[0][0], rows)
    assert np.array_equal(out[0][1][1], out[1][1][1]) and out[0][1][2] == out[1][1][2]
    # a spectrum shifted by one index is not the same impedance
    bad = E.ImpedanceTable(table.f, np.roll(table.Zreal, 1), table.Zimag)
    r = rows.copy()
    o = O.Oracle("emath")
    o.set_macro_charge(q)
    o.zlongit(r, E.ZLONGIT_TABLE(bad).p.zlongit, PO, 0)
    assert not np.array_equal(r, out[0][0])


def test_zlongit_table_dc_and_nyquist_terms_are_doubled():
    """a pure DC impedance Z(0) = R: the packed spectrum's DC slot holds 2R
    (zlongit.c:686-688), so the voltage is flat over the window at 2 R times
    the mean current, V = 2 R q n / (N dt) for n binned particles of charge q
    in N bins of dt (undoubled, or with another normalisation, it would be
    R q n / (N dt) or N times off)"""
    n = 1025
    f = np.arange(n) * 1e9
    zr = np.zeros(n)
    zr[0] = 50.0
    t = E.ZLONGIT_TABLE(E.ImpedanceTable(f, zr))
    rows = wake_bunch(3000, 20e-12, seed=42)
    q = 1e-12
    o = O.Oracle("emath")
    o.set_macro_charge(q)
    r = rows.copy()
    o.zlongit(r, t.p.zlongit, PO, 0)
    h, v, tmin, dt = o.last_hist()
    nb = 2 * (n - 1)
    assert h.size == nb and dt == 1.0 / (nb * 1e9)
    # flat voltage, proportional to the doubled DC term
    assert np.allclose(v, v[0], rtol=1e-12, atol=0) and v[0] != 0
    expect = 2 * 50.0 * q / dt * h.sum() / nb
    assert abs(v[0] / expect - 1) < 1e-12


@pytest.mark.parametrize("mutate", ["spacing", "dc_imag", "count", "none"])
def test_zlongit_table_validation(mutate):
    n = 65
    f = np.arange(n) * 1e9
    zr = np.linspace(0, 10, n)
    zi = np.linspace(0, 5, n)
    if mutate == "spacing":
        f[10] += 0.5e9
    elif mutate == "dc_imag":
        zi[0] = 1.0
    elif mutate == "count":
        f, zr, zi = f[:64], zr[:64], zi[:64]
    t = E.ZLONGIT_TABLE(E.ImpedanceTable(f, zr, zi), ALLOW_LONG_BEAM=1)
    rows = wake_bunch(500, 20e-12, seed=43)
    o = O.Oracle("emath")
    o.set_macro_charge(1e-12)
    if mutate == "none":
        o.zlongit(rows, t.p.zlongit, PO, 0)
    else:
        with pytest.raises(RuntimeError):
            o.zlongit(rows, t.p.zlongit, PO, 0)


def test_wake_change_p0_matches_the_mean_momentum():
    """WAKE with CHANGE_P0 (wake.c:288-289): after the wake kick the central
    momentum becomes the KahanPlus mean momentum of the rows and every delta
    is re-expressed against it, so each particle's momentum is the one the
    same wake without CHANGE_P0 gives, and the new central momentum is
    their mean"""
    table = Lt.resonator_wake()
    rows = wake_bunch(20000, 20e-12, seed=45)
    out = {}
    for cp in (0, 1):
        bl = E.Beamline([E.CHARGE("Q", TOTAL=5e-9), E.WAKE(table, "WK", CHANGE_P0=cp)], keep=[table])
        r = rows.copy()
        o = O.Oracle("emath")
        o.do_tracking(r, bl, PO, 1)
        out[cp] = (r, o.P_central)
    (ra, Pa), (rb, Pb) = out[0], out[1]
    assert Pa == PO and Pb < PO  # the wake takes energy: the central momentum drops
    pa = Pa * (1 + ra[:, 5])
    pb = Pb * (1 + rb[:, 5])
    assert np.allclose(pb, pa, rtol=2e-15, atol=0)
    mean = float(sum(Fraction(x) for x in pa) / len(pa))
    assert abs(Pb / mean - 1) < 1e-15
    assert np.array_equal(ra[:, [0, 1, 2, 3, 4]], rb[:, [0, 1, 2, 3, 4]])


def test_malign_offsets_beam_as_offset_beam():
    """MALIGN (offset_beam, do_tracking.c:3243-3306): transverse offsets
    add; DZ is an exact drift; DP and DE change the momentum at constant
    arrival time, so s scales with the velocity; DT shifts the time; the
    element applies on ON_PASS only (-1: every pass)"""
    pts = [[1e-3, 2e-4, -3e-4, 5e-5, 0.2, 1e-3], [-2e-3, -1e-4, 4e-4, 1e-4, -0.1, -2e-3]]
    rows = rows_of(pts)
    r, n, _ = track([E.MALIGN(DX=1e-4, DXP=-2e-5, DY=3e-4, DYP=1e-5)], rows)
    for p, q in zip(pts, r):
        assert q[0] == p[0] + 1e-4 and q[1] == p[1] + -2e-5 and q[2] == p[2] + 3e-4 and q[3] == p[3] + 1e-5
        assert q[4] == p[4] and q[5] == p[5]
    r, n, _ = track([E.MALIGN(DZ=0.3)], rows)
    for p, q in zip(pts, r):
        assert q[0] == p[0] + 0.3 * p[1] and q[2] == p[2] + 0.3 * p[3]
        assert abs(q[4] - (p[4] + 0.3 * math.sqrt(1 + p[1] ** 2 + p[3] ** 2))) < 1e-15

    def beta(P):
        return P / math.sqrt(P * P + 1)

    for kw in ({"DP": 2e-3}, {"DE": -1e-3}, {"DT": 3e-12}, {"DP": 1e-3, "DT": -2e-12}):
        r, n, _ = track([E.MALIGN(**kw)], rows)
        for p, q in zip(pts, r):
            P0 = PO * (1 + p[5])
            if "DE" in kw:
                g1 = math.sqrt(1 + P0 * P0) * (1 + kw["DE"])
                P1 = math.sqrt(g1 * g1 - 1)
            else:
                P1 = PO * (1 + p[5] + kw.get("DP", 0.0))
            assert abs(PO * (1 + q[5]) / P1 - 1) < 1e-14
            t = p[4] / (beta(P0) * 2.99792458e8) + kw.get("DT", 0.0)
            assert abs(q[4] - t * beta(P1) * 2.99792458e8) < 1e-15 * max(1.0, abs(q[4]))
    bl = [E.MALIGN(DX=1e-3, ON_PASS=1), E.DRIF(L=1.0)]
    r, n, _ = track(bl, rows, passes=3)
    expect = rows.copy()
    for k in range(3):
        if k == 1:
            expect[:, 0] += 1e-3
        expect[:, 0] += 1.0 * expect[:, 1]
    assert np.allclose(r[:, 0], expect[:, 0], rtol=0, atol=1e-15)


def test_monitors_are_drifts_without_an_aperture_check():
    """MONI/HMON/VMON: the drift matrix of their length and order
    (compute_matrices.c:1317-1328) with NO_APERTURE (track_data.c:3928-3930):
    where a DRIF of the same length loses a particle beyond the MAXAMP
    aperture, the monitor keeps it; the NaN cull still follows them"""
    # row 1 starts inside the 20 mm aperture and drifts out of it
    pts = [[1e-3, 2e-4, -3e-4, 5e-5, 0.2, 1e-3], [0.019, 1e-2, 0, 0, 0, 0]]
    ap = E.MAXAMP(X_MAX=0.02, Y_MAX=0.02)
    for mon in (E.MONI(L=0.7), E.HMON(L=0.7), E.VMON(L=0.7)):
        r, n, _ = track([ap, mon], rows_of(pts))
        d, nd, _ = track([ap, E.DRIF(L=0.7)], rows_of(pts))
        assert n == 2 and nd == 1
        assert np.array_equal(r[0], d[0])  # the same drift
        assert r[1, 0] == 0.019 + 0.7 * 1e-2 and r[1, 0] > 0.02  # beyond the aperture, kept
    # an infinite slope makes the drifted x infinite: the NaN cull takes it
    r, n, _ = track([E.MONI(L=0.7)], rows_of([[0.0, np.inf, 0, 0, 0, 0], pts[0]]))
    assert n == 1 and r[1, 6] == 1
    r1, _, _ = track([E.MONI(L=0.7, ORDER=1)], rows_of(pts[:1]))
    d1, _, _ = track([E.DRIF(L=0.7, ORDER=1)], rows_of(pts[:1]))
    assert np.array_equal(r1, d1)


def test_zero_strength_correctors_are_drifts_without_extrapolation():
    """HKICK/VKICK with no kick: corrector_matrix returns the drift matrix
    (bend_matrix.c:314-317), and the aperture check after them does not
    extrapolate the loss point back as it does after a DRIF
    (do_tracking.c:2463); a kick is refused"""
    pts = [[1e-3, 2e-4, -3e-4, 5e-5, 0.2, 1e-3], [0.019, 1e-2, 0, 0, 0, 0]]
    ap = E.MAXAMP(X_MAX=0.02, Y_MAX=0.02)
    for k in (E.HKICK(L=0.7, TILT=0.1), E.VKICK(L=0.7, KICK=1e-3, CALIBRATION=0.0)):
        r, n, _ = track([ap, k], rows_of(pts))
        d, nd, _ = track([ap, E.DRIF(L=0.7)], rows_of(pts))
        assert n == nd == 1
        assert np.array_equal(r[0], d[0])
        # the lost row: drifted to the end, s at the element's end, not extrapolated back
        assert r[1, 0] == 0.019 + 0.7 * 1e-2 and r[1, 4] == 0.7
        assert d[1, 4] < 0.7
    with pytest.raises(RuntimeError):
        track([E.HKICK(L=0.1, KICK=1e-4)], rows_of(pts))


def test_thin_corrector_kicks():
    """a zero-length HKICK/VKICK with a kick: corrector_matrix's C[1] = kick,
    R[1][5] = -sin(kick) (bend_matrix.c:323-330) applied by track_particles
    (matrix.c:430-458): x' -> (kick - sin(kick) dp) + x' exactly, VKICK the
    same in y through tilt_matrices' exact quarter turn; a tilt splits the
    kick by cos and sin (a wrong rotation sign fails the y half); calibration
    scales it; a thick corrector with a kick stays refused"""
    pts = [[1e-3, 2e-4, -3e-4, 5e-5, 0.2, 1e-3], [-2e-3, -1e-4, 4e-4, 0.0, -0.1, -2.5e-3]]
    rows = rows_of(pts)
    k = 3e-4
    r, n, _ = track([E.HKICK(KICK=k)], rows)
    assert n == 2
    for a, b in zip(r, rows):
        assert a[1] == (k - math.sin(k) * b[5]) + b[1]
        assert np.array_equal(a[[0, 2, 3, 4, 5]], b[[0, 2, 3, 4, 5]])
    r, n, _ = track([E.VKICK(KICK=k / 2, CALIBRATION=2.0)], rows)
    for a, b in zip(r, rows):
        assert a[3] == (k - math.sin(k) * b[5]) + b[3]
        assert np.array_equal(a[[0, 1, 2, 4, 5]], b[[0, 1, 2, 4, 5]])
    th = 0.4
    r, n, _ = track([E.HKICK(KICK=k, TILT=th)], rows)
    for a, b in zip(r, rows):
        g = k - math.sin(k) * b[5]
        assert abs((a[1] - b[1]) - math.cos(th) * g) < 1e-18
        assert abs((a[3] - b[3]) - math.sin(th) * g) < 1e-18
        assert np.allclose(a[[0, 2, 4, 5]], b[[0, 2, 4, 5]], rtol=1e-15, atol=1e-20)
    with pytest.raises(RuntimeError):
        track([E.VKICK(L=0.2, KICK=k)], rows)


def test_elements_before_recirc_act_on_the_first_pass_only():
    """RECIRC (do_tracking.c:459-465, 1382-1387): passes after pass 0 start
    at the recirculation point, so an offset placed before it acts once"""
    pts = [[1e-4, 2e-5, -3e-5, 1e-5, 0.0, 0.0]]
    r, n, _ = track([E.MALIGN(DX=1e-3), E.RECIRC(), E.DRIF(L=1.5)], rows_of(pts), passes=3)
    x = pts[0][0] + 1e-3
    for _ in range(3):
        x = x + 1.5 * pts[0][1]
    assert abs(r[0, 0] - x) < 1e-18 and abs(r[0, 4] - 3 * (1.5 + 0.75 * (2e-5 ** 2 + 1e-5 ** 2))) < 1e-15
    r2, _, _ = track([E.MALIGN(DX=1e-3), E.DRIF(L=1.5)], rows_of(pts), passes=3)
    assert abs(r2[0, 0] - (x + 2e-3)) < 1e-17  # without RECIRC it acts every pass
    # a later call (pass offset > 0) starts at the recirculation point at once
    bl = E.Beamline([E.MALIGN(DX=1e-3), E.RECIRC(), E.DRIF(L=1.5)])
    o = O.Oracle("glibc")
    r3 = rows_of(pts)
    o.do_tracking(r3, bl, PO, 2, pass_offset=5)
    assert abs(r3[0, 0] - (pts[0][0] + 3.0 * pts[0][1])) < 1e-18


def test_rectangular_collimator_losses():
    """RCOL (rectangular_collimator, limit_amplitudes.c:32-185): a particle
    outside the jaws at the entrance is lost at the element's start; one
    that leaves through a jaw inside the collimator is lost where its
    straight orbit meets the jaw; an open side spares particles leaving
    through it (without placing them); INVERT swaps lost and kept; without
    jaws the element is an exact drift"""
    L, xm = 0.5, 0.01
    pts = [[0.02, 0.0, 0, 0, 0, 0],        # outside at the entrance
           [0.009, 0.004, 0, 0, 0, 0],     # hits the +x jaw at z = (0.01-0.009)/0.004 = 0.25
           [-0.009, -0.01, 0, 0, 0, 0],    # hits the -x jaw at z = 0.1
           [0.001, 1e-4, 2e-3, 1e-4, 0.3, 0]]  # passes
    # the same particles one metre of drift upstream, so that they reach the
    # collimator (at z = 1) as listed
    up = [[p[0] - 1.0 * p[1], p[1], p[2] - 1.0 * p[3], p[3], p[4], p[5]] for p in pts]
    r, n, _ = track([E.DRIF(L=1.0), E.RCOL(L=L, X_MAX=xm, Y_MAX=0.02)], rows_of(up))
    assert n == 1 and r[0, 6] == 4
    lost = {int(row[6]): row for row in r[1:]}
    assert lost[1][4] == 1.0
    assert abs(lost[2][4] - 1.25) < 1e-15 and abs(lost[2][0] - 0.01) < 1e-15
    assert abs(lost[3][4] - 1.1) < 1e-15 and abs(lost[3][0] + 0.01) < 1e-15
    assert abs(r[0, 4] - (0.3 + 1.0 + (1e-4 ** 2 + 1e-4 ** 2) / 2 + L * math.sqrt(1 + 2e-8))) < 1e-15
    r, n, _ = track([E.RCOL(L=L, X_MAX=xm, OPEN_SIDE="+x")], rows_of(pts))
    kept = set(r[:n, 6].astype(int))
    assert kept == {1, 2, 4}  # the +x side is open: rows 1 and 2 leave through it
    r, n, _ = track([E.RCOL(L=0.0, X_MAX=xm, INVERT=1)], rows_of(pts))
    assert set(r[:n, 6].astype(int)) == {1}
    r, n, _ = track([E.RCOL(L=L)], rows_of(pts))
    e, _, _ = track([E.EDRIFT(L=L)], rows_of(pts))
    assert n == 4 and np.array_equal(r, e)
