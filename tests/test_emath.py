"""Deterministic math (elegant_amd/csrc/emath.h): accuracy against an
independent high-precision reference, agreement statistics with glibc, and
bit-identity between the CPU build and the GPU build."""
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle_ctypes as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

FN = {"sin": 0, "cos": 1, "asin": 2, "log": 3, "tan": 4, "sqrt": 5, "div": 6}


def ulp_err(got, ref_longdouble):
    ref = ref_longdouble.astype(np.float64)
    spacing = np.spacing(np.abs(ref))
    return np.abs((got.astype(np.longdouble) - ref_longdouble) / spacing.astype(np.longdouble)).astype(np.float64)


def args(fn, n, seed=5):
    rng = np.random.default_rng(seed)
    if fn in ("sin", "cos"):
        return np.concatenate([rng.uniform(-0.8, 0.8, n), rng.uniform(-40, 40, n), rng.uniform(-3e5, 3e5, n)])
    if fn == "tan":
        return rng.uniform(-1.2, 1.2, n)
    if fn == "asin":
        return np.concatenate([rng.uniform(-0.01, 0.01, n), rng.uniform(-0.5, 0.5, n), rng.uniform(-1, 1, n)])
    if fn == "log":
        return np.concatenate([np.exp(rng.uniform(-30, 30, n)), 1 + rng.uniform(-1e-3, 1e-3, n)])
    return rng.uniform(1e-3, 10, n)


@pytest.mark.parametrize("fn,bound", [("sin", 0.5005), ("cos", 0.5005), ("asin", 1.2), ("log", 0.5005),
                                      ("tan", 0.5005)])
def test_emath_accuracy_vs_long_double(fn, bound):
    x = args(fn, 100000)
    got = O.math_eval(FN[fn], x, mode="emath")
    ld = np.asarray(x, dtype=np.longdouble)
    ref = {"sin": np.sin, "cos": np.cos, "asin": np.arcsin, "log": np.log, "tan": np.tan}[fn](ld)
    err = ulp_err(got, ref)
    assert np.nanmax(err) < bound, (fn, np.nanmax(err))


def test_emath_matches_glibc_nearly_always():
    for fn in ("sin", "cos", "asin", "log"):
        x = args(fn, 50000)
        a = O.math_eval(FN[fn], x, mode="emath")
        b = O.math_eval(FN[fn], x, mode="glibc")
        differ = np.mean(a != b)
        assert differ < 0.06, (fn, differ)
        # when they differ it is by one unit in the last place
        d = np.abs(a.view(np.int64) - b.view(np.int64))
        assert d.max() <= 1, fn


def test_emath_special_values():
    x = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e300])
    s = O.math_eval(0, x)
    assert s[0] == 0 and not np.signbit(s[0]) and np.signbit(s[1])
    assert np.isnan(s[2:5]).all()
    assert s[5] == np.float64(-0.8178819121159085)  # sin(1e300), 1200-bit reference
    c = O.math_eval(1, x)
    assert c[0] == 1 and c[1] == 1 and np.isnan(c[2:5]).all()
    t = O.math_eval(4, x)
    assert t[0] == 0 and np.signbit(t[1]) and np.isnan(t[2:5]).all()
    assert np.isnan(O.math_eval(2, np.array([1.5, -1.0000001]))).all()
    assert O.math_eval(2, np.array([1.0]))[0] == np.float64(math.pi / 2)
    lg = O.math_eval(3, np.array([1.0, 0.0, -1.0, np.inf, 5e-324]))
    assert lg[0] == 0 and lg[1] == -np.inf and np.isnan(lg[2]) and lg[3] == np.inf
    assert abs(lg[4] - (-744.4400719213812)) < 1e-12


def test_reduction_exact_near_multiples_of_pi_over_2():
    # k*pi/2 computed in exact rational arithmetic; sin of the nearest double
    # is the (tiny) distance to the multiple, known to high relative accuracy
    from fractions import Fraction as F
    pi = F(245850922, 78256779)  # coarse; only used to pick candidates
    ks = np.arange(1, 2000)
    xs = np.array([float(k * math.pi / 2) for k in ks])
    s = O.math_eval(0, xs)
    c = O.math_eval(1, xs)
    ref_s = np.sin(xs.astype(np.longdouble))
    ref_c = np.cos(xs.astype(np.longdouble))
    assert np.max(ulp_err(s, ref_s)) < 1.0
    assert np.max(ulp_err(c, ref_c)) < 1.0


def wide_operands(n, seed):
    """doubles of every exponent and both signs, the special values, the
    edges of the unscaled ranges of em_div / em_sqrt and their neighbours,
    and dividend/divisor pairs whose quotient lies next to a rounding
    midpoint (a = b (q + ulp/2) rounded, the hardest cases for a quotient)"""
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 2**63, n, dtype=np.int64) | (rng.integers(0, 2, n, dtype=np.int64) << 63)
    a = bits.view(np.float64).copy()
    b = (rng.integers(0, 2**63, n, dtype=np.int64) | (rng.integers(0, 2, n, dtype=np.int64) << 63)).view(np.float64).copy()
    # moderate magnitudes (where the unscaled sequences run) for half of them
    m = rng.random(n) < 0.5
    a[m] = rng.uniform(1, 2, m.sum()) * np.exp2(rng.integers(-650, 650, m.sum()))
    b[m] = rng.uniform(1, 2, m.sum()) * np.exp2(rng.integers(-70, 70, m.sum()))
    # quotients next to a midpoint
    k = n // 8
    q = rng.uniform(1, 2, k)
    ulp = np.spacing(q)
    bb = rng.uniform(0.5, 2, k)
    a[:k] = bb * (q + ulp / 2)
    b[:k] = bb
    edges = []
    for e in (-1074, -1022, -767, -700, -600, -64, 0, 64, 600, 1000, 1023):
        v = np.float64(2.0) ** e if e > -1074 else np.float64(5e-324)
        edges += [np.nextafter(v, 0), v, np.nextafter(v, np.inf)]
    edges = np.array(edges + [0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0])
    ea, eb = np.meshgrid(edges, np.concatenate([edges, -edges]))
    return np.concatenate([a, ea.ravel()]), np.concatenate([b, eb.ravel()])


def _round_p(x, p):
    """x (a Fraction) rounded to nearest-even in p-bit binary floating point"""
    if x == 0:
        return Fraction(0)
    sgn = -1 if x < 0 else 1
    x = abs(x)
    e = x.numerator.bit_length() - x.denominator.bit_length()
    while Fraction(2) ** e > x:
        e -= 1
    while Fraction(2) ** (e + 1) <= x:
        e += 1
    ulp = Fraction(2) ** (e - p + 1)
    m = x / ulp
    fl = m.numerator // m.denominator
    rem = m - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2):
        fl += 1
    return sgn * fl * ulp


@pytest.mark.parametrize("p", [6, 8])
def test_one_correction_quotient_exhaustive_low_precision(p):
    """The quotient step em_div runs with a reciprocal rounded to nearest
    (q = RN(a y), r = RN(a - b q), RN(q + r y) with y = RN(1/b)) is
    correctly rounded for every pair of p-bit significands, dividends over
    two binades: the same step in double precision is what the bend radii's
    host-rounded reciprocals feed."""
    sig = [Fraction(1) + Fraction(k, 2 ** (p - 1)) for k in range(2 ** (p - 1))]
    for b in sig:
        y = _round_p(1 / b, p)
        for a0 in sig:
            for a in (a0, 2 * a0):
                q = _round_p(a * y, p)
                r = _round_p(a - b * q, p)
                assert _round_p(q + r * y, p) == _round_p(a / b, p), (p, a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("fn,host", [("sqrt_unscaled", "sqrt"), ("div_shared_recip", "div"),
                                     ("div_rounded_recip", "div")])
def test_device_unscaled_sequences_match_ieee(tracker, fn, host):
    """em_sqrt and em_div (emath.h) against the host's correctly rounded
    operators on every kind of operand: bit-identical, NaN for NaN.  em_div
    runs with the device's refined reciprocal (divisions by a particle's
    1 + delta and by fresh denominators) and with 1/b rounded to nearest
    (the bend radii, whose reciprocals the host rounds)"""
    x, y = wide_operands(1 << 20, seed=23)
    if fn == "sqrt_unscaled":
        x = np.abs(x)
        x[::7] = np.negative(x[::7])  # NaN operands stay NaN
    gpu = tracker.math_probe({"sqrt_unscaled": 7, "div_shared_recip": 8, "div_rounded_recip": 9}[fn], x, y)
    cpu = O.math_eval(FN[host], x, y, mode="emath")
    same = (gpu.view(np.int64) == cpu.view(np.int64)) | (np.isnan(gpu) & np.isnan(cpu))
    assert same.all(), (fn, int((~same).sum()), x[~same][:4], y[~same][:4])


@pytest.mark.gpu
@pytest.mark.parametrize("fn", ["sin", "cos", "asin", "log", "tan", "sqrt", "div"])
def test_device_math_bit_identical_to_cpu(tracker, fn):
    """the device build of emath against the CPU build on the same
    arguments, including (sin, cos, tan) huge ones, reduced by Payne-Hanek,
    and enough arguments that the correctly rounded functions take their
    double-double paths thousands of times"""
    x = args(fn, 200000, seed=11)
    if fn in ("sin", "cos", "tan"):
        rng = np.random.default_rng(12)
        x = np.concatenate([x, rng.uniform(1, 2, 50000) * np.exp2(rng.integers(20, 1024, 50000)),
                            -rng.uniform(1, 2, 1000) * np.exp2(rng.integers(20, 1024, 1000)),
                            np.array([6381956970095103.0 * 2.0 ** 797, 1e300, 2.0 ** 28, -2.0 ** 28])])
    y = None
    if fn == "div":
        y = np.random.default_rng(3).uniform(-7, 7, x.size)
    gpu = tracker.math_probe(FN[fn], x, y)
    cpu = O.math_eval(FN[fn], x, y, mode="emath")
    same = (gpu == cpu) | (np.isnan(gpu) & np.isnan(cpu))
    assert same.all(), (fn, int((~same).sum()), x[~same][:5])


@pytest.mark.gpu
def test_device_aperture_power_matches_checker(tracker):
    """the aperture's integer power (the unrolled recursion of the device)
    against the checker's recursive one, exponents 0..254, bit for bit"""
    rng = np.random.default_rng(41)
    x = np.concatenate([rng.uniform(-2, 2, 1 << 16), rng.uniform(-1e-3, 1e-3, 1 << 12),
                        np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 5e-324])])
    p = rng.integers(1, 255, x.size).astype(np.float64)
    p[-8:] = [2, 2, 254, 253, 2, 3, 4, 2]
    p[:16] = 0  # x^0 = 1
    gpu = tracker.math_probe(10, x, p)
    cpu = O.math_eval(7, x, p, mode="emath")
    same = (gpu.view(np.int64) == cpu.view(np.int64)) | (np.isnan(gpu) & np.isnan(cpu))
    assert same.all(), (int((~same).sum()), x[~same][:4], p[~same][:4])


def _mp():
    mp = pytest.importorskip("mpmath")
    return mp


def _cr(mp, fn, xs, prec=200):
    f = {"sin": mp.sin, "cos": mp.cos, "tan": mp.tan, "log": mp.log, "asin": mp.asin}[fn]
    with mp.workprec(prec):
        return np.array([float(f(mp.mpf(float(v)))) for v in xs])


def cr_args(fn, n, seed):
    rng = np.random.default_rng(seed)
    huge = rng.uniform(1, 2, n) * np.exp2(rng.integers(28, 1020, n)) * rng.choice([-1.0, 1.0], n)
    if fn in ("sin", "cos"):
        return [("band edge and beyond", rng.uniform(-0.8, 0.8, n)), ("RF phases", rng.uniform(-3e6, 3e6, n)),
                ("beyond 2^28", huge)]
    if fn == "tan":
        return [("|x| < 1.6", rng.uniform(-1.6, 1.6, n)), ("moderate", rng.uniform(-1e4, 1e4, n)),
                ("beyond 2^28", huge)]
    return [("all exponents", np.exp(rng.uniform(-740, 709, n))), ("near 1", 1 + rng.uniform(-1e-2, 1e-2, n)),
            ("RF matrix 1 + dP/P", 1 + rng.uniform(-1e-4, 1e-4, n))]


@pytest.mark.parametrize("fn", ["sin", "cos", "tan", "log"])
def test_correctly_rounded_against_high_precision(fn):
    """sin and cos (|x| >= 2^-5), tan and log return the double nearest the
    exact value: every sampled argument against a 200-bit evaluation (1,200
    bits beyond 2^28, where reducing modulo pi/2 needs the extra bits), which
    exercises the fast paths and, on the arguments whose fast result is too
    close to a midpoint, the double-double ones"""
    mp = _mp()
    for what, x in cr_args(fn, 1500, seed=hash(fn) % 1000):
        if fn in ("sin", "cos"):
            x = x[np.abs(x) >= 2.0 ** -5]
        got = O.math_eval(FN[fn], x, mode="emath")
        ref = _cr(mp, fn, x, 1200 if np.abs(x).max() > 2.0 ** 28 else 200)
        bad = got != ref
        assert not bad.any(), (fn, what, int(bad.sum()), x[bad][:3], got[bad][:3], ref[bad][:3])


def test_reduction_of_the_hardest_double():
    """6381956970095103 * 2^797 lies closest of all doubles to a multiple of
    pi/2 (about 2^-61 away), so its reduction needs the full Payne-Hanek
    product; sin, cos and tan of it and of its neighbours are correctly
    rounded, as is the RF phase of a storage ring after 10^5 turns (omega t
    past 2^28, where the Cody-Waite reduction stops)"""
    mp = _mp()
    x0 = 6381956970095103.0 * 2.0 ** 797
    xs = np.array([x0, np.nextafter(x0, 0), np.nextafter(x0, np.inf), -x0, 2626.0 * 1.2e5 + 0.37, 2.0 ** 28,
                   np.nextafter(2.0 ** 28, 0), 1e22, 5e-324 * 0 + 1.0e300])
    for fn in ("sin", "cos", "tan"):
        got = O.math_eval(FN[fn], xs, mode="emath")
        ref = _cr(mp, fn, xs, 1500)
        assert np.array_equal(got, ref), (fn, xs[got != ref], got[got != ref], ref[got != ref])


def test_small_argument_bands_and_glibc_asin_misrounding():
    """On the bend's arguments (|x| in [1e-3, 3e-2]) the band polynomials
    of sin, cos and asin that the branch-free integrators use are correct
    roundings except on a few in 10^5 arguments; glibc's asin misrounds a
    few in 10^3 there, which is what keeps the glibc and emath checkers from
    agreeing bit for bit on the storage ring (measured on 10^5 arguments:
    emath 4 / 5 / 2 misrounded, glibc 2 / 3 / 298)"""
    mp = _mp()
    rng = np.random.default_rng(7)
    x = rng.uniform(1e-3, 3e-2, 20000) * rng.choice([-1.0, 1.0], 20000)
    rates = {}
    for fn in ("sin", "cos", "asin"):
        ref = _cr(mp, fn, x)
        rates[fn] = (float(np.mean(O.math_eval(FN[fn], x, mode="emath") != ref)),
                     float(np.mean(O.math_eval(FN[fn], x, mode="glibc") != ref)))
    assert all(r[0] <= 4e-4 for r in rates.values()), rates
    assert rates["asin"][1] >= 1e-3, rates


def test_tables_against_high_precision():
    """emath_tables.h (written by gen_emath_tables.py in integer arithmetic)
    against mpmath: the bits of 2/pi, sin and cos of j pi/64, the log table,
    1/k! and 1/k pairs and the split constants"""
    import re
    mp = _mp()
    text = open(os.path.join(ROOT, "elegant_amd", "csrc", "emath_tables.h")).read()

    def arr(name):
        body = re.search(r"%s\[[^\]]*\](?:\[2\])? = \{(.*?)\};" % name, text, re.S).group(1)
        return [v for v in re.findall(r"(-?0x[0-9a-fA-F.p+\-]+u?)", body)]

    with mp.workprec(1400):
        words = [int(w.rstrip("u"), 16) for w in arr("em_two_over_pi_bits")]
        big = int(mp.floor(2 / mp.pi * mp.mpf(2) ** (32 * len(words))))
        assert big == sum(w << (32 * (len(words) - 1 - k)) for k, w in enumerate(words))
        for name_hi, name_lo, f in (("em_sin_tab_hi", "em_sin_tab_lo", mp.sin),
                                    ("em_cos_tab_hi", "em_cos_tab_lo", mp.cos)):
            hi, lo = arr(name_hi), arr(name_lo)
            for j in range(17):
                exact = f(j * mp.pi / 64)
                assert float.fromhex(hi[j]) == float(exact)
                assert abs(mp.mpf(float.fromhex(hi[j])) + mp.mpf(float.fromhex(lo[j])) - exact) <= \
                    abs(exact) * mp.mpf(2) ** -105 + mp.mpf(2) ** -1000
        r, lh, ll = arr("em_log_r"), arr("em_log_hi"), arr("em_log_lo")
        for k, i in enumerate(range(91, 182)):
            rv = float.fromhex(r[k])
            assert rv == 128.0 / i
            exact = -mp.log(mp.mpf(rv))
            assert abs(mp.mpf(float.fromhex(lh[k])) + mp.mpf(float.fromhex(ll[k])) - exact) <= \
                abs(exact) * mp.mpf(2) ** -105 + mp.mpf(2) ** -1000
        consts = dict(re.findall(r"#define (EM_\w+) (-?0x[0-9a-fA-F.p+\-]+)", text))
        c = {k: mp.mpf(float.fromhex(v)) for k, v in consts.items()}
        assert abs(c["EM_PIO2_1"] + c["EM_PIO2_2"] + c["EM_PIO2_3"] - mp.pi / 2) < mp.mpf(2) ** -160
        assert abs(c["EM_LN2_1"] + c["EM_LN2_2"] + c["EM_LN2_3"] - mp.log(2)) < mp.mpf(2) ** -150
        assert abs(c["EM_PIO64_A"] + c["EM_PIO64_B"] + c["EM_PIO64_C"] - mp.pi / 64) < mp.mpf(2) ** -160
        # A in [2^-5, 2^-4) with 48 significant bits: j A exact for |j| <= 32
        assert mp.isint(c["EM_PIO64_A"] * 2 ** 52) and 2 ** 47 <= c["EM_PIO64_A"] * 2 ** 52 < 2 ** 48
        # ln 2's first term in [1/2, 1) with 42 bits: e EM_LN2_1 exact for |e| < 2048
        assert mp.isint(c["EM_LN2_1"] * 2 ** 42)
