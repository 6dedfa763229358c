"""Seeded generators (host code of the library, no GPU).  DLARAN is pinned
exactly against the 48-bit congruential recurrence it implements; the
Gaussian routines (polar method; the toolkit's gauss_rn is not shipped with
the reference, so bit parity with elegant is unpinned) against the moments
and cuts they must produce."""
import math

import numpy as np

from elegant_amd import _abi as A
from elegant_amd import generators as G

A48 = 33952834046453
M48 = 1 << 48


def test_dlaran_is_the_48_bit_congruential_generator():
    for iseed in (987654321, 12345, 2):
        r = G.Rng(iseed)
        x = (abs(iseed) // 2) * 2 + 1          # odd seed (drand_oag.c:102)
        x &= M48 - 1
        for _ in range(2000):
            x = (x * A48) % M48
            assert r.uniform() == x / M48      # the limb arithmetic and conversion are exact


def test_seeds_give_distinct_reproducible_streams():
    a = G.Rng(7).uniform(100)
    b = G.Rng(7).uniform(100)
    c = G.Rng(9).uniform(100)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.all((a > 0) & (a < 1))


def test_gaussian_moments_and_cutoff():
    r = G.Rng(99)
    z = r.gauss(200000)
    assert abs(z.mean()) < 4 / math.sqrt(z.size)
    assert abs(z.var() - 1) < 4 * math.sqrt(2 / z.size)
    zc = r.gauss(100000, limit=2.0)
    assert np.abs(zc).max() <= 2.0
    # variance of the unit normal truncated at +-2
    phi = math.exp(-2) / math.sqrt(2 * math.pi)
    var = 1 - 2 * 2 * phi / math.erf(2 / math.sqrt(2))
    assert abs(zc.var() - var) < 4 * math.sqrt(2 / zc.size) * var


def test_perturbation_types():
    r = G.Rng(5)
    u = np.array([r.perturbation(2e-3, 0, G.UNIFORM) for _ in range(20000)])
    assert np.abs(u).max() <= 2e-3 and abs(u.std() - 2e-3 / math.sqrt(3)) < 2e-5
    pm = {r.perturbation(1e-4, 0, G.PLUS_OR_MINUS) for _ in range(200)}
    assert pm == {1e-4, -1e-4}
    g = np.array([r.perturbation(30e-6, 3.0, G.GAUSSIAN) for _ in range(20000)])
    assert np.abs(g).max() <= 3 * 30e-6


def test_symmetrized_distribution_holds_mirror_images():
    r = G.Rng(3)
    rows = r.gaussian_distribution(400, 1.0, 0.5, 3.0, symmetrize=True)
    q = rows.reshape(100, 4, 2)
    assert np.array_equal(q[:, 1, 0], q[:, 0, 0]) and np.array_equal(q[:, 1, 1], -q[:, 0, 1])
    assert np.array_equal(q[:, 2, 0], -q[:, 0, 0]) and np.array_equal(q[:, 3, 1], -q[:, 0, 1])
    assert np.all(np.abs(rows[:, 0]) <= 3.0) and np.all(np.abs(rows[:, 1]) <= 1.5)


def test_gaussian_bunch_has_the_requested_twiss_and_emittance():
    n = 100000
    ex, bx, ax = 1e-9, 8.0, -1.2
    ey, by, ay = 1e-11, 3.0, 0.7
    rows = G.gaussian_bunch(n, 2024, (ex, bx, ax, 6.0), (ey, by, ay, 6.0), (3e-3, 1e-3, 0.0, 6.0))
    for (e, b, a), c in (((ex, bx, ax), 0), ((ey, by, ay), 2)):
        S = np.cov(rows[:, c], rows[:, c + 1])
        emit = math.sqrt(np.linalg.det(S))
        assert abs(emit / e - 1) < 0.02
        assert abs(S[0, 0] / emit / b - 1) < 0.02
        assert abs(-S[0, 1] / emit - a) < 0.03 * max(1, abs(a))
    assert abs(rows[:, 4].std() / 3e-3 - 1) < 0.02 and abs(rows[:, 5].std() / 1e-3 - 1) < 0.02
    assert np.array_equal(rows[:, 6], np.arange(1, n + 1)) and np.all(rows[:, 7] == -1)


def test_perturbations_vector_matches_successive_calls():
    from elegant_amd import generators as G
    a, b = G.Rng(987654321), G.Rng(987654321)
    for kind in (G.UNIFORM, G.GAUSSIAN, G.PLUS_OR_MINUS):
        x = [a.perturbation(2.5e-4, 2.0, kind) for _ in range(101)]
        y = b.perturbation(2.5e-4, 2.0, kind, n=101)
        assert np.array_equal(np.array(x), y)


def test_error_variants_are_steps_of_one_run():
    """variant k carries the draws that follow those of variants 0..k-1 on
    the run's stream: FSE, DX, DY, TILT in turn over every multipole, each
    a Gaussian perturbation cut at 3 sigma"""
    from elegant_amd import errors
    from elegant_amd import generators as G
    from elegant_amd import lattices as Lt
    bl, Po = Lt.storage_ring(seed=1)
    mults = [e for e in bl.elements if e.type in (A.EH_T_KQUAD, A.EH_T_KSEXT)]
    r = G.Rng(errors.RUN_SEED)
    r.perturbation(1.0, 3.0, G.GAUSSIAN, n=2 * 4 * len(mults))  # steps 0 and 1
    z = r.perturbation(1.0, 3.0, G.GAUSSIAN, n=4 * len(mults)).reshape(4, len(mults))
    v, feats = errors.perturb(bl, 2)
    vm = [e for e in v.elements if e.type in (A.EH_T_KQUAD, A.EH_T_KSEXT)]
    for j, (e0, e1) in enumerate(zip(mults, vm)):
        p0 = e0.p.kquad if e0.type == A.EH_T_KQUAD else e0.p.ksext
        p1 = e1.p.kquad if e1.type == A.EH_T_KQUAD else e1.p.ksext
        assert p1.fse == p0.fse + 1e-3 * z[0, j] and p1.dx == p0.dx + 30e-6 * z[1, j]
        assert p1.dy == p0.dy + 30e-6 * z[2, j] and p1.tilt == p0.tilt + 100e-6 * z[3, j]
    assert np.max(np.abs(z)) <= 3.0
    # distinct, reproducible, uncorrelated from one step to the next
    fs = np.array([errors.perturb(bl, k)[1] for k in range(64)])
    assert len({tuple(f) for f in fs}) == 64
    assert np.array_equal(errors.perturb(bl, 5)[1], fs[5])
    assert abs(np.corrcoef(fs[:-1, 0], fs[1:, 0])[0, 1]) < 0.4
