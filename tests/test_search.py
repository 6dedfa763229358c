"""Host logic of the grid aperture searches (no GPU): the tracker is
replaced by a survival rule known in closed form, so the plain grid's answer
is known and the guided and bisection searches must reproduce it."""
import numpy as np
import pytest

from elegant_amd import da
from elegant_amd import learned
from elegant_amd import studies


class FakeVariant:
    def __init__(self, nx, radius):
        self.nx = nx
        self.radius = radius  # survival radius (in grid fractions) per offset index
        self.features = np.zeros(16)
        self.beamline = None
        self.Po = 1.0


def install(monkeypatch, rays, deltas, holes=()):
    """replace the batched tracker: a point survives inside the variant's
    radius; effort counts every pass for a survivor and one for a loss"""
    calls = {"n": 0}

    def fake(tracker, vs, sets, passes, tune=None):
        calls["n"] += 1
        lps = []
        effort = 0
        for v, rows in zip(vs, sets):
            di = np.searchsorted(np.asarray(deltas), rows[:, 5])
            r = np.hypot(rows[:, 0] / rays.xmax, rows[:, 2] / rays.ymax)
            ok = r <= v.radius[di] + 1e-12
            for (hx, hy) in holes:
                ok &= ~((np.abs(rows[:, 0] - hx) < 1e-12) & (np.abs(rows[:, 2] - hy) < 1e-12))
            lps.append(np.where(ok, -1, 0).astype(np.int32))
            effort += int(np.sum(np.where(ok, passes, 1)))
        return lps, effort, 0.0, None

    monkeypatch.setattr(studies, "track_sets", fake)
    return calls


def reference_limits(rays, deltas, v):
    frac = np.arange(v.nx) / (v.nx - 1)
    out = np.empty((len(deltas), rays.lines), dtype=int)
    for di in range(len(deltas)):
        ok = frac <= v.radius[di] + 1e-12
        out[di] = (np.argmax(~ok) if (~ok).any() else v.nx) - 1
    return out


@pytest.mark.parametrize("guess_kind", ["none", "exact", "off_by_3", "far", "origin"])
def test_index_search_finds_the_grid_boundary(monkeypatch, guess_kind):
    rays = da.Rays(lines=9, xmax=0.01, ymax=0.01)
    deltas = (-0.02, 0.0, 0.02)
    vs = [FakeVariant(nx, np.array(rad)) for nx, rad in
          [(2, [0.5, 1.2, -0.1]), (17, [0.31, 0.77, 0.0]), (64, [0.9, 0.05, 1.0]), (145, [0.44, 0.62, 0.991])]]
    install(monkeypatch, rays, deltas)
    ref = [reference_limits(rays, deltas, v) for v in vs]
    g = None
    if guess_kind == "exact":
        g = [r.copy() for r in ref]
    elif guess_kind == "off_by_3":
        g = [np.clip(r + 3, -1, v.nx - 1) for r, v in zip(ref, vs)]
    elif guess_kind == "far":
        g = [np.full_like(r, v.nx - 1) for r, v in zip(ref, vs)]
    elif guess_kind == "origin":
        g = [np.full_like(r, -1) for r in ref]
    for kary, ladder in ((1, ()), (4, ()), (16, ()), (16, (-12, -4, 4, 12)), (2, (-1, 3))):
        lim, effort, rounds, _, first = learned.index_search(None, vs, 10, rays, deltas, g, kary=kary, ladder=ladder)
        for a, b in zip(lim, ref):
            assert np.array_equal(a, b), (kary, ladder)
    # a sweep of the brackets from the second round settles every ray in it
    for sweep in (0, 1, 2):
        lim, effort, rounds, _, first = learned.index_search(None, vs, 10, rays, deltas, g, kary=16,
                                                             ladder=(-4, 4), sweep_from=sweep)
        for a, b in zip(lim, ref):
            assert np.array_equal(a, b), sweep
        assert rounds <= sweep + 1
    if guess_kind == "exact":
        assert rounds == 1 and first == 1.0
    if guess_kind == "none":
        assert rounds <= int(np.ceil(np.log2(146))) + 1


def test_a_good_guess_costs_fewer_points_than_bisection(monkeypatch):
    rays = da.Rays(lines=41, xmax=0.01, ymax=0.01)
    deltas = tuple(np.linspace(-0.03, 0.03, 11))
    rng = np.random.default_rng(1)
    vs = [FakeVariant(145, rng.uniform(0.2, 0.9, 11)) for _ in range(4)]
    install(monkeypatch, rays, deltas)
    ref = [reference_limits(rays, deltas, v) for v in vs]
    near = [np.clip(r + rng.integers(-1, 2, r.shape), -1, 144) for r in ref]
    _, e_guess, _, _, _ = learned.index_search(None, vs, 10, rays, deltas, near, kary=4)
    _, e_bis, _, _, _ = learned.index_search(None, vs, 10, rays, deltas, None)
    assert e_guess < 0.6 * e_bis


def test_sweep_reports_the_first_loss_like_the_grid(monkeypatch):
    """on a ray with a hole (a lost point below survivors) the sweep of a
    bracket finds the first loss inside it, as the plain grid's outward scan
    does"""
    rays = da.Rays(lines=3, xmax=0.01, ymax=0.01)
    deltas = (0.0,)
    v = FakeVariant(20, np.array([0.8]))
    frac = np.arange(20) / 19
    hole = (frac[5] * rays.xmax * rays.dxf[1], frac[5] * rays.ymax * rays.dyf[1])
    install(monkeypatch, rays, deltas, holes=[hole])
    # on the middle ray the guess (12) and guess + 1 survive, the ladder's
    # -8 point (index 4) survives and its +4 point (16) is lost; round two
    # sweeps (13, 16), and the bracket below the guess is never reopened
    lim, _, rounds, _, _ = learned.index_search(None, [v], 10, rays, deltas, [np.full((1, 3), 12)], kary=16,
                                                ladder=(-8, 4), sweep_from=1)
    assert rounds <= 2
    assert lim[0][0].tolist() == [15, 15, 15]
    # with no guess the sweep from the first round is the plain grid itself
    lim, _, rounds, _, _ = learned.index_search(None, [v], 10, rays, deltas, None, sweep_from=0)
    assert rounds == 1 and lim[0][0].tolist() == [15, 4, 15]


def test_grid_limits_and_sequential_cost():
    surv = np.array([[1, 1, 0, 1], [0, 1, 1, 1], [1, 1, 1, 1]], dtype=bool)
    lp = np.where(surv, -1, np.array([[0, 0, 3, 0], [5, 0, 0, 0], [0, 0, 0, 0]]))
    lim, cost = studies.grid_limits(surv, lp, 1, 3, 4, 10)
    assert lim.tolist() == [[1, -1, 3]]
    # passes completed: ray 0: 10 + 10 + 3; ray 1: 5; ray 2: 4 x 10
    assert cost == 23 + 5 + 40


def test_non_monotone_ray_keeps_a_confirmed_boundary(monkeypatch):
    """a hole inside the aperture: the plain grid stops at the hole, the
    searches report some confirmed survivor/loss pair"""
    rays = da.Rays(lines=3, xmax=0.01, ymax=0.01)
    deltas = (0.0,)
    v = FakeVariant(33, np.array([0.8]))
    hx, hy = rays.xy(np.full(3, 0.25))
    install(monkeypatch, rays, deltas, holes=[(hx[1], hy[1])])
    lim, _, _, _, _ = learned.index_search(None, [v], 10, rays, deltas, None)
    assert lim[0][0, 1] in (7, int(0.8 * 32))
