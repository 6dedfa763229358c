"""Batched dynamic-aperture search over lattice variants and momentum offsets.

Restates the n-line search of do_aperture_search_line
(aperture_search.c:910-1240): rays in the upper half plane at angles
-pi/2 + k dtheta (dxFactor = sin, dyFactor = cos, values below 1e-6 set to 0),
points at index*dx along each ray, the aperture on a ray is the last point
before the first lost one, and the area comes from trimApertureSearchResult.
Instead of one do_tracking call per point, every point of every (variant,
momentum offset, ray) is tracked in one batched GPU run; the bisection and
learned searches run one batched round per refinement step.
"""
import math

import numpy as np

from . import native


class Rays:
    def __init__(self, lines=41, xmax=0.015, ymax=0.015):
        self.lines = lines
        self.xmax = xmax
        self.ymax = ymax
        dtheta = math.pi / (lines - 1)
        k = np.arange(lines)
        self.dxf = np.sin(-math.pi / 2 + dtheta * k)
        self.dyf = np.cos(-math.pi / 2 + dtheta * k)
        self.dxf[np.abs(self.dxf) < 1e-6] = 0
        self.dyf[np.abs(self.dyf) < 1e-6] = 0

    def xy(self, frac):
        """points at fraction frac (0..1) of the full extent along each ray;
        frac broadcasts against the ray index"""
        return frac * self.xmax * self.dxf, frac * self.ymax * self.dyf


def _rows(x, y, delta, first_id):
    n = x.size
    r = np.zeros((n, 8))
    r[:, 0] = x
    r[:, 2] = y
    r[:, 5] = delta
    r[:, 6] = np.arange(first_id, first_id + n)
    r[:, 7] = -1
    return r


def area(rays, frac_limit):
    """area from the last surviving fraction on each ray (trimApertureSearchResult)"""
    x, y = rays.xy(frac_limit)
    a, _, _ = native.da_trim_area(x, y, rays.dxf, rays.dyf)
    return a


def run_points(tracker, variants, point_sets, passes, tune=None):
    """Track, for every variant v, the rows point_sets[v] (array (n, 8)) for
    `passes` passes in one batch.  Returns (survived flags per variant,
    effort, stats)."""
    b = native.Batch(tracker)
    for (bl, Po), rows in zip(variants, point_sets):
        b.add_variant(bl, Po, rows)
    if tune == "auto":
        b.autotune(4)
    elif tune is not None:
        b.set_tuning(*tune)
    st = b.track(0, passes)
    out = [b.loss_passes(v) < 0 for v in range(len(variants))]
    tun = b.tuning()
    b.close()
    return out, int(st.effort), st, tun


def plain_grid(tracker, variants, rays, deltas, nx, passes, tune=None):
    """Uniform grid of nx points per ray (index 0 = origin).  Returns
    {limit fractions [variant][delta][line], areas [variant][delta],
    effort (batched, every point), effort_reference (points up to the first
    loss on each ray, as the sequential search tracks them)}."""
    nd = len(deltas)
    frac = np.arange(nx) / (nx - 1)
    sets = []
    for v in range(len(variants)):
        fx, fl = np.meshgrid(frac, np.arange(rays.lines))  # (lines, nx)
        xs, ys = [], []
        for d in deltas:
            x = fx * rays.xmax * rays.dxf[fl]
            y = fx * rays.ymax * rays.dyf[fl]
            xs.append(x.ravel())
            ys.append(y.ravel())
        x = np.concatenate(xs)
        y = np.concatenate(ys)
        dl = np.repeat(np.asarray(deltas, dtype=float), rays.lines * nx)
        sets.append(_rows(x, y, dl, 1))
    surv, effort, st, tun = run_points(tracker, variants, sets, passes, tune)
    limits = np.zeros((len(variants), nd, rays.lines))
    areas = np.zeros((len(variants), nd))
    ref_effort = 0
    for v in range(len(variants)):
        s = surv[v].reshape(nd, rays.lines, nx)
        for di in range(nd):
            for li in range(rays.lines):
                lost = np.nonzero(~s[di, li])[0]
                first = lost[0] if lost.size else nx
                limits[v, di, li] = frac[first - 1] if first > 0 else -1.0 / (nx - 1)
                ref_effort += passes * min(first + 1, nx)  # upper bound: the lost point counted fully
            origin = s[di, :, 0].all()
            areas[v, di] = area(rays, np.maximum(limits[v, di], 0)) if origin else 0.0
    return {"limits": limits, "areas": areas, "effort": effort, "effort_reference": ref_effort, "stats": st,
            "tuning": tun}


def refine(tracker, variants, rays, deltas, lo, hi, passes, tol, guess=None, max_rounds=20, tune=None):
    """Per-ray bisection between a known survivor fraction lo and a known (or
    assumed) loss fraction hi, batched over all variants, offsets and rays.
    With a guess (fractions from a surrogate), the first round tests
    guess -+ tol and narrows the interval to it when both are confirmed.
    Returns (lo, hi, effort, rounds)."""
    nv, nd, nl = lo.shape
    lo = lo.copy()
    hi = hi.copy()
    effort = 0
    rounds = 0
    if guess is not None:
        a = np.clip(guess - tol, 0, 1)
        b = np.clip(guess + tol, 0, 1)
        both = np.stack([a, b], axis=-1)  # (nv, nd, nl, 2)
        sets = []
        for v in range(nv):
            x, y = rays.xy(both[v].transpose(0, 2, 1))  # (nd, 2, nl) broadcast against ray index
            dl = np.repeat(np.asarray(deltas, dtype=float), 2 * nl)
            sets.append(_rows(x.ravel(), y.ravel(), dl, 1))
        surv, e, st, _ = run_points(tracker, variants, sets, passes, tune)
        effort += e
        rounds += 1
        for v in range(nv):
            s = surv[v].reshape(nd, 2, nl)
            sa, sb = s[:, 0, :], s[:, 1, :]
            # a survived: raise lo; b lost: lower hi
            lo[v] = np.where(sa, np.maximum(lo[v], a[v]), lo[v])
            hi[v] = np.where(~sa, np.minimum(hi[v], a[v]), hi[v])
            lo[v] = np.where(sb, np.maximum(lo[v], b[v]), lo[v])
            hi[v] = np.where(~sb, np.minimum(hi[v], b[v]), hi[v])
    while rounds < max_rounds:
        active = (hi - lo) > 2 * tol + 1e-12
        if not active.any():
            break
        mid = 0.5 * (lo + hi)
        sets = []
        idx = []
        for v in range(nv):
            m = np.nonzero(active[v])
            fr = mid[v][m]
            li = m[1]
            x = fr * rays.xmax * rays.dxf[li]
            y = fr * rays.ymax * rays.dyf[li]
            dl = np.asarray(deltas, dtype=float)[m[0]]
            sets.append(_rows(x, y, dl, 1) if x.size else _rows(np.zeros(0), np.zeros(0), np.zeros(0), 1))
            idx.append(m)
        keep = [k for k in range(nv) if sets[k].shape[0] > 0]
        surv, e, st, _ = run_points(tracker, [variants[k] for k in keep], [sets[k] for k in keep], passes, tune)
        effort += e
        rounds += 1
        for j, v in enumerate(keep):
            m = idx[v]
            s = surv[j]
            lo[v][m] = np.where(s, mid[v][m], lo[v][m])
            hi[v][m] = np.where(~s, mid[v][m], hi[v][m])
    return lo, hi, effort, rounds
