"""Error-seed dynamic-aperture studies over many lattice variants.

The workload: variants of the 40-cell storage ring, each with seeded FSE,
misalignment and roll errors on every KQUAD and KSEXT (relative strength
sigma 1e-3, 30 um, 100 urad), and each carrying its own aperture grid of
41 rays x 11 momentum offsets in [-3 %, +3 %], with a log-uniform number of
particles between 1,024 and 65,536 drawn by seed (points per ray =
particles // (rays x offsets), at least 2).  All variants run as one batch.
"""
import math

import numpy as np

from . import da
from . import errors
from . import lattices as Lt

RAYS = 41
DELTAS = tuple(np.linspace(-0.03, 0.03, 11))


class Variant:
    __slots__ = ("seed", "beamline", "Po", "features", "particles", "nx")


def grid_particles(seed, lo=1024, hi=65536):
    """log-uniform particle count for a variant's aperture grid"""
    u = Lt.Rng(0xDA000000 + seed).uniform(1)[0]
    return int(round(math.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))))


def variants(seeds, lattice_seed=1, rays=RAYS, deltas=DELTAS):
    bl, Po = Lt.storage_ring(seed=lattice_seed)
    out = []
    for s in seeds:
        v = Variant()
        v.seed = s
        v.beamline, v.features = errors.perturb(bl, s)
        v.Po = Po
        v.particles = grid_particles(s)
        v.nx = max(2, v.particles // (rays * len(deltas)))
        out.append(v)
    return out


def grid_rows(rays, deltas, nx):
    """aperture grid rows (n, 8): for each offset, each ray, points
    index*dx outward from the origin (index 0..nx-1)"""
    frac = np.arange(nx) / (nx - 1)
    fx = np.broadcast_to(frac, (rays.lines, nx))
    x = (fx * rays.xmax * rays.dxf[:, None]).ravel()
    y = (fx * rays.ymax * rays.dyf[:, None]).ravel()
    nd = len(deltas)
    return da._rows(np.tile(x, nd), np.tile(y, nd), np.repeat(np.asarray(deltas, dtype=float), x.size), 1)


def grid_limits(surv, loss_pass, nd, lines, nx, passes):
    """limit index per (offset, ray): last point before the first lost one
    (-1 when the origin is lost), and the particle-turns a sequential
    outward search (aperture_search.c:1019-1120 with no splits) spends to
    find it: every point up to and including the first lost one, each for
    the passes it completes, as elegant's effort counts survivors at the end
    of each pass (do_tracking.c:2758-2765)."""
    s = surv.reshape(nd, lines, nx)
    lp = loss_pass.reshape(nd, lines, nx)
    cost = np.where(s, passes, lp)
    first = np.where(s.all(axis=2), nx, np.argmax(~s, axis=2))
    upto = np.arange(nx)[None, None, :] <= first[:, :, None]
    return first - 1, int(np.sum(np.where(upto, cost, 0)))


def areas(rays, nx, limits):
    """DA area per offset from limit indices (zero when the origin is lost)"""
    out = np.zeros(limits.shape[0])
    for di in range(limits.shape[0]):
        if np.all(limits[di] >= 0):
            out[di] = da.area(rays, limits[di] / (nx - 1))
    return out


def track_sets(tracker, vs, sets, passes, tune=None):
    """one batched run: rows sets[k] on variant vs[k]; returns (loss pass
    per row for each variant (-1 = survived), effort, wall seconds, tuning)"""
    import time
    from . import native
    b = native.Batch(tracker)
    try:
        for v, rows in zip(vs, sets):
            b.add_variant(v.beamline, v.Po, rows)
        if tune == "auto":
            b.autotune(4)
        elif tune is not None:
            b.set_tuning(*tune)
        t0 = time.perf_counter()
        st = b.track(0, passes)
        wall = time.perf_counter() - t0
        lps = [b.loss_passes(k) for k in range(len(vs))]
        return lps, int(st.effort), wall, b.tuning()
    finally:
        b.close()


def plain_grid(tracker, vs, passes, rays=None, deltas=DELTAS, tune=None):
    """every variant's full aperture grid in one batch; returns per-variant
    limit indices (offsets x rays), survival flags per grid point, areas,
    the batched effort, the sequential-search effort and wall seconds"""
    rays = rays or da.Rays(RAYS)
    sets = [grid_rows(rays, deltas, v.nx) for v in vs]
    lps, effort, wall, tun = track_sets(tracker, vs, sets, passes, tune)
    out = {"limits": [], "survived": [], "areas": [], "effort": effort, "effort_sequential": 0, "wall": wall,
           "tuning": tun}
    for v, lp in zip(vs, lps):
        lim, cost = grid_limits(lp < 0, lp, len(deltas), rays.lines, v.nx, passes)
        out["limits"].append(lim)
        out["survived"].append((lp < 0).reshape(len(deltas), rays.lines, v.nx))
        out["areas"].append(areas(rays, v.nx, lim))
        out["effort_sequential"] += cost
    return out
