"""Learned dynamic-aperture boundary proposals.

A perceptron (19 -> 256 -> 256 -> 256 -> 1, trained and evaluated by the
library's own GPU kernels) maps (x0, y0, delta) and a variant's 16 error
summaries to a survival probability.  Its proposed boundary on each ray is
then confirmed by FP64 tracking of two points straddling it; rays where the
proposal is wrong fall back to bisection.  The FP64 tracker is the judge of
every boundary; the network only decides where to look.
"""
import ctypes as C

import numpy as np

from . import da
from . import native

PF = C.POINTER(C.c_float)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class MLP:
    def __init__(self, tracker, n_in=19, hidden=256, seed=1):
        self.tracker = tracker
        self._m = C.c_void_p()
        native.check(tracker._ctx, native.lib().eh_mlp_create(tracker._ctx, n_in, hidden, seed, C.byref(self._m)))
        tracker.adopt(self)
        self.n_in = n_in
        self.hidden = hidden

    def close(self):
        if self._m:
            native.lib().eh_mlp_destroy(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def num_params(self):
        return int(native.lib().eh_mlp_num_params(self._m))

    def params(self):
        p = np.zeros(self.num_params(), dtype=np.float32)
        native.check(self.tracker._ctx, native.lib().eh_mlp_get_params(self._m, p.ctypes.data_as(PF)))
        return p

    def set_params(self, p):
        p = _f(p)
        native.check(self.tracker._ctx, native.lib().eh_mlp_set_params(self._m, p.ctypes.data_as(PF)))

    def gradient(self, X, y):
        X = _f(X)
        y = _f(y)
        g = np.zeros(self.num_params(), dtype=np.float32)
        loss = C.c_float()
        native.check(self.tracker._ctx, native.lib().eh_mlp_gradient(self._m, X.ctypes.data_as(PF),
                                                                     y.ctypes.data_as(PF), X.shape[0],
                                                                     g.ctypes.data_as(PF), C.byref(loss)))
        return g, float(loss.value)

    def train(self, X, y, epochs=8, batch=4096, lr=1e-3, seed=7):
        X = _f(X)
        y = _f(y)
        losses = np.zeros(epochs, dtype=np.float32)
        native.check(self.tracker._ctx, native.lib().eh_mlp_train(self._m, X.ctypes.data_as(PF), y.ctypes.data_as(PF),
                                                                  X.shape[0], epochs, batch, lr, seed,
                                                                  losses.ctypes.data_as(PF)))
        return losses

    def predict(self, X):
        X = _f(X)
        p = np.zeros(X.shape[0], dtype=np.float32)
        native.check(self.tracker._ctx, native.lib().eh_mlp_predict(self._m, X.ctypes.data_as(PF), X.shape[0],
                                                                    p.ctypes.data_as(PF)))
        return p


def unpack(params, n_in=19, hidden=256):
    """views (W1, b1, W2, b2, W3, b3, W4, b4) of the flat parameter array"""
    dims = [n_in, hidden, hidden, hidden, 1]
    out = []
    off = 0
    for l in range(4):
        w = params[off:off + dims[l] * dims[l + 1]].reshape(dims[l + 1], dims[l])
        off += w.size
        b = params[off:off + dims[l + 1]]
        off += b.size
        out += [w, b]
    return out


def grid_features(rays, deltas, nx, feats):
    """network inputs for every grid point of one variant, ordered
    (offset, ray, index): x0/xmax, y0/ymax, delta/0.03, then the variant's
    16 normalized error summaries"""
    frac = np.arange(nx) / (nx - 1)
    nd = len(deltas)
    X = np.empty((nd, rays.lines, nx, 3 + len(feats)), dtype=np.float32)
    X[..., 0] = frac[None, None, :] * rays.dxf[None, :, None]
    X[..., 1] = frac[None, None, :] * rays.dyf[None, :, None]
    X[..., 2] = (np.asarray(deltas, dtype=float) / 0.03)[:, None, None]
    X[..., 3:] = np.asarray(feats, dtype=np.float32)
    return X.reshape(-1, 3 + len(feats))


def training_set(vs, grid, rays, deltas):
    """(X, y): every tracked grid point of the training variants with the
    tracker's own verdict (1 = survived all passes)"""
    Xs = [grid_features(rays, deltas, v.nx, v.features) for v in vs]
    ys = [s.astype(np.float32).ravel() for s in grid["survived"]]
    return np.concatenate(Xs), np.concatenate(ys)


def propose(mlp, vs, rays, deltas, confidence=False):
    """guessed limit index per variant (offsets x rays): the end of the first
    run of points with survival probability >= 0.5 going outward.  With
    confidence, also a mask of the guesses the network is unsure of: the
    probability does not step from >= 0.8 at the guess to <= 0.2 just past
    it."""
    out, unsure = [], []
    for v in vs:
        p = mlp.predict(grid_features(rays, deltas, v.nx, v.features)).reshape(len(deltas), rays.lines, v.nx)
        bad = p < 0.5
        g = np.where(bad.any(axis=2), np.argmax(bad, axis=2), v.nx) - 1
        out.append(g)
        if confidence:
            pin = np.take_along_axis(p, np.maximum(g, 0)[..., None], axis=2)[..., 0]
            pout = np.take_along_axis(p, np.minimum(g + 1, v.nx - 1)[..., None], axis=2)[..., 0]
            sure = ((g < 0) | (pin >= 0.8)) & ((g + 1 >= v.nx) | (pout <= 0.2))
            unsure.append(~sure)
    return (out, unsure) if confidence else out


def index_search(tracker, vs, passes, rays, deltas, guess=None, tune=None, kary=1, max_rounds=64, ladder=(),
                 ladder_rays=None, sweep_from=None):
    """Find, on every ray of every (variant, offset), a grid index lo that
    survives with lo + 1 lost (or lo = nx - 1 surviving, or lo = -1 with the
    origin lost), assuming survival is monotone along the ray.  Each round is
    one batched run of every unresolved ray of every variant.  A ray with the
    open interval (lo, hi) of untested indices tracks up to `kary` points
    spread evenly over it; kary = 1 is bisection (the interval splitting of
    aperture_search.c:1027-1050 with split_fraction = 0.5, on grid indices).
    With a guess, the first round tracks only guess and guess + 1, which
    settles the ray when the guess is right; a refuted guess then tracks up
    to `kary` points stepping away from the refuted side in doubling steps,
    so points beyond the boundary (lost within a few passes, cheap) rather
    than inside it (tracked for every pass) take most of the tests.  `ladder`
    adds points at these offsets from the guess to the first round: they
    bracket a wrong guess at once, trading particle-turns for rounds (each
    round costs the latency of `passes` passes however few particles it
    holds); `ladder_rays` (per variant, offsets x rays) limits it to the
    rays whose guess is uncertain.  From round `sweep_from` on, every
    untested index of an unsettled ray is tracked at once, which settles it in
    that round: the search then costs at most sweep_from + 1 rounds, each the
    latency of `passes` passes, and the sweep reports the first loss inside
    the bracket as the plain grid does.  Returns (limits per variant, effort,
    rounds, wall seconds, fraction of rays settled by the first round)."""
    nd, nl = len(deltas), rays.lines
    nx = np.concatenate([np.full(nd * nl, v.nx) for v in vs])
    R = nx.size
    didx = np.tile(np.repeat(np.arange(nd), nl), len(vs))
    lidx = np.tile(np.arange(nl), nd * len(vs))
    lo = np.full(R, -1)
    hi = nx.copy()
    g = None if guess is None else np.clip(np.concatenate([gv.ravel() for gv in guess]), -1, nx - 1)
    # rays that get the ladder (all of them unless a mask of unsure guesses is given)
    lm = np.ones(R, dtype=bool) if ladder_rays is None else np.concatenate([m.ravel() for m in ladder_rays])
    step = np.ones(R, dtype=np.int64)
    effort = 0
    wall = 0.0
    rounds = 0
    settled_first = None
    dl = np.asarray(deltas, dtype=float)
    from .studies import track_sets
    while rounds < max_rounds:
        active = hi - lo > 1
        if not active.any():
            break
        if sweep_from is not None and rounds >= sweep_from:
            gap = hi - lo - 1                      # untested indices lo+1 .. hi-1
            j = np.arange(1, int(gap[active].max()) + 1)[None, :]
            T = np.where(active[:, None] & (j <= gap[:, None]), lo[:, None] + j, -1)
        elif rounds == 0 and g is not None:
            offs = (0, 1) + tuple(ladder)
            T = np.stack([np.where(active & (g + o >= 0) & (g + o < nx) & (lm | (k < 2)), g + o, -1)
                          for k, o in enumerate(offs)], axis=1)
            # a ladder point that clips onto another test is dropped
            for a in range(1, len(offs)):
                for b2 in range(a):
                    T[:, a] = np.where(T[:, a] == T[:, b2], -1, T[:, a])
        else:
            gap = hi - lo - 1                      # untested indices lo+1 .. hi-1
            k = np.minimum(gap, kary)
            j = np.arange(1, kary + 1)[None, :]
            T = lo[:, None] + (j * (gap[:, None] + 1)) // (k[:, None] + 1)
            if g is not None:
                # a refuted guess: a window of points next to the refuted side,
                # its spacing growing by (window + 1) each round until the
                # other side is found.  Upward the window holds `kary` points
                # (points past the boundary are lost early and cost little);
                # downward only 2 (points below the boundary survive every pass)
                up = (lo > g) & (hi == nx)   # no loss found above the guess yet
                dn = (hi <= g) & (lo < 0)    # no survivor found below it yet
                w = np.where(up, kary, np.minimum(kary, 2))
                off = step[:, None] * j
                T = np.where(up[:, None], np.minimum(lo[:, None] + off, hi[:, None] - 1), T)
                T = np.where(dn[:, None], np.maximum(hi[:, None] - off, lo[:, None] + 1), T)
                T = np.where((up | dn)[:, None] & (j > w[:, None]), -1, T)
                step = np.where(up | dn, step * (w + 1), step)
                # steps that clip to the interval's end repeat it; keep one
                T = np.where((j > 1) & (T == np.roll(T, 1, axis=1)), -1, T)
            T = np.where(active[:, None] & (j <= k[:, None]), T, -1)
        sets, where = [], []
        for v in range(len(vs)):
            sel = slice(v * nd * nl, (v + 1) * nd * nl)
            r, c = np.nonzero(T[sel] >= 0)
            m = r + v * nd * nl
            t = T[m, c]
            fr = t / (nx[m] - 1)
            sets.append(da._rows(fr * rays.xmax * rays.dxf[lidx[m]], fr * rays.ymax * rays.dyf[lidx[m]],
                                 dl[didx[m]], 1))
            where.append((m, t))
        keep = [v for v in range(len(vs)) if sets[v].shape[0]]
        lps, e, w, _ = track_sets(tracker, [vs[v] for v in keep], [sets[v] for v in keep], passes, tune)
        effort += e
        wall += w
        rounds += 1
        ms, ts, ss = [], [], []
        for j, v in enumerate(keep):
            m, t = where[v]
            ms.append(m)
            ts.append(t)
            ss.append(lps[j] < 0)
        m = np.concatenate(ms)
        t = np.concatenate(ts)
        sv = np.concatenate(ss)
        # losses first: the first loss bounds the ray; then the survivors below it
        np.minimum.at(hi, m[~sv], t[~sv])
        ok = sv & (t < hi[m])
        np.maximum.at(lo, m[ok], t[ok])
        # a survivor recorded earlier above a new loss says nothing about the first loss below it
        lo[lo >= hi] = -1
        if settled_first is None:
            settled_first = float(np.mean(hi - lo <= 1))
    limits = [lo[v * nd * nl:(v + 1) * nd * nl].reshape(nd, nl) for v in range(len(vs))]
    return limits, effort, rounds, wall, settled_first

def study(tracker, train_seeds, eval_seeds, passes=1024, epochs=6, batch=4096, lr=1e-3, tune=None, log=None,
          rays=None, deltas=None, kary=4, ladder=(-12, -4, 4, 12)):
    """Train on the survival verdicts of the training variants' full grids,
    then find the aperture of the evaluation variants three ways: the plain
    grid (the reference answer), bisection alone, and the network's guess
    confirmed by tracking with galloping/bisection fallback.  Returns a
    report of particle-turns, wall time and area agreement."""
    import time
    from . import studies
    rays = rays or da.Rays(studies.RAYS)
    deltas = studies.DELTAS if deltas is None else deltas
    say = log or (lambda *a: None)
    tv = studies.variants(train_seeds)
    ev = studies.variants(eval_seeds)
    t0 = time.perf_counter()
    gtrain = studies.plain_grid(tracker, tv, passes, rays, deltas, tune)
    say("training grids tracked: %d particle-turns in %.1f s" % (gtrain["effort"], gtrain["wall"]))
    X, y = training_set(tv, gtrain, rays, deltas)
    mlp = MLP(tracker, n_in=X.shape[1])
    t1 = time.perf_counter()
    losses = mlp.train(X, y, epochs=epochs, batch=min(batch, X.shape[0]), lr=lr)
    t_train = time.perf_counter() - t1
    say("trained on %d samples: loss %.4f -> %.4f in %.1f s" % (X.shape[0], losses[0], losses[-1], t_train))
    t_label = time.perf_counter() - t0

    gref = studies.plain_grid(tracker, ev, passes, rays, deltas, tune)
    say("evaluation grids: %d particle-turns in %.1f s" % (gref["effort"], gref["wall"]))
    t2 = time.perf_counter()
    guess, unsure = propose(mlp, ev, rays, deltas, confidence=True)
    t_prop = time.perf_counter() - t2
    lim_l, eff_l, rounds_l, wall_l, first_l = index_search(tracker, ev, passes, rays, deltas, guess, tune,
                                                           kary=kary)
    say("learned search: %d particle-turns, %d rounds, %.1f s" % (eff_l, rounds_l, wall_l))
    lim_f, eff_f, rounds_f, wall_f, first_f = index_search(tracker, ev, passes, rays, deltas, guess, tune,
                                                           kary=16, ladder=ladder, ladder_rays=unsure)
    say("learned search, ladder %s: %d particle-turns, %d rounds, %.1f s" % (ladder, eff_f, rounds_f, wall_f))
    lim_s, eff_s, rounds_s, wall_s, first_s = index_search(tracker, ev, passes, rays, deltas, guess, tune,
                                                           kary=16, ladder=ladder, ladder_rays=unsure, sweep_from=1)
    say("learned search, ladder %s, then a sweep of the brackets: %d particle-turns, %d rounds, %.1f s" % (
        ladder, eff_s, rounds_s, wall_s))
    lim_b, eff_b, rounds_b, wall_b, _ = index_search(tracker, ev, passes, rays, deltas, None, tune)
    say("bisection search: %d particle-turns, %d rounds, %.1f s" % (eff_b, rounds_b, wall_b))
    mlp.close()

    def compare(lims):
        same = np.mean([np.mean(a == b) for a, b in zip(lims, gref["limits"])])
        errs = []
        for v, lim, aref in zip(ev, lims, gref["areas"]):
            a = studies.areas(rays, v.nx, lim)
            for x, r in zip(a, aref):
                if r > 0:
                    errs.append(abs(x - r) / r)
                elif x > 0:
                    errs.append(1.0)
        errs = np.array(errs) if errs else np.zeros(1)
        return {"rays_identical": float(same), "area_rel_err_mean": float(errs.mean()),
                "area_rel_err_max": float(errs.max())}

    guess_hit = np.mean([np.mean(a == b) for a, b in zip(guess, gref["limits"])])
    maps = len(ev) * len(deltas)
    rep = {
        "variants_train": len(tv), "variants_eval": len(ev), "passes": passes,
        "grid_points_eval": int(sum(len(deltas) * rays.lines * v.nx for v in ev)),
        "training": {"samples": int(X.shape[0]), "effort": gtrain["effort"], "track_s": gtrain["wall"],
                     "train_s": t_train, "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
                     "total_s": t_label},
        "plain_grid": {"effort_batched": gref["effort"], "effort_sequential": gref["effort_sequential"],
                       "wall_s": gref["wall"], "maps_per_hour": maps / gref["wall"] * 3600},
        "bisection": dict(effort=eff_b, rounds=rounds_b, wall_s=wall_b, maps_per_hour=maps / wall_b * 3600,
                          **compare(lim_b)),
        "learned": dict(effort=eff_l, rounds=rounds_l, wall_s=wall_l + t_prop, propose_s=t_prop,
                        maps_per_hour=maps / (wall_l + t_prop) * 3600,
                        maps_per_hour_incl_training=maps / (wall_l + t_prop + gtrain["wall"] + t_train) * 3600,
                        guess_exact=float(guess_hit),
                        settled_first_round=first_l, **compare(lim_l)),
        "learned_ladder": dict(ladder=list(ladder), unsure_rays=float(np.mean(np.concatenate(
                                   [u.ravel() for u in unsure]))), effort=eff_f, rounds=rounds_f,
                               wall_s=wall_f + t_prop,
                               maps_per_hour=maps / (wall_f + t_prop) * 3600,
                               maps_per_hour_incl_training=maps / (wall_f + t_prop + gtrain["wall"] + t_train) * 3600,
                               settled_first_round=first_f, **compare(lim_f)),
        "learned_sweep": dict(ladder=list(ladder), sweep_from=1, effort=eff_s, rounds=rounds_s,
                              wall_s=wall_s + t_prop, maps_per_hour=maps / (wall_s + t_prop) * 3600,
                              maps_per_hour_incl_training=maps / (wall_s + t_prop + gtrain["wall"] + t_train) * 3600,
                              settled_first_round=first_s, **compare(lim_s)),
    }
    return rep
