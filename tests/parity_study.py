"""Staged tracking and divergence measures for the comparison of the GPU
(whose math is emath's) with the CPU checker built on glibc's libm, the
reference's own math library: the same particles are tracked to a list of
pass counts on both sides and compared particle by particle (by ID).

Particles of these rings are independent apart from the RF cavity's TMEAN
fiducial time, which is fixed on the first pass from all particles
(simple_rfca.c:412-471); the checker's shards are given the fiducial phase
the GPU fixed, so sharding the CPU work over threads changes nothing."""
import threading

import numpy as np

from elegant_amd import _abi as A
from oracle import oracle_ctypes as O

COLS = 6


def rf_index(beamline):
    idx = [i for i, e in enumerate(beamline.elements) if e.type == A.EH_T_RFCA]
    return idx[0] if idx else None


def by_id(rows):
    """rows sorted by particle ID (column 6); IDs are 1..n"""
    out = np.empty_like(rows)
    out[rows[:, 6].astype(np.int64) - 1] = rows
    return out


def staged_gpu(tracker, rows, beamline, Po, checkpoints):
    """the GPU's rows (sorted by ID) after each pass count in checkpoints,
    and the RF fiducial phase it fixed on the first pass"""
    tracker.load(beamline, Po)
    tracker.upload(rows)
    snaps, done = {}, 0
    for cp in checkpoints:
        tracker.track(done, cp - done)
        done = cp
        g, _ = tracker.download()
        snaps[cp] = by_id(g)
    rf = rf_index(beamline)
    phase = tracker.rfca_fiducial(rf) if rf is not None else None
    return snaps, phase


def staged_checker(rows, beamline, Po, checkpoints, mode, phase, threads):
    """the checker (mode "glibc" or "emath") on `threads` shards of rows,
    the RF fiducial phase fixed to `phase`; rows sorted by ID per checkpoint"""
    rf = rf_index(beamline)
    shards = [np.array(s, dtype=np.float64, order="C") for s in np.array_split(rows, threads)]  # copies
    snaps = {cp: [None] * threads for cp in checkpoints}
    errors = []

    def work(k):
        try:
            o = O.Oracle(mode)
            if rf is not None:
                o.set_fiducial(rf, phase)
            sh, done, left = shards[k], 0, shards[k].shape[0]
            for cp in checkpoints:
                if left:
                    left = o.do_tracking(sh[:left], beamline, Po, cp - done, done)
                done = cp
                snaps[cp][k] = sh.copy()
        except Exception as e:  # reported in the calling thread
            errors.append(e)

    O.lib(mode)
    ts = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    if errors:
        raise errors[0]
    return {cp: by_id(np.concatenate(snaps[cp])) for cp in checkpoints}


def divergence(g, c, scale):
    """per particle, for rows alive on both sides: the largest relative
    difference over the six coordinates, |g - c| / |c| (0 where both are 0),
    and the largest difference in units of the beam's initial rms size per
    coordinate; alive mask on each side"""
    ag, ac = g[:, 7] < 0, c[:, 7] < 0
    both = ag & ac
    d = np.abs(g[:, :COLS] - c[:, :COLS])
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.where(d == 0, 0.0, d / np.abs(c[:, :COLS]))
    rel = np.where(both[:, None], rel, 0.0).max(axis=1)
    beam = np.where(both[:, None], d / scale[None, :], 0.0).max(axis=1)
    return rel, beam, ag, ac
