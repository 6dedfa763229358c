"""Batched variants on the GPU: each variant's rows, in elegant's order,
bit-identical to tracking that variant alone on the CPU checker, for every
tuning the autotuner may pick."""
import numpy as np
import pytest

from elegant_amd import errors
from elegant_amd import lattices as Lt
from elegant_amd import native
from oracle import oracle_ctypes as O
from tests.helpers import assert_rows_identical

pytestmark = pytest.mark.gpu


def variants():
    bl, Po = Lt.storage_ring(seed=2)
    out = []
    for seed, n in [(1, 700), (2, 300), (3, 1100)]:
        v, feats = errors.perturb(bl, seed)
        rows = Lt.matched_beam(v, n, 1e-9, 1e-11, 1e-3, 10e-12, 100 + seed, Po)
        rows[: n // 4, 0] += 0.0135  # halo near the 15 mm aperture
        rows[: n // 8, 5] += 0.02
        out.append((v, Po, rows))
    return out


@pytest.mark.parametrize("tuning", [(1, 256, 64, 1, 0), (2, 64, 1, 0, 0), (2, 128, 2, 1, 0), (1, 64, 4, 1, 1),
                                    (2, 64, 2, 0, 1), (1, 128, 4, 1, 2), (2, 64, 2, 0, 2), (1, 256, 1, 1, 2)])
def test_batch_bit_exact_per_variant(tracker, tuning):
    """compact 1 rebuilds the queue from the survivors on the host between
    launches (1, 2, ... passes per launch), compact 2 on the device inside
    the captured graph (use_graph 1) or between plain launches"""
    vs = variants()
    b = native.Batch(tracker)
    for bl, Po, rows in vs:
        b.add_variant(bl, Po, rows)
    b.set_tuning(*tuning)
    passes = 7 if tuning[4] else 3
    st = b.track(0, passes)
    total_lost = 0
    for k, (bl, Po, rows) in enumerate(vs):
        g, ng = b.download(k)
        c = rows.copy()
        o = O.Oracle("emath")
        nc = o.do_tracking(c, bl, Po, passes)
        assert ng == nc
        total_lost += rows.shape[0] - ng
        assert_rows_identical(g, c, "variant %d" % k)
    assert total_lost > 50
    b.close()


def test_batch_tmean_fiducial_from_the_first_row_that_reaches_the_cavity(tracker):
    """variants keeping the TMEAN fiducial: as in elegant's aperture search,
    where the first grid point tracked alone fixes the cavity's phase, each
    variant's phase comes from its first row that reaches the cavity on the
    first pass (in one variant rows 0 and 1 are lost at the entrance
    aperture, so row 2 sets it); rows bit-identical to the checker with that
    phase"""
    from elegant_amd import _abi as A
    bl0, Po = Lt.storage_ring(seed=2)
    vs = []
    for seed, n in [(4, 500), (5, 400)]:
        v, feats = errors.perturb(bl0, seed, rf_reference=None)
        rows = Lt.matched_beam(v, n, 1e-9, 1e-11, 1e-3, 10e-12, 200 + seed, Po)
        rows[0, :6] = 0.0  # the origin point first, as a grid scan starts
        rows[n // 2:, 0] += 0.0135  # a halo in the later rows
        if seed == 5:
            rows[0:2, 0] = 0.03  # beyond the 15 mm aperture at the ring's start
        vs.append((v, Po, rows))
    b = native.Batch(tracker)
    for bl, Po_, rows in vs:
        b.add_variant(bl, Po_, rows)
    b.track(0, 4)
    for k, (bl, Po_, rows) in enumerate(vs):
        ei = [i for i, e in enumerate(bl.elements) if e.type == A.EH_T_RFCA][0]
        assert bl.elements[ei].p.rfca.tReference == -1
        first = None
        for r in range(rows.shape[0]):
            o = O.Oracle("emath")
            o.do_tracking(rows[r:r + 1].copy(), bl, Po_, 1)  # a copy: the slice is a view
            if not np.isnan(o.fiducial(ei)):
                first, ph = r, o.fiducial(ei)
                break
        assert first == (2 if k == 1 else 0)
        g, ng = b.download(k)
        c = rows.copy()
        o = O.Oracle("emath")
        o.set_fiducial(ei, ph)
        nc = o.do_tracking(c, bl, Po_, 4)
        assert ng == nc and rows.shape[0] - ng > 20
        assert_rows_identical(g, c, "TMEAN variant %d" % k)
    b.close()


def test_autotune_keeps_results(tracker):
    vs = variants()
    b = native.Batch(tracker)
    for bl, Po, rows in vs:
        b.add_variant(bl, Po, rows)
    t = b.autotune(1)
    assert t.particles_per_lane in (1, 2) and t.block in (64, 128, 256) and t.passes_per_launch >= 1
    assert t.compact in (0, 1, 2)
    print("autotune picked ppl %d block %d passes/launch %d graph %d compact %d" % (
        t.particles_per_lane, t.block, t.passes_per_launch, t.use_graph, t.compact))
    b.track(0, 2)
    bl, Po, rows = vs[1]
    g, ng = b.download(1)
    c = rows.copy()
    nc = O.Oracle("emath").do_tracking(c, bl, Po, 2)
    assert ng == nc
    assert_rows_identical(g, c, "after autotune")
    b.close()


@pytest.mark.parametrize("ppl", [1, 2])
def test_batch_with_shaped_apertures(tracker, ppl):
    """variants whose MAXAMP is elliptical (one with an open side) next to
    a rectangular one: the batch runs on the kernel instance that evaluates
    the shapes, every variant bit-identical to the checker"""
    from elegant_amd import _abi as A
    from elegant_amd import elements as E
    vs = []
    for k, (bl, Po, rows) in enumerate(variants()):
        if k < 2:
            els = list(bl.elements)
            assert els[0].type == A.EH_T_MAXAMP
            els[0] = E.MAXAMP("AP", X_MAX=0.015, Y_MAX=0.008, ELLIPTICAL=1, EXPONENT=2 + 2 * k,
                              OPEN_SIDE="-x" if k else "")
            bl = E.Beamline(els, keep=bl.keep)
        vs.append((bl, Po, rows))
    b = native.Batch(tracker)
    for bl, Po, rows in vs:
        b.add_variant(bl, Po, rows)
    b.set_tuning(ppl, 128, 2, 1, 0)
    b.track(0, 3)
    for k, (bl, Po, rows) in enumerate(vs):
        g, ng = b.download(k)
        c = rows.copy()
        o = O.Oracle("emath")
        nc = o.do_tracking(c, bl, Po, 3)
        assert ng == nc and ng < rows.shape[0]
        assert_rows_identical(g, c, "shaped variant %d" % k)
    b.close()


@pytest.mark.parametrize("compact,graph", [(1, 0), (2, 1), (2, 0)])
def test_batch_many_launches(tracker, compact, graph):
    """a long run in launches of at most one pass (70 launches, the doubling
    schedule capped at once), the queue rebuilt between them by the host or
    on the device: every variant bit-identical to the checker"""
    from elegant_amd import elements as E
    bl, Po, rows = Lt.fodo_workload(n=600, seed=5)
    rows[:200, 0] *= 200.0  # a halo, partly lost over the run
    vs = [(bl, Po, rows[:350].copy()), (bl, Po, rows[350:].copy())]
    b = native.Batch(tracker)
    for v in vs:
        b.add_variant(*v)
    b.set_tuning(1, 64, 1, graph, compact)
    passes = 70
    st = b.track(0, passes)
    assert st.kernel_launches == passes
    lost = 0
    for k, (bl_, Po_, r) in enumerate(vs):
        g, ng = b.download(k)
        c = r.copy()
        nc = O.Oracle("emath").do_tracking(c, bl_, Po_, passes)
        assert ng == nc
        lost += r.shape[0] - ng
        assert_rows_identical(g, c, "variant %d, %d launches" % (k, passes))
    assert lost > 20
    b.close()
