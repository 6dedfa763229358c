"""Seeded lattice errors for error-seed studies.

The errors come from the library's restatement of elegant's error
generator: perturbation() (error.c:496-534) with Gaussian errors cut at
`cutoff` sigma (gauss_rn_lim), drawing from the DLARAN uniform stream of
random_1_elegant (drand_oag.c:94-180; generators.Rng, host/eh_gen.c).  As
in an elegant error study, variant k is step k of one run: the run's stream
is seeded once (RUN_SEED, 987654321, the seed elegant uses without
permuting its bits) and each step draws its errors after those of the steps
before it.  The errors of a step are asserted as four error_element
definitions in this order: FSE, then DX, then DY, then TILT, each over
every KQUAD and KSEXT in beamline order, one draw per element occurrence
(error.c, assert_errors).  The Gaussian deviate itself follows
the published polar method that the SDDS toolkit's gauss_rn implements; the
toolkit is not part of the reference, so the sequences are elegant's where
the reference ships the code and parity-unpinned for the Gaussian step.
"""
import numpy as np

from . import _abi as A
from . import elements as E
from . import generators as G

QUAD_F = ("QF1", "QF2", "QF3", "QF4")
QUAD_D = ("QD1", "QD2")
SEXT_F = ("SF1", "SF2")
SEXT_D = ("SD1", "SD2")
FAMILIES = (QUAD_F, QUAD_D, SEXT_F, SEXT_D)
RUN_SEED = 987654321

# stream states at the start of each step, per (run seed, multipoles per
# step, cutoff): the draws of a step depend on nothing else
_STEPS = {}


def _stream_at(run_seed, step, n_mult, cutoff):
    states = _STEPS.setdefault((run_seed, n_mult, cutoff), [G.Rng(run_seed)])
    while len(states) <= step:
        r = states[-1].copy()
        r.perturbation(1.0, cutoff, G.GAUSSIAN, n=4 * n_mult)
        states.append(r)
    return states[step].copy()


def perturb(beamline, seed, fse=1e-3, dxy=30e-6, tilt=100e-6, cutoff=3.0, rf_reference=0.0, run_seed=RUN_SEED):
    """A copy of the beamline with the errors of step `seed` of the error
    run (FSE, DX, DY and TILT on every KQUAD and KSEXT), plus the 16
    error-summary features (per family F/D quadrupoles and F/D sextupoles:
    RMS of FSE, DX, DY, TILT).  RFCA elements get T_REFERENCE = rf_reference
    (None keeps their fiducial mode).  A study that tracks a variant's grid
    in several rounds needs one phase per variant in every round: with the
    cavity at the end of a ring of h RF wavelengths, T_REFERENCE = 0 is the
    phase the on-orbit particle's TMEAN fiducial gives, modulo 2 pi."""
    els = [A.eh_element.from_buffer_copy(e) for e in beamline.elements]
    mults = [c for c in els if c.type in (A.EH_T_KQUAD, A.EH_T_KSEXT)]
    rng = _stream_at(run_seed, seed, len(mults), cutoff)

    def par(c):
        return c.p.kquad if c.type == A.EH_T_KQUAD else c.p.ksext

    for field, amp in (("fse", fse), ("dx", dxy), ("dy", dxy), ("tilt", tilt)):
        for c, v in zip(mults, rng.perturbation(amp, cutoff, G.GAUSSIAN, n=len(mults))):
            p = par(c)
            setattr(p, field, getattr(p, field) + v)
    per_family = {k: [] for k in range(4)}
    for c in els:
        if c.type in (A.EH_T_KQUAD, A.EH_T_KSEXT):
            p = par(c)
            name = c.name.decode()
            for k, fam in enumerate(FAMILIES):
                if name in fam:
                    per_family[k].append((p.fse, p.dx, p.dy, p.tilt))
        elif c.type == A.EH_T_RFCA and rf_reference is not None:
            c.p.rfca.tReference = rf_reference
    feats = []
    for k in range(4):
        arr = np.array(per_family[k]) if per_family[k] else np.zeros((1, 4))
        feats.extend(np.sqrt(np.mean(arr ** 2, axis=0)).tolist())
    scale = np.array([fse, dxy, dxy, tilt] * 4)
    out = E.Beamline(els, keep=beamline.keep)
    return out, np.array(feats) / scale
