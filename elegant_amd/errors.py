"""Seeded lattice errors for error-seed studies.

Restates the Gaussian branch of perturbation() (error.c:496-503): each error
is amplitude * z with z a unit Gaussian redrawn until |z| <= cutoff
(gauss_rn_lim, default cutoff 3, error.nl:23).  The toolkit's generator
(random_1_elegant / gauss_rn_lim, drand_oag.c) is not vendored, so the draws
come from the project's SplitMix64 stream: seeds are reproducible here but do
not reproduce elegant's sequences.
"""
import numpy as np

from . import _abi as A
from . import elements as E
from .lattices import Rng

QUAD_F = ("QF1", "QF2", "QF3", "QF4")
QUAD_D = ("QD1", "QD2")
SEXT_F = ("SF1", "SF2")
SEXT_D = ("SD1", "SD2")
FAMILIES = (QUAD_F, QUAD_D, SEXT_F, SEXT_D)


def gauss_cut(rng, n, cutoff=3.0):
    out = np.empty(n)
    filled = 0
    while filled < n:
        z = rng.normal(2 * (n - filled) + 8)
        z = z[np.abs(z) <= cutoff]
        take = min(n - filled, z.size)
        out[filled:filled + take] = z[:take]
        filled += take
    return out


def perturb(beamline, seed, fse=1e-3, dxy=30e-6, tilt=100e-6, cutoff=3.0, rf_reference=0.0):
    """A copy of the beamline with FSE, DX, DY and TILT errors on every KQUAD
    and KSEXT, plus the 16 error-summary features (per family F/D quadrupoles
    and F/D sextupoles: RMS of FSE, DX, DY, TILT).  RFCA elements get
    T_REFERENCE = rf_reference so the variant needs no fiducial pass."""
    rng = Rng(0x5EED0000 + seed)
    els = []
    per_family = {k: [] for k in range(4)}
    for e in beamline.elements:
        c = A.eh_element.from_buffer_copy(e)
        if c.type in (A.EH_T_KQUAD, A.EH_T_KSEXT):
            p = c.p.kquad if c.type == A.EH_T_KQUAD else c.p.ksext
            z = gauss_cut(rng, 4, cutoff)
            p.fse += fse * z[0]
            p.dx += dxy * z[1]
            p.dy += dxy * z[2]
            p.tilt += tilt * z[3]
            name = c.name.decode()
            for k, fam in enumerate(FAMILIES):
                if name in fam:
                    per_family[k].append((p.fse, p.dx, p.dy, p.tilt))
        elif c.type == A.EH_T_RFCA and rf_reference is not None:
            c.p.rfca.tReference = rf_reference
        els.append(c)
    feats = []
    for k in range(4):
        arr = np.array(per_family[k]) if per_family[k] else np.zeros((1, 4))
        feats.extend(np.sqrt(np.mean(arr ** 2, axis=0)).tolist())
    scale = np.array([fse, dxy, dxy, tilt] * 4)
    out = E.Beamline(els, keep=beamline.keep)
    return out, np.array(feats) / scale
