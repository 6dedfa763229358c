"""Benchmark lattices, built programmatically with the reference's element
parameters, plus the linear optics needed to match a beam to them.

Workloads:
  fodo_ring            100 FODO cells (DRIF 2.0, KQUAD 0.5 K1=+1.2 N_KICKS=4,
                       KSEXT 0.2 K2=+20, DRIF 2.0, KQUAD 0.5 K1=-1.2, KSEXT 0.2
                       K2=-35), MAXAMP +-20 mm, p = 6000 MeV/c.
  storage_ring         40 cells of 2 CSBEND (1.0 m, angle 2pi/80, N_KICKS=16,
                       edges on), 6 KQUAD with |K1| drawn from [0.5, 2.0] m^-2 by
                       seed, 4 KSEXT (|K2| <= 60 m^-3), 1 KOCT and drifts; one
                       RFCA at 352 MHz, 6 MV, L=0; MAXAMP +-15 mm.
  storage_ring_wake    storage_ring + CHARGE 5 nC + WAKE (broad-band resonator
                       Green function) + ZLONGIT (same resonator, 8192 bins).
"""
import math

import numpy as np

from . import _abi as A
from . import elements as E

ME_MEV = 0.51099895000
C_MKS = 2.99792458e8


def momentum(p_mev):
    """Po = p/(m c) for electrons."""
    return p_mev / ME_MEV


# ------------------------------------------------------------------ optics

def _drift2(L):
    return np.array([[1.0, L], [0.0, 1.0]])


def _quad2(L, k):
    if k > 0:
        s = math.sqrt(k)
        return np.array([[math.cos(s * L), math.sin(s * L) / s], [-s * math.sin(s * L), math.cos(s * L)]])
    if k < 0:
        s = math.sqrt(-k)
        return np.array([[math.cosh(s * L), math.sinh(s * L) / s], [s * math.sinh(s * L), math.cosh(s * L)]])
    return _drift2(L)


def _sbend2(L, angle):
    rho = L / angle
    th = angle
    return np.array([[math.cos(th), rho * math.sin(th)], [-math.sin(th) / rho, math.cos(th)]])


def linear_maps(beamline):
    """(Mx, My): 2x2 one-turn matrices at the start, linear element optics
    (hard-edge thick quads, sector dipoles with E1=E2=0, drifts elsewhere)."""
    Mx = np.eye(2)
    My = np.eye(2)
    for e in beamline.elements:
        t = e.type
        if t in (A.EH_T_DRIF, A.EH_T_EDRIFT, A.EH_T_KSEXT, A.EH_T_KOCT, A.EH_T_RFCA):
            L = getattr(e.p, A.PARAM_STRUCT[t][0]).length
            mx = my = _drift2(L)
        elif t == A.EH_T_KQUAD:
            q = e.p.kquad
            k = q.k1 * (1 + q.fse)
            mx, my = _quad2(q.length, k), _quad2(q.length, -k)
        elif t == A.EH_T_CSBEND:
            b = e.p.csbend
            if b.angle:
                mx, my = _sbend2(b.length, b.angle), _drift2(b.length)
            else:
                mx = my = _drift2(b.length)
        else:
            continue
        Mx = mx @ Mx
        My = my @ My
    return Mx, My


def twiss_from_map(M):
    """(beta, alpha, tune) of a stable 2x2 one-turn map."""
    c = 0.5 * (M[0, 0] + M[1, 1])
    if abs(c) >= 1:
        raise ValueError("unstable one-turn map (|cos mu| = %.4f)" % abs(c))
    s = math.copysign(math.sqrt(1 - c * c), M[0, 1])
    beta = M[0, 1] / s
    alpha = (M[0, 0] - M[1, 1]) / (2 * s)
    mu = math.atan2(s, c) % (2 * math.pi)
    return beta, alpha, mu / (2 * math.pi)


# -------------------------------------------------------------- lattices

def fodo_ring(cells=100, x_max=0.02):
    els = [E.MAXAMP("AP", X_MAX=x_max, Y_MAX=x_max)]
    for i in range(cells):
        els += [E.DRIF("D1", L=2.0), E.KQUAD("QF", L=0.5, K1=1.2, N_KICKS=4), E.KSEXT("SF", L=0.2, K2=20.0),
                E.DRIF("D2", L=2.0), E.KQUAD("QD", L=0.5, K1=-1.2, N_KICKS=4), E.KSEXT("SD", L=0.2, K2=-35.0)]
    return E.Beamline(els)


class Rng:
    """SplitMix64 stream (Steele, Lea & Flood 2014) -> uniform doubles in [0, 1)."""

    def __init__(self, seed):
        self.state = np.uint64(seed)

    def next_u64(self, n):
        with np.errstate(over="ignore"):
            idx = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + idx * np.uint64(0x9E3779B97F4A7C15)
            self.state = self.state + np.uint64(n) * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n):
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n):
        """Box-Muller pairs."""
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]


def _ring_cell(rng, cell_index, filler):
    """one storage-ring cell: 6 KQUAD, 4 KSEXT, 1 KOCT, 2 CSBEND, 15 drifts"""
    k = 0.5 + 1.5 * rng.uniform(6)            # |K1| in [0.5, 2.0]
    s = 60.0 * (2 * rng.uniform(4) - 1)        # |K2| <= 60
    k3 = 200.0 * (2 * rng.uniform(1)[0] - 1)   # |K3| <= 200 m^-4
    ang = 2 * math.pi / 80
    q = lambda n, kk: E.KQUAD(n, L=0.3, K1=kk)
    b = lambda n: E.CSBEND(n, L=1.0, ANGLE=ang, N_KICKS=16)
    return [
        E.DRIF("D0", L=0.5 + filler), q("QF1", k[0]), E.DRIF("D1", L=0.2), E.KSEXT("SF1", L=0.2, K2=abs(s[0])),
        E.DRIF("D2", L=0.2), q("QD1", -k[1]), E.DRIF("D3", L=0.4), b("B1"), E.DRIF("D4", L=0.4),
        q("QF2", k[2]), E.DRIF("D5", L=0.2), E.KSEXT("SD1", L=0.2, K2=-abs(s[1])), E.DRIF("D6", L=0.2),
        E.KOCT("O1", L=0.1, K3=k3), E.DRIF("D7", L=0.2), E.KSEXT("SD2", L=0.2, K2=-abs(s[2])), E.DRIF("D8", L=0.2),
        q("QF3", k[3]), E.DRIF("D9", L=0.4), b("B2"), E.DRIF("D10", L=0.4), q("QD2", -k[4]), E.DRIF("D11", L=0.2),
        E.KSEXT("SF2", L=0.2, K2=abs(s[3])), E.DRIF("D12", L=0.2), q("QF4", k[5]), E.DRIF("D13", L=0.5 + filler),
    ]


def storage_ring(seed=1, cells=40, x_max=0.015, volt=6e6, freq=352e6, p_mev=6000.0, sext_scale=None):
    """A stable 40-cell ring: quadrupole strengths are drawn by seed and
    re-drawn until the linear one-turn map is stable in both planes; drift
    fillers make the circumference an integer number of RF wavelengths."""
    Po = momentum(p_mev)
    beta = Po / math.sqrt(Po * Po + 1)
    rng = Rng(seed)
    for attempt in range(10000):
        state = rng.state
        cell_rng = Rng(int(state))
        cell = _ring_cell(cell_rng, 0, 0.0)
        bl = E.Beamline(cell)
        Mx, My = linear_maps(bl)
        ok = abs(0.5 * np.trace(Mx)) < 0.95 and abs(0.5 * np.trace(My)) < 0.95
        rng.next_u64(1)
        if ok:
            break
    else:
        raise RuntimeError("no stable cell found")
    # same cell (same draws) repeated; fillers set the harmonic number
    c0 = E.Beamline(_ring_cell(Rng(int(state)), 0, 0.0)).length() * cells
    lam = beta * C_MKS / freq
    h = math.ceil(c0 / lam)
    filler = (h * lam - c0) / (2 * cells)
    els = [E.MAXAMP("AP", X_MAX=x_max, Y_MAX=x_max)]
    for i in range(cells):
        cell = _ring_cell(Rng(int(state)), i, filler)
        if sext_scale is not None:
            for e in cell:
                if e.type == A.EH_T_KSEXT:
                    e.p.ksext.k2 *= sext_scale
        els += cell
    els.append(E.RFCA("RF", L=0.0, VOLT=volt, FREQ=freq, PHASE=180.0))
    bl = E.Beamline(els)
    bl.harmonic = h
    bl.seed_state = int(state)
    return bl, Po


def matched_beam(beamline, n, emit_x, emit_y, sigma_delta, sigma_t, seed, Po):
    """6D Gaussian rows (n, 8): transverse planes matched to the linear
    one-turn map at the start, delta and s = c beta t uncorrelated,
    particleID = 1..n, column 7 (pass of loss) = -1.  Drawn by the library's
    restatement of elegant's bunch generator (eh_gaussian_bunch: DLARAN
    stream seeded with seed, gaussian_distribution per plane, the
    normalized-coordinate transform; generate_bunch.c), every plane cut at
    10 sigma."""
    from . import generators
    Mx, My = linear_maps(beamline)
    bx, ax, _ = twiss_from_map(Mx)
    by, ay, _ = twiss_from_map(My)
    beta = Po / math.sqrt(Po * Po + 1)
    return generators.gaussian_bunch(n, seed, (emit_x, bx, ax, 10.0), (emit_y, by, ay, 10.0),
                                     (C_MKS * beta * sigma_t, sigma_delta, 0.0, 10.0))


def fodo_workload(n=10000, seed=12345):
    """The FODO ring with its 10k-particle beam: ex = ey = 1 nm,
    sigma_delta = 1e-3, sigma_t = 10 ps."""
    bl = fodo_ring()
    Po = momentum(6000.0)
    rows = matched_beam(bl, n, 1e-9, 1e-9, 1e-3, 10e-12, seed, Po)
    return bl, Po, rows


def storage_ring_workload(n=1_000_000, seed=2024, lattice_seed=1):
    """The storage ring with a matched beam (ex = 1 nm, ey = 10 pm,
    sigma_delta = 1e-3, sigma_t = 10 ps)."""
    bl, Po = storage_ring(seed=lattice_seed)
    rows = matched_beam(bl, n, 1e-9, 1e-11, 1e-3, 10e-12, seed, Po)
    return bl, Po, rows


# ------------------------------------------------------- FLOP accounting

LIBM = 20  # one sin/cos/asin/tan/log call


def _multipole_flops_alg(order, slices, misaligned, tilted):
    """Algorithmic FLOPs of multipole_tracking2 with the 4th-order
    integrator (multipole.c:693-1410) by the per-element convention:
    26 + S (243 + 15 m) + 15 for S slices and multipole order m (KQUAD
    m = 1: 299, KSEXT 314, KOCT 329 at one slice), +24 for a DX/DY/DZ
    offset and +24 for a TILT roll (entry and exit transforms)."""
    f = 26 + slices * (243 + 15 * order) + 15
    if misaligned:
        f += 24
    if tilted:
        f += 24
    return f


def _csbend_flops_alg(slices):
    """Algorithmic FLOPs of track_through_csbend + integrate_csbend_ordn
    (csbend.c:665-1693), 4th order: 113 + 277 S arithmetic and 16 S + 4 libm
    calls (asin, sin, cos, cos per drift substep; the edges' tan) at LIBM
    each; S = 16 gives 9,745."""
    return 113 + 277 * slices + (16 * slices + 4) * LIBM


def _rfca_flops_alg():
    """RFCA, one kick (simple_rfca.c:565-941): 85 arithmetic + 3 libm = 145."""
    return 85 + 3 * LIBM


def flops_per_particle_turn(beamline):
    """Algorithmic FLOPs per surviving particle per pass: +, -, *, /, sqrt
    count 1, each libm call LIBM = 20, integer work and compares 0, by the
    per-element formulas of the reference routines' operation counts: DRIF
    10, EDRIFT 11 (exactDrift, csbend.c:3767-3790), multipoles and CSBEND as
    above, RFCA 145, WAKE/ZLONGIT 45 per particle (time, bin index and the
    energy kick; the O(bins) convolution or FFT is left out)."""
    total = 0
    for e in beamline.elements:
        t = e.type
        if t in (A.EH_T_DRIF, A.EH_T_MONI, A.EH_T_HMON, A.EH_T_VMON, A.EH_T_HKICK, A.EH_T_VKICK, A.EH_T_RCOL):
            total += 10
        elif t == A.EH_T_EDRIFT:
            total += 11
        elif t in (A.EH_T_KQUAD, A.EH_T_KSEXT, A.EH_T_KOCT):
            p = getattr(e.p, A.PARAM_STRUCT[t][0])
            order = {A.EH_T_KQUAD: 1, A.EH_T_KSEXT: 2, A.EH_T_KOCT: 3}[t]
            nk = p.n_kicks
            S = max(1, math.ceil(nk / p.integration_order)) if nk > 0 else p.nSlices
            total += _multipole_flops_alg(order, S, bool(p.dx or p.dy or p.dz), bool(p.tilt))
            if t == A.EH_T_KQUAD and p.lEffective > 0:
                total += 22  # the two exact end drifts of LEFFECTIVE
        elif t == A.EH_T_CSBEND:
            b = e.p.csbend
            total += _csbend_flops_alg(b.nSlices) if b.angle else 11
        elif t == A.EH_T_RFCA:
            total += _rfca_flops_alg() if e.p.rfca.volt else 11
        elif t in (A.EH_T_WAKE, A.EH_T_ZLONGIT):
            total += 45
    return total


def _multipole_flops(order, slices, misaligned, tilted):
    """An itemised count of the same routine as written (multipole.c:
    693-1410), 4th-order integrator: entry slopes->momenta (11) and
    momenta->slopes (9); per drift substep dsh, x, y, s (with sqrt) and
    dzLoss (13); per kick the power arrays, the field sums and q updates
    (14/19/24 for order 1/2/3), momenta->slopes (9) and the q/(1+dp)*(1+dp)
    round trip with its second conversion (17); aperture offsets x+dx, y+dy
    (2 per check); exit conversion and s update (10)."""
    kick = {1: 14, 2: 19, 3: 24}[order] + 9 + 17
    per_slice = 4 * 13 + 3 * kick + 2
    f = 20 + slices * per_slice + 2 + 10
    if misaligned:
        f += 26
    if tilted:
        f += 24
    return f


def _csbend_flops(slices, edge1, edge2):
    """Itemised count of track_through_csbend + integrate_csbend_ordn
    (csbend.c:665-1693), pure dipole, 4th order: entry transform 25,
    canonical conversion 9, per drift substep 109 (asin, sin, cos, cos = 4
    libm calls + 29 arithmetic), per kick 21 (field kick 12, round trip 8,
    ds 1), exit conversion 10, velocity rescaling 15, exit transform 29;
    each edge (edge order 1) 31 incl. one tan."""
    per_slice = 4 * (29 + 4 * LIBM) + 3 * 21
    return 25 + 9 + slices * per_slice + 2 + 10 + 15 + 29 + 31 * (edge1 + edge2)


def flops_per_particle_turn_itemised(beamline):
    """The itemised count of the operations each routine performs as written
    (reported beside the algorithmic figure; the convention is the same:
    +, -, *, /, sqrt 1, libm LIBM)."""
    total = 0
    for e in beamline.elements:
        t = e.type
        if t == A.EH_T_DRIF:
            total += 10
        elif t in (A.EH_T_KQUAD, A.EH_T_KSEXT, A.EH_T_KOCT):
            p = getattr(e.p, A.PARAM_STRUCT[t][0])
            order = {A.EH_T_KQUAD: 1, A.EH_T_KSEXT: 2, A.EH_T_KOCT: 3}[t]
            nk = p.n_kicks
            S = max(1, math.ceil(nk / p.integration_order)) if nk > 0 else p.nSlices
            total += _multipole_flops(order, S, bool(p.dx or p.dy or p.dz), bool(p.tilt))
        elif t == A.EH_T_CSBEND:
            b = e.p.csbend
            total += _csbend_flops(b.nSlices, b.edge1_effects, b.edge2_effects) if b.angle else 12
        elif t == A.EH_T_RFCA:
            total += 10 + 5 + LIBM + 2 + 28 + 4 + 4
        elif t == A.EH_T_EDRIFT:
            total += 12
    return total


def collective_bytes_per_particle():
    """Algorithmic HBM bytes per particle of one WAKE or ZLONGIT step: read
    (s, delta) once for time and histogram (16 B), read and write
    (x', y', s, delta) for the kick (64 B)."""
    return 80


# ------------------------------------------------------ collective effects

def resonator_wake(fr=20e9, Q=1.0, Rs=1e4, points=10000, dt=0.1e-12):
    """Green function of a broad-band resonator, W(t) = 2k exp(-a t) (cos wb t
    - (a/wb) sin wb t), k = wr Rs/(2Q), a = wr/(2Q), wb = sqrt(wr^2 - a^2),
    sampled at t = 0, dt, ..., (points-1) dt (V/C).  W(0) is set to k, half of
    the jump at t = 0+ (the fundamental theorem of beam loading), so that the
    causal sum of convolveArrays counts a bin's own charge once."""
    wr = 2 * math.pi * fr
    k = wr * Rs / (2 * Q)
    a = wr / (2 * Q)
    wb = math.sqrt(wr * wr - a * a)
    t = np.arange(points, dtype=np.float64) * dt
    W = 2 * k * np.exp(-a * t) * (np.cos(wb * t) - (a / wb) * np.sin(wb * t))
    W[0] = k
    return E.WakeTable(t, W)


def storage_ring_wake(lattice_seed=1, charge=5e-9, wake_table=None, zlongit=True, wake=True):
    """The storage ring with CHARGE, a WAKE (resonator Green function) and a
    ZLONGIT (same resonator, 8192 bins of 2 ps) after the RF cavity."""
    bl, Po = storage_ring(seed=lattice_seed)
    table = wake_table if wake_table is not None else resonator_wake()
    els = [E.CHARGE("Q", TOTAL=charge)] + bl.elements
    if wake:
        els.append(E.WAKE(table, "WK"))
    if zlongit:
        els.append(E.ZLONGIT("ZL", BROAD_BAND=1, RS=1e4, Q=1.0, FREQ=20e9, N_BINS=8192, BIN_SIZE=2e-12))
    out = E.Beamline(els, keep=[table])
    out.harmonic = bl.harmonic
    return out, Po


def storage_ring_wake_workload(n=2_000_000, seed=77, lattice_seed=1):
    """2M-particle, 5 nC bunch with sigma_t = 20 ps in the ring with wakes."""
    bl, Po = storage_ring_wake(lattice_seed=lattice_seed)
    rows = matched_beam(bl, n, 1e-9, 1e-11, 1e-3, 20e-12, seed, Po)
    return bl, Po, rows
