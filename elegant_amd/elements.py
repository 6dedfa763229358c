"""Beamline elements with elegant's parameter names and defaults.

Each constructor returns an ``eh_element`` (ctypes) filled with the defaults
of the reference's parameter tables (track_data.c: KQUAD :1222-1283, KSEXT
:1120-1160, KOCT :1164-1187, CSBEND :1893-1980, RFCA :576-597, WAKE
:2396-2414, ZLONGIT :2251-2283, MAXAMP :990-997) and then the keyword
arguments given, using the reference's parameter names (L, K1, N_KICKS, ...).
"""
import ctypes as C

import numpy as np

from . import _abi as A

# (python keyword) -> (struct field) for each type; keywords are elegant's names
_NAMES = {
    A.EH_T_DRIF: {"L": "length", "ORDER": "order"},
    A.EH_T_EDRIFT: {"L": "length"},
    A.EH_T_KQUAD: {"L": "length", "K1": "k1", "TILT": "tilt", "DX": "dx", "DY": "dy", "DZ": "dz",
                   "FSE": "fse", "HKICK": "xkick", "VKICK": "ykick", "HCALIBRATION": "xKickCalibration",
                   "VCALIBRATION": "yKickCalibration", "BORE": "bore", "B": "B", "LEFFECTIVE": "lEffective",
                   "N_KICKS": "n_kicks", "N_SLICES": "nSlices", "INTEGRATION_ORDER": "integration_order",
                   "EXPAND_HAMILTONIAN": "expandHamiltonian", "SYNCH_RAD": "synch_rad", "ISR": "isr",
                   "EDGE1_EFFECTS": "edge1_effects", "EDGE2_EFFECTS": "edge2_effects", "RADIAL": "radial",
                   "MALIGN_METHOD": "malignMethod"},
    A.EH_T_KSEXT: {"L": "length", "K2": "k2", "K1": "k1", "J1": "j1", "TILT": "tilt", "DX": "dx", "DY": "dy",
                   "DZ": "dz", "FSE": "fse", "HKICK": "xkick", "VKICK": "ykick", "HCALIBRATION": "xKickCalibration",
                   "VCALIBRATION": "yKickCalibration", "BORE": "bore", "B": "B", "N_KICKS": "n_kicks",
                   "N_SLICES": "nSlices", "INTEGRATION_ORDER": "integration_order",
                   "EXPAND_HAMILTONIAN": "expandHamiltonian", "SYNCH_RAD": "synch_rad", "ISR": "isr",
                   "MALIGN_METHOD": "malignMethod"},
    A.EH_T_KOCT: {"L": "length", "K3": "k3", "TILT": "tilt", "DX": "dx", "DY": "dy", "DZ": "dz", "FSE": "fse",
                  "BORE": "bore", "B": "B", "N_KICKS": "n_kicks", "N_SLICES": "nSlices",
                  "INTEGRATION_ORDER": "integration_order", "EXPAND_HAMILTONIAN": "expandHamiltonian",
                  "SYNCH_RAD": "synch_rad", "ISR": "isr", "MALIGN_METHOD": "malignMethod"},
    A.EH_T_CSBEND: {"L": "length", "ANGLE": "angle", "K1": "k1", "K2": "k2", "K3": "k3", "K4": "k4", "E1": "e1",
                    "E2": "e2", "TILT": "tilt", "H1": "h1", "H2": "h2", "HGAP": "hgap", "FINT": "fint",
                    "FINT1": "fint1", "FINT2": "fint2", "DX": "dx", "DY": "dy", "DZ": "dz", "XKICK": "xKick",
                    "YKICK": "yKick", "FSE": "fse", "FSE_DIPOLE": "fseDipole", "FSE_QUADRUPOLE": "fseQuadrupole",
                    "ETILT": "etilt", "EDGE1_KICK_LIMIT": "edge1_kick_limit",
                    "EDGE2_KICK_LIMIT": "edge2_kick_limit", "N_SLICES": "nSlices", "N_KICKS": "nSlices",
                    "ETILT_SIGN": "etiltSign", "NONLINEAR": "nonlinear", "SYNCH_RAD": "synch_rad",
                    "EDGE1_EFFECTS": "edge1_effects", "EDGE2_EFFECTS": "edge2_effects", "EDGE_ORDER": "edge_order",
                    "INTEGRATION_ORDER": "integration_order", "EXPAND_HAMILTONIAN": "expandHamiltonian",
                    "ISR": "isr", "MALIGN_METHOD": "malignMethod", "REFERENCE_CORRECTION": "referenceCorrection",
                    "KICK_LIMIT_SCALING": "kick_limit_scaling", "FSE_CORRECTION": "fseCorrection"},
    A.EH_T_RFCA: {"L": "length", "VOLT": "volt", "PHASE": "phase", "FREQ": "freq", "Q": "Q", "DX": "dx",
                  "DY": "dy", "T_REFERENCE": "tReference", "N_KICKS": "nKicks", "CHANGE_P0": "change_p0",
                  "CHANGE_T": "change_t", "END1_FOCUS": "end1Focus", "END2_FOCUS": "end2Focus",
                  "STANDING_WAVE": "standingWave", "LINEARIZE": "linearize", "LOCK_PHASE": "lockPhase",
                  "FIDUCIAL": "fiducialMode"},
    A.EH_T_WAKE: {"FACTOR": "factor", "N_BINS": "n_bins", "INTERPOLATE": "interpolate",
                  "RAMP_PASSES": "rampPasses", "ALLOW_LONG_BEAM": "allowLongBeam", "SMOOTHING": "smoothing",
                  "CHANGE_P0": "change_p0", "BUNCHED_BEAM_MODE": "bunchedBeamMode", "START_BUNCH": "startBunch",
                  "END_BUNCH": "endBunch"},
    A.EH_T_ZLONGIT: {"RA": "Ra", "RS": "Rs", "Q": "Q", "FREQ": "freq", "BIN_SIZE": "bin_size", "FACTOR": "factor",
                     "BROAD_BAND": "broad_band", "N_BINS": "n_bins", "MAX_N_BINS": "max_n_bins",
                     "INTERPOLATE": "interpolate", "AREA_WEIGHT": "area_weight", "START_ON_PASS": "startOnPass",
                     "RAMP_PASSES": "rampPasses", "REVERSE_TIME_ORDER": "reverseTimeOrder",
                     "SMOOTHING": "smoothing", "ALLOW_LONG_BEAM": "allowLongBeam",
                     "BUNCHED_BEAM_MODE": "bunchedBeamMode", "START_BUNCH": "startBunch", "END_BUNCH": "endBunch"},
    A.EH_T_MAXAMP: {"X_MAX": "x_max", "Y_MAX": "y_max", "ELLIPTICAL": "elliptical", "EXPONENT": "exponent",
                    "YEXPONENT": "yExponent", "OPEN_SIDE": "openSide"},
    A.EH_T_CHARGE: {"TOTAL": "charge", "PER_PARTICLE": "chargePerParticle"},
    A.EH_T_MONI: {"L": "length", "DX": "dx", "DY": "dy", "WEIGHT": "weight", "TILT": "tilt",
                  "XCALIBRATION": "xcalibration", "YCALIBRATION": "ycalibration", "XSETPOINT": "xsetpoint",
                  "YSETPOINT": "ysetpoint", "ORDER": "order"},
    A.EH_T_HMON: {"L": "length", "DX": "dx", "DY": "dy", "WEIGHT": "weight", "TILT": "tilt",
                  "CALIBRATION": "xcalibration", "SETPOINT": "xsetpoint", "ORDER": "order"},
    A.EH_T_VMON: {"L": "length", "DX": "dx", "DY": "dy", "WEIGHT": "weight", "TILT": "tilt",
                  "CALIBRATION": "ycalibration", "SETPOINT": "ysetpoint", "ORDER": "order"},
    A.EH_T_HKICK: {"L": "length", "KICK": "kick", "TILT": "tilt", "B2": "b2", "CALIBRATION": "calibration",
                   "EDGE_EFFECTS": "edge_effects", "ORDER": "order", "STEERING": "steering",
                   "SYNCH_RAD": "synchRad", "ISR": "isr", "LERAD": "lEffRad"},
    A.EH_T_VKICK: {"L": "length", "KICK": "kick", "TILT": "tilt", "B2": "b2", "CALIBRATION": "calibration",
                   "EDGE_EFFECTS": "edge_effects", "ORDER": "order", "STEERING": "steering",
                   "SYNCH_RAD": "synchRad", "ISR": "isr", "LERAD": "lEffRad"},
    A.EH_T_RCOL: {"L": "length", "X_MAX": "x_max", "Y_MAX": "y_max", "DX": "dx", "DY": "dy",
                  "OPEN_SIDE": "openSide", "INVERT": "invert"},
    A.EH_T_MALIGN: {"DXP": "dxp", "DYP": "dyp", "DX": "dx", "DY": "dy", "DZ": "dz", "DT": "dt", "DP": "dp",
                    "DE": "de", "ON_PASS": "on_pass", "START_PID": "startPID", "END_PID": "endPID"},
    A.EH_T_MARK: {},
    A.EH_T_RECIRC: {},
}

# defaults that differ from zero (track_data.c; IS_LONG/IS_SHORT defaults come
# from the table's integer field, mad_parse.c:937)
_DEFAULTS = {
    A.EH_T_KQUAD: {"xKickCalibration": 1.0, "yKickCalibration": 1.0, "nSlices": 1, "integration_order": 4},
    A.EH_T_KSEXT: {"xKickCalibration": 1.0, "yKickCalibration": 1.0, "nSlices": 1, "integration_order": 4},
    A.EH_T_KOCT: {"n_kicks": 4, "nSlices": 4, "integration_order": 4},
    A.EH_T_CSBEND: {"fint": 0.5, "fint1": -1.0, "fint2": -1.0, "nSlices": 4, "etiltSign": 1, "nonlinear": 1,
                    "edge1_effects": 1, "edge2_effects": 1, "edge_order": 1, "integration_order": 4},
    A.EH_T_RFCA: {"freq": 500e6, "tReference": -1.0, "fiducialMode": A.EH_FID_TMEAN},
    A.EH_T_WAKE: {"factor": 1.0, "bunchedBeamMode": 1, "startBunch": -1, "endBunch": -1},
    A.EH_T_ZLONGIT: {"n_bins": 128, "factor": 1.0, "bunchedBeamMode": 1, "startBunch": -1, "endBunch": -1},
    A.EH_T_MAXAMP: {"exponent": 2},
    A.EH_T_MALIGN: {"on_pass": -1, "startPID": -1, "endPID": -1},
    A.EH_T_MONI: {"weight": 1.0, "xcalibration": 1.0, "ycalibration": 1.0},
    A.EH_T_HKICK: {"calibration": 1.0, "steering": 1},
    A.EH_T_VKICK: {"calibration": 1.0, "steering": 1},
    A.EH_T_HMON: {"weight": 1.0, "xcalibration": 1.0, "ycalibration": 1.0},
    A.EH_T_VMON: {"weight": 1.0, "xcalibration": 1.0, "ycalibration": 1.0},
}


# determineOpenSideCode (limit_amplitudes.c:1130-1162; codes track.h:1515-1518)
OPEN_SIDE_CODES = {"+x": A.EH_OPEN_PLUS_X, "+y": A.EH_OPEN_PLUS_Y, "-x": A.EH_OPEN_MINUS_X, "-y": A.EH_OPEN_MINUS_Y}


def open_side_code(text):
    """MAXAMP OPEN_SIDE string -> code: "+x"/"+h", "+y"/"+v", "-x"/"-h",
    "-y"/"-v" in either case; "" -> 0 (no open side); anything else is an
    error, as in the reference"""
    if text == "":
        return 0
    t = text.lower().replace("h", "x").replace("v", "y")
    if len(text) != 2 or t not in OPEN_SIDE_CODES:
        raise ValueError("open_side=%s is not valid" % text)
    return OPEN_SIDE_CODES[t]


def make(type_code, name="", **kw):
    """An element of the given EH_T_* type with defaults, then keyword overrides."""
    e = A.eh_element()
    C.memset(C.addressof(e), 0, C.sizeof(e))
    e.type = type_code
    e.name = name.encode()[:31]
    if type_code in A.PARAM_STRUCT:
        field, _ = A.PARAM_STRUCT[type_code]
        sub = getattr(e.p, field)
        for k, v in _DEFAULTS.get(type_code, {}).items():
            setattr(sub, k, v)
        names = _NAMES[type_code]
        for k, v in kw.items():
            key = k.upper()
            if key not in names:
                raise KeyError("%s has no parameter %s" % (A.TYPE_NAMES[type_code], k))
            if key == "FIDUCIAL" and isinstance(v, str):
                v = {"tmean": A.EH_FID_TMEAN, "light": A.EH_FID_LIGHT}[v.lower()]
            if key == "OPEN_SIDE" and isinstance(v, str):
                v = open_side_code(v)
            setattr(sub, names[key], v)
    elif kw:
        raise KeyError("%s takes no parameters" % A.TYPE_NAMES[type_code])
    return e


def RCOL(name="", **kw):
    return make(A.EH_T_RCOL, name, **kw)


def RECIRC(name=""):
    return make(A.EH_T_RECIRC, name)


def HKICK(name="", **kw):
    return make(A.EH_T_HKICK, name, **kw)


def VKICK(name="", **kw):
    return make(A.EH_T_VKICK, name, **kw)


def MONI(name="", **kw):
    return make(A.EH_T_MONI, name, **kw)


def HMON(name="", **kw):
    return make(A.EH_T_HMON, name, **kw)


def VMON(name="", **kw):
    return make(A.EH_T_VMON, name, **kw)


def MALIGN(name="", **kw):
    return make(A.EH_T_MALIGN, name, **kw)


def DRIF(name="", **kw):
    return make(A.EH_T_DRIF, name, **kw)


def EDRIFT(name="", **kw):
    return make(A.EH_T_EDRIFT, name, **kw)


def KQUAD(name="", **kw):
    return make(A.EH_T_KQUAD, name, **kw)


def KSEXT(name="", **kw):
    return make(A.EH_T_KSEXT, name, **kw)


def KOCT(name="", **kw):
    return make(A.EH_T_KOCT, name, **kw)


def CSBEND(name="", **kw):
    return make(A.EH_T_CSBEND, name, **kw)


def RFCA(name="", **kw):
    return make(A.EH_T_RFCA, name, **kw)


def MAXAMP(name="", **kw):
    return make(A.EH_T_MAXAMP, name, **kw)


def MARK(name=""):
    return make(A.EH_T_MARK, name)


def CHARGE(name="", **kw):
    return make(A.EH_T_CHARGE, name, **kw)


def ZLONGIT(name="", **kw):
    return make(A.EH_T_ZLONGIT, name, **kw)


class WakeTable:
    """Owns the t[], W[] arrays a WAKE element points to."""

    def __init__(self, t, W):
        self.t = np.ascontiguousarray(t, dtype=np.float64)
        self.W = np.ascontiguousarray(W, dtype=np.float64)
        if self.t.shape != self.W.shape or self.t.ndim != 1:
            raise ValueError("t and W must be 1-D arrays of equal length")


class ImpedanceTable:
    """Owns the f[], Zreal[], Zimag[] arrays a ZLONGIT with BROAD_BAND=0
    points to (zlongit.c:634-700): n_spect equally spaced frequencies from
    0 (Hz), n_spect - 1 a power of 2; Zreal or Zimag None for zero."""

    def __init__(self, f, Zreal=None, Zimag=None):
        self.f = np.ascontiguousarray(f, dtype=np.float64)
        self.Zreal = None if Zreal is None else np.ascontiguousarray(Zreal, dtype=np.float64)
        self.Zimag = None if Zimag is None else np.ascontiguousarray(Zimag, dtype=np.float64)
        for z in (self.Zreal, self.Zimag):
            if z is not None and z.shape != self.f.shape:
                raise ValueError("impedance samples and frequencies differ in length")


def ZLONGIT_TABLE(table, name="", **kw):
    """ZLONGIT with an impedance table (BROAD_BAND=0); keep `table` alive
    while the element is in use (Beamline(..., keep=[table]))"""
    e = make(A.EH_T_ZLONGIT, name, **kw)
    z = e.p.zlongit
    z.broad_band = 0
    P = C.POINTER(C.c_double)
    z.f = table.f.ctypes.data_as(P)
    z.Zreal = table.Zreal.ctypes.data_as(P) if table.Zreal is not None else None
    z.Zimag = table.Zimag.ctypes.data_as(P) if table.Zimag is not None else None
    z.n_spect = table.f.size
    return e


def WAKE(table, name="", **kw):
    e = make(A.EH_T_WAKE, name, **kw)
    e.p.wake.t = table.t.ctypes.data_as(C.POINTER(C.c_double))
    e.p.wake.W = table.W.ctypes.data_as(C.POINTER(C.c_double))
    e.p.wake.wakePoints = table.t.size
    return e


class Beamline:
    """A flattened beamline: a ctypes array of eh_element plus owned data."""

    def __init__(self, elements, keep=()):
        self.elements = list(elements)
        self.keep = list(keep)
        self.array = (A.eh_element * max(len(self.elements), 1))(*self.elements)

    def __len__(self):
        return len(self.elements)

    def pointer(self):
        return C.cast(self.array, C.POINTER(A.eh_element))

    def length(self):
        total = 0.0
        for e in self.elements:
            if e.type in (A.EH_T_DRIF, A.EH_T_EDRIFT, A.EH_T_KQUAD, A.EH_T_KSEXT, A.EH_T_KOCT,
                          A.EH_T_CSBEND, A.EH_T_RFCA, A.EH_T_MONI, A.EH_T_HMON, A.EH_T_VMON,
                          A.EH_T_HKICK, A.EH_T_VKICK, A.EH_T_RCOL):
                total += getattr(e.p, A.PARAM_STRUCT[e.type][0]).length
        return total

    def count(self, type_code):
        return sum(1 for e in self.elements if e.type == type_code)
