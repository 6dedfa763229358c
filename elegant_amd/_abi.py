"""ctypes mirror of include/elegant_hip.h (structs and constants only).

The layouts are checked against the library's own ``eh_struct_size`` by the
test suite, so a drift between this file and the header fails loudly.
"""
import ctypes as C

EH_OK = 0
EH_ERR_INVALID = -1
EH_ERR_UNSUPPORTED = -2
EH_ERR_HIP = -3
EH_ERR_NOMEM = -4
EH_ERR_STATE = -5
EH_ERR_NODEVICE = -6
EH_ERR_LONGBEAM = -7
EH_ERR_FIDUCIAL = -8

EH_NOALLOC_WAKE_POINTS = 16384  # most wake points of eh_track_through_wake_noalloc

EH_T_MARK = 1
EH_T_DRIF = 2
EH_T_EDRIFT = 3
EH_T_KQUAD = 4
EH_T_KSEXT = 5
EH_T_KOCT = 6
EH_T_CSBEND = 7
EH_T_RFCA = 8
EH_T_WAKE = 9
EH_T_ZLONGIT = 10
EH_T_MAXAMP = 11
EH_T_CHARGE = 12
EH_T_MALIGN = 13
EH_T_MONI = 14
EH_T_HMON = 15
EH_T_VMON = 16
EH_T_HKICK = 17
EH_T_VKICK = 18
EH_T_RECIRC = 19
EH_T_RCOL = 20

TYPE_NAMES = {
    EH_T_MARK: "MARK", EH_T_DRIF: "DRIF", EH_T_EDRIFT: "EDRIFT", EH_T_KQUAD: "KQUAD",
    EH_T_KSEXT: "KSEXT", EH_T_KOCT: "KOCT", EH_T_CSBEND: "CSBEND", EH_T_RFCA: "RFCA",
    EH_T_WAKE: "WAKE", EH_T_ZLONGIT: "ZLONGIT", EH_T_MAXAMP: "MAXAMP", EH_T_CHARGE: "CHARGE", EH_T_MALIGN: "MALIGN", EH_T_MONI: "MONI", EH_T_HMON: "HMON", EH_T_VMON: "VMON",
    EH_T_HKICK: "HKICK", EH_T_VKICK: "VKICK", EH_T_RECIRC: "RECIRC", EH_T_RCOL: "RCOL",
}

EH_OPEN_PLUS_X, EH_OPEN_PLUS_Y, EH_OPEN_MINUS_X, EH_OPEN_MINUS_Y = 1, 2, 3, 4
EH_FID_TMEAN = 0
EH_FID_LIGHT = 1

D = C.c_double
I = C.c_int32
L = C.c_int64


def _fields(doubles, ints):
    return [(n, D) for n in doubles] + [(n, I) for n in ints]


class EH_DRIFT(C.Structure):
    _fields_ = [("length", D), ("order", I), ("pad", I)]


class EH_EDRIFT(C.Structure):
    _fields_ = [("length", D)]


class EH_KQUAD(C.Structure):
    _fields_ = _fields(
        ["length", "k1", "tilt", "dx", "dy", "dz", "fse", "xkick", "ykick", "xKickCalibration",
         "yKickCalibration", "bore", "B", "lEffective"],
        ["n_kicks", "nSlices", "integration_order", "expandHamiltonian", "synch_rad", "isr",
         "edge1_effects", "edge2_effects", "radial", "malignMethod"])


class EH_KSEXT(C.Structure):
    _fields_ = _fields(
        ["length", "k2", "k1", "j1", "tilt", "dx", "dy", "dz", "fse", "xkick", "ykick",
         "xKickCalibration", "yKickCalibration", "bore", "B"],
        ["n_kicks", "nSlices", "integration_order", "expandHamiltonian", "synch_rad", "isr",
         "malignMethod", "pad"])


class EH_KOCT(C.Structure):
    _fields_ = _fields(
        ["length", "k3", "tilt", "dx", "dy", "dz", "fse", "bore", "B"],
        ["n_kicks", "nSlices", "integration_order", "expandHamiltonian", "synch_rad", "isr",
         "malignMethod", "pad"])


class EH_CSBEND(C.Structure):
    _fields_ = _fields(
        ["length", "angle", "k1", "k2", "k3", "k4", "e1", "e2", "tilt", "h1", "h2", "hgap", "fint",
         "fint1", "fint2", "dx", "dy", "dz", "xKick", "yKick", "fse", "fseDipole", "fseQuadrupole",
         "etilt", "edge1_kick_limit", "edge2_kick_limit"],
        ["nSlices", "etiltSign", "nonlinear", "synch_rad", "edge1_effects", "edge2_effects",
         "edge_order", "integration_order", "expandHamiltonian", "isr", "malignMethod",
         "referenceCorrection", "kick_limit_scaling", "fseCorrection"])


class EH_RFCA(C.Structure):
    _fields_ = _fields(
        ["length", "volt", "phase", "freq", "Q", "dx", "dy", "tReference"],
        ["nKicks", "change_p0", "change_t", "end1Focus", "end2Focus", "standingWave", "linearize",
         "lockPhase", "fiducialMode", "bodyFocusModel"])


class EH_WAKE(C.Structure):
    _fields_ = [("t", C.POINTER(D)), ("W", C.POINTER(D)), ("wakePoints", L), ("factor", D)] + [
        (n, I) for n in ["n_bins", "interpolate", "rampPasses", "allowLongBeam", "smoothing",
                         "change_p0", "bunchedBeamMode", "startBunch", "endBunch", "pad"]]


class EH_ZLONGIT(C.Structure):
    _fields_ = _fields(
        ["Ra", "Rs", "Q", "freq", "bin_size", "factor"],
        ["broad_band", "n_bins", "max_n_bins", "interpolate", "area_weight", "startOnPass",
         "rampPasses", "reverseTimeOrder", "smoothing", "allowLongBeam", "bunchedBeamMode", "startBunch",
         "endBunch", "pad"]) + [("f", C.POINTER(D)), ("Zreal", C.POINTER(D)), ("Zimag", C.POINTER(D)),
                                ("n_spect", L)]


class EH_MAXAMP(C.Structure):
    _fields_ = [("x_max", D), ("y_max", D), ("elliptical", I), ("exponent", I), ("yExponent", I),
                ("openSide", I)]


class EH_CHARGE(C.Structure):
    _fields_ = [("charge", D), ("chargePerParticle", D)]


class EH_MONI(C.Structure):
    _fields_ = _fields(["length", "dx", "dy", "weight", "tilt", "xcalibration", "ycalibration", "xsetpoint",
                        "ysetpoint"], ["order", "pad"])


class EH_CORRECTOR(C.Structure):
    _fields_ = _fields(["length", "kick", "tilt", "b2", "calibration", "lEffRad"],
                       ["edge_effects", "order", "steering", "synchRad", "isr", "pad"])


class EH_RCOL(C.Structure):
    _fields_ = _fields(["length", "x_max", "y_max", "dx", "dy"], ["openSide", "invert"])


class EH_MALIGN(C.Structure):
    _fields_ = _fields(["dxp", "dyp", "dx", "dy", "dz", "dt", "dp", "de"], ["on_pass", "startPID", "endPID", "pad"])


class _P(C.Union):
    _fields_ = [("drift", EH_DRIFT), ("edrift", EH_EDRIFT), ("kquad", EH_KQUAD), ("ksext", EH_KSEXT),
                ("koct", EH_KOCT), ("csbend", EH_CSBEND), ("rfca", EH_RFCA), ("wake", EH_WAKE),
                ("zlongit", EH_ZLONGIT), ("maxamp", EH_MAXAMP), ("charge", EH_CHARGE), ("malign", EH_MALIGN),
                ("moni", EH_MONI), ("corrector", EH_CORRECTOR), ("rcol", EH_RCOL)]


class eh_element(C.Structure):
    _fields_ = [("type", I), ("pad", I), ("name", C.c_char * 32), ("p", _P)]


class eh_species(C.Structure):
    _fields_ = [("mass", D), ("charge", D), ("massMV", D), ("radius", D), ("relSign", D)]


class eh_stats(C.Structure):
    _fields_ = [("effort", L), ("n_left", L), ("passes_done", L), ("kernel_launches", L),
                ("track_ms", D), ("collective_ms", D)]


class eh_batch_tuning(C.Structure):
    _fields_ = [("particles_per_lane", I), ("block", I), ("passes_per_launch", I), ("use_graph", I),
                ("compact", I), ("pad", I), ("sample_ms", D)]


class eh_rng(C.Structure):
    _fields_ = [("seed", C.c_int64 * 4), ("have_saved", I), ("pad", I), ("saved", D)]


PARAM_STRUCT = {
    EH_T_DRIF: ("drift", EH_DRIFT), EH_T_EDRIFT: ("edrift", EH_EDRIFT), EH_T_KQUAD: ("kquad", EH_KQUAD),
    EH_T_KSEXT: ("ksext", EH_KSEXT), EH_T_KOCT: ("koct", EH_KOCT), EH_T_CSBEND: ("csbend", EH_CSBEND),
    EH_T_RFCA: ("rfca", EH_RFCA), EH_T_WAKE: ("wake", EH_WAKE), EH_T_ZLONGIT: ("zlongit", EH_ZLONGIT),
    EH_T_MAXAMP: ("maxamp", EH_MAXAMP), EH_T_CHARGE: ("charge", EH_CHARGE), EH_T_MALIGN: ("malign", EH_MALIGN),
    EH_T_MONI: ("moni", EH_MONI), EH_T_HMON: ("moni", EH_MONI), EH_T_VMON: ("moni", EH_MONI),
    EH_T_HKICK: ("corrector", EH_CORRECTOR), EH_T_VKICK: ("corrector", EH_CORRECTOR),
    EH_T_RCOL: ("rcol", EH_RCOL),
}


def rows_pointer(arr):
    """double** over the rows of a C-contiguous float64 (n, m) array.

    Returns (pointer, keepalive): keep the second object alive while the
    pointer is in use."""
    import numpy as np
    if arr.dtype != np.float64 or not arr.flags.c_contiguous or arr.ndim != 2:
        raise ValueError("rows must be a C-contiguous float64 (n, m) array")
    n, m = arr.shape
    ptrs = np.arange(max(n, 1), dtype=np.uint64) * np.uint64(8 * m) + np.uint64(arr.ctypes.data)
    return ptrs.ctypes.data_as(C.POINTER(C.POINTER(D))), ptrs
