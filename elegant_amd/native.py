"""Loads libelegant_hip.so (built in-tree by elegant_amd/build.py) and wraps
its C ABI.  There is no fallback: if the library is missing, or no GPU is
present when a context is created, the call raises."""
import ctypes as C
import weakref
import os

import numpy as np

from . import _abi as A

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libelegant_hip.so")
_lib = None

P = C.c_void_p
PP = C.POINTER(C.POINTER(C.c_double))
PE = C.POINTER(A.eh_element)
PD = C.POINTER(C.c_double)
PF = C.POINTER(C.c_float)


class EHError(RuntimeError):
    def __init__(self, code, message=""):
        self.code = code
        super().__init__("elegant_hip error %d (%s)%s" % (code, ERRNAMES.get(code, "?"),
                                                           (": " + message) if message else ""))


ERRNAMES = {A.EH_ERR_INVALID: "invalid", A.EH_ERR_UNSUPPORTED: "unsupported", A.EH_ERR_HIP: "hip",
            A.EH_ERR_NOMEM: "nomem", A.EH_ERR_STATE: "state", A.EH_ERR_NODEVICE: "nodevice",
            A.EH_ERR_LONGBEAM: "longbeam", A.EH_ERR_FIDUCIAL: "fiducial"}

# every symbol include/elegant_hip.h declares: (name, restype, argtypes)
SIGNATURES = [
    ("eh_element_defaults", C.c_int, [PE, C.c_int32]),
    ("eh_struct_size", C.c_int64, [C.c_int32]),
    ("eh_ctx_create", C.c_int, [C.c_int, C.POINTER(P)]),
    ("eh_ctx_destroy", C.c_int, [P]),
    ("eh_strerror", C.c_char_p, [C.c_int]),
    ("eh_last_error", C.c_char_p, [P]),
    ("eh_load_program", C.c_int, [P, PE, C.c_int64, C.c_double]),
    ("eh_upload", C.c_int, [P, PP, C.c_int64, C.c_int32]),
    ("eh_track", C.c_int, [P, C.c_int64, C.c_int64, PD, C.POINTER(A.eh_stats)]),
    ("eh_download", C.c_int, [P, PP, C.c_int64, C.c_int32, C.POINTER(C.c_int64)]),
    ("eh_set_particle_properties", C.c_int, [P, C.c_int32]),
    ("eh_set_id_slots_per_bunch", C.c_int, [P, C.c_int64]),
    ("eh_species_named", C.c_int, [C.c_char_p, C.c_double, C.c_double, C.POINTER(A.eh_species)]),
    ("eh_set_species", C.c_int, [P, C.POINTER(A.eh_species)]),
    ("eh_set_particle_limit", C.c_int, [P, C.c_int64]),
    ("eh_set_interrupt_file", C.c_int, [P, C.c_char_p]),
    ("eh_set_interrupt_flag", C.c_int, [P, C.POINTER(C.c_int32)]),
    ("eh_do_tracking", C.c_int64, [P, PP, C.c_int64, C.POINTER(C.c_int64), PE, C.c_int64, PD, PP, C.c_int64,
                                   C.c_int64]),
    ("eh_multipole_tracking2", C.c_int64, [P, PP, C.c_int64, PE, C.c_double, PP, C.c_double,
                                           C.POINTER(A.EH_MAXAMP)]),
    ("eh_track_through_csbend", C.c_int64, [P, PP, C.c_int64, C.POINTER(A.EH_CSBEND), C.c_double, PP, C.c_double,
                                            C.POINTER(A.EH_MAXAMP)]),
    ("eh_rfca_fiducial", C.c_int, [P, C.c_int64, PD]),
    ("eh_simple_rf_cavity", C.c_int64, [P, PP, C.c_int64, C.POINTER(A.EH_RFCA), PP, PD, C.c_double, PD]),
    ("eh_track_particles_drift", C.c_int64, [P, PP, C.c_int64, C.POINTER(A.EH_DRIFT)]),
    ("eh_limit_amplitudes", C.c_int64, [P, PP, C.c_double, C.c_double, C.c_int64, PP, C.c_double, C.c_double,
                                        C.c_int32, C.c_int32]),
    ("eh_elimit_amplitudes", C.c_int64, [P, PP, C.c_double, C.c_double, C.c_int64, PP, C.c_double, C.c_double,
                                         C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    ("eh_track_through_wake", C.c_int64, [P, PP, C.c_int64, C.POINTER(A.EH_WAKE), C.c_double, C.c_int64,
                                          C.c_double]),
    ("eh_track_through_zlongit", C.c_int64, [P, PP, C.c_int64, C.POINTER(A.EH_ZLONGIT), C.c_double,
                                             C.c_int64, C.c_double]),
    ("eh_track_through_wake_noalloc", C.c_int64, [P, C.c_void_p * 6, C.c_int64, C.POINTER(A.EH_WAKE), C.c_void_p,
                                                  C.c_double, C.c_int64, C.c_double]),
    ("eh_last_kernel_timing", C.c_int, [P, C.POINTER(C.c_int64), PD]),
    ("eh_batch_create", C.c_int, [P, C.POINTER(P)]),
    ("eh_batch_destroy", C.c_int, [P]),
    ("eh_batch_add_variant", C.c_int, [P, PE, C.c_int64, C.c_double, PP, C.c_int64, C.c_int32]),
    ("eh_batch_autotune", C.c_int, [P, C.c_int64, C.POINTER(A.eh_batch_tuning)]),
    ("eh_batch_set_tuning", C.c_int, [P, C.POINTER(A.eh_batch_tuning)]),
    ("eh_batch_get_tuning", C.c_int, [P, C.POINTER(A.eh_batch_tuning)]),
    ("eh_batch_track", C.c_int, [P, C.c_int64, C.c_int64, C.POINTER(A.eh_stats)]),
    ("eh_batch_download", C.c_int, [P, C.c_int32, PP, C.c_int64, C.c_int32, C.POINTER(C.c_int64)]),
    ("eh_batch_loss_passes", C.c_int, [P, C.c_int32, C.POINTER(C.c_int32), C.c_int64]),
    ("eh_batch_variants", C.c_int32, [P]),
    ("eh_da_trim_area", C.c_double, [C.c_int64, PD, PD, PD, PD, C.c_int32]),
    ("eh_mlp_create", C.c_int, [P, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(P)]),
    ("eh_mlp_destroy", C.c_int, [P]),
    ("eh_mlp_num_params", C.c_int64, [P]),
    ("eh_mlp_get_params", C.c_int, [P, PF]),
    ("eh_mlp_set_params", C.c_int, [P, PF]),
    ("eh_mlp_gradient", C.c_int, [P, PF, PF, C.c_int32, PF, PF]),
    ("eh_mlp_train", C.c_int, [P, PF, PF, C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_uint64, PF]),
    ("eh_mlp_predict", C.c_int, [P, PF, C.c_int64, PF]),
    ("eh_rng_seed", C.c_int, [C.POINTER(A.eh_rng), C.c_int64]),
    ("eh_rng_uniform", C.c_double, [C.POINTER(A.eh_rng)]),
    ("eh_gauss_rn", C.c_double, [C.POINTER(A.eh_rng)]),
    ("eh_gauss_rn_lim", C.c_double, [C.POINTER(A.eh_rng), C.c_double, C.c_double, C.c_double]),
    ("eh_perturbation", C.c_double, [C.POINTER(A.eh_rng), C.c_double, C.c_double, C.c_int32]),
    ("eh_perturbations", C.c_int, [C.POINTER(A.eh_rng), C.c_double, C.c_double, C.c_int32, PD, C.c_int64]),
    ("eh_gaussian_distribution", C.c_int, [C.POINTER(A.eh_rng), PP, C.c_int64, C.c_int32, C.c_double, C.c_double,
                                           C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32]),
    ("eh_transform_from_normalized", None, [PP, C.c_int64, C.c_int32, C.c_double, C.c_double]),
    ("eh_gaussian_bunch", C.c_int, [C.POINTER(A.eh_rng), PP, C.c_int64, C.c_int32, PD, PD, PD, C.c_int64]),
]
# helpers outside the public header (measurement and tests)
EXTRA = [
    ("eh_get_loss_records", C.c_int, [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("eh_math_probe", C.c_int, [P, C.c_int32, PD, PD, PD, C.c_int64]),
    ("eh_test_buffer", C.c_int, [P, C.c_int32, C.POINTER(C.c_void_p), C.c_void_p, C.c_int64]),
    ("eh_test_fold", C.c_int32, [PE, C.c_int64, C.c_double, C.POINTER(C.c_int32), C.c_int32]),
    ("eh_test_program_runs", C.c_int32, [P]),
    ("eh_set_chunk_passes", None, [P, C.c_int32]),
    ("eh_set_block", None, [P, C.c_int32]),
    ("eh_collective_probe", C.c_int, [P, PP, C.c_int64, PE, C.c_double, C.c_int64, C.c_double, PD, PD,
                                      C.POINTER(C.c_int64), PD]),
]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libelegant_hip.so is not built: run `python -m elegant_amd.build` "
                               "(or __graft_entry__.build())")
        _lib = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES + EXTRA:
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def check(ctx, rc):
    if rc < 0:
        msg = ""
        if ctx:
            m = lib().eh_last_error(ctx)
            msg = m.decode() if m else ""
        raise EHError(int(rc), msg or lib().eh_strerror(int(rc)).decode())
    return rc


def _dptr(a):
    return a.ctypes.data_as(PD)


def _acc_ptr(accepted, rows):
    """row pointers of an accepted array shaped like rows (or NULL)"""
    if accepted is None:
        return None, None
    if accepted.dtype != np.float64 or not accepted.flags.c_contiguous or accepted.shape != rows.shape:
        raise ValueError("accepted must be C-contiguous float64 shaped like the rows")
    return A.rows_pointer(accepted)


class Batch:
    """Many lattice variants, each with its own particles, tracked together
    (eh_batch_* of the C ABI).  Keeps a reference to its Tracker's context."""

    def __init__(self, tracker):
        self.tracker = tracker
        self._b = P()
        check(tracker._ctx, lib().eh_batch_create(tracker._ctx, C.byref(self._b)))
        tracker.adopt(self)
        self.sizes = []
        self._keep = []

    def close(self):
        if self._b:
            lib().eh_batch_destroy(self._b)
            self._b = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_variant(self, beamline, Po, rows):
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        ptr, keep = A.rows_pointer(rows)
        self._keep.append(beamline)
        v = check(self.tracker._ctx, lib().eh_batch_add_variant(self._b, beamline.pointer(), len(beamline), Po, ptr,
                                                                 rows.shape[0], rows.shape[1]))
        self.sizes.append(rows.shape[0])
        return v

    def autotune(self, sample_passes=4):
        t = A.eh_batch_tuning()
        check(self.tracker._ctx, lib().eh_batch_autotune(self._b, sample_passes, C.byref(t)))
        return t

    def set_tuning(self, ppl, block, passes_per_launch, use_graph=1, compact=0):
        t = A.eh_batch_tuning(ppl, block, passes_per_launch, use_graph, compact, 0, 0.0)
        check(self.tracker._ctx, lib().eh_batch_set_tuning(self._b, C.byref(t)))

    def tuning(self):
        t = A.eh_batch_tuning()
        check(self.tracker._ctx, lib().eh_batch_get_tuning(self._b, C.byref(t)))
        return t

    def track(self, pass0, npasses):
        st = A.eh_stats()
        check(self.tracker._ctx, lib().eh_batch_track(self._b, pass0, npasses, C.byref(st)))
        return st

    def download(self, v):
        n = self.sizes[v]
        out = np.zeros((n, 8))
        ptr, keep = A.rows_pointer(out)
        nl = C.c_int64()
        check(self.tracker._ctx, lib().eh_batch_download(self._b, v, ptr, n, 8, C.byref(nl)))
        return out, int(nl.value)

    def loss_passes(self, v):
        n = self.sizes[v]
        lp = np.zeros(n, dtype=np.int32)
        check(self.tracker._ctx, lib().eh_batch_loss_passes(self._b, v, lp.ctypes.data_as(C.POINTER(C.c_int32)), n))
        return lp


def da_trim_area(x, y, dxf, dyf, full_plane=0):
    """trimApertureSearchResult on copies of the limits; returns (area, x, y)."""
    x = np.array(x, dtype=np.float64)
    y = np.array(y, dtype=np.float64)
    dxf = np.ascontiguousarray(dxf, dtype=np.float64)
    dyf = np.ascontiguousarray(dyf, dtype=np.float64)
    a = lib().eh_da_trim_area(x.size, _dptr(x), _dptr(y), _dptr(dxf), _dptr(dyf), full_plane)
    return a, x, y


def species(name, mass_ratio=0.0, charge_ratio=0.0):
    """change_particle's species NAME ("electron", "positron", "proton",
    "muon", or "custom" with MASS_RATIO and CHARGE_RATIO in electron units)"""
    sp = A.eh_species()
    rc = lib().eh_species_named(name.encode(), float(mass_ratio), float(charge_ratio), C.byref(sp))
    if rc:
        raise EHError(rc, "eh_species_named(%r)" % name)
    return sp


def program_runs(beamline, Po):
    """The runs the tracking kernel walks for this beamline, as (first, len,
    reps, flat) tuples: repeated cells fold into one body of records."""
    cap = max(len(beamline), 1) * 2 + 16
    out = (C.c_int32 * (4 * cap))()
    nr = lib().eh_test_fold(beamline.pointer(), len(beamline), Po, out, cap)
    if nr < 0:
        raise EHError(nr, "eh_test_fold")
    return [tuple(out[4 * k:4 * k + 4]) for k in range(min(nr, cap))]


class Tracker:
    """One device context: a loaded beamline and a particle population.

    Particle arrays are float64 (n, 8): x, x', y, y', s, delta, particleID,
    pass of loss (track.h:81-101)."""

    def __init__(self, device=0):
        self._ctx = P()
        check(None, lib().eh_ctx_create(device, C.byref(self._ctx)))
        self._beamline = None
        self.Po = None
        self._children = weakref.WeakSet()

    def adopt(self, child):
        """register an object (Batch, MLP) whose device state lives in this
        context: closing the context closes it first"""
        self._children.add(child)

    def close(self):
        if self._ctx:
            for c in list(self._children):
                c.close()
            lib().eh_ctx_destroy(self._ctx)
            self._ctx = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, beamline, Po):
        self._beamline = beamline
        self.Po = float(Po)
        check(self._ctx, lib().eh_load_program(self._ctx, beamline.pointer(), len(beamline), self.Po))

    def upload(self, rows):
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        ptr, keep = A.rows_pointer(rows)
        check(self._ctx, lib().eh_upload(self._ctx, ptr, rows.shape[0], rows.shape[1]))
        self.n = rows.shape[0]

    def track(self, pass0, npasses):
        st = A.eh_stats()
        pc = C.c_double(self.Po)
        check(self._ctx, lib().eh_track(self._ctx, pass0, npasses, C.byref(pc), C.byref(st)))
        return st

    def download(self):
        out = np.zeros((self.n, 8), dtype=np.float64)
        ptr, keep = A.rows_pointer(out)
        nleft = C.c_int64()
        check(self._ctx, lib().eh_download(self._ctx, ptr, self.n, 8, C.byref(nleft)))
        return out, int(nleft.value)

    def set_particle_properties(self, nprops):
        """columns of every particle row (elegant's totalPropertiesPerParticle)"""
        check(self._ctx, lib().eh_set_particle_properties(self._ctx, int(nprops)))

    def set_species(self, sp):
        """the particle species (an eh_species, e.g. from species("proton"))"""
        check(self._ctx, lib().eh_set_species(self._ctx, C.byref(sp)))

    def set_id_slots_per_bunch(self, n):
        """charge->idSlotsPerBunch: with n > 0 and rows of at least 9
        columns, WAKE/ZLONGIT in bunched-beam mode act per bunch index
        (column 8)"""
        check(self._ctx, lib().eh_set_id_slots_per_bunch(self._ctx, int(n)))

    def set_particle_limit(self, limit):
        check(self._ctx, lib().eh_set_particle_limit(self._ctx, int(limit)))

    def set_interrupt_file(self, path):
        check(self._ctx, lib().eh_set_interrupt_file(self._ctx, path.encode() if path else None))

    def set_interrupt_flag(self, flag):
        """flag: a ctypes c_int32 kept alive by the caller (None turns it off)"""
        self._flag = flag
        check(self._ctx, lib().eh_set_interrupt_flag(self._ctx, C.byref(flag) if flag is not None else None))

    def do_tracking(self, rows, beamline, Po, n_passes, pass_offset=0, accepted=None):
        """do_tracking on host rows (n, >=8), in place; returns (nLeft, effort).
        accepted: None or an array shaped like rows, filled with each
        particle's initial row at its final row."""
        if rows.dtype != np.float64 or not rows.flags.c_contiguous or rows.shape[1] < 8:
            raise ValueError("rows must be C-contiguous float64 (n, >=8)")
        ptr, keep = A.rows_pointer(rows)
        aptr, akeep = _acc_ptr(accepted, rows)
        eff = C.c_int64(0)
        pc = C.c_double(Po)
        self._beamline = beamline
        self.Po = float(Po)
        r = lib().eh_do_tracking(self._ctx, ptr, rows.shape[0], C.byref(eff), beamline.pointer(), len(beamline),
                                 C.byref(pc), aptr, n_passes, pass_offset)
        check(self._ctx, r)
        self.n = rows.shape[0]
        self.P_central = pc.value  # moved by CHANGE_P0 cavities
        return int(r), int(eff.value)

    def loss_records(self):
        p = np.zeros(self.n, dtype=np.int32)
        c = np.zeros(self.n, dtype=np.int32)
        check(self._ctx, lib().eh_get_loss_records(self._ctx, p.ctypes.data_as(C.POINTER(C.c_int32)),
                                                   c.ctypes.data_as(C.POINTER(C.c_int32))))
        return p, c

    def program_runs(self):
        """runs of the tracking program now on the device"""
        return int(lib().eh_test_program_runs(self._ctx))

    def kernel_timing(self):
        n = C.c_int64()
        ms = C.c_double()
        check(self._ctx, lib().eh_last_kernel_timing(self._ctx, C.byref(n), C.byref(ms)))
        return int(n.value), float(ms.value)

    def set_chunk_passes(self, k):
        lib().eh_set_chunk_passes(self._ctx, int(k))

    def set_block(self, b):
        lib().eh_set_block(self._ctx, int(b))

    def math_probe(self, fn, x, y=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = x if y is None else np.ascontiguousarray(y, dtype=np.float64)
        out = np.empty_like(x)
        check(self._ctx, lib().eh_math_probe(self._ctx, fn, _dptr(x), _dptr(y), _dptr(out), x.size))
        return out

    def collective_probe(self, rows, elem, Po, i_pass, macro, cap=1 << 17):
        """one WAKE/ZLONGIT step on rows (in place); returns (hist, volt, tmin)"""
        ptr, keep = A.rows_pointer(rows)
        h = np.zeros(cap)
        v = np.zeros(cap)
        nb = C.c_int64()
        tmin = C.c_double()
        check(self._ctx, lib().eh_collective_probe(self._ctx, ptr, rows.shape[0], C.byref(elem), Po, i_pass, macro,
                                                   _dptr(h), _dptr(v), C.byref(nb), C.byref(tmin)))
        return h[:nb.value], v[:nb.value], tmin.value

    # element-by-element entry points (rows: float64 (n, 8), mutated in place;
    # accepted: None or an array shaped like rows, permuted with them)
    def multipole_tracking2(self, rows, elem, Po, z_start, maxamp=None, accepted=None):
        ptr, keep = A.rows_pointer(rows)
        aptr, akeep = _acc_ptr(accepted, rows)
        mp = C.byref(maxamp) if maxamp is not None else None
        return check(self._ctx, lib().eh_multipole_tracking2(self._ctx, ptr, rows.shape[0], C.byref(elem), Po,
                                                             aptr, z_start, mp))

    def track_through_csbend(self, rows, csbend, Po, z_start, maxamp=None, accepted=None):
        ptr, keep = A.rows_pointer(rows)
        aptr, akeep = _acc_ptr(accepted, rows)
        mp = C.byref(maxamp) if maxamp is not None else None
        return check(self._ctx, lib().eh_track_through_csbend(self._ctx, ptr, rows.shape[0], C.byref(csbend), Po,
                                                              aptr, z_start, mp))

    def rfca_fiducial(self, elem_index):
        """the fiducial phase the last fused run fixed for the RFCA at elem_index (NaN if unset)"""
        ph = C.c_double()
        check(self._ctx, lib().eh_rfca_fiducial(self._ctx, elem_index, C.byref(ph)))
        return ph.value

    def simple_rf_cavity(self, rows, rfca, Po, zEnd, phase_fiducial=float("nan"), accepted=None):
        ptr, keep = A.rows_pointer(rows)
        aptr, akeep = _acc_ptr(accepted, rows)
        pc = C.c_double(Po)
        ph = C.c_double(phase_fiducial)
        r = check(self._ctx, lib().eh_simple_rf_cavity(self._ctx, ptr, rows.shape[0], C.byref(rfca), aptr,
                                                       C.byref(pc), zEnd, C.byref(ph)))
        self.P_central = pc.value  # moved by CHANGE_P0
        return r, ph.value

    def track_particles_drift(self, rows, drift):
        ptr, keep = A.rows_pointer(rows)
        return check(self._ctx, lib().eh_track_particles_drift(self._ctx, ptr, rows.shape[0], C.byref(drift)))

    def limit_amplitudes(self, rows, xmax, ymax, z, Po, extrapolate_z, open_code=0, accepted=None):
        ptr, keep = A.rows_pointer(rows)
        aptr, akeep = _acc_ptr(accepted, rows)
        return check(self._ctx, lib().eh_limit_amplitudes(self._ctx, ptr, xmax, ymax, rows.shape[0], aptr, z, Po,
                                                          extrapolate_z, open_code))

    def elimit_amplitudes(self, rows, xmax, ymax, z, Po, extrapolate_z, open_code=0, exponent=2, yexponent=0,
                          accepted=None):
        ptr, keep = A.rows_pointer(rows)
        aptr, akeep = _acc_ptr(accepted, rows)
        return check(self._ctx, lib().eh_elimit_amplitudes(self._ctx, ptr, xmax, ymax, rows.shape[0], aptr, z, Po,
                                                           extrapolate_z, open_code, exponent, yexponent))

    def track_through_wake(self, rows, wake, Po, i_pass, macro_charge):
        ptr, keep = A.rows_pointer(rows)
        return check(self._ctx, lib().eh_track_through_wake(self._ctx, ptr, rows.shape[0], C.byref(wake), Po,
                                                            i_pass, macro_charge))

    def track_through_wake_noalloc(self, coord_ptrs, n, wake, W_ptr, Po, i_pass, macro_charge):
        """the no-allocation WAKE step on device arrays: coord_ptrs are six
        device addresses (x, x', y, y', s, delta), W_ptr the Green function's"""
        arr = (C.c_void_p * 6)(*[C.c_void_p(int(p)) for p in coord_ptrs])
        return check(self._ctx, lib().eh_track_through_wake_noalloc(self._ctx, arr, n, C.byref(wake),
                                                                    C.c_void_p(int(W_ptr)), Po, i_pass,
                                                                    macro_charge))

    def device_array(self, host):
        """a device copy of a float64 numpy array (tests of the device-pointer
        entry points); returns its address; free with free_device_array"""
        h = np.ascontiguousarray(host, dtype=np.float64)
        ptr = C.c_void_p()
        check(self._ctx, lib().eh_test_buffer(self._ctx, 0, C.byref(ptr), None, h.nbytes))
        check(self._ctx, lib().eh_test_buffer(self._ctx, 2, C.byref(ptr), h.ctypes.data_as(C.c_void_p), h.nbytes))
        return ptr.value

    def read_device_array(self, addr, n):
        out = np.empty(n, dtype=np.float64)
        ptr = C.c_void_p(addr)
        check(self._ctx, lib().eh_test_buffer(self._ctx, 3, C.byref(ptr), out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def free_device_array(self, addr):
        ptr = C.c_void_p(addr)
        check(self._ctx, lib().eh_test_buffer(self._ctx, 1, C.byref(ptr), None, 0))

    def track_through_zlongit(self, rows, zl, Po, i_pass, macro_charge):
        ptr, keep = A.rows_pointer(rows)
        return check(self._ctx, lib().eh_track_through_zlongit(self._ctx, ptr, rows.shape[0], C.byref(zl), Po,
                                                               i_pass, macro_charge))
