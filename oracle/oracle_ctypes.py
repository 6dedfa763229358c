"""ctypes wrapper of the CPU checker (TEST INFRASTRUCTURE ONLY).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg;
never by the product package."""
import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

from elegant_amd import _abi as A

HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}

PP = C.POINTER(C.POINTER(C.c_double))
PE = C.POINTER(A.eh_element)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(mode="glibc"):
    if mode not in _libs:
        path = os.path.join(HERE, "liboracle_%s.so" % mode)
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_ctx_new.restype = C.c_void_p
        L.or_ctx_free.argtypes = [C.c_void_p]
        L.or_math_mode.restype = C.c_int
        L.or_ctx_effort.restype = C.c_int64
        L.or_ctx_effort.argtypes = [C.c_void_p]
        L.or_math_eval.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.c_long]
        L.or_do_tracking.restype = C.c_long
        L.or_do_tracking.argtypes = [C.c_void_p, PP, C.c_long, PE, C.c_long, C.POINTER(C.c_double), C.c_long,
                                     C.c_long]
        L.or_multipole_tracking2.restype = C.c_long
        L.or_multipole_tracking2.argtypes = [PP, C.c_long, PE, C.c_double, C.c_double, C.POINTER(A.EH_MAXAMP)]
        L.or_track_through_csbend.restype = C.c_long
        L.or_track_through_csbend.argtypes = [PP, C.c_long, C.POINTER(A.EH_CSBEND), C.c_double, C.c_double,
                                              C.POINTER(A.EH_MAXAMP)]
        L.or_track_particles_drift.argtypes = [PP, C.c_long, C.c_double, C.c_int]
        L.or_exact_drift.argtypes = [PP, C.c_long, C.c_double]
        L.or_limit_amplitudes.restype = C.c_long
        L.or_limit_amplitudes.argtypes = [PP, C.c_double, C.c_double, C.c_long, C.c_double, C.c_double, C.c_long,
                                          C.c_long]
        L.or_elimit_amplitudes.restype = C.c_long
        L.or_elimit_amplitudes.argtypes = [PP, C.c_double, C.c_double, C.c_long, C.c_double, C.c_double, C.c_long,
                                           C.c_long, C.c_long, C.c_long]
        L.or_add_to_particle_energy.argtypes = [C.POINTER(C.c_double), C.c_double, C.c_double, C.c_double]
        L.or_kahan_plus.restype = C.c_double
        L.or_kahan_plus.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double)]
        L.or_do_match_energy.argtypes = [PP, C.c_long, C.POINTER(C.c_double)]
        L.or_etilt_offsets.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)]
        L.or_set_row_width.argtypes = [C.c_long]
        L.or_row_width.restype = C.c_long
        L.or_set_species.argtypes = [C.c_double] * 5
        L.or_set_id_slots_per_bunch.argtypes = [C.c_void_p, C.c_long]
        L.or_real_fft.argtypes = [C.POINTER(C.c_double), C.c_long, C.c_int]
        L.or_simple_rf_cavity.restype = C.c_long
        L.or_simple_rf_cavity.argtypes = [C.c_void_p, C.c_long, PP, C.c_long, C.POINTER(A.EH_RFCA),
                                          C.POINTER(C.c_double), C.c_double]
        L.or_track_through_wake.restype = C.c_long
        L.or_track_through_wake.argtypes = [C.c_void_p, PP, C.c_long, C.POINTER(A.EH_WAKE), C.c_double, C.c_long]
        L.or_track_through_zlongit.restype = C.c_long
        L.or_track_through_zlongit.argtypes = [C.c_void_p, C.c_long, PP, C.c_long, C.POINTER(A.EH_ZLONGIT),
                                               C.c_double, C.c_long]
        PD = C.POINTER(C.c_double)
        L.or_last_hist.restype = C.c_int64
        L.or_last_hist.argtypes = [C.c_void_p, PD, PD, C.c_int64, PD, PD]
        L.or_compute_average.restype = C.c_double
        L.or_compute_average.argtypes = [PD, C.c_long]
        L.or_set_macro_charge.argtypes = [C.c_void_p, C.c_double]
        L.or_set_fiducial.argtypes = [C.c_void_p, C.c_long, C.c_double]
        L.or_get_fiducial.argtypes = [C.c_void_p, C.c_long]
        L.or_get_fiducial.restype = C.c_double
        L.or_set_particle_limit.argtypes = [C.c_void_p, C.c_long]
        _libs[mode] = L
    return _libs[mode]


@contextlib.contextmanager
def species(sp, mode="emath"):
    """the oracle library's particle species (an eh_species) inside the
    block; electrons again afterwards"""
    L = lib(mode)
    L.or_set_species(sp.mass, sp.charge, sp.massMV, sp.radius, sp.relSign)
    try:
        yield
    finally:
        L.or_set_species(9.1093837015e-31, 1.602176634e-19, 0.51099895000, 2.8179403262e-15, 1.0)


class Oracle:
    """or_ctx plus its library (math mode 'glibc' or 'emath')."""

    def __init__(self, mode="glibc"):
        self.L = lib(mode)
        self.ctx = self.L.or_ctx_new()
        self.row_width = 8

    def set_row_width(self, nprops):
        """columns per particle row (totalPropertiesPerParticle) that move
        with a particle; the rows passed in must have at least this many"""
        self.row_width = int(nprops)

    def set_id_slots_per_bunch(self, n):
        """charge->idSlotsPerBunch: > 0 makes bunched-beam WAKE/ZLONGIT act
        per bunch index (column 8)"""
        self.L.or_set_id_slots_per_bunch(self.ctx, int(n))

    @contextlib.contextmanager
    def _rows(self, rows):
        """the row pointers, with the library's row width (a global of the
        checker, as totalPropertiesPerParticle is elegant's) set to this
        context's for the call and put back to 8 after it, so that direct
        calls of the library on 8-column rows stay in bounds"""
        if rows.shape[1] < self.row_width:
            raise ValueError("rows narrower than the row width %d" % self.row_width)
        self.L.or_set_row_width(self.row_width)
        try:
            yield A.rows_pointer(rows)
        finally:
            self.L.or_set_row_width(8)

    def __del__(self):
        try:
            self.L.or_ctx_free(self.ctx)
        except Exception:
            pass

    def do_tracking(self, rows, beamline, Po, n_passes, pass_offset=0):
        """rows: float64 (n, >=8) C-contiguous, modified in place.  Returns nLeft."""
        pc = C.c_double(Po)
        with self._rows(rows) as (ptr, keep):
            r = self.L.or_do_tracking(self.ctx, ptr, rows.shape[0], beamline.pointer(), len(beamline), C.byref(pc),
                                      n_passes, pass_offset)
        if r < 0:
            raise RuntimeError("oracle do_tracking failed: %d" % r)
        self.P_central = pc.value  # changed by CHANGE_P0 cavities
        return int(r)

    def match_energy(self, rows, np_, Po):
        """do_match_energy on rows[:np_] (in place); returns the new P_central"""
        ptr, keep = A.rows_pointer(rows)
        pc = C.c_double(Po)
        self.L.or_do_match_energy(ptr, np_, C.byref(pc))
        return pc.value

    def last_hist(self):
        PD = C.POINTER(C.c_double)
        nb = self.L.or_last_hist(self.ctx, None, None, 0, None, None)
        h = np.zeros(max(nb, 1))
        v = np.zeros(max(nb, 1))
        tmin = C.c_double()
        dt = C.c_double()
        self.L.or_last_hist(self.ctx, h.ctypes.data_as(PD), v.ctypes.data_as(PD), nb, C.byref(tmin), C.byref(dt))
        return h[:nb], v[:nb], tmin.value, dt.value

    def set_macro_charge(self, q):
        self.L.or_set_macro_charge(self.ctx, q)

    def set_particle_limit(self, limit):
        """do_tracking stops before a pass that would start with fewer particles"""
        self.L.or_set_particle_limit(self.ctx, int(limit))

    def set_fiducial(self, elem_index, phase_fiducial):
        """the RFCA at elem_index keeps this fiducial phase (as after its first pass)"""
        self.L.or_set_fiducial(self.ctx, elem_index, phase_fiducial)

    def fiducial(self, elem_index):
        """the fiducial phase this context fixed for the RFCA at elem_index (NaN if none)"""
        return float(self.L.or_get_fiducial(self.ctx, elem_index))

    def wake(self, rows, wake, Po, i_pass):
        with self._rows(rows) as (ptr, keep):
            r = self.L.or_track_through_wake(self.ctx, ptr, rows.shape[0], C.byref(wake), Po, i_pass)
        if r < 0:
            raise RuntimeError("oracle wake failed: %d" % r)
        return r

    def zlongit(self, rows, zl, Po, i_pass, elem_index=0):
        with self._rows(rows) as (ptr, keep):
            r = self.L.or_track_through_zlongit(self.ctx, elem_index, ptr, rows.shape[0], C.byref(zl), Po, i_pass)
        if r < 0:
            raise RuntimeError("oracle zlongit failed: %d" % r)
        return r

    def effort(self):
        return int(self.L.or_ctx_effort(self.ctx))


def track(rows, beamline, Po, n_passes, mode="glibc"):
    """Convenience: copy rows to (n, 8), run the oracle, return (rows, nLeft)."""
    r = np.zeros((rows.shape[0], 8), dtype=np.float64)
    r[:, :rows.shape[1]] = rows[:, :8]
    o = Oracle(mode)
    n = o.do_tracking(r, beamline, Po, n_passes)
    return r, n


def math_eval(fn, x, y=None, mode="emath"):
    """Per-particle math of the given oracle build on an array."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = x if y is None else np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty_like(x)
    pd = C.POINTER(C.c_double)
    lib(mode).or_math_eval(fn, x.ctypes.data_as(pd), y.ctypes.data_as(pd), out.ctypes.data_as(pd), x.size)
    return out


def real_fft(x, inverse=False, mode="glibc"):
    d = np.array(x, dtype=np.float64)
    lib(mode).or_real_fft(d.ctypes.data_as(C.POINTER(C.c_double)), d.size, 1 if inverse else 0)
    return d


def compute_average(v, mode="glibc"):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib(mode).or_compute_average(v.ctypes.data_as(C.POINTER(C.c_double)), v.size)
