"""Seeded input generators through the library's host routines (eh_gen.c):
elegant's DLARAN uniform stream, Gaussian deviates with a cutoff, element
error perturbations and Gaussian bunches in Twiss coordinates."""
import ctypes as C

import numpy as np

from . import _abi as A
from . import native

UNIFORM, GAUSSIAN, PLUS_OR_MINUS = 0, 1, 2


class Rng:
    """random_1_elegant-style stream (drand_oag.c:94-180) seeded with iseed."""

    def __init__(self, iseed):
        self._r = A.eh_rng()
        native.check(None, native.lib().eh_rng_seed(C.byref(self._r), int(iseed)))

    def copy(self):
        """an independent generator at the same point of the stream"""
        r = Rng.__new__(Rng)
        r._r = A.eh_rng.from_buffer_copy(self._r)
        return r

    def uniform(self, n=None):
        f = native.lib().eh_rng_uniform
        if n is None:
            return f(C.byref(self._r))
        return np.array([f(C.byref(self._r)) for _ in range(n)])

    def gauss(self, n=None, mean=0.0, sigma=1.0, limit=0.0):
        f = native.lib().eh_gauss_rn_lim
        if n is None:
            return f(C.byref(self._r), mean, sigma, limit)
        return np.array([f(C.byref(self._r), mean, sigma, limit) for _ in range(n)])

    def perturbation(self, amplitude, cutoff=3.0, error_type=GAUSSIAN, n=None):
        if n is None:
            return native.lib().eh_perturbation(C.byref(self._r), amplitude, cutoff, error_type)
        out = np.zeros(n)
        native.check(None, native.lib().eh_perturbations(C.byref(self._r), amplitude, cutoff, error_type,
                                                         native._dptr(out), n))
        return out

    def gaussian_distribution(self, n, s1, s2, limit, symmetrize=False, limit_invar=0.0, beta=0.0, halo=False):
        rows = np.zeros((n, 2))
        ptr, keep = A.rows_pointer(rows)
        native.check(None, native.lib().eh_gaussian_distribution(C.byref(self._r), ptr, n, 0, s1, s2,
                                                                 int(symmetrize), limit, limit_invar, beta,
                                                                 int(halo)))
        return rows


def gaussian_bunch(n, seed, x_plane, y_plane, longit, first_id=1):
    """(n, 8) rows: x_plane/y_plane = (emit, beta, alpha, cutoff), longit =
    (sigma_s, sigma_dp, alpha_z, cutoff), IDs first_id.., column 7 = -1."""
    r = Rng(seed)
    rows = np.zeros((n, 8))
    ptr, keep = A.rows_pointer(rows)
    xa = np.ascontiguousarray(x_plane, dtype=np.float64)
    ya = np.ascontiguousarray(y_plane, dtype=np.float64)
    la = np.ascontiguousarray(longit, dtype=np.float64)
    native.check(None, native.lib().eh_gaussian_bunch(C.byref(r._r), ptr, n, 8, native._dptr(xa), native._dptr(ya),
                                                      native._dptr(la), first_id))
    return rows
