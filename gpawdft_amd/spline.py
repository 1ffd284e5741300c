"""Radial cubic splines (host FP64 set-up).

Drop-in for gpaw/spline.py:11-45: `Spline(l, rmax, f_g)` with the last
value forced to zero, `.get_cutoff()`, `.get_angular_momentum_number()`,
`__call__(r)`, `.map(r_x)`.  Backed by gpd_spline_* of libgpawdft_amd.so.
"""
import numpy as np

from gpawdft_amd import _lib
from gpawdft_amd._lib import check, lib


class Spline:
    def __init__(self, l, rmax, f_g):
        assert 0.0 < rmax
        f_g = np.array(f_g, float)
        f_g[-1] = 0.0            # copy, so the caller's array is unchanged
        self.l = int(l)
        self.rmax = float(rmax)
        self.f_g = f_g
        h = _lib.ctypes.c_void_p()
        check(lib().gpd_spline_create(self.l, self.rmax, _lib.ptr_double(f_g),
                                      len(f_g), _lib.ctypes.byref(h)),
              'Spline')
        self._handle = h

    def __del__(self, _release=_lib.release):
        _release('gpd_spline_destroy', getattr(self, '_handle', None))
        self._handle = None

    @property
    def handle(self):
        return self._handle

    def get_cutoff(self):
        return lib().gpd_spline_get_cutoff(self._handle)

    def get_angular_momentum_number(self):
        return lib().gpd_spline_get_angular_momentum_number(self._handle)

    def __call__(self, r):
        assert r >= 0.0
        return lib().gpd_spline_value(self._handle, float(r))

    def map(self, r_x):
        return np.vectorize(self, [float])(r_x)
