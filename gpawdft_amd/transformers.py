"""Restriction / interpolation between grids of 2:1 spacing.

Drop-in for gpaw/transformers.py: `Transformer(gdin, gdout, nn=1,
dtype=float)` with `.apply(input, output=None, phases=None) -> output`.
Paddings, neighbour paddings and skip flags follow
gpaw/transformers.py:20-76; the work runs in the fused gfx950 kernels of
libgpawdft_amd.so (k = 2 nn taps).
"""
import numpy as np

from gpawdft_amd import _lib
from gpawdft_amd._lib import check, lib
from gpawdft_amd.device import DeviceView, stream_ptr


def transformer_paddings(gdin, gdout, nn):
    """Paddings, neighbour paddings, skip flags and kind of a transformer
    between gdin and gdout (gpaw/transformers.py:27-55)."""
    pad_cd = np.empty((3, 2), int)
    neighborpad_cd = np.empty((3, 2), int)
    skip_cd = np.zeros((3, 2), int)
    if (gdin.N_c == 2 * gdout.N_c).all():
        # restriction
        pad_cd[:, 0] = 2 * nn - 1 - 2 * gdout.beg_c + gdin.beg_c
        pad_cd[:, 1] = 2 * nn - 2 + 2 * gdout.end_c - gdin.end_c
        neighborpad_cd[:, 0] = 2 * nn - 2 + 2 * gdout.beg_c - gdin.beg_c
        neighborpad_cd[:, 1] = 2 * nn - 1 - 2 * gdout.end_c + gdin.end_c
        interpolate = False
    else:
        assert (gdout.N_c == 2 * gdin.N_c).all()
        # interpolation
        pad_cd[:, 0] = nn - 1 - gdout.beg_c // 2 + gdin.beg_c
        pad_cd[:, 1] = nn + gdout.end_c // 2 - gdin.end_c
        neighborpad_cd[:, 0] = nn + gdout.beg_c // 2 - gdin.beg_c
        neighborpad_cd[:, 1] = nn - 1 - gdout.end_c // 2 + gdin.end_c
        skip_cd[:, 0] = gdout.beg_c % 2
        skip_cd[:, 1] = gdout.end_c % 2
        interpolate = True
        inpoints = (gdin.n_c[0] + 2 * nn - 1) * (gdin.n_c[1] + 2 * nn - 1)
        outpoints = gdout.n_c[0] * gdout.n_c[1]
        if inpoints > outpoints:
            points = ' x '.join([str(N) for N in gdin.N_c])
            raise ValueError('Cannot construct interpolator.  Grid %s '
                             'may be too small' % points)
    return pad_cd, neighborpad_cd, skip_cd, interpolate


class _Transformer:
    def __init__(self, gdin, gdout, nn=1, dtype=float):
        self.gdin = gdin
        self.gdout = gdout
        self.nn = nn
        assert 1 <= nn <= 4
        self.dtype = dtype

        pad_cd, neighborpad_cd, skip_cd, self.interpolate = \
            transformer_paddings(gdin, gdout, nn)
        assert (pad_cd.ravel() >= 0).all()
        self.ngpin = tuple(int(x) for x in gdin.n_c)
        self.ngpout = tuple(int(x) for x in gdout.n_c)
        self.pad_cd = pad_cd
        self.neighborpad_cd = neighborpad_cd
        self.skip_cd = skip_cd

        h = _lib.ctypes.c_void_p()
        args = [np.ascontiguousarray(x, dtype=np.int64) for x in
                (gdin.n_c, gdout.n_c, pad_cd, neighborpad_cd, skip_cd,
                 gdin.neighbor_cd)]
        check(lib().gpd_transformer_create(
            _lib.ptr_long(args[0]), _lib.ptr_long(args[1]), 2 * nn,
            _lib.ptr_long(args[2]), _lib.ptr_long(args[3]),
            _lib.ptr_long(args[4]), _lib.ptr_long(args[5]), 1,
            int(self.interpolate), _lib.ctypes.byref(h)), 'Transformer')
        self._handle = h

    def __del__(self, _release=_lib.release):
        _release('gpd_transformer_destroy', getattr(self, '_handle', None))
        self._handle = None

    @property
    def handle(self):
        return self._handle

    def apply(self, input, output=None, phases=None):
        """Restrict or interpolate a batch of grids.  For complex128 grids
        phases ((3, 2) complex, [c][0] for the periodic image above, [c][1]
        below) gives the Bloch boundary condition; None means Gamma."""
        if tuple(input.shape[-3:]) != self.ngpin:
            raise ValueError('input grid %s does not match %s'
                             % (tuple(input.shape[-3:]), self.ngpin))
        if output is None:
            if isinstance(input, np.ndarray):
                output = np.empty(tuple(input.shape[:-3]) + self.ngpout,
                                  input.dtype)
            else:
                output = self.gdout.empty(tuple(input.shape[:-3]),
                                          dtype=input.dtype,
                                          device=input.device)
        if tuple(output.shape[-3:]) != self.ngpout:
            raise ValueError('output grid does not match')
        vin = DeviceView(input)
        vout = DeviceView(output, copy_in=False)
        if vin.code != vout.code:
            raise TypeError('in/out dtypes differ')
        nin = int(np.prod(input.shape[:-3], dtype=np.int64))
        ph = None
        if vin.code == _lib.GPD_C128 and phases is not None:
            ph = np.ascontiguousarray(np.asarray(phases, np.complex128)
                                      .reshape(3, 2)).view(np.float64)
        check(lib().gpd_transformer_apply(
            self._handle, vin.code, nin, vin.ptr, vout.ptr,
            None if ph is None else ph.ctypes.data, stream_ptr()),
            'Transformer.apply')
        vout.finish()
        return output

    def get_async_sizes(self):
        return (0, 1)


def Transformer(gdin, gdout, nn=1, dtype=float):
    if nn != 9:
        return _Transformer(gdin, gdout, nn, dtype)

    class T:
        nn = 1

        def apply(self, input, output, phases=None):
            output[:] = input
            return output

    return T()
