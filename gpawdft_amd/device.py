"""Device-array plumbing.

Grids live in HBM as torch tensors (C-contiguous, (..., n0, n1, n2), last
axis contiguous -- the reference's numpy layout).  torch supplies the
allocator and the current HIP stream; all arithmetic is done by the HIP
kernels of libgpawdft_amd.so.

Host numpy arrays are accepted wherever the reference takes numpy arrays:
they are copied to the device, processed there, and the results copied
back into the caller's array (a PCIe round trip; keep data resident for
speed).
"""
import numpy as np

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None

from gpawdft_amd._lib import GPD_C128, GPD_F32, GPD_F64

DEFAULT_DEVICE = 'cuda'


def _require_torch():
    if torch is None:
        raise ImportError('gpawdft_amd needs torch (ROCm) for device memory')


def dtype_code(dt):
    """Map float/np.float64/np.float32/torch dtypes to the ABI dtype code."""
    if torch is not None and isinstance(dt, torch.dtype):
        if dt == torch.float64:
            return GPD_F64
        if dt == torch.float32:
            return GPD_F32
        if dt == torch.complex128:
            return GPD_C128
        raise TypeError('unsupported dtype %s' % dt)
    dt = np.dtype(dt)
    if dt == np.float64:
        return GPD_F64
    if dt == np.float32:
        return GPD_F32
    if dt == np.complex128:
        return GPD_C128
    raise TypeError('unsupported dtype %s' % dt)


def torch_dtype(dt):
    _require_torch()
    if isinstance(dt, torch.dtype):
        return dt
    return {GPD_F64: torch.float64, GPD_F32: torch.float32,
            GPD_C128: torch.complex128}[dtype_code(dt)]


def stream_ptr():
    """The current HIP stream of torch, as a void* for the C ABI."""
    return torch.cuda.current_stream().cuda_stream


def is_device_array(a):
    return torch is not None and isinstance(a, torch.Tensor) and a.is_cuda


def require_device(a, name='array'):
    if not is_device_array(a):
        raise TypeError('%s must be a device (cuda) tensor' % name)
    if not a.is_contiguous():
        raise ValueError('%s must be C-contiguous' % name)
    return a


class DeviceView:
    """Pair a caller array with its device tensor.

    For a device tensor the view is the tensor itself.  For a numpy array a
    device copy is made (if `copy_in`), and `finish()` writes the device
    result back into the numpy array.
    """

    def __init__(self, a, copy_in=True):
        _require_torch()
        self.host = None
        if is_device_array(a):
            if not a.is_contiguous():
                raise ValueError('device arrays must be C-contiguous')
            self.t = a
        elif isinstance(a, np.ndarray):
            if not a.flags.c_contiguous:
                raise ValueError('arrays must be C-contiguous')
            self.host = a
            if copy_in:
                self.t = torch.from_numpy(a).to(DEFAULT_DEVICE)
            else:
                self.t = torch.empty(a.shape, dtype=torch_dtype(a.dtype),
                                     device=DEFAULT_DEVICE)
        else:
            raise TypeError('expected a cuda tensor or a numpy array, got %r'
                            % type(a))
        self.code = dtype_code(self.t.dtype)

    @property
    def ptr(self):
        return self.t.data_ptr()

    def finish(self):
        if self.host is not None:
            self.host[...] = self.t.cpu().numpy()


def empty(shape, dtype=float, device=None):
    _require_torch()
    return torch.empty(tuple(shape), dtype=torch_dtype(dtype),
                       device=device or DEFAULT_DEVICE)


def zeros(shape, dtype=float, device=None):
    _require_torch()
    return torch.zeros(tuple(shape), dtype=torch_dtype(dtype),
                       device=device or DEFAULT_DEVICE)


def to_device(a, dtype=None):
    """Copy a numpy array (or tensor) to a contiguous device tensor."""
    _require_torch()
    if isinstance(a, torch.Tensor):
        t = a.to(DEFAULT_DEVICE)
    else:
        t = torch.from_numpy(np.ascontiguousarray(a)).to(DEFAULT_DEVICE)
    if dtype is not None:
        t = t.to(torch_dtype(dtype))
    return t.contiguous()
