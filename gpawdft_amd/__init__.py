"""gpawdft_amd -- MI355X-native real-space grid Hamiltonian path.

Drop-in device implementations of GPAW's grid operators (FD stencil,
multigrid transformers and preconditioner, PAW projector LFC, subspace
matrices), computed by the hand-written gfx950 kernels of
libgpawdft_amd.so through its C ABI (include/gpawdft_amd.h).
"""
__version__ = '0.1.0'

from gpawdft_amd._lib import lib, LIB_PATH  # noqa: F401


def library_loaded():
    """True once libgpawdft_amd.so is loaded (raises if it is missing)."""
    return lib() is not None
