"""MI355X (gfx950) implementation of TAN's time-varying multichannel
convolution path behind the TANContext / TANConvolution / TANFFT interfaces.

The compute runs in the in-tree HIP library tan_amd/lib/libtan_mi355x.so
(C ABI: include/tan_mi355x.h); this package is the host-side mirror of the
reference's C++ interface.
"""
from .tan import *  # noqa: F401,F403
from . import signals  # noqa: F401
from ._lib import LIB_PATH, load as load_library  # noqa: F401
