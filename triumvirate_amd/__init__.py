"""MI355X (gfx950) engine for Triumvirate's clustering estimators:
bispectrum and three-point correlation function, and the two-point
statistics (power spectrum, correlation function, window).

The estimators run in hand-written HIP kernels (``libtrvgpu.so``) behind a
C ABI (``include/trv_gpu.h``); this package is the Python host layer with
the reference's public call surface.  The FFTLog Hankel and spherical
Bessel transforms used downstream of the estimators are the submodules
``fftlog`` and ``transforms``.
"""

__version__ = '0.1.0'

from .catalogue import ParticleCatalogue  # noqa: F401
from .dataobjs import Binning  # noqa: F401
from .parameters import ParameterSet, fetch_paramset_template  # noqa: F401
from .threept import (  # noqa: F401
    compute_3pcf,
    compute_3pcf_in_gpp_box,
    compute_3pcf_window,
    compute_bispec,
    compute_bispec_in_gpp_box,
    compute_bispec_multipoles,
)
from .twopt import (  # noqa: F401
    compute_corrfunc,
    compute_corrfunc_in_gpp_box,
    compute_corrfunc_window,
    compute_powspec,
    compute_powspec_in_gpp_box,
)
