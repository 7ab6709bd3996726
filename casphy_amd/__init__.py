"""casphy_amd -- MI355X-native CasPhy bit-packed phylogeny hot path.

Hand-written HIP kernels for gfx950 behind a C ABI (include/casphy_hip.h,
libcasphy_hip.so), with this package as the host-side mirror of the
reference's Rcpp / executable interface.  See api.py.
"""
from .api import *  # noqa: F401,F403
from .api import __all__  # noqa: F401
