"""MI355X-native (gfx950, HIP, FP64) compressible finite-volume right-hand
side and SSP-RK3 time step of anthony-firn/emergence's `compressible` crate:
WCNS (characteristic WENO5-Z midpoint interpolation) + HLLC + positivity
limiter, hybrid KEEP-PE central flux, CFL reduction and boundary ghosts.

The compute path is the C-ABI library `libemhip.so` (include/em_hip.h);
this package is its Python plumbing.
"""
from ._lib import DeviceArray, EmError, load  # noqa: F401
from .solver import BC, MODE, HipSolver, case_init, padded_shape  # noqa: F401
from .amr import AMR_CASE, HipAmr, HipAmrGroup  # noqa: F401
from .ensemble import HipEnsemble, post_shock, shock_bubble_2d  # noqa: F401
from .fiveeq import HipFiveEq  # noqa: F401
