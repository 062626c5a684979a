"""MI355X-native many-sphere Stokes boundary-integral mobility operator.

The compute path is libsphstk.so (hand-written HIP kernels for gfx950 behind
the C ABI of include/sphstk.h); this package holds its Python binding, a
mirror of the reference's BIE classes, the seeded input generators and the
XYZ files and Euler step on either side of the solve.
"""
from . import configs, system  # noqa: F401
from ._lib import LIB_PATH, SphstkError, load  # noqa: F401
from .bie import (FLAG_TIMING, PREC_FP64, PREC_MIXED, SphereSTKMobMat, SphereSTKSHOperator,  # noqa: F401
                  solver_params)

__version__ = "0.1.0"
