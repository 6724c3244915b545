"""ptq2: MI355X-native matrix-free Q2-P1disc Stokes operators (pTatin3D's operator path).

The compute lives in libptq2.so (hand-written gfx950 HIP kernels behind the C ABI of
include/ptq2.h); this package is the host-side mirror of the reference's operator
objects and of the solver that consumes them.
"""
from . import mesh  # noqa: F401
from ._lib import (ADD, DEVICE, FP32, MATMULT, PTQ2Error, LIB_PATH, SIGNATURES)  # noqa: F401
from .operators import Context, MatA11MF, MatStokesMF, a11_batch, register_a11_op  # noqa: F401
