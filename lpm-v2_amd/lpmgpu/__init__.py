"""lpmgpu -- host side of the MI355X direct-sum library for lpm-v2.

The compute path is liblpm_direct.so (hand-written HIP for gfx950, C ABI in
include/lpm_direct.h); this package binds it with ctypes and provides the
mesh / initial-condition generators that feed it.
"""
from ._lib import LpmError, PREC_FP32_KAHAN, PREC_FP64  # noqa: F401
from .api import BVEBatch, BVESolver, Context, PlaneSolver, load_balance, pse_eps  # noqa: F401
