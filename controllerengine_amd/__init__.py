"""ControllerEngine for AMD Instinct MI355X: the candidate-evaluation chain of ControllerEngine as
hand-written gfx950 HIP kernels behind the reference's C ABI (see include/ControllerEngine.h)."""
from ._abi import PRECISION_FP64, PRECISION_MIXED, lib  # noqa: F401
from .problem import Problem, encode  # noqa: F401
