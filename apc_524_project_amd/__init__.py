"""emopic-mi355x: EMOOPIC's per-timestep particle-in-cell loop on MI355X.

The compute path is libemopic.so (hand-written HIP for gfx950) behind the C
ABI in include/emo_api.h; this package binds it with ctypes for tests,
benchmarks and the autograd wrapper.
"""
from ._lib import (  # noqa: F401
    BC_EXTERNAL, BC_PERIODIC, BC_REFLECTING, FP32_CHOSEN, FP64_PLAIN, Context, EmoConfig,
    EmoError, EmoLoader, lib,
)
from .workloads import Workload, make_workload  # noqa: F401

__all__ = ["Context", "EmoConfig", "EmoLoader", "EmoError", "lib", "Workload", "make_workload"]
