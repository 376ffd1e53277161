"""MI355X-native (gfx950) SCFT string-method hot path.

Implements SCFT_string's solve_field path (density::props, solve_field0, solve_field,
FreeE) as hand-written HIP kernels behind the C-ABI in include/scft_api.h. The Python
layer is a thin ctypes mirror used by tests and the benchmark.
"""
from .capi import Context, ScftError, LIB_PATH, iterate_concurrent  # noqa: F401

__all__ = ["Context", "ScftError", "LIB_PATH", "iterate_concurrent"]
