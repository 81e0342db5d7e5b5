"""lex_amd: MI355X-native (gfx950) kernels for LEX's per-step dynamics.

Modules mirror the reference's (advection, pressure_equations, pressure_gradient_coriolis,
turbulence, boundary_conditions, one_step_integration, one_sprint_functions, setup_lex)
with the same function names, argument order and state tuples, operating on float32
device tensors in the reference's padded layout.  All arithmetic runs in the HIP
library liblexhip.so (C ABI: include/lex_capi.h).
"""
__version__ = "0.1.0"

from . import _lib  # noqa: F401
