"""Exception types mirroring the Julia errors the reference throws at its public boundary.

``ArgumentError`` and ``DimensionMismatch`` are raised before any O(N^2) work, as in
src/Calculations/dispatch.jl:7-13, src/Calculations/shapes.jl:51-105 and
src/Calculations/serial.jl:111-120.
"""


class ArgumentError(ValueError):
    """Julia ``ArgumentError``."""


class DimensionMismatch(ValueError):
    """Julia ``DimensionMismatch``."""


class BackendError(RuntimeError):
    """The gfx950 library reported a device failure or is not loadable."""
