"""Exception types of the reference's error paths (Julia Base exceptions)."""


class ArgumentError(ValueError):
    """Julia ArgumentError (bad border string, Fill convert, window spec...)."""


class DimensionMismatch(ValueError):
    """Julia DimensionMismatch (output axes, kernel vs padded input, IIR length)."""


class MethodError(TypeError):
    """Julia MethodError: a call shape the reference does not define."""


class InexactError(ValueError):
    """Julia InexactError (a result that does not fit the output eltype)."""
