"""Element types of the filtering path and Julia's promotion rules for them.

numpy dtypes stand for Julia's Float16/32/64 and UInt8.  N0f8 (FixedPointNumbers,
UInt8 storage read as code/255) and BFloat16 (UInt16 storage) are marked with
ndarray subclasses, since numpy has no such scalar types.
"""
import numpy as np

from . import _lib


class Eltype:
    def __init__(self, name, code, storage, kind):
        self.name, self.code, self.storage, self.kind = name, code, np.dtype(storage), kind

    def __repr__(self):
        return self.name

    @property
    def itemsize(self):
        return self.storage.itemsize


UInt8 = Eltype("UInt8", _lib.U8, np.uint8, "int")
N0f8 = Eltype("N0f8", _lib.N0F8, np.uint8, "fixed")
Float16 = Eltype("Float16", _lib.F16, np.float16, "float")
BFloat16 = Eltype("BFloat16", _lib.BF16, np.uint16, "float")
Float32 = Eltype("Float32", _lib.F32, np.float32, "float")
Float64 = Eltype("Float64", _lib.F64, np.float64, "float")
# Integer eltypes other than UInt8 are not library storage types: images of
# them enter the filters as Float64, and integer results leave through
# ifm_convert_int (an exact conversion, or InexactError).
Int8 = Eltype("Int8", None, np.int8, "int")
Int16 = Eltype("Int16", None, np.int16, "int")
Int32 = Eltype("Int32", None, np.int32, "int")
Int64 = Eltype("Int64", None, np.int64, "int")
UInt16 = Eltype("UInt16", None, np.uint16, "int")
UInt32 = Eltype("UInt32", None, np.uint32, "int")
UInt64 = Eltype("UInt64", None, np.uint64, "int")
Bool = Eltype("Bool", None, np.bool_, "int")

_BY_DTYPE = {np.dtype(np.uint8): UInt8, np.dtype(np.float16): Float16, np.dtype(np.float32): Float32,
             np.dtype(np.float64): Float64, np.dtype(np.int64): Int64, np.dtype(np.int32): Int32,
             np.dtype(np.int16): Int16, np.dtype(np.int8): Int8, np.dtype(np.uint16): UInt16,
             np.dtype(np.uint32): UInt32, np.dtype(np.uint64): UInt64, np.dtype(np.bool_): Bool}

# (bits, signed) of the integer eltypes
INT_INFO = {"Bool": (1, False), "Int8": (8, True), "Int16": (16, True), "Int32": (32, True), "Int64": (64, True),
            "UInt8": (8, False), "UInt16": (16, False), "UInt32": (32, False), "UInt64": (64, False)}
_INT_BY = {v: k for k, v in INT_INFO.items() if k != "Bool"}


def _int_eltype(bits, signed):
    return globals()[_INT_BY[(bits, signed)]]


def promote_int(a, b):
    """promote_type for two integer eltypes: Bool yields to the other; of a
    signed and an unsigned type the larger wins, the unsigned one on a tie."""
    if a is b:
        return a
    if a is Bool:
        return b
    if b is Bool:
        return a
    (ba, sa), (bb, sb) = INT_INFO[a.name], INT_INFO[b.name]
    if sa == sb:
        return a if ba >= bb else b
    s, u = (a, b) if sa else (b, a)
    return s if INT_INFO[s.name][0] > INT_INFO[u.name][0] else u


def signed_type(S):
    """signed_type for the Laplacian's output (src/imfilter.jl:1143-1146)."""
    return {"UInt8": Int16, "UInt16": Int32, "UInt32": Int64}.get(S.name, Int64)


class N0f8Array(np.ndarray):
    """UInt8 codes read as N0f8 (code/255) by linear filters."""


class BF16Array(np.ndarray):
    """UInt16 bit patterns of BFloat16 values."""


def n0f8(codes):
    """Tag a uint8 array as N0f8 (like reinterpret(N0f8, codes))."""
    return np.asarray(codes, dtype=np.uint8).view(N0f8Array)


def bf16(values):
    """Round float values to BFloat16 (nearest even) and tag the bits."""
    a = np.array(values, dtype=np.float32, ndmin=1)
    shape = np.shape(values)
    u = a.view(np.uint32).astype(np.uint64)
    r = np.where(np.isnan(a), (u >> 16) | 0x40, (u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r.reshape(shape).view(BF16Array)


def bf16_to_float32(bits):
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def eltype_of(x):
    """Element type of an array (numpy, tagged numpy, DeviceArray) or kernel."""
    et = getattr(x, "eltype", None)
    if isinstance(et, Eltype):
        return et
    if isinstance(x, N0f8Array):
        return N0f8
    if isinstance(x, BF16Array):
        return BFloat16
    dt = np.asarray(x).dtype
    if dt in _BY_DTYPE:
        return _BY_DTYPE[dt]
    raise TypeError("unsupported element type %s" % dt)


def as_eltype(T):
    """Accept an Eltype, a numpy dtype or scalar type (e.g. np.float32)."""
    if isinstance(T, Eltype):
        return T
    try:
        dt = np.dtype(T)
    except TypeError:
        raise TypeError("not an element type: %r" % (T,))
    if dt in _BY_DTYPE:
        return _BY_DTYPE[dt]
    raise TypeError("unsupported element type %s" % dt)


def is_eltype_spec(x):
    if isinstance(x, Eltype):
        return True
    if isinstance(x, type) and issubclass(x, np.generic):
        return True
    return isinstance(x, np.dtype)


_FLOATRANK = {"Float16": 1, "BFloat16": 1, "Float32": 2, "Float64": 3}


def promote_float(a, b):
    """promote_type for two float eltypes (BFloat16 with Float16 -> Float32)."""
    if a is b:
        return a
    ra, rb = _FLOATRANK[a.name], _FLOATRANK[b.name]
    if ra == rb:
        return Float32
    return a if ra > rb else b


def mul_type(S, T):
    """typeof(zero(S) * zero(T)) for a pixel eltype S and a tap eltype T
    (src/imfilter.jl:1131-1134)."""
    if T.kind == "float":
        if S.kind == "float":
            return promote_float(S, T)
        if S is N0f8:  # FixedPointNumbers promotes N0f8 with a float to floattype(N0f8)=Float32 first
            return promote_float(Float32, T)
        return T       # UInt8 * Float -> Float
    # integer / Bool taps
    if S.kind == "float":
        return S
    if S is N0f8:
        return Float32
    # zero(S) * zero(T) + zero(S) * zero(T): Bool + Bool is Int
    R = promote_int(S, T)
    return Int64 if R is Bool else R


def promote(a, b):
    if a is b:
        return a
    if a.kind == "float" and b.kind == "float":
        return promote_float(a, b)
    if a.kind == "float":
        return a
    if b.kind == "float":
        return b
    if a.kind == "fixed" or b.kind == "fixed":
        return Float32
    return promote_int(a, b)
