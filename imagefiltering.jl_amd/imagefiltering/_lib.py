"""ctypes binding of libifm.so (the C ABI in include/ifm.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
imagefiltering.jl_amd/csrc``).  There is no fallback: every filtering call
of this package runs through libifm on the GPU, and a missing library or GPU
raises.
"""
import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.normpath(os.path.join(_HERE, "..", "libifm.so"))

MAXD = 4

# ifm_status
OK, E_ARG, E_DIM, E_NOMEM, E_DEVICE, E_UNSUPPORTED, E_BUDGET = range(7)
# ifm_dtype
U8, N0F8, F32, F64, BF16, F16 = range(6)
# ifm_style
REPLICATE, CIRCULAR, SYMMETRIC, REFLECT, FILL = range(5)
# ifm_precision
PREC_AUTO, PREC_FP32, PREC_F16X3 = range(3)

i64x4 = C.c_int64 * MAXD


class IfmArray(C.Structure):
    _fields_ = [("ndim", C.c_int32), ("dtype", C.c_int32), ("size", i64x4), ("stride", i64x4),
                ("data", C.c_void_p), ("device", C.c_int32), ("reserved", C.c_int32)]


class IfmBorder(C.Structure):
    _fields_ = [("style", C.c_int32), ("reserved", C.c_int32), ("fill", C.c_double),
                ("lo", i64x4), ("hi", i64x4)]


class IfmFactor(C.Structure):
    _fields_ = [("axis", C.c_int32), ("len", C.c_int32), ("first", C.c_int64),
                ("taps", C.POINTER(C.c_double))]


class IfmKernel(C.Structure):
    _fields_ = [("ndim", C.c_int32), ("reserved", C.c_int32), ("size", i64x4), ("first", i64x4),
                ("taps", C.POINTER(C.c_double))]


class IfmIir(C.Structure):
    _fields_ = [("a", C.c_double * 3), ("b", C.c_double * 3), ("scale", C.c_double),
                ("M", C.c_double * 9), ("axis", C.c_int32), ("reserved", C.c_int32)]


# ifm_op
OP_SEPARABLE, OP_DENSE, OP_MEDIAN = range(3)


class IfmRequest(C.Structure):
    _fields_ = [("op", C.c_int32), ("nfactors", C.c_int32), ("in_", IfmArray), ("out", IfmArray),
                ("border", IfmBorder), ("factors", IfmFactor * 2), ("kernel", IfmKernel), ("window", C.c_int64),
                ("nfactors2", C.c_int32), ("reserved", C.c_int32), ("out2", IfmArray),
                ("factors2", IfmFactor * 2)]


class IfmServerConfig(C.Structure):
    _fields_ = [("budget_bytes", C.c_uint64), ("max_batch", C.c_int32), ("max_inflight", C.c_int32)]


class IfmServeConfig(C.Structure):
    _fields_ = [("budget_bytes", C.c_uint64), ("mode", C.c_int32), ("max_batch", C.c_int32),
                ("max_inflight", C.c_int32), ("reserved", C.c_int32)]


class IfmTileRule(C.Structure):
    _fields_ = [("op", C.c_int32), ("dtype", C.c_int32), ("kmax", C.c_int32), ("tile_h", C.c_int32),
                ("max_pixels", C.c_int64)]


TILE_SEP2D = 1
MAX_TILE_RULES = 256

EXPORTS = [
    "ifm_ctx_create", "ifm_ctx_destroy", "ifm_ctx_set_stream", "ifm_ctx_synchronize", "ifm_last_error",
    "ifm_version", "ifm_padarray", "ifm_imfilter_fir", "ifm_imfilter_fir_noalloc", "ifm_imfilter_dense", "ifm_all_finite", "ifm_na_prepare", "ifm_divide",
    "ifm_copy_into", "ifm_local_extrema", "ifm_convert_int", "ifm_imfilter_fft",
    "ifm_imfilter_iir", "ifm_mapwindow_median", "ifm_mapwindow_minmax", "ifm_extrema_filter", "ifm_mapwindow_generic",
    "ifm_batch_run", "ifm_serve", "ifm_fill_uniform",
    "ifm_server_create", "ifm_server_submit", "ifm_server_poll", "ifm_server_wait", "ifm_server_destroy",
    "ifm_device_alloc", "ifm_device_free", "ifm_memcpy_h2d", "ifm_memcpy_d2h", "ifm_memset_d",
    "ifm_timer_start", "ifm_timer_stop", "ifm_device_count", "ifm_set_tile_rules", "ifm_tile_height",
]

_lib = None
_ctx = None
_lock = threading.Lock()


def load(path=LIBPATH):
    """Load libifm.so (no GPU needed); raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libifm.so is not built at %s (run __graft_entry__.build())" % path)
    L = C.CDLL(path)
    P = C.POINTER
    vp = C.c_void_p
    L.ifm_ctx_create.argtypes = [C.c_int, vp, P(vp)]
    L.ifm_ctx_destroy.argtypes = [vp]
    L.ifm_ctx_set_stream.argtypes = [vp, vp]
    L.ifm_ctx_synchronize.argtypes = [vp]
    L.ifm_last_error.restype = C.c_char_p
    L.ifm_last_error.argtypes = [vp]
    L.ifm_version.restype = C.c_char_p
    L.ifm_padarray.argtypes = [vp, P(IfmArray), P(IfmArray), P(IfmBorder)]
    L.ifm_imfilter_fir.argtypes = [vp, P(IfmArray), P(IfmArray), P(C.c_int64), P(IfmFactor), C.c_int,
                                   P(IfmBorder)]
    L.ifm_imfilter_fir_noalloc.argtypes = [vp, P(IfmArray), P(IfmArray), P(C.c_int64), P(IfmFactor), C.c_int,
                                           P(IfmBorder)]
    L.ifm_imfilter_dense.argtypes = [vp, P(IfmArray), P(IfmArray), P(C.c_int64), P(IfmKernel), P(IfmBorder),
                                     C.c_int]
    L.ifm_imfilter_fft.argtypes = [vp, P(IfmArray), P(IfmArray), P(C.c_int64), P(IfmKernel), P(IfmBorder)]
    L.ifm_all_finite.argtypes = [vp, P(IfmArray), P(C.c_int)]
    L.ifm_na_prepare.argtypes = [vp, P(IfmArray), P(IfmArray), P(IfmArray), P(C.c_int)]
    L.ifm_divide.argtypes = [vp, P(IfmArray), P(IfmArray)]
    L.ifm_copy_into.argtypes = [vp, P(IfmArray), P(IfmArray), P(C.c_int64)]
    L.ifm_convert_int.argtypes = [vp, P(IfmArray), vp, C.c_int, C.c_int, C.c_int, C.c_double, P(C.c_int)]
    L.ifm_local_extrema.argtypes = [vp, P(IfmArray), P(IfmArray), P(C.c_int64), P(C.c_int32), C.c_int]
    L.ifm_imfilter_iir.argtypes = [vp, P(IfmArray), P(IfmArray), P(IfmIir), C.c_int, P(IfmBorder)]
    L.ifm_mapwindow_median.argtypes = [vp, P(IfmArray), P(IfmArray), P(C.c_int64), P(C.c_int64), P(C.c_int64),
                                       P(IfmBorder)]
    L.ifm_mapwindow_minmax.argtypes = [vp, P(IfmArray), P(IfmArray), P(IfmArray), P(C.c_int64), P(C.c_int64),
                                       P(C.c_int64), P(IfmBorder)]
    L.ifm_extrema_filter.argtypes = [vp, P(IfmArray), P(IfmArray), P(IfmArray), P(C.c_int64)]
    L.ifm_mapwindow_generic.argtypes = [vp, P(IfmArray), P(IfmArray), C.c_int, C.c_double, P(C.c_int64),
                                        P(C.c_int64), P(C.c_int64), P(C.c_int64), P(IfmBorder), P(C.c_int)]
    L.ifm_batch_run.argtypes = [vp, P(IfmRequest), C.c_int]
    L.ifm_serve.argtypes = [vp, P(IfmRequest), P(C.c_double), C.c_int, P(IfmServeConfig), P(C.c_double),
                            P(C.c_double)]
    L.ifm_fill_uniform.argtypes = [vp, P(IfmArray), C.c_uint64]
    L.ifm_server_create.argtypes = [vp, P(IfmServerConfig), P(vp)]
    L.ifm_server_submit.argtypes = [vp, P(IfmRequest), P(C.c_uint64)]
    L.ifm_server_poll.argtypes = [vp, C.c_uint64, P(C.c_int)]
    L.ifm_server_wait.argtypes = [vp, C.c_uint64]
    L.ifm_server_destroy.argtypes = [vp]
    L.ifm_device_alloc.argtypes = [vp, C.c_uint64, P(vp)]
    L.ifm_device_free.argtypes = [vp, vp]
    L.ifm_memcpy_h2d.argtypes = [vp, vp, vp, C.c_uint64]
    L.ifm_memcpy_d2h.argtypes = [vp, vp, vp, C.c_uint64]
    L.ifm_memset_d.argtypes = [vp, vp, C.c_int, C.c_uint64]
    L.ifm_timer_start.argtypes = [vp]
    L.ifm_timer_stop.argtypes = [vp, P(C.c_float)]
    L.ifm_device_count.argtypes = [P(C.c_int)]
    L.ifm_set_tile_rules.argtypes = [vp, P(IfmTileRule), C.c_int]
    L.ifm_tile_height.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int64, P(C.c_int)]
    _lib = L
    return L


def ctx():
    """The process-wide context on device 0 (created on first use)."""
    global _ctx
    with _lock:
        if _ctx is None:
            L = load()
            h = C.c_void_p()
            rc = L.ifm_ctx_create(0, None, C.byref(h))
            if rc != OK:
                raise RuntimeError("libifm: no usable GPU (ifm_ctx_create returned %d)" % rc)
            _ctx = h
            new = True
        else:
            new = False
    if new:
        from . import tiles
        tiles.load()  # tuned tile shapes of this machine, when a tuning result is cached
    return _ctx


def check(rc):
    """Map an ifm_status to the reference's exception types."""
    if rc == OK:
        return
    from .errors import ArgumentError, DimensionMismatch
    msg = load().ifm_last_error(ctx()).decode(errors="replace")
    if rc == E_ARG:
        raise ArgumentError(msg)
    if rc == E_DIM:
        raise DimensionMismatch(msg)
    if rc == E_UNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == E_NOMEM:
        raise MemoryError(msg)
    raise RuntimeError("libifm error %d: %s" % (rc, msg))


def i64(vals, fill=0):
    v = list(vals) + [fill] * (MAXD - len(vals))
    return i64x4(*[int(x) for x in v[:MAXD]])
