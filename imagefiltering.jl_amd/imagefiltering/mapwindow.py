"""mapwindow / mapwindow! (src/mapwindow.jl), on the GPU through libifm.
Julia's `mapwindow!` is `mapwindow_`, and `median!` is `median_`.

Window functions are named like the reference's: ``median_`` (Statistics.
median!, rewritten to median_fast!), ``median`` (Statistics.median),
``minimum``, ``maximum``, ``extrema``, ``mean``, ``sum_`` (Base.sum), ``first``
and ``quantile(p)`` (the window function v -> Statistics.quantile(v, p)).
Python's and numpy's equivalents are accepted too (``sum``, ``min``, ``max``,
``numpy.mean``, ``numpy.median``, ...).  Arbitrary functions are outside this
package's GPU path and raise.  extrema returns an array of shape img.shape +
(2,) holding (min, max) pairs.

``indices`` takes one range per axis (0-based, any positive step; the
strided subsets of src/mapwindow.jl:156-189): the output holds f at those
window centres, with the ranges' lengths as its shape.
"""
import builtins
import statistics

import numpy as np

from . import _lib
from .border import STYLES, Fill, Inner, NA, NoPad, Pad, borderinstance, convert_fill
from .device import DeviceArray, empty_like, host_dense, ifm_array
from .eltypes import BFloat16, Float16, Float32, Float64, N0f8, UInt8, eltype_of
from .errors import ArgumentError, MethodError
from .offsetarrays import OffsetArray


def median_(v):
    """Statistics.median! on a window buffer (rewritten to median_fast!)."""
    raise TypeError("median_ is a window function for mapwindow")


def median(v):
    """Statistics.median on a window buffer."""
    raise TypeError("median is a window function for mapwindow")


def minimum(v):
    raise TypeError("minimum is a window function for mapwindow")


def maximum(v):
    raise TypeError("maximum is a window function for mapwindow")


def extrema(v):
    raise TypeError("extrema is a window function for mapwindow")


def mean(v):
    """Statistics.mean on a window buffer."""
    raise TypeError("mean is a window function for mapwindow")


def sum_(v):
    """Base.sum on a window buffer."""
    raise TypeError("sum_ is a window function for mapwindow")


def first(v):
    """Base.first on a window buffer (its first element, column-major)."""
    raise TypeError("first is a window function for mapwindow")


class quantile:
    """The window function v -> Statistics.quantile(v, p) (type-7 linear
    interpolation between order statistics)."""

    def __init__(self, p):
        p = float(p)
        if not 0.0 <= p <= 1.0:
            raise ArgumentError("quantile p must lie in [0, 1]")
        self.p = p

    def __call__(self, v):
        raise TypeError("quantile(p) is a window function for mapwindow")

    def __repr__(self):
        return "quantile(%r)" % self.p


sum_.__name__ = "sum"

# the library's window-function codes (include/ifm.h ifm_window_op)
_WIN = {"sum": 0, "mean": 1, "minimum": 2, "maximum": 3, "first": 4, "median!": 5, "median": 6, "quantile": 7}


def _canonical_f(f):
    """The reference function a window function stands for (Python's and
    numpy's equivalents included), or None."""
    if isinstance(f, quantile):
        return "quantile"
    table = [((median_,), "median!"), ((median, np.median, statistics.median), "median"),
             ((minimum, builtins.min, np.min, np.amin), "minimum"),
             ((maximum, builtins.max, np.max, np.amax), "maximum"),
             ((extrema,), "extrema"), ((mean, np.mean, statistics.mean), "mean"),
             ((sum_, builtins.sum, np.sum), "sum"), ((first,), "first")]
    for fs, name in table:
        if any(f is g for g in fs):
            return name
    return None


median_.__name__ = "median!"


def resolve_window(window):
    """src/mapwindow.jl:133-152: odd Dims or an Integer -> centred ranges;
    ranges are taken as they are."""
    if isinstance(window, range):
        return (window,)
    if isinstance(window, (int, np.integer)):
        w = int(window)
        if w % 2 != 1:
            raise RuntimeError("window must be odd, got %d" % w)
        return (range(-(w >> 1), (w >> 1) + 1),)
    window = tuple(window)
    if len(window) == 0:
        raise ArgumentError("empty window")
    if all(isinstance(w, range) for w in window):
        return window
    if all(isinstance(w, (int, np.integer)) for w in window):
        if not all(w % 2 == 1 for w in window):
            raise RuntimeError("entries in window must be odd, got %r" % (window,))
        return tuple(range(-(w >> 1), (w >> 1) + 1) for w in window)
    return tuple(w if isinstance(w, range) else resolve_window(w)[0] for w in window)


def _median_eltype(S):
    """typeof(middle(x)) (Statistics: AbstractFloat -> itself, integers ->
    Float64, other Numbers -> (x + zero(x))/1, which FixedPointNumbers
    promotes to Float32 for N0f8)."""
    if S in (Float16, BFloat16, Float32, Float64):
        return S
    if S is N0f8:
        return Float32
    return Float64


def _border(border, S):
    b = _lib.IfmBorder()
    if isinstance(border, Fill):
        b.style = _lib.FILL
        b.fill = convert_fill(border.value, S)
    elif isinstance(border, Pad):
        b.style = STYLES[border.style]
    elif isinstance(border, Inner):
        b.style = _lib.REPLICATE
    else:
        raise NotImplementedError("mapwindow border %r" % (border,))
    return b


def _region(img, win, border, indices):
    """Window centres: (first, shape, step) per axis in 0-based image indices."""
    N = len(img.shape)
    if indices is not None:
        if isinstance(indices, range):
            indices = (indices,)
        if len(indices) != N:
            raise ArgumentError("indices must give one range per dimension")
        first, shape, step = [], [], []
        for d, r in enumerate(indices):
            if not isinstance(r, range) or r.step < 1:
                raise ArgumentError("indices must be ranges with positive steps")
            if len(r) and (r[0] < 0 or r[-1] >= img.shape[d]):
                raise ArgumentError("indices lie outside the image axes")
            if isinstance(border, Inner) and len(r):
                # the reference has no edge map for Inner(): every centre
                # must keep its window inside the image
                if r[0] + win[d].start < 0 or r[-1] + win[d].stop - 1 > img.shape[d] - 1:
                    raise MethodError("Inner() border with indices whose windows leave the image")
            first.append(r.start if len(r) else 0)
            shape.append(len(r))
            step.append(r.step)
        return first, shape, step
    if isinstance(border, Inner):
        first, shape = [], []
        for d in range(N):
            lo = max(0, -win[d].start)
            hi = img.shape[d] - 1 - max(0, win[d].stop - 1)
            first.append(lo)
            shape.append(max(0, hi - lo + 1))
        return first, shape, [1] * N
    return [0] * N, list(img.shape), [1] * N


def mapwindow(f, img, window, border=None, indices=None, callmode="copy!"):
    """mapwindow(f, img, window; border="replicate", indices=axes(img)) (src/mapwindow.jl:75-85)."""
    if callmode != "copy!":
        raise RuntimeError("Only callmode=:copy! is currently supported")
    img = img if isinstance(img, DeviceArray) else host_dense(img)
    S = eltype_of(img)
    fname = _canonical_f(f)
    if fname is None:
        raise NotImplementedError("mapwindow with window function %r runs on the CPU only in the reference; this "
                                  "GPU package provides median!, median, minimum, maximum, extrema, mean, sum, "
                                  "first and quantile(p)" % (f,))
    if fname == "extrema" and border is None and indices is None and not any(
            isinstance(w, range) for w in (window if isinstance(window, tuple) else (window,))):
        return _extrema_filter(img, window)
    win = resolve_window(window)
    N = len(img.shape)
    if len(win) != N:
        raise DimensionMismatch_("window and image dimensionality differ")
    border = borderinstance("replicate" if border is None else border)
    if isinstance(border, (NA, NoPad)):
        raise NotImplementedError("mapwindow border %r" % (border,))
    first, shape, step = _region(img, win, border, indices)
    if isinstance(border, Inner):
        for d in range(N):
            if len(win[d]) > img.shape[d]:
                from .errors import DimensionMismatch
                raise DimensionMismatch("window larger than the image with Inner()")
    b = _border(border, S)
    wf = _lib.i64([w.start for w in win])
    wl = _lib.i64([len(w) for w in win], 1)
    of = _lib.i64(first)
    L = _lib.load()
    ctx = _lib.ctx()
    Lw = int(np.prod([len(w) for w in win]))
    strided = any(s != 1 for s in step)
    # Statistics.median is NaN for a window holding NaN (median_fast! orders
    # NaN last instead): images with NaN take the generic kernel
    stats_nan = fname == "median" and S.kind == "float" and not _all_finite(img)
    if strided or stats_nan or fname in ("mean", "sum", "first", "quantile") or (fname == "median" and Lw % 2 == 0):
        return _wrap(_generic(fname, f, img, S, win, b, first, shape, step), first, border, indices)
    if fname in ("median!", "median"):
        # the kernel writes typeof(middle(x)) directly (Float32 for N0f8,
        # Float64 for UInt8, the eltype itself for floats)
        T = _median_eltype(S)
        res = empty_like(img, T, shape)
        _lib.check(L.ifm_mapwindow_median(ctx, ifm_array(img), ifm_array(res), of, wf, wl, b))
        return _wrap(res, first, border, indices)
    omin = empty_like(img, S, shape) if fname in ("minimum", "extrema") else None
    omax = empty_like(img, S, shape) if fname in ("maximum", "extrema") else None
    _lib.check(L.ifm_mapwindow_minmax(ctx, ifm_array(img), ifm_array(omin) if omin is not None else None,
                                      ifm_array(omax) if omax is not None else None, of, wf, wl, b))
    if fname == "minimum":
        return _wrap(omin, first, border, indices)
    if fname == "maximum":
        return _wrap(omax, first, border, indices)
    return _wrap(_pair(omin, omax), first, border, indices)


def _all_finite(img):
    from .imfilter import _all_finite as af
    return af(img)


def _generic_eltype(fname, S):
    """typeof(f(window)) of the window functions the generic kernel runs
    (Statistics / Base rules: floats keep their type, UInt8 gives Float64
    for sum's widened mean, median and quantile; quantile promotes with its
    Float64 p); None where the GPU path does not provide the pair."""
    if fname in ("minimum", "maximum", "first", "extrema"):
        return S
    if fname == "median!":
        return _median_eltype(S)
    if fname == "quantile":
        return Float64 if S in (Float32, Float64, UInt8) else None
    if S in (Float32, Float64):
        return S
    if S is UInt8 and fname in ("mean", "median", "sum"):
        return Float64
    if S in (BFloat16, Float16) and fname == "median":
        return S
    return None


def _generic(fname, f, img, S, win, b, first, shape, step):
    """One call of ifm_mapwindow_generic (two for extrema)."""
    import ctypes
    L, ctx = _lib.load(), _lib.ctx()
    N = len(img.shape)
    names = ["minimum", "maximum"] if fname == "extrema" else [fname]
    outs = []
    for nm in names:
        T = _generic_eltype(nm, S)
        if T is None:
            raise NotImplementedError("mapwindow(%s, ...) on %s images is not provided on the GPU" % (nm, S))
        out = empty_like(img, T, shape)
        flag = ctypes.c_int(0)
        _lib.check(L.ifm_mapwindow_generic(ctx, ifm_array(img), ifm_array(out), _WIN[nm],
                                           f.p if nm == "quantile" else 0.0, _lib.i64(first), _lib.i64(step, 1),
                                           _lib.i64([w.start for w in win]), _lib.i64([len(w) for w in win], 1),
                                           b, ctypes.byref(flag)))
        if flag.value:
            raise ArgumentError("quantiles are undefined in presence of NaNs or missing values")
        outs.append(out)
    if fname == "sum" and S is UInt8:
        # Base.sum widens UInt8 to UInt64; the FP64 sums are exact integers
        from .imfilter import _to_int
        from .eltypes import UInt64
        return _to_int(outs[0], UInt64)
    if fname == "extrema":
        return _pair(outs[0], outs[1])
    return outs[0]


def DimensionMismatch_(msg):
    from .errors import DimensionMismatch
    return DimensionMismatch(msg)


def _wrap(res, first, border, indices):
    """Explicit indices give an array with the ranges' lengths as its axes
    (compute_output_indices, src/mapwindow.jl:191-201); Inner()'s default
    indices keep the image's indices of the interior (self_offset)."""
    if indices is None and isinstance(border, Inner):
        return OffsetArray(res, tuple(first))
    return res


def _pair(omin, omax):
    """(min, max) pairs as a trailing axis of length 2, in the eltype of the
    results (N0f8 and BFloat16 keep their tags)."""
    if isinstance(omin, DeviceArray):
        return (omin, omax)
    out = np.stack([np.asarray(omin), np.asarray(omax)], axis=-1)
    return out.view(type(omin)) if type(omin) is not np.ndarray else out


def _extrema_filter(img, window):
    """mapwindow(extrema, A, window::Dims) -> extrema_filter: truncated
    windows per axis, even widths allowed (src/mapwindow.jl:337-417)."""
    N = len(img.shape)
    if isinstance(window, (int, np.integer)):
        window = (int(window),) * N
    window = tuple(int(w) for w in window)
    if len(window) != N:
        raise RuntimeError("`window` must have the same number of entries as dimensions of `A`")
    S = eltype_of(img)
    omin, omax = empty_like(img, S), empty_like(img, S)
    _lib.check(_lib.load().ifm_extrema_filter(_lib.ctx(), ifm_array(img), ifm_array(omin), ifm_array(omax),
                                              _lib.i64(window, 1)))
    return _pair(omin, omax)


def mapwindow_(f, out, img, window, border=None, indices=None):
    """mapwindow!(f, out, img, window; border, indices) (src/mapwindow.jl:101-121)."""
    res = mapwindow(f, img, window, border=border, indices=indices)
    res = res.data if isinstance(res, OffsetArray) else res
    target = out.data if isinstance(out, OffsetArray) else out
    if isinstance(target, DeviceArray) or isinstance(res, DeviceArray):
        raise NotImplementedError("mapwindow! into a DeviceArray: use mapwindow")
    if np.shape(target) != np.shape(res):
        from .errors import DimensionMismatch
        raise DimensionMismatch("out has the wrong size")
    np.copyto(np.asarray(target), np.asarray(res), casting="unsafe")
    return out
