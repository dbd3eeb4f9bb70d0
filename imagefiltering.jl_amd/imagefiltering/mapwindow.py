"""mapwindow / mapwindow! for the median, minimum, maximum and extrema filters
(src/mapwindow.jl), on the GPU through libifm.  Julia's `mapwindow!` is
`mapwindow_`, and `median!` is `median_`.

Window functions are named like the reference's: ``median_`` (Statistics.
median!, rewritten to median_fast!), ``median``, ``minimum``, ``maximum`` and
``extrema``.  Other functions are outside this package's GPU path and raise.
extrema returns an array of shape img.shape + (2,) holding (min, max) pairs.
"""
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
    """Output region (first, shape) in 0-based image indices."""
    N = len(img.shape)
    if indices is not None:
        if isinstance(indices, range):
            indices = (indices,)
        if len(indices) != N:
            raise ArgumentError("indices must give one range per dimension")
        for r in indices:
            if not isinstance(r, range) or r.step != 1:
                raise NotImplementedError("strided indices are not implemented")
        return [r.start for r in indices], [len(r) for r in indices]
    if isinstance(border, Inner):
        first, shape = [], []
        for d in range(N):
            lo = max(0, -win[d].start)
            hi = img.shape[d] - 1 - max(0, win[d].stop - 1)
            first.append(lo)
            shape.append(max(0, hi - lo + 1))
        return first, shape
    return [0] * N, list(img.shape)


def mapwindow(f, img, window, border=None, indices=None, callmode="copy!"):
    """mapwindow(f, img, window; border="replicate", indices=axes(img)) (src/mapwindow.jl:75-85)."""
    if callmode != "copy!":
        raise RuntimeError("Only callmode=:copy! is currently supported")
    img = img if isinstance(img, DeviceArray) else host_dense(img)
    S = eltype_of(img)
    if f is extrema and border is None and indices is None and not any(isinstance(w, range) for w in
                                                                         (window if isinstance(window, tuple) else (window,))):
        return _extrema_filter(img, window)
    win = resolve_window(window)
    N = len(img.shape)
    if len(win) != N:
        raise DimensionMismatch_("window and image dimensionality differ")
    border = borderinstance("replicate" if border is None else border)
    if isinstance(border, (NA, NoPad)):
        raise NotImplementedError("mapwindow border %r" % (border,))
    first, shape = _region(img, win, border, indices)
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
    if f in (median_, median):
        Lw = int(np.prod([len(w) for w in win]))
        if f is median and Lw % 2 == 0:
            raise NotImplementedError("Statistics.median over an even window (mean of two ranks)")
        # the kernel writes typeof(middle(x)) directly (Float32 for N0f8,
        # Float64 for UInt8, the eltype itself for floats)
        T = _median_eltype(S)
        res = empty_like(img, T, shape)
        _lib.check(L.ifm_mapwindow_median(ctx, ifm_array(img), ifm_array(res), of, wf, wl, b))
        return _wrap(res, first, border, indices)
    if f in (minimum, maximum, extrema):
        omin = empty_like(img, S, shape) if f in (minimum, extrema) else None
        omax = empty_like(img, S, shape) if f in (maximum, extrema) else None
        _lib.check(L.ifm_mapwindow_minmax(ctx, ifm_array(img), ifm_array(omin) if omin is not None else None,
                                          ifm_array(omax) if omax is not None else None, of, wf, wl, b))
        if f is minimum:
            return _wrap(omin, first, border, indices)
        if f is maximum:
            return _wrap(omax, first, border, indices)
        return _wrap(_pair(omin, omax), first, border, indices)
    raise NotImplementedError("mapwindow with window function %r runs on the CPU only in the reference; "
                              "this GPU package provides median!, median, minimum, maximum and extrema" % (f,))


def DimensionMismatch_(msg):
    from .errors import DimensionMismatch
    return DimensionMismatch(msg)


def _wrap(res, first, border, indices):
    if isinstance(border, Inner) or (indices is not None and any(first)):
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
