"""Border specifications and index maps (src/border.jl, src/borderarray.jl).

Pad / Fill / Inner / NA / NoPad mirror the reference's types; padindex and the
lazy BorderArray are host-side index logic; padarray materialises on the GPU
through ifm_padarray.
"""
import math

import numpy as np

from . import _lib
from .device import DeviceArray, empty_like, host_dense, ifm_array
from .eltypes import BFloat16, N0f8, UInt8, as_eltype, eltype_of, is_eltype_spec
from .errors import ArgumentError
from .offsetarrays import OffsetArray

STYLES = {"replicate": _lib.REPLICATE, "circular": _lib.CIRCULAR, "symmetric": _lib.SYMMETRIC,
          "reflect": _lib.REFLECT}
valid_borders = ("replicate", "circular", "reflect", "symmetric")


def _style_name(s):
    s = str(s)
    return s[1:] if s.startswith(":") else s


def _range(r):
    """(first, last) of a range-like (Python range, or a (first, last) pair)."""
    if isinstance(r, range):
        return (r.start, r.stop - 1)
    return (int(r[0]), int(r[1]))


def _lo(r):
    f, _ = _range(r)
    return max(-f, 0)


def _hi(r):
    _, l = _range(r)
    return max(l, 0)


class AbstractBorder:
    pass


class NoPad(AbstractBorder):
    """NoPad(border): input already padded; remembers `border` (src/border.jl:7-30)."""

    def __init__(self, border=None):
        self.border = border

    def __getitem__(self, _):
        return self.border


class Pad(AbstractBorder):
    """Pad(style, lo, hi) (src/border.jl:136-218).  lo/hi are None for Pad{0}
    (padding still to be computed from the kernel)."""

    def __init__(self, *args):
        style = "replicate"
        if args and isinstance(args[0], str):
            style, args = _style_name(args[0]), args[1:]
        self.style = style
        self.lo = self.hi = None
        if not args:
            return
        if all(isinstance(a, (int, np.integer)) for a in args):          # Pad(style, m, n, ...)
            self.lo = self.hi = tuple(int(a) for a in args)
        elif len(args) == 1:
            a = args[0]
            if len(a) and isinstance(a[0], (range, tuple, list)) and not all(isinstance(x, (int, np.integer)) for x in a):
                self.lo = tuple(_lo(r) for r in a)                        # Pad(style, inds)
                self.hi = tuple(_hi(r) for r in a)
            else:
                self.lo = self.hi = tuple(int(x) for x in a)              # Pad(style, both::Dims)
        elif len(args) == 2:
            lo, hi = tuple(int(x) for x in args[0]), tuple(int(x) for x in args[1])
            if len(lo) == 0 and len(hi) > 0:
                lo = (0,) * len(hi)
            if len(hi) == 0 and len(lo) > 0:
                hi = (0,) * len(lo)
            if len(lo) != len(hi):
                raise ArgumentError("lo and hi must have the same length")
            self.lo, self.hi = lo, hi
        else:
            raise ArgumentError("bad Pad arguments %r" % (args,))

    @property
    def ndims(self):
        return 0 if self.lo is None else len(self.lo)

    def __call__(self, kernel, img=None, alg=None):
        """Pad{0}(kernel[, img, alg]) (src/border.jl:236-248)."""
        inds = calculate_padding(kernel)
        if alg is not None and type(alg).__name__ == "FFT":
            inds = tuple(padfft(r, n) for r, n in zip(inds, img.shape))
        return Pad(self.style, inds)

    def __eq__(self, other):
        return isinstance(other, Pad) and (self.style, self.lo, self.hi) == (other.style, other.lo, other.hi)

    def __repr__(self):
        return "Pad(%r, %r, %r)" % (self.style, self.lo, self.hi)


class Fill(AbstractBorder):
    """Fill(value[, lo, hi]) (src/border.jl:465-555)."""

    def __init__(self, value, *args):
        self.value = value
        self.lo = self.hi = None
        if not args:
            return
        if len(args) == 1:
            a = args[0]
            if hasattr(a, "shape") or isinstance(a, OffsetArray) or type(a).__name__ in ("ReshapedOneD", "Laplacian"):
                inds = calculate_padding(a)                             # Fill(value, kernel)
                self.lo = tuple(_lo(r) for r in inds)
                self.hi = tuple(_hi(r) for r in inds)
            elif len(a) and isinstance(a[0], (range, tuple, list)) and not all(isinstance(x, (int, np.integer)) for x in a):
                self.lo = tuple(_lo(r) for r in a)
                self.hi = tuple(_hi(r) for r in a)
            elif isinstance(a, tuple) and a and type(a[0]).__name__ in ("ReshapedOneD", "OffsetArray"):
                inds = calculate_padding(a)
                self.lo = tuple(_lo(r) for r in inds)
                self.hi = tuple(_hi(r) for r in inds)
            else:
                self.lo = self.hi = tuple(int(x) for x in a)
        elif len(args) == 2:
            self.lo, self.hi = tuple(int(x) for x in args[0]), tuple(int(x) for x in args[1])
        else:
            raise ArgumentError("bad Fill arguments")

    @property
    def ndims(self):
        return 0 if self.lo is None else len(self.lo)

    def __call__(self, kernel, img=None, alg=None):
        inds = calculate_padding(kernel)
        if alg is not None and type(alg).__name__ == "FFT":
            inds = tuple(padfft(r, n) for r, n in zip(inds, img.shape))
        return Fill(self.value, inds)

    def __eq__(self, other):
        return isinstance(other, Fill) and (self.value, self.lo, self.hi) == (other.value, other.lo, other.hi)

    def __repr__(self):
        return "Fill(%r, %r, %r)" % (self.value, self.lo, self.hi)


class Inner(AbstractBorder):
    """Inner(lo, hi): keep only the interior (src/border.jl:368-386)."""

    def __init__(self, *args):
        self.lo = self.hi = None
        if not args:
            return
        if all(isinstance(a, (int, np.integer)) for a in args):
            self.lo = self.hi = tuple(int(a) for a in args)
        elif len(args) == 1:
            a = args[0]
            if len(a) and isinstance(a[0], (range, tuple, list)):
                self.lo = tuple(_lo(r) for r in a)
                self.hi = tuple(_hi(r) for r in a)
            else:
                self.lo = self.hi = tuple(int(x) for x in a)
        else:
            lo, hi = tuple(int(x) for x in args[0]), tuple(int(x) for x in args[1])
            if len(lo) == 0:
                lo = (0,) * len(hi)
            if len(hi) == 0:
                hi = (0,) * len(lo)
            self.lo, self.hi = lo, hi

    @property
    def ndims(self):
        return 0 if self.lo is None else len(self.lo)

    def __call__(self, kernel, img=None, alg=None):
        return Inner(calculate_padding(kernel))

    def __eq__(self, other):
        return isinstance(other, Inner) and (self.lo, self.hi) == (other.lo, other.hi)

    def __repr__(self):
        return "Inner(%r, %r)" % (self.lo, self.hi)


class NA(AbstractBorder):
    """NA(na=isnan): normalise by the number of available values (src/border.jl:403-405)."""

    def __init__(self, na=None):
        self.na = na


def borderinstance(border):
    """src/border.jl:147-156."""
    if isinstance(border, AbstractBorder):
        return border
    if isinstance(border, str):
        if border in valid_borders:
            return Pad(border)
        if border == "inner":
            raise ArgumentError("specifying Inner as a string is deprecated, use `imfilter(img, kern, Inner())` instead")
        raise ArgumentError("%s not a recognized border" % border)
    raise ArgumentError("%r not a recognized border" % (border,))


def padfft(indk, l):
    """range(first(indk), length=nextprod([2,3], l+lk) - l + 1) (src/border.jl:245-248)."""
    f, last = _range(indk)
    lk = last - f + 1
    return (f, f + nextprod23(l + lk) - l)


def nextprod23(n):
    best = None
    p2 = 1
    while p2 < 2 * n + 2:
        p = p2
        while p < 2 * n + 2:
            if p >= n and (best is None or p < best):
                best = p
            p *= 3
        p2 *= 2
    return best


# ---- kernel extents ----------------------------------------------------------------
def kernel_axes(k):
    """axes(kernel) as (first, last) pairs."""
    if hasattr(k, "kernel_axes"):
        return k.kernel_axes()
    if isinstance(k, OffsetArray):
        return tuple((f, f + n - 1) for f, n in zip(k.first, k.shape))
    a = np.asarray(k)
    return tuple((1, n) for n in a.shape)       # a plain (1-based) array has axes 1:n


def _expand(inds, pad):
    n = max(len(inds), len(pad))
    out = []
    for d in range(n):
        if d < len(inds) and d < len(pad):
            out.append((inds[d][0] + pad[d][0], inds[d][1] + pad[d][1]))
        elif d < len(inds):
            out.append(inds[d])
        else:
            out.append(pad[d])
    return tuple(out)


def calculate_padding(kernel):
    """src/border.jl:614-621: accumulated axes of a factored kernel, doubled
    when IIR and FIR factors are mixed."""
    if isinstance(kernel, tuple):
        if not kernel:
            return ()
        inds = kernel_axes(kernel[0])
        for k in kernel[1:]:
            inds = _expand(inds, kernel_axes(k))
        isiir = [getattr(k, "is_iir", False) for k in kernel]
        if any(isiir) and not all(isiir):
            inds = tuple((2 * f if f < 0 else f, 2 * l if l > 0 else l) for f, l in inds)
        return inds
    return kernel_axes(kernel)


# ---- index maps ------------------------------------------------------------------------
def padindex(border, lo, inds, hi):
    """Index vector for padding one axis (src/border.jl:564-590), 0-based:
    element t is the source index of padded position t - lo."""
    n = inds if isinstance(inds, (int, np.integer)) else len(inds)
    first = 0 if isinstance(inds, (int, np.integer)) else inds.start
    style = border.style if isinstance(border, Pad) else _style_name(border)
    i = np.arange(-lo, n + hi, dtype=np.int64)
    if style == "replicate":
        j = np.clip(i, 0, n - 1)
    elif style == "circular":
        j = np.mod(i, n)
    elif style == "symmetric":
        j = np.mod(i, 2 * n)
        j = np.where(j < n, j, 2 * n - 1 - j)
    elif style == "reflect":
        if n == 1:
            raise ZeroDivisionError("reflect padding of an axis of length 1")
        j = np.mod(i, 2 * n - 2)
        j = np.where(j < n, j, 2 * n - 2 - j)
    else:
        raise ValueError("border style %s unrecognized" % style)
    return j + first


def _padding_of(img_ndim, border):
    if border.lo is None or len(border.lo) != img_ndim:
        raise ArgumentError("%r lacks the proper padding sizes for an array with %d dimensions" % (border, img_ndim))
    return border.lo, border.hi


def convert_fill(value, et):
    """convert(eltype(inner), border.value) or ArgumentError (src/borderarray.jl:11-24)."""
    try:
        if value is None:
            raise TypeError
        v = float(value)
        if et is UInt8:
            if v != int(v) or not 0 <= v <= 255:
                raise ValueError
            return v
        if et is N0f8:
            if not 0.0 <= v <= 1.0:
                raise ValueError
            return float(round(v * 255)) / 255.0   # the value code/255 (the library recovers the code)
        if et is BFloat16:
            from .eltypes import bf16, bf16_to_float32
            return float(bf16_to_float32(bf16(np.float32(v)))[()])
        if et.name == "Float16":
            return float(np.float16(v))
        if et.name == "Float32":
            return float(np.float32(v))
        return v
    except (TypeError, ValueError, OverflowError):
        raise ArgumentError("Cannot convert elements of border=%r to eltype(inner)=%s." % (value, et))


def padarray(*args):
    """padarray([T], img, border) (src/border.jl:324-331), on the GPU.
    Returns an OffsetArray whose first indices are -lo (0-based image axes)."""
    if len(args) == 3:
        T, img, border = as_eltype(args[0]), args[1], args[2]
    elif len(args) == 2:
        img, border = args
        T = eltype_of(img)
    else:
        raise TypeError("padarray([T], img, border)")
    if isinstance(border, type) and issubclass(border, AbstractBorder):
        raise ArgumentError("must supply padding sizes to %s" % border.__name__)
    img = host_dense(img) if not isinstance(img, DeviceArray) else img
    S = eltype_of(img)
    if isinstance(border, Inner):
        out = empty_like(img, T)
        _pad_call(img, out, Pad("replicate", (0,) * len(img.shape), (0,) * len(img.shape)), S)
        return out
    if isinstance(border, str):
        border = borderinstance(border)
    lo, hi = _padding_of(len(img.shape), border)
    shape = tuple(l + n + h for l, n, h in zip(lo, img.shape, hi))
    out = empty_like(img, T, shape)
    _pad_call(img, out, border, S)
    return OffsetArray(out, tuple(-l for l in lo))


def _pad_call(img, out, border, S):
    b = _lib.IfmBorder()
    if isinstance(border, Fill):
        b.style = _lib.FILL
        b.fill = convert_fill(border.value, S)
    else:
        if border.style not in STYLES:
            raise ValueError("border style %s unrecognized" % border.style)
        b.style = STYLES[border.style]
    for d in range(len(border.lo)):
        b.lo[d], b.hi[d] = border.lo[d], border.hi[d]
    ai, ao = ifm_array(img), ifm_array(out)
    _lib.check(_lib.load().ifm_padarray(_lib.ctx(), ai, ao, b))


class BorderArray:
    """Lazy padded view (src/borderarray.jl:59-149).  Indices are 0-based and
    may lie in the padding; a Pad border reflects once (valid while pad < n),
    as the reference's getindex does.  copy() materialises with padarray."""

    def __init__(self, inner, border):
        if isinstance(border, str):
            border = borderinstance(border)
        if not isinstance(border, Inner):
            _padding_of(np.ndim(inner) if not isinstance(inner, DeviceArray) else inner.ndim, border)
        self.inner = inner
        self._padding = border
        if isinstance(border, Fill):
            et = eltype_of(inner)
            v = convert_fill(border.value, et)
            # indexing returns stored elements: the code for N0f8
            border = Fill(int(round(v * 255)) if et is N0f8 else v, border.lo, border.hi)
        self.border = border

    def axes(self):
        shape = np.shape(self.inner)
        if isinstance(self.border, Inner):
            return tuple(range(0, n) for n in shape)
        return tuple(range(-l, n + h) for l, n, h in zip(self.border.lo, shape, self.border.hi))

    @property
    def shape(self):
        return tuple(len(r) for r in self.axes())

    def __getitem__(self, idx):
        if not isinstance(idx, tuple):
            idx = (idx,)
        a = np.asarray(self.inner)
        shape = a.shape
        if all(0 <= i < n for i, n in zip(idx, shape)):
            return a[idx]
        if not all(r.start <= i < r.stop for i, r in zip(idx, self.axes())):
            raise IndexError(idx)
        if isinstance(self.border, Fill):
            return self.border.value
        s = self.border.style
        out = []
        for i, n in zip(idx, shape):
            if s == "replicate":
                i = min(max(i, 0), n - 1)
            elif s == "circular":
                i = i % n
            elif s == "reflect":
                i = 2 * (n - 1) - i if i > n - 1 else (-i if i < 0 else i)
            elif s == "symmetric":
                i = 2 * (n - 1) + 1 - i if i > n - 1 else (-1 - i if i < 0 else i)
            out.append(i)
        return a[tuple(out)]

    def copy(self):
        if isinstance(self.border, Inner):
            return np.array(self.inner, copy=True)
        return padarray(self.inner, self._padding)
