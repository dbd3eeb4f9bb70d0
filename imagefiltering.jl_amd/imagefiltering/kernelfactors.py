"""KernelFactors: separable kernels stored as 1-D factors (src/kernelfactors.jl)."""
import math

import numpy as np

from .offsetarrays import OffsetArray, centered


class ReshapedOneD:
    """A 1-D factor extended along axis Npre of an N-dimensional kernel
    (src/kernelfactors.jl:52-131).  `data` is a 1-D OffsetArray or a TriggsSdika."""

    def __init__(self, N, Npre, data):
        if isinstance(data, OffsetArray):
            if data.ndim != 1:
                raise ValueError("must be one dimensional, got %d" % data.ndim)
        elif not isinstance(data, TriggsSdika):
            data = OffsetArray(np.asarray(data).ravel(), (1,))   # a plain vector: axes 1:n
        self.N, self.Npre, self.data = N, Npre, data

    @property
    def is_iir(self):
        return isinstance(self.data, TriggsSdika)

    @property
    def dtype(self):
        return np.dtype(np.float64) if self.is_iir else self.data.dtype

    def kernel_axes(self):
        ax = self.data.kernel_axes()[0] if self.is_iir else (self.data.first[0],
                                                             self.data.first[0] + self.data.shape[0] - 1)
        return tuple(ax if d == self.Npre else (0, 0) for d in range(self.N))

    def __len__(self):
        return 1 if self.is_iir else self.data.shape[0]

    def to_dense(self):
        """convert(AbstractArray, A): the N-D array with singleton axes."""
        shape = [1] * self.N
        shape[self.Npre] = self.data.shape[0]
        first = [0] * self.N
        first[self.Npre] = self.data.first[0]
        return OffsetArray(np.asarray(self.data.data).reshape(shape), tuple(first))

    def __repr__(self):
        return "ReshapedOneD{%d,%d}(%r)" % (self.N, self.Npre, self.data)


def kernelfactors(factors):
    """src/kernelfactors.jl:586-602: 1-D vectors become ReshapedOneDs along
    successive axes; N-D arrays are extended with trailing singleton axes."""
    factors = tuple(factors)
    N = len(factors)
    if all(isinstance(f, TriggsSdika) or (np.ndim(f.data if isinstance(f, OffsetArray) else f) == 1)
           for f in factors):
        return tuple(ReshapedOneD(N, i, f) for i, f in enumerate(factors))
    out = []
    for f in factors:
        a = f if isinstance(f, OffsetArray) else OffsetArray(np.asarray(f), (1,) * np.ndim(f))
        extra = N - a.ndim
        out.append(OffsetArray(a.data.reshape(a.shape + (1,) * extra), a.first + (0,) * extra))
    return tuple(out)


def box(*sz):
    if len(sz) == 1 and isinstance(sz[0], tuple):
        sz = sz[0]
    if not all(s % 2 == 1 for s in sz):
        from .errors import ArgumentError
        raise ArgumentError("kernel dimensions must be odd, got %r" % (sz,))
    return kernelfactors(tuple(centered(np.full(s, 1.0 / s)) for s in sz))


def _kdim(keep, k):
    return centered(np.asarray(k, dtype=np.float64)) if keep else OffsetArray(np.array([1.0]), (0,))


def gradfactors(extended, d, k1, k2):
    """src/kernelfactors.jl:603-605 (d is 0-based here)."""
    return kernelfactors(tuple(_kdim(extended[i], k1 if i == d else k2) for i in range(len(extended))))


def _pair(f1, f2):
    f1, f2 = centered(np.asarray(f1, dtype=np.float64)), centered(np.asarray(f2, dtype=np.float64))
    return kernelfactors((f2, f1)), kernelfactors((f1, f2))


def sobel(extended=None, d=None):
    if extended is None:
        return _pair(np.array([1.0, 2.0, 1.0]) / 4, np.array([-1.0, 0.0, 1.0]) / 2)
    return gradfactors(extended, d, np.array([-1.0, 0.0, 1.0]) / 2, np.array([1.0, 2.0, 1.0]) / 4)


def prewitt(extended=None, d=None):
    if extended is None:
        return _pair(np.array([1.0, 1.0, 1.0]) / 3, np.array([-1.0, 0.0, 1.0]) / 2)
    return gradfactors(extended, d, np.array([-1.0, 0.0, 1.0]) / 2, np.array([1.0, 1.0, 1.0]) / 3)


def scharr(extended=None, d=None):
    f1 = np.array([3.0 / 32.0, 5.0 / 16.0, 3.0 / 32.0]) * 2
    if extended is None:
        return _pair(f1, np.array([-1.0, 0.0, 1.0]) / 2)
    return gradfactors(extended, d, np.array([-1.0, 0.0, 1.0]) / 2, f1)


def bickley(extended=None, d=None):
    f1 = np.array([1.0 / 12.0, 1.0 / 3.0, 1.0 / 12.0]) * 2
    if extended is None:
        return _pair(f1, np.array([-1.0, 0.0, 1.0]) / 2)
    return gradfactors(extended, d, np.array([-1.0, 0.0, 1.0]) / 2, f1)


def ando3(extended=None, d=None):
    f1 = 2 * np.array([0.112737, 0.274526, 0.112737])
    if extended is None:
        return _pair(f1, np.array([-1.0, 0.0, 1.0]) / 2)
    return gradfactors(extended, d, np.array([-1.0, 0.0, 1.0]) / 2, f1)


def ando4(extended=None, d=None):
    f1 = np.array([0.0919833, 0.408017, 0.408017, 0.0919833])
    f2 = 1.46205884 * np.array([-0.0919833, -0.408017, 0.408017, 0.0919833])
    pair = _pair(f1, f2)
    if extended is None:
        return pair
    if len(extended) == 2 and all(extended):
        return pair[d]
    raise ValueError("dimensions other than 2 are not yet supported")


def ando5(extended=None, d=None):
    f1 = np.array([0.0357338, 0.248861, 0.43081, 0.248861, 0.0357338])
    f2 = 0.784406 * np.array([-0.137424, -0.362576, 0.0, 0.362576, 0.137424])
    pair = _pair(f1, f2)
    if extended is None:
        return pair
    if len(extended) == 2 and all(extended):
        return pair[d]
    raise ValueError("dimensions other than 2 are not yet supported")


def gaussian(sigma, l=None):
    """gaussian(σ[, l]) (src/kernelfactors.jl:438-459): exp(-x²/2σ²)/Σ on
    -w:w with l = 4⌈σ⌉+1; a tuple of σ gives one factor per dimension."""
    if isinstance(sigma, (tuple, list)):
        ls = l if l is not None else [None] * len(sigma)
        return kernelfactors(tuple(gaussian(s, li) for s, li in zip(sigma, ls)))
    if l is None:
        l = 4 * int(math.ceil(sigma)) + 1
    if l % 2 != 1:
        from .errors import ArgumentError
        raise ArgumentError("length must be odd")
    w = l >> 1
    dt = np.float32 if isinstance(sigma, np.float32) else np.float64
    if sigma == 0:
        g = np.array([math.exp(0.0)], dtype=dt)
    else:
        x = np.arange(-w, w + 1, dtype=np.float64)
        g = np.exp(-x ** 2 / (2 * float(sigma) ** 2)).astype(dt)
    return centered(g / g.sum())


class TriggsSdika:
    """IIR factor with forward a, backward b, scale and boundary matrix M
    (src/kernelfactors.jl:463-513)."""

    is_iir = True

    def __init__(self, a, b=None, scale=None, M=None):
        a = np.asarray(a, dtype=np.float64)
        if b is not None and scale is None and M is None:
            scale, b = b, None         # TriggsSdika(ab, scale)
        if b is None:
            a1, a2, a3 = a
            Mdenom = (1 + a1 - a2 + a3) * (1 - a1 - a2 - a3) * (1 + a2 + (a1 - a3) * a3)
            M = np.array([[-a3 * a1 + 1 - a3 ** 2 - a2, (a3 + a1) * (a2 + a3 * a1), a3 * (a1 + a3 * a2)],
                          [a1 + a3 * a2, -(a2 - 1) * (a2 + a3 * a1), -(a3 * a1 + a3 ** 2 + a2 - 1) * a3],
                          [a3 * a1 + a2 + a1 ** 2 - a2 ** 2,
                           a1 * a2 + a3 * a2 ** 2 - a1 * a3 ** 2 - a3 ** 3 - a3 * a2 + a3,
                           a3 * (a1 + a3 * a2)]]) / Mdenom
            b = a
        self.a = np.asarray(a, dtype=np.float64)
        self.b = np.asarray(b, dtype=np.float64)
        self.scale = float(scale)
        self.M = np.asarray(M, dtype=np.float64)
        self.asum = float(self.a.sum())
        self.bsum = float(self.b.sum())

    dtype = np.dtype(np.float64)

    def kernel_axes(self):
        return ((0, 0),)

    def iscopy(self):
        return bool(np.all(self.a == 0) and np.all(self.b == 0) and self.scale == 1)


def IIRGaussian(sigma, emit_warning=True):
    """Young-van Vliet IIR Gaussian as a TriggsSdika filter
    (src/kernelfactors.jl:533-569); a tuple of σ gives one factor per axis."""
    if isinstance(sigma, (tuple, list)):
        N = len(sigma)
        return tuple(ReshapedOneD(N, i, IIRGaussian(s, emit_warning)) for i, s in enumerate(sigma))
    if emit_warning and sigma < 1 and sigma != 0:
        import warnings
        warnings.warn("σ is too small for accuracy")
    m0, m1, m2 = 1.16680, 1.10783, 1.40586
    q = 1.31564 * (math.sqrt(1 + 0.490811 * sigma * sigma) - 1)
    ascale = (m0 + q) * (m1 * m1 + m2 * m2 + 2 * m1 * q + q * q)
    B = (m0 * (m1 * m1 + m2 * m2) / ascale) ** 2
    a1 = q * (2 * m0 * m1 + m1 * m1 + m2 * m2 + (2 * m0 + 4 * m1) * q + 3 * q * q) / ascale
    a2 = -q * q * (m0 + 2 * m1 + 3 * q) / ascale
    a3 = q * q * q / ascale
    return TriggsSdika(np.array([a1, a2, a3]), B)
