"""Kernel: dense correlation kernels (src/kernel.jl)."""
import math

import numpy as np

from . import kernelfactors as KF
from .offsetarrays import OffsetArray, centered


def _broadcast_product(factors):
    """broadcast(*, factors...) of ReshapedOneDs / singleton-extended arrays."""
    dense = [f.to_dense() if isinstance(f, KF.ReshapedOneD) else f for f in factors]
    N = max(d.ndim for d in dense)
    data, first = None, [0] * N
    for d in dense:
        a = d.data.reshape(d.shape + (1,) * (N - d.ndim))
        f = d.first + (0,) * (N - d.ndim)
        for i in range(N):
            if a.shape[i] > 1:
                first[i] = f[i]
        data = a if data is None else data * a
    return OffsetArray(np.asarray(data), tuple(first))


def _product2d(kf):
    k1, k2 = kf
    return _broadcast_product(k1), _broadcast_product(k2)


def sobel(extended=None, d=None):
    if extended is None:
        return _product2d(KF.sobel())
    return (_broadcast_product(KF.sobel(extended, d)),)


def prewitt(extended=None, d=None):
    if extended is None:
        return _product2d(KF.prewitt())
    return (_broadcast_product(KF.prewitt(extended, d)),)


def scharr(extended=None, d=None):
    if extended is None:
        return _product2d(KF.scharr())
    return (_broadcast_product(KF.scharr(extended, d)),)


def bickley(extended=None, d=None):
    if extended is None:
        return _product2d(KF.bickley())
    return (_broadcast_product(KF.bickley(extended, d)),)


def ando3(extended=None, d=None):
    if extended is None:
        return _product2d(KF.ando3())
    return (_broadcast_product(KF.ando3(extended, d)),)


def box(sz):
    return _broadcast_product(KF.box(tuple(sz)))


def gaussian(sigma, l=None):
    """Kernel.gaussian (src/kernel.jl:275-287); a scalar σ gives a 2-D kernel."""
    if not isinstance(sigma, (tuple, list)):
        sigma = (sigma, sigma)
        if l is not None and not isinstance(l, (tuple, list)):
            l = (l, l)
    if len(sigma) == 0:
        return OffsetArray(np.array(1.0).reshape(()), ())
    if len(sigma) == 1:
        return KF.gaussian(sigma[0], None if l is None else l[0])
    return _broadcast_product(KF.gaussian(tuple(sigma), l))


def DoG(sp, sm=None, l=None):
    """Difference of Gaussians (src/kernel.jl:305-313)."""
    if not isinstance(sp, (tuple, list)):
        sp = (sp, sp)
    if sm is None:
        sm = tuple(s * math.sqrt(2) for s in sp)
        neg = gaussian(tuple(sm))
        l = neg.shape
        pos = gaussian(tuple(sp), l)
        return OffsetArray(pos.data - neg.data, pos.first)
    a, b = gaussian(tuple(sp), l), gaussian(tuple(sm), l)
    return OffsetArray(a.data - b.data, a.first)


def LoG(sigma):
    """Laplacian of Gaussian (src/kernel.jl:325-339): C·(Σx²/σ⁴ - Σ1/σ²)·exp(-Σx²/2σ²)
    on -w:w with w = ⌈8.5σ⌉>>1, C = 1/(∏σ·(2π)^(N/2))."""
    if not isinstance(sigma, (tuple, list)):
        sigma = (sigma, sigma)
    s = np.asarray(sigma, dtype=np.float64)
    N = len(s)
    ws = [int(math.ceil(8.5 * x)) >> 1 for x in sigma]
    C = 1.0 / (np.prod(s) * (2 * math.pi) ** (N / 2))
    s2 = s ** 2
    s2i = np.sum(1.0 / s2)
    grids = np.meshgrid(*[np.arange(-w, w + 1, dtype=np.float64) for w in ws], indexing="ij")
    xs = sum(g ** 2 / s2[i] for i, g in enumerate(grids))
    xs2 = sum(g ** 2 / s2[i] / s2[i] for i, g in enumerate(grids))
    return OffsetArray(C * (xs2 - s2i) * np.exp(-xs / 2), tuple(-w for w in ws))


class Laplacian:
    """The 2N-point Laplacian stencil along the flagged axes (src/kernel.jl:341-391)."""

    def __init__(self, flags=(True, True), N=None):
        if N is not None:          # Laplacian(dims, N)
            dims = flags
            flags = tuple((i + 1) in dims for i in range(N))
        self.flags = tuple(bool(f) for f in flags)

    def kernel_axes(self):
        return tuple((-1, 1) if f else (0, 0) for f in self.flags)

    @property
    def dtype(self):
        return np.dtype(np.int64)

    def to_array(self):
        """convert(AbstractArray, L)."""
        N = len(self.flags)
        shape = tuple(3 if f else 1 for f in self.flags)
        a = np.zeros(shape, dtype=np.int64)
        center = tuple(1 if f else 0 for f in self.flags)
        for i, f in enumerate(self.flags):
            if f:
                for s in (-1, 1):
                    idx = list(center)
                    idx[i] += s
                    a[tuple(idx)] = 1
        a[center] = -2 * sum(self.flags)
        return OffsetArray(a, tuple(-1 if f else 0 for f in self.flags))


def laplacian2d(alpha=0):
    """Weighted discrete 2-D Laplacian (src/kernel.jl:421-436)."""
    lc = alpha / (1 + alpha)
    lb = (1 - alpha) / (1 + alpha)
    lm = -4 / (1 + alpha)
    return centered(np.array([[lc, lb, lc], [lb, lm, lb], [lc, lb, lc]]))


def reflect(kernel):
    """Reflected kernel: k[-i] (src/kernel.jl:514-529)."""
    k = kernel if isinstance(kernel, OffsetArray) else centered(kernel)
    data = k.data[tuple(slice(None, None, -1) for _ in range(k.ndim))]
    return OffsetArray(data.copy(), tuple(-(f + n - 1) for f, n in zip(k.first, k.shape)))
