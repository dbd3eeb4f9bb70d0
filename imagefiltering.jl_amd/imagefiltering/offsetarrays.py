"""Arrays with shifted axes (OffsetArrays.jl's role for kernels and padded
images).  ``first[d]`` is the index of ``data``'s element 0 along axis d:
kernel offsets use the reference's own numbers (``centered`` kernels start at
-(n-1)÷2), and image indices are 0-based (a padded array starts at -lo)."""
import warnings

import numpy as np


class OffsetArray:
    def __init__(self, data, *first):
        self.data = data
        if len(first) == 1 and isinstance(first[0], (tuple, list)):
            first = tuple(first[0])
        first = list(first)
        # ranges are accepted like Julia's OffsetArray(A, -1:0, 1:2)
        for i, f in enumerate(first):
            if isinstance(f, range):
                if len(f) != data.shape[i]:
                    raise ValueError("axis %d: range length %d != size %d" % (i, len(f), data.shape[i]))
                first[i] = f.start
        if len(first) != np.ndim(data):
            raise ValueError("need one offset per dimension")
        self.first = tuple(int(f) for f in first)

    @property
    def shape(self):
        return self.data.shape

    @property
    def ndim(self):
        return self.data.ndim

    @property
    def dtype(self):
        return self.data.dtype

    def axes(self):
        return tuple(range(f, f + n) for f, n in zip(self.first, self.data.shape))

    def __getitem__(self, idx):
        if not isinstance(idx, tuple):
            idx = (idx,)
        return self.data[tuple(i - f for i, f in zip(idx, self.first))]

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self.data, dtype=dtype)

    def __repr__(self):
        return "OffsetArray(%r, first=%r)" % (self.data, self.first)

    def __eq__(self, other):
        if isinstance(other, OffsetArray):
            return self.first == other.first and np.array_equal(self.data, other.data)
        return NotImplemented


def centered(a):
    """OffsetArrays.centered: axis d of length n runs from 1-(n+1)÷2."""
    if isinstance(a, OffsetArray):
        a = a.data
    a = np.asarray(a)
    return OffsetArray(a, tuple(1 - (n + 1) // 2 for n in a.shape))


def kernelshift(k):
    """A kernel given as a plain (1-based) array is assumed centred, with the
    reference's deprecation warning (src/imfilter.jl:1186-1196)."""
    if isinstance(k, OffsetArray):
        return k
    warnings.warn("assuming that the origin is at the center of the kernel; to avoid this warning, "
                  "call `centered(kernel)` or use an OffsetArray", DeprecationWarning, stacklevel=3)
    return centered(k)
