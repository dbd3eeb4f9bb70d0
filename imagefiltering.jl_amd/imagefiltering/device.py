"""Device-resident arrays in HBM, allocated through libifm.

A DeviceArray keeps the element type and a dense layout (column-major by
default, like Julia arrays; axis d is Julia dimension d+1).  Filtering calls
given DeviceArrays run without host round trips and return DeviceArrays.
"""
import ctypes as C

import numpy as np

from . import _lib
from .eltypes import BF16Array, BFloat16, N0f8, N0f8Array, eltype_of


# Caching allocator: freed device blocks are kept by rounded size and handed to
# the next array of that size, so repeated calls (imfilter allocates its
# output) do not pay a hipMalloc / hipFree each.  Every library call runs on
# the one context stream, so a block reused while earlier work on it is still
# queued is ordered after that work.
_POOL = {}
_POOL_BYTES = [0]
POOL_LIMIT = 16 << 30


def _rounded(nbytes):
    nbytes = max(int(nbytes), 16)
    gran = 2 << 20 if nbytes >= (1 << 20) else 4096
    return (nbytes + gran - 1) // gran * gran


def _pool_take(size):
    blocks = _POOL.get(size)
    if blocks:
        _POOL_BYTES[0] -= size
        return blocks.pop()
    return None


def empty_cache():
    """Frees every cached block (torch.cuda.empty_cache's role)."""
    L = _lib.load()
    for size, blocks in list(_POOL.items()):
        for p in blocks:
            L.ifm_device_free(_lib.ctx(), C.c_void_p(p))
    _POOL.clear()
    _POOL_BYTES[0] = 0


class DeviceArray:
    def __init__(self, shape, eltype, order="F"):
        self.shape = tuple(int(s) for s in shape)
        self.eltype = eltype
        self.order = order
        n = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = n * eltype.itemsize
        self._size = _rounded(self.nbytes)
        self._ptr = C.c_void_p()
        p = _pool_take(self._size)
        if p is not None:
            self._ptr = C.c_void_p(p)
            return
        L = _lib.load()
        rc = L.ifm_device_alloc(_lib.ctx(), self._size, C.byref(self._ptr))
        if rc == _lib.E_NOMEM and _POOL_BYTES[0]:
            empty_cache()
            rc = L.ifm_device_alloc(_lib.ctx(), self._size, C.byref(self._ptr))
        _lib.check(rc)

    @property
    def ptr(self):
        return self._ptr.value

    @property
    def ndim(self):
        return len(self.shape)

    def strides_elems(self):
        st, acc = [0] * len(self.shape), 1
        axes = range(len(self.shape)) if self.order == "F" else reversed(range(len(self.shape)))
        for d in axes:
            st[d] = acc
            acc *= self.shape[d]
        return st

    @classmethod
    def from_host(cls, a, order=None):
        et = eltype_of(a)
        arr = np.asarray(a)
        if order is None:
            order = "F" if (arr.flags.f_contiguous and not arr.flags.c_contiguous) or arr.ndim <= 1 else "C"
            if arr.ndim > 1 and arr.flags.f_contiguous:
                order = "F"
        buf = np.asfortranarray(arr) if order == "F" else np.ascontiguousarray(arr)
        d = cls(arr.shape, et, order)
        _lib.check(_lib.load().ifm_memcpy_h2d(_lib.ctx(), d.ptr, buf.ctypes.data, d.nbytes))
        return d

    def to_host(self):
        buf = np.empty(self.shape, dtype=self.eltype.storage, order=self.order)
        _lib.check(_lib.load().ifm_memcpy_d2h(_lib.ctx(), buf.ctypes.data, self.ptr, self.nbytes))
        if self.eltype is N0f8:
            return buf.view(N0f8Array)
        if self.eltype is BFloat16:
            return buf.view(BF16Array)
        return buf

    def similar(self, eltype=None, shape=None):
        return DeviceArray(self.shape if shape is None else shape, eltype or self.eltype, self.order)

    def __del__(self):
        try:
            if self._ptr and self._ptr.value and _lib._lib is not None and _lib._ctx is not None:
                if _POOL_BYTES[0] + self._size <= POOL_LIMIT:
                    _POOL.setdefault(self._size, []).append(self._ptr.value)
                    _POOL_BYTES[0] += self._size
                else:
                    _lib._lib.ifm_device_free(_lib._ctx, self._ptr)
                self._ptr = C.c_void_p()
        except Exception:
            pass

    def __repr__(self):
        return "DeviceArray(shape=%r, eltype=%r)" % (self.shape, self.eltype)


def ifm_array(a, eltype=None):
    """Describe a numpy array or DeviceArray for the C ABI (no copies for
    dense arrays; the caller keeps `a` alive during the call)."""
    st = _lib.IfmArray()
    if isinstance(a, DeviceArray):
        st.ndim = a.ndim
        st.dtype = (eltype or a.eltype).code
        for d in range(a.ndim):
            st.size[d] = a.shape[d]
        for d, s in enumerate(a.strides_elems()):
            st.stride[d] = s
        st.data = a.ptr
        st.device = 1
        return st
    arr = np.asarray(a)
    et = eltype or eltype_of(a)
    st.ndim = arr.ndim
    st.dtype = et.code
    for d in range(arr.ndim):
        st.size[d] = arr.shape[d]
        st.stride[d] = arr.strides[d] // arr.itemsize if arr.shape[d] > 1 else 1
    st.data = arr.ctypes.data
    st.device = 0
    return st


def host_dense(a):
    """A dense (some axis order) version of a host array, keeping its tag."""
    arr = np.asarray(a)
    if arr.flags.c_contiguous or arr.flags.f_contiguous:
        ok = all(s >= 0 for s in arr.strides)
        if ok:
            return a
    out = np.ascontiguousarray(arr)
    if isinstance(a, N0f8Array):
        return out.view(N0f8Array)
    if isinstance(a, BF16Array):
        return out.view(BF16Array)
    return out


def empty_like(a, eltype, shape=None):
    """A new output with the same layout (and location) as `a`."""
    shape = a.shape if shape is None else tuple(shape)
    if isinstance(a, DeviceArray):
        return DeviceArray(shape, eltype, a.order)
    arr = np.asarray(a)
    order = "F" if (arr.ndim > 1 and arr.flags.f_contiguous and not arr.flags.c_contiguous) else "C"
    out = np.empty(shape, dtype=eltype.storage, order=order)
    if eltype is N0f8:
        return out.view(N0f8Array)
    if eltype is BFloat16:
        return out.view(BF16Array)
    return out
