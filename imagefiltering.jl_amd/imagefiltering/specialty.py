"""Callers one step above imfilter: imgradients (src/specialty.jl:19-53),
findlocalmaxima / findlocalminima and blob_LoG / multiLoG (src/extrema.jl).
Every filter and comparison runs on the GPU through libifm; the host only
collects the flagged indices into a list, as the reference builds its Vector."""
import ctypes as C

import numpy as np

from . import _lib
from . import kernel as Kernel
from .device import DeviceArray, ifm_array
from .eltypes import Float32, Float64, UInt8, eltype_of
from .imfilter import imfilter
from .offsetarrays import OffsetArray


def imgradients(img, kernelfun, border="replicate"):
    """imgradients(img, kernelfun=KernelFactors.ando3, border="replicate")
    (src/specialty.jl:39-53): one filtered array per dimension, gimg_d =
    imfilter(img, kernelfun(extended, d), border), extended[d] = size(img, d) > 1."""
    shape = img.shape
    extended = tuple(n > 1 for n in shape)
    return tuple(imfilter(img, kernelfun(extended, d), border) for d in range(len(shape)))


def _default_window(img):
    return (3,) * len(img.shape)


def _local_extrema(img, window, edges, sense):
    N = len(img.shape)
    window = _default_window(img) if window is None else tuple(int(w) for w in window)
    if isinstance(edges, (bool, np.bool_)):
        edges = (bool(edges),) * N
    if len(window) != N or len(edges) != N:
        raise ValueError("window and edges need one entry per dimension")
    src = img if isinstance(img, DeviceArray) else np.asarray(img)
    if not isinstance(src, DeviceArray) and eltype_of(src).code is None:
        src = src.astype(np.float64)
    flags = DeviceArray(src.shape, UInt8, src.order if isinstance(src, DeviceArray) else
                        ("F" if src.ndim > 1 and src.flags.f_contiguous and not src.flags.c_contiguous else "C"))
    e = (C.c_int32 * N)(*[1 if x else 0 for x in edges])
    _lib.check(_lib.load().ifm_local_extrema(_lib.ctx(), ifm_array(src), ifm_array(flags), _lib.i64(window), e,
                                             sense))
    f = np.asarray(flags.to_host())
    # CartesianIndex order: column-major (the first index varies fastest)
    idx = np.argwhere(f.transpose(tuple(reversed(range(N)))))[:, ::-1] if N > 1 else np.argwhere(f)
    return [tuple(int(v) for v in row) for row in idx]


def findlocalmaxima(img, window=None, edges=True):
    """findlocalmaxima(img; window=default_window(img), edges=true)
    (src/extrema.jl:122-123): 0-based indices of elements larger than every
    neighbour in the window, in column-major order."""
    return _local_extrema(img, window, edges, 1)


def findlocalminima(img, window=None, edges=True):
    """findlocalminima (src/extrema.jl:130-131)."""
    return _local_extrema(img, window, edges, -1)


class BlobLoG:
    """BlobLoG(location, σ, amplitude) (src/extrema.jl:13-17)."""

    def __init__(self, location, sigma, amplitude):
        self.location, self.sigma, self.amplitude = tuple(location), tuple(sigma), float(amplitude)

    def __repr__(self):
        return "BlobLoG(location=%r, σ=%r, amplitude=%r)" % (self.location, self.sigma, self.amplitude)


def multiLoG(img, sigmas, sigma_shape=None):
    """multiLoG(img, sigmas, σshape) (src/extrema.jl:87-97): the -σ-scaled
    LoG(σ·σshape) responses stacked along a new first axis, each computed on
    the GPU with border "reflect" (the -σ scale is folded into the taps)."""
    sigmas = list(sigmas)
    if sigmas != sorted(sigmas):
        raise RuntimeError("sigmas must be sorted")
    N = len(img.shape)
    sigma_shape = (1,) * N if sigma_shape is None else tuple(sigma_shape)
    # the output eltype float(eltype(img)): floats keep theirs, N0f8 gives Float32, integers Float64
    S = eltype_of(img)
    T = S if S.kind == "float" else (Float32 if S.kind == "fixed" else Float64)
    outs = []
    for s in sigmas:
        k = Kernel.LoG(tuple(s * sigma_shape[i] for i in range(N)))
        ks = OffsetArray(np.asarray(k.data, dtype=np.float64) * (-s), k.first)
        r = imfilter(T, img, ks, "reflect")
        outs.append(np.asarray(r.to_host() if isinstance(r, DeviceArray) else r))
    return np.stack(outs, axis=0)


def blob_LoG(img, sigma_scales, edges=None, sigma_shape=None, rthresh=1e-3):
    """blob_LoG(img, σscales; edges=(true, false, ...), σshape=(1, ...),
    rthresh=1/1000) (src/extrema.jl:63-85)."""
    N = len(img.shape)
    if edges is None:
        edges = (True,) + (False,) * N
    elif isinstance(edges, (bool, np.bool_)):
        edges = (bool(edges),) * (N + 1)
    sigma_shape = (1,) * N if sigma_shape is None else tuple(sigma_shape)
    sigmas = sorted(sigma_scales)
    L = multiLoG(img, sigmas, sigma_shape)
    maxima = findlocalmaxima(L, edges=tuple(edges))
    out = []
    if rthresh:
        a = img.to_host() if isinstance(img, DeviceArray) else img
        imgmax = float(np.max(np.abs(np.asarray(a, dtype=np.float64))))
        athresh = [rthresh / (s ** N * float(np.prod(sigma_shape))) for s in sigmas]
    for x in maxima:
        amp = L[x]
        if rthresh and not amp > athresh[x[0]] * imgmax:
            continue
        out.append(BlobLoG(x[1:], tuple(sigmas[x[0]] * sh for sh in sigma_shape), amp))
    return out
