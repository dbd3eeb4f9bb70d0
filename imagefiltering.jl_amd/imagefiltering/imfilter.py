"""imfilter / imfilter! (src/imfilter.jl:1-254) on the MI355X.

The dispatch steps follow the reference: (1) output eltype (filter_type),
(2) kernel factoring (factorkernel, SVD rank test), (3) default border
"replicate", (4) output allocation (Inner -> interior axes), (5) algorithm
choice (filter_algorithm) and resource.  Every algorithm then runs on the GPU
through libifm; there is no CPU compute path.  Julia's `imfilter!` is
`imfilter_` here.
"""
import numpy as np

from . import _lib
from .border import (STYLES, AbstractBorder, Fill, Inner, NA, NoPad, Pad, borderinstance, calculate_padding,
                     convert_fill)
from .device import DeviceArray, empty_like, host_dense, ifm_array
from .eltypes import (INT_INFO, BFloat16, Bool, Float16, Float32, Float64, Int8, Int64, N0f8, UInt8, as_eltype,
                      eltype_of, is_eltype_spec, mul_type, promote, signed_type)
from .errors import ArgumentError, DimensionMismatch, InexactError, MethodError
from .kernel import Laplacian
from .kernelfactors import ReshapedOneD, TriggsSdika
from .offsetarrays import OffsetArray, kernelshift


# ---- Algorithm tags and resources (src/ImageFiltering.jl:52-68) -------------------
class Alg:
    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self).__name__)

    def __repr__(self):
        return "Algorithm.%s()" % type(self).__name__


class FFT(Alg):
    """Filter using the Fast Fourier Transform."""


class FIR(Alg):
    """Filter using a direct algorithm."""


class FIRTiled(Alg):
    """Cache-efficient filtering using tiles (one fused GPU pass here)."""

    def __init__(self, tilesize=()):
        self.tilesize = tuple(tilesize)


class IIR(Alg):
    """Filter with an Infinite Impulse Response filter."""


class Mixed(Alg):
    """Filter with a cascade of mixed types (IIR, FIR)."""


class Algorithm:
    FFT, FIR, FIRTiled, IIR, Mixed = FFT, FIR, FIRTiled, IIR, Mixed


class AbstractResource:
    def __init__(self, settings=None):
        self.settings = settings if settings is not None else FIR()

    def __repr__(self):
        return "%s(%r)" % (type(self).__name__, self.settings)


class HIPDevice(AbstractResource):
    """The MI355X (libifm).  This is where every call runs."""


class CPU1(AbstractResource):
    """Accepted for drop-in compatibility; the work runs on the GPU."""


class CPUThreads(AbstractResource):
    """Accepted for drop-in compatibility; the work runs on the GPU."""


# ---- kernel canonicalisation ---------------------------------------------------------
def _as_offset(k):
    if isinstance(k, OffsetArray):
        return k
    a = np.asarray(k)
    return OffsetArray(a, (1,) * a.ndim)   # a plain array inside a tuple keeps its 1-based axes


def factorkernel(kernel):
    """src/imfilter.jl:1156-1184: a 2-D kernel is tested for rank 1 with
    svd (tolerance sqrt(eps)); rank 1 -> two factors scaled by sqrt(s)."""
    if isinstance(kernel, Laplacian):
        return (kernel,)
    if isinstance(kernel, (ReshapedOneD, TriggsSdika)):
        return (kernel,)
    k = kernelshift(kernel) if not isinstance(kernel, OffsetArray) else kernel
    if k.ndim != 2:
        return (k,)
    data = np.asarray(k.data)
    m, n = data.shape
    if data.dtype.kind in "iub":
        data = data.astype(np.float64)
    U, S, Vt = np.linalg.svd(data.astype(np.float64) if data.dtype != np.float32 else data)
    eps = np.sqrt(np.finfo(S.dtype).eps)
    separable = all(abs(S[i]) < eps for i in range(1, len(S)))
    if not separable:
        return (k,)   # the reference's leading dummykernel is a copy and is skipped
    ss = np.sqrt(S[0])
    u = (ss * U[:, 0:1]).astype(data.dtype)
    v = (ss * Vt[0:1, :]).astype(data.dtype)
    return (OffsetArray(u, (k.first[0], 0)), OffsetArray(v, (0, k.first[1])))


def _kernel_dtype(k):
    if isinstance(k, Laplacian):
        return None
    if isinstance(k, TriggsSdika):
        return Float64
    if isinstance(k, ReshapedOneD):
        return Float64 if k.is_iir else eltype_of(np.asarray(k.data.data))
    return eltype_of(np.asarray(k.data if isinstance(k, OffsetArray) else k))


def filter_type(S, kernel):
    """Output eltype (src/imfilter.jl:1131-1154)."""
    ks = kernel if isinstance(kernel, tuple) else (kernel,)
    T = None
    for k in ks:
        if isinstance(k, Laplacian):          # src/imfilter.jl:1136-1147
            if S.kind == "float":
                t = S
            elif S is N0f8:
                t = Float32
            elif S is Bool:
                t = Int8
            elif INT_INFO[S.name][1]:
                t = S
            else:
                t = signed_type(S)
        else:
            t = mul_type(S, _kernel_dtype(k))
        T = t if T is None else promote(T, t)
    return T


class _Comp:
    """One stage of a processed kernel on an N-dimensional image."""

    def __init__(self, kind, axis=None, taps=None, first=None, dense=None, iir=None):
        self.kind, self.axis, self.taps, self.first, self.dense, self.iir = kind, axis, taps, first, dense, iir


def _components(kernel, N):
    comps = []
    # a tuple of bare TriggsSdika filters dimension d with kernel[d]
    # (_imfilter_inplace_tuple!, src/imfilter.jl:891-945)
    bare_iir = bool(kernel) and all(isinstance(k, TriggsSdika) for k in kernel)
    if bare_iir and len(kernel) > N:
        raise DimensionMismatch("cannot have more kernels than dimensions")
    for pos, k in enumerate(kernel):
        if isinstance(k, Laplacian):
            if len(k.flags) != N:
                raise DimensionMismatch("Laplacian dimensionality must match the image")
            comps.append(_Comp("dense", dense=_extend(k.to_array(), N)))
            continue
        if isinstance(k, TriggsSdika):
            if N < 1:
                raise DimensionMismatch("kernel dimensionality")
            if not k.iscopy():
                comps.append(_Comp("iir", axis=pos if bare_iir else 0, iir=k))
            continue
        if isinstance(k, ReshapedOneD):
            if k.N > N:
                raise DimensionMismatch("kernel has more dimensions than the image")
            if k.is_iir:
                if not k.data.iscopy():
                    comps.append(_Comp("iir", axis=k.Npre, iir=k.data))
                continue
            taps = np.asarray(k.data.data, dtype=np.float64)
            if len(taps) == 1 and k.data.first[0] == 0 and taps[0] == 1:
                continue   # iscopy
            comps.append(_Comp("fir", axis=k.Npre, taps=taps, first=k.data.first[0]))
            continue
        a = _extend(_as_offset(k), N)
        ext = [d for d in range(N) if a.shape[d] > 1]
        if all(s == 1 for s in a.shape) and all(f == 0 for f in a.first) and np.asarray(a.data).ravel()[0] == 1:
            continue   # iscopy
        shifts = [d for d in range(N) if a.shape[d] == 1 and a.first[d] != 0]
        if len(ext) <= 1 and not shifts:
            ax = ext[0] if ext else 0
            comps.append(_Comp("fir", axis=ax, taps=np.asarray(a.data, dtype=np.float64).ravel(), first=a.first[ax]))
        else:
            comps.append(_Comp("dense", dense=a))
    return comps


def _extend(k, N):
    if k.ndim > N:
        raise DimensionMismatch("kernel has more dimensions (%d) than the image (%d)" % (k.ndim, N))
    extra = N - k.ndim
    return OffsetArray(np.asarray(k.data).reshape(k.shape + (1,) * extra), k.first + (0,) * extra)


def _comp_dense(c, N):
    if c.kind == "dense":
        return OffsetArray(np.asarray(c.dense.data, dtype=np.float64), c.dense.first)
    shape = [1] * N
    shape[c.axis] = len(c.taps)
    first = [0] * N
    first[c.axis] = c.first
    return OffsetArray(c.taps.reshape(shape), tuple(first))


def kernelconv(comps, N):
    """Full convolution of all FIR stages into one dense kernel
    (src/imfilter.jl:1257-1270)."""
    out = None
    for c in comps:
        k = _comp_dense(c, N)
        if out is None:
            out = k
            continue
        shape = tuple(a + b - 1 for a, b in zip(out.shape, k.shape))
        acc = np.zeros(shape)
        for idx in np.ndindex(*k.shape):
            sl = tuple(slice(i, i + s) for i, s in zip(idx, out.shape))
            acc[sl] += out.data * k.data[idx]
        out = OffsetArray(acc, tuple(a + b for a, b in zip(out.first, k.first)))
    return out


def maxlen(kernel):
    """src/imfilter.jl:1211-1216 (a dense array counts all its elements)."""
    L = 0
    for k in kernel:
        if isinstance(k, ReshapedOneD):
            L = max(L, len(k))
        elif isinstance(k, (TriggsSdika, Laplacian)):
            L = max(L, 1)
        else:
            L = max(L, int(np.asarray(k.data if isinstance(k, OffsetArray) else k).size))
    return L


def _all_finite(img):
    if isinstance(img, DeviceArray):
        return _device_all_finite(img)
    et = eltype_of(img)
    if et.kind != "float":
        return True
    a = np.asarray(img)
    if et is BFloat16:
        return bool(np.all((a & 0x7F80) != 0x7F80))
    return bool(np.isfinite(a).all())


def _device_all_finite(img):
    if img.eltype.kind != "float":
        return True
    # one streaming read on the device; only the answer crosses to the host
    import ctypes
    r = ctypes.c_int(0)
    _lib.check(_lib.load().ifm_all_finite(_lib.ctx(), ifm_array(img), ctypes.byref(r)))
    return bool(r.value)


def filter_algorithm(img, kernel):
    """src/imfilter.jl:1200-1209."""
    iir = [isinstance(k, TriggsSdika) or (isinstance(k, ReshapedOneD) and k.is_iir) for k in kernel]
    if kernel and all(iir):
        return IIR()
    if any(iir):
        return Mixed()
    if maxlen(kernel) > 30 and _all_finite(img):
        return FFT()
    return FIRTiled() if len(kernel) > 1 else FIR()


# ---- the GPU calls -----------------------------------------------------------------
def _border_struct(border, S):
    b = _lib.IfmBorder()
    if isinstance(border, Fill):
        b.style = _lib.FILL
        b.fill = convert_fill(border.value, S)
    elif isinstance(border, Pad):
        if border.style not in STYLES:
            raise ValueError("border style %s unrecognized" % border.style)
        b.style = STYLES[border.style]
    else:
        b.style = _lib.REPLICATE
    return b


def _required(kernel, N):
    inds = calculate_padding(kernel)
    inds = tuple(inds) + ((0, 0),) * (N - len(inds))
    return inds


def _check_explicit_padding(border, kernel, N):
    if border.lo is None:
        return
    if len(border.lo) != N:
        raise ArgumentError("%r lacks the proper padding sizes for an array with %d dimensions" % (border, N))
    for d, (f, l) in enumerate(_required(kernel, N)):
        if border.lo[d] < max(-f, 0) or border.hi[d] < max(l, 0):
            raise DimensionMismatch("requested indices and kernel indices do not agree with indices of padded input")


def _run(alg, out, img, kernel, border, out_first):
    N = len(img.shape)
    S = eltype_of(img)
    comps = _components(kernel, N)
    b = _border_struct(border, S)
    of = _lib.i64(out_first) if out_first is not None else None
    ai, ao = ifm_array(img), ifm_array(out)
    L = _lib.load()
    ctx = _lib.ctx()
    if not comps:
        fa = (_lib.IfmFactor * 1)()
        _lib.check(L.ifm_imfilter_fir(ctx, ai, ao, of, fa, 0, b))
        return out
    kinds = {c.kind for c in comps}
    if "iir" in kinds and kinds != {"iir"}:
        return _run_mixed(out, img, comps, border, out_first)
    if isinstance(alg, IIR) or kinds == {"iir"}:
        if kinds != {"iir"}:
            raise ArgumentError("Algorithm.IIR() needs IIR kernels only")
        if isinstance(border, Fill):
            # leftborder!/rightborder! convert the Fill value to the eltype of
            # the padded array, i.e. the output eltype (src/imfilter.jl:1027-1029)
            b.fill = convert_fill(border.value, eltype_of(out))
        _lib.check(L.ifm_imfilter_iir(ctx, ai, ao, _iir_struct(comps), len(comps), b))
        return out
    if isinstance(alg, FFT):
        k = kernelconv(comps, N)
        kk = _kernel_struct(k)
        _lib.check(L.ifm_imfilter_fft(ctx, ai, ao, of, kk, b))
        return out
    if kinds == {"fir"}:
        fa = (_lib.IfmFactor * len(comps))()
        keep = []
        for i, c in enumerate(comps):
            t = np.ascontiguousarray(c.taps, dtype=np.float64)
            keep.append(t)
            fa[i].axis, fa[i].len, fa[i].first = c.axis, len(t), c.first
            fa[i].taps = t.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double))
        rc = L.ifm_imfilter_fir(ctx, ai, ao, of, fa, len(comps), b)
        if rc == _lib.E_UNSUPPORTED:       # a factor longer than the FIR kernels take
            k = kernelconv(comps, N)
            _lib.check(L.ifm_imfilter_fft(ctx, ai, ao, of, _kernel_struct(k), b))
            return out
        _lib.check(rc)
        return out
    k = kernelconv(comps, N)
    _lib.check(L.ifm_imfilter_dense(ctx, ai, ao, of, _kernel_struct(k), b, _lib.PREC_AUTO))
    return out


def _run_na(alg, out, img, kernel, border):
    """imfilter(img, kernel, NA()): filter the image with its NaNs set to zero
    and the mask of valid pixels, both with Fill(0), and divide
    (imfilter_na_inseparable!, src/imfilter.jl:1109-1122).  Without NaNs and a
    separable kernel the reference divides by per-axis filtered ones instead
    (imfilter_na_separable!, :1124-1128); the filtered all-ones mask is the
    product of those factors, so the result agrees up to rounding."""
    if border.na is not None and border.na is not np.isnan:
        raise NotImplementedError("NA with a predicate other than isnan")
    import ctypes
    T = eltype_of(out)
    L = _lib.load()
    dimg = img if isinstance(img, DeviceArray) else DeviceArray.from_host(img)
    tmp = DeviceArray(img.shape, T, dimg.order)
    valid = DeviceArray(img.shape, T, dimg.order)
    hasna = ctypes.c_int(0)
    _lib.check(L.ifm_na_prepare(_lib.ctx(), ifm_array(dimg), ifm_array(tmp), ifm_array(valid), ctypes.byref(hasna)))
    _run(alg, out, tmp, kernel, Fill(0), None)
    vp = DeviceArray(img.shape, T, dimg.order)
    _run(alg, vp, valid, kernel, Fill(0), None)
    _lib.check(L.ifm_divide(_lib.ctx(), ifm_array(out), ifm_array(vp)))
    return out


def _iir_struct(comps):
    arr = (_lib.IfmIir * len(comps))()
    for i, c in enumerate(comps):
        k = c.iir
        for j in range(3):
            arr[i].a[j], arr[i].b[j] = float(k.a[j]), float(k.b[j])
        arr[i].scale = k.scale
        for j in range(9):
            arr[i].M[j] = float(np.asarray(k.M).ravel()[j])
        arr[i].axis = c.axis
    return arr


def _stage_axes(c, N):
    """axes(kernel) of one cascade stage as (first, last) per dimension; an
    IIR stage has axes 0:0 (src/kernelfactors.jl:508-509)."""
    ax = [(0, 0)] * N
    if c.kind == "fir":
        ax[c.axis] = (c.first, c.first + len(c.taps) - 1)
    elif c.kind == "dense":
        ax = [(c.dense.first[d], c.dense.first[d] + c.dense.shape[d] - 1) for d in range(N)]
    return ax


def _run_mixed(out, img, comps, border, out_first):
    """A cascade of IIR and FIR stages (Algorithm.Mixed), scheduled as the
    reference's generic cascade does (src/imfilter.jl:372-457): the image is
    padded once with the accumulated padding, doubled because IIR and FIR are
    mixed (src/border.jl:614-619); each stage then filters the previous
    stage's region shrunk by its own axes, an IIR stage over the whole region
    it is given with its edge initialised from that region's ends (or the
    Fill value), and the last stage is cut to the output axes.  Every stage
    runs on the GPU; intermediates stay in HBM in the output eltype."""
    N = len(img.shape)
    T = eltype_of(out)
    L = _lib.load()
    ctx = _lib.ctx()
    of = list(out_first) if out_first is not None else [0] * N
    on = list(np.shape(out.data if isinstance(out, OffsetArray) else out))
    inds = [(of[d], of[d] + on[d] - 1) for d in range(N)]
    axes = [_stage_axes(c, N) for c in comps]
    pf = [sum(a[d][0] for a in axes) for d in range(N)]
    pl = [sum(a[d][1] for a in axes) for d in range(N)]
    pf = [2 * f if f < 0 else f for f in pf]                    # doublepadding
    pl = [2 * l if l > 0 else l for l in pl]
    lo = [max(-f, 0) for f in pf]
    hi = [max(l, 0) for l in pl]
    dimg = img if isinstance(img, DeviceArray) else DeviceArray.from_host(img)
    if isinstance(border, NoPad):
        # imfilter!(r, out, A, kernel, NoPad(), inds): A is already padded and
        # out_first is in A's own indices
        A = dimg
        cov = [(0, img.shape[d] - 1) for d in range(N)]
        border = border.border if border.border is not None else Pad("replicate")
    else:
        if isinstance(border, Fill):
            pb = Fill(border.value, tuple(lo), tuple(hi))
        else:
            pb = Pad(border.style if isinstance(border, Pad) else "replicate", tuple(lo), tuple(hi))
        from .border import padarray
        A = padarray(T, dimg, pb).data                         # covers [-lo, n + hi)
        cov = [(-lo[d], img.shape[d] + hi[d] - 1) for d in range(N)]
    region = [(inds[d][0] + pf[d], inds[d][1] + pl[d]) for d in range(N)]   # expand(inds, padding)
    b_iir = _border_struct(border, T)
    if isinstance(border, Fill):
        b_iir.fill = convert_fill(border.value, T)
    b_rep = _lib.IfmBorder()
    b_rep.style = _lib.REPLICATE                               # no read leaves the buffer
    nst = len(comps)
    for s, c in enumerate(comps):
        last = s == nst - 1
        region = [(region[d][0] - axes[s][d][0], region[d][1] - axes[s][d][1]) for d in range(N)]   # shrink
        if c.kind == "iir":
            src = A
            if region != cov:                                  # the stage filters its own region only
                src = DeviceArray([r1 - r0 + 1 for r0, r1 in region], T, A.order)
                fa0 = (_lib.IfmFactor * 1)()
                _lib.check(L.ifm_imfilter_fir(ctx, ifm_array(A), ifm_array(src),
                                              _lib.i64([region[d][0] - cov[d][0] for d in range(N)]), fa0, 0,
                                              b_rep))
            dst = DeviceArray(src.shape, T, A.order)
            _lib.check(L.ifm_imfilter_iir(ctx, ifm_array(src), ifm_array(dst), _iir_struct([c]), 1, b_iir))
            A, cov = dst, region
            if last:
                fa0 = (_lib.IfmFactor * 1)()
                _lib.check(L.ifm_imfilter_fir(ctx, ifm_array(A), ifm_array(out),
                                              _lib.i64([inds[d][0] - cov[d][0] for d in range(N)]), fa0, 0, b_rep))
            continue
        target = [(max(region[d][0], inds[d][0]), min(region[d][1], inds[d][1])) for d in range(N)] if last else region
        dst = out if last else DeviceArray([r1 - r0 + 1 for r0, r1 in target], T, A.order)
        ofs = _lib.i64([target[d][0] - cov[d][0] for d in range(N)])
        if c.kind == "fir":
            fa = (_lib.IfmFactor * 1)()
            t = np.ascontiguousarray(c.taps, dtype=np.float64)
            fa[0].axis, fa[0].len, fa[0].first = c.axis, len(t), c.first
            fa[0].taps = t.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double))
            _lib.check(L.ifm_imfilter_fir(ctx, ifm_array(A), ifm_array(dst), ofs, fa, 1, b_rep))
        else:
            _lib.check(L.ifm_imfilter_dense(ctx, ifm_array(A), ifm_array(dst), ofs,
                                            _kernel_struct(_comp_dense(c, N)), b_rep, _lib.PREC_AUTO))
        A, cov = dst, target
    return out


def _kernel_struct(k):
    kk = _lib.IfmKernel()
    data = np.asfortranarray(np.asarray(k.data, dtype=np.float64))
    kk.ndim = data.ndim
    for d in range(data.ndim):
        kk.size[d] = data.shape[d]
        kk.first[d] = k.first[d]
    flat = np.ascontiguousarray(data.ravel(order="F"))
    kk.taps = flat.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double))
    kk._keep = flat   # the taps live as long as the struct
    return kk


# ---- public API ------------------------------------------------------------------------
def _parse_imfilter(args):
    args = list(args)
    r = args.pop(0) if args and isinstance(args[0], AbstractResource) else None
    T = as_eltype(args.pop(0)) if args and is_eltype_spec(args[0]) else None
    if len(args) < 2:
        raise MethodError("imfilter([r], [T], img, kernel, [border], [alg])")
    img, kernel, rest = args[0], args[1], args[2:]
    border = rest[0] if len(rest) >= 1 else "replicate"
    alg = rest[1] if len(rest) >= 2 else None
    if len(rest) > 2 or (r is not None and alg is not None):
        raise MethodError("no method matching this imfilter call")
    if alg is not None and not isinstance(alg, Alg):
        raise MethodError("the algorithm must be an Algorithm.* instance")
    return r, T, img, kernel, border, alg


def _processed(kernel):
    if isinstance(kernel, tuple):
        return tuple(_as_offset(k) if isinstance(k, np.ndarray) else k for k in kernel)
    return factorkernel(kernel)


def imfilter(*args):
    """imfilter([r], [T], img, kernel, [border="replicate"], [alg]) (src/imfilter.jl:1-49)."""
    r, T, img, kernel, border, alg = _parse_imfilter(args)
    if isinstance(kernel, TriggsSdika):
        kernel = (kernel,)
    kernel = _processed(kernel)
    if isinstance(img, ColorImage):
        return _imfilter_color(r, T, img, kernel, border, alg)
    img = img if isinstance(img, DeviceArray) else host_dense(img)
    S = eltype_of(img)
    if T is None:
        T = filter_type(S, kernel)
    if T.kind != "float":
        return _imfilter_int(r, T, img, kernel, border, alg)
    img = _library_input(img)
    border = borderinstance(border)
    N = len(img.shape)
    if isinstance(border, NA):
        out = empty_like(img, T)
        return _run_na(_choose(r, alg, img, kernel), out, img, kernel, border)
    if isinstance(border, Inner):
        lo, hi = _inner_padding(border, kernel, N)
        shape = tuple(n - l - h for n, l, h in zip(img.shape, lo, hi))
        if any(s < 0 for s in shape):
            raise DimensionMismatch("kernel larger than the image for Inner()")
        out = empty_like(img, T, shape)
        _run(_choose(r, alg, img, kernel), out, img, kernel, Pad("replicate"), lo)
        return OffsetArray(out, tuple(lo))
    if isinstance(border, (Pad, Fill)):
        _check_explicit_padding(border, kernel, N)
    out = empty_like(img, T)
    return _run(_choose(r, alg, img, kernel), out, img, kernel, border, None)


def _inner_padding(border, kernel, N):
    if border.lo is None:
        inds = _required(kernel, N)
        return [max(-f, 0) for f, _ in inds], [max(l, 0) for _, l in inds]
    if len(border.lo) != N:
        raise DimensionMismatch("dimensionality of img and the border must agree, got %d and %d" % (N, len(border.lo)))
    for d, (f, l) in enumerate(_required(kernel, N)):
        if border.lo[d] < max(-f, 0) or border.hi[d] < max(l, 0):
            raise DimensionMismatch("requested indices and kernel indices do not agree with indices of padded input")
    return list(border.lo), list(border.hi)


def _choose(r, alg, img, kernel):
    if alg is None and r is not None:
        alg = r.settings
    if alg is None:
        alg = filter_algorithm(img, kernel)
    return alg


def imfilter_(*args):
    """imfilter!(out, img, kernel, [border], [alg]), imfilter!(r, out, img, kernel, [border]),
    imfilter!(r, out, img, kernel::TriggsSdika, dim, border) and
    imfilter!(r, out, A, kernel, NoPad(), [inds]) (src/imfilter.jl:204-254, :922-932)."""
    args = list(args)
    r = args.pop(0) if args and isinstance(args[0], AbstractResource) else None
    if len(args) < 3:
        raise MethodError("imfilter!([r], out, img, kernel, ...)")
    out, img, kernel, rest = args[0], args[1], args[2], args[3:]
    T_out = eltype_of(_outdata(out))
    if T_out.kind != "float" and not (isinstance(kernel, TriggsSdika) and rest and isinstance(rest[0], (int, np.integer))):
        # an integer out: filter into a Float64 block of its layout, then convert
        tmp = empty_like(_outdata(out), Float64)
        targs = ([r] if r is not None else []) + [OffsetArray(tmp, out.first) if isinstance(out, OffsetArray) else tmp,
                                                  img, kernel] + list(rest)
        imfilter_(*targs)
        _to_int(tmp, T_out, _outdata(out))
        return out
    if isinstance(kernel, TriggsSdika) and rest and isinstance(rest[0], (int, np.integer)):
        dim = int(rest[0])                                  # 1-based like the reference
        border = rest[1] if len(rest) > 1 else "replicate"
        N = len(img.shape)
        comp = ReshapedOneD(N, dim - 1, kernel)
        return _run(IIR(), out, host_dense(img) if not isinstance(img, DeviceArray) else img, (comp,),
                    borderinstance(border) if isinstance(border, str) else border, None)
    border = rest[0] if len(rest) >= 1 else "replicate"
    alg = rest[1] if len(rest) >= 2 else None
    if r is not None and len(rest) >= 2 and not isinstance(border, NoPad):
        raise MethodError("no method matching imfilter!(r, out, img, kernel, border, alg)")
    if isinstance(kernel, TriggsSdika):
        kernel = (kernel,)
    kernel = _processed(kernel)
    img = _library_input(img if isinstance(img, DeviceArray) else host_dense(img))
    N = len(img.shape)
    if isinstance(border, NoPad):
        return _nopad(out, img, kernel, rest[1] if len(rest) >= 2 else None, r)
    border = borderinstance(border)
    if isinstance(border, NA):
        if tuple(np.shape(_outdata(out))) != tuple(img.shape):
            raise DimensionMismatch("output axes %r disagree with the image axes %r" % (np.shape(out), img.shape))
        _run_na(_choose(r, alg, img, kernel), _outdata(out), img, kernel, border)
        return out
    if isinstance(border, Inner):
        lo, hi = _inner_padding(border, kernel, N)
        _run(_choose(r, alg, img, kernel), _outdata(out), img, kernel, Pad("replicate"), lo)
        return out
    if isinstance(border, (Pad, Fill)):
        _check_explicit_padding(border, kernel, N)
    if tuple(np.shape(_outdata(out))) != tuple(img.shape):
        raise DimensionMismatch("output axes %r disagree with the image axes %r" % (np.shape(out), img.shape))
    _run(_choose(r, alg, img, kernel), _outdata(out), img, kernel, border, None)
    return out


def _integer_taps(kernel):
    for k in kernel:
        if isinstance(k, Laplacian):
            continue
        if isinstance(k, TriggsSdika) or (isinstance(k, ReshapedOneD) and k.is_iir):
            return None
        d = np.asarray(k.data.data if isinstance(k, ReshapedOneD) else (k.data if isinstance(k, OffsetArray) else k),
                       dtype=np.float64)
        if not np.all(d == np.round(d)):
            return None
    return True


def _tap_l1(k):
    if isinstance(k, Laplacian):
        return 4.0 * sum(k.flags)
    d = k.data.data if isinstance(k, ReshapedOneD) else (k.data if isinstance(k, OffsetArray) else k)
    return float(np.sum(np.abs(np.asarray(d, dtype=np.float64))))


def _imfilter_int(r, T, img, kernel, border, alg):
    """An integer output eltype (filter_type of integer images and taps, or an
    explicit T): the filter runs in floating point on the GPU and the results
    are converted on the device, InexactError where a value is not an integer
    of T's range (src/imfilter.jl:233-254).  With integer taps the FP32 sums
    are exact while every partial sum stays below 2^24; calls that could pass
    it are refused rather than rounded."""
    if T is Bool:
        raise NotImplementedError("Bool output eltype")
    S = eltype_of(img)
    if S.kind in ("int", "fixed") and _integer_taps(kernel):
        if isinstance(img, DeviceArray) or S is N0f8:
            amax = 255.0 if S in (UInt8, N0f8) else float(2 ** (INT_INFO[S.name][0] - (1 if INT_INFO[S.name][1] else 0)))
        else:
            a = np.asarray(img)
            amax = float(np.max(np.abs(a.astype(np.float64)))) if a.size else 0.0
        bound = amax
        for k in kernel:
            bound *= _tap_l1(k)
        if bound >= 2.0 ** 24:
            raise NotImplementedError("integer result with partial sums up to %g: exact only below 2^24" % bound)
    args = ([r] if r is not None else []) + [Float64, img, kernel, border] + ([alg] if alg is not None else [])
    res = imfilter(*args)
    data = res.data if isinstance(res, OffsetArray) else res
    out = _to_int(data, T)
    return OffsetArray(out, res.first) if isinstance(res, OffsetArray) else out


def _to_int(data, T, out=None):
    import ctypes
    if T is N0f8:
        bits, signed, scale = 8, False, 255.0
    else:
        bits, signed = INT_INFO[T.name]
        scale = 1.0
    if out is None:
        out = empty_like(data, T)
    inexact = ctypes.c_int(0)
    if isinstance(out, DeviceArray):
        dst, dev = out.ptr, 1
    else:
        dst, dev = np.asarray(out).ctypes.data, 0
    _lib.check(_lib.load().ifm_convert_int(_lib.ctx(), ifm_array(data), dst, dev, bits, 1 if signed else 0, scale,
                                           ctypes.byref(inexact)))
    if inexact.value:
        raise InexactError("convert(%s, x): a filtered value is not an integer of %s's range "
                           "(consider imfilter(Float64, img, kernel, ...))" % (T, T))
    return out


class ColorImage:
    """A colour image in channelview form: `channels` holds the colour
    channels along a leading axis (RGB: 3), the image axes after it.  Every
    channel is filtered with the same kernel, as imfilter does for Colorant
    arrays (arithmetic per channel; filtfft's channelview, src/imfilter.jl:868-883)."""

    def __init__(self, channels, colorant="RGB"):
        self.channels = channels
        self.colorant = colorant

    @property
    def shape(self):
        return tuple(self.channels.shape[1:])

    @property
    def ndim(self):
        return len(self.shape)

    def __repr__(self):
        return "ColorImage(%s, %s)" % (self.colorant, self.shape)


def colorview(colorant, channels):
    """colorview(RGB, A): the channel-first array A as a colour image."""
    return ColorImage(channels, colorant)


def channelview(img):
    return img.channels


def _prepend_axis(kernel, N):
    out = []
    bare = all(isinstance(k, TriggsSdika) for k in kernel)
    for pos, k in enumerate(kernel):
        if isinstance(k, Laplacian):
            out.append(Laplacian((False,) + tuple(k.flags)))
        elif isinstance(k, ReshapedOneD):
            out.append(ReshapedOneD(k.N + 1, k.Npre + 1, k.data))
        elif isinstance(k, TriggsSdika):
            out.append(ReshapedOneD(N + 1, 1 + (pos if bare else 0), k))
        else:
            a = _extend(_as_offset(k), N)
            out.append(OffsetArray(np.asarray(a.data).reshape((1,) + a.shape), (0,) + tuple(a.first)))
    return tuple(out)


def _imfilter_color(r, T, img, kernel, border, alg):
    N = len(img.shape)
    k2 = _prepend_axis(kernel, N)
    b = borderinstance(border)
    if isinstance(b, (Pad, Fill, Inner)) and getattr(b, "lo", None) is not None:
        lo, hi = (0,) + tuple(b.lo), (0,) + tuple(b.hi)
        b = Fill(b.value, lo, hi) if isinstance(b, Fill) else (Pad(b.style, lo, hi) if isinstance(b, Pad)
                                                              else Inner(lo, hi))
    if isinstance(b, NA):
        raise NotImplementedError("NA() on colour images")
    args = ([r] if r is not None else []) + ([T] if T is not None else []) + [img.channels, k2, b] + \
        ([alg] if alg is not None else [])
    res = imfilter(*args)
    if isinstance(res, OffsetArray):
        return OffsetArray(ColorImage(res.data, img.colorant), res.first[1:])
    return ColorImage(res, img.colorant)


def _library_input(img):
    """Integer images other than UInt8 (Int8 ... Int64, UInt16, UInt32, Bool)
    enter the library as Float64, the type their products with floating-point
    taps are computed in (an exact conversion below 2^53)."""
    if isinstance(img, (DeviceArray, OffsetArray)):
        return img
    if eltype_of(img).code is None:
        return np.asarray(img, dtype=np.float64)
    return img


def _outdata(out):
    return out.data if isinstance(out, OffsetArray) else out


def _nopad(out, A, kernel, inds, r):
    """imfilter!(r, out, A, kernel, NoPad(), inds): A is already padded; every
    output index must be computable in bounds (src/imfilter.jl:592-617)."""
    a_first = A.first if isinstance(A, OffsetArray) else (0,) * len(np.shape(A))
    A_data = A.data if isinstance(A, OffsetArray) else A
    o_first = out.first if isinstance(out, OffsetArray) else (0,) * len(np.shape(out))
    o_data = _outdata(out)
    N = len(np.shape(A_data))
    if inds is None:
        inds = tuple(range(f, f + n) for f, n in zip(o_first, np.shape(o_data)))
    for d in range(N):
        if inds[d].start < o_first[d] or inds[d].stop > o_first[d] + np.shape(o_data)[d]:
            raise DimensionMismatch("output indices disagree with requested indices")
    req = _required(kernel, N)
    for d, (f, l) in enumerate(req):
        if a_first[d] > inds[d].start + f or a_first[d] + np.shape(A_data)[d] - 1 < inds[d].stop - 1 + l:
            raise DimensionMismatch("requested indices and kernel indices do not agree with indices of padded input")
    full = all(inds[d].start == o_first[d] and len(inds[d]) == np.shape(o_data)[d] for d in range(N))
    of = [inds[d].start - a_first[d] for d in range(N)]
    alg = r.settings if r is not None else FIR()
    src = host_dense(A_data) if not isinstance(A_data, DeviceArray) else A_data
    if full:
        _run(alg, o_data, src, kernel, NoPad(Pad("replicate")), of)
        return out
    # only inds is written: filter into a block of that size, then copyto!
    # it into out (the rest of out keeps its values)
    T = eltype_of(o_data)
    blk = empty_like(o_data, T, [len(r_) for r_ in inds])
    _run(alg, blk, src, kernel, NoPad(Pad("replicate")), of)
    _lib.check(_lib.load().ifm_copy_into(_lib.ctx(), ifm_array(blk), ifm_array(o_data),
                                         _lib.i64([inds[d].start - o_first[d] for d in range(N)])))
    return out
